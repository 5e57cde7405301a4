/*
 * FP64 CPU ORACLE for the bundle-adjustment hot path.  TEST INFRASTRUCTURE
 * ONLY: the tests, the smoke check and the CPU-baseline leg of the benchmark
 * call it as the checker.  Nothing in the product library links or calls it.
 *
 * It restates, in plain C, the arithmetic the reference runs inside Ceres
 * Solver 1.10 (vendored under 3rdparty/ceres) for the libs/sfm BA call:
 *   - residual functor: libs/sfm/include/mvg/sfm/pinhole_brown_rt_ceres_functor.h:60-107
 *     with ceres::AngleAxisRotatePoint (include/ceres/rotation.h:565-625) and
 *     ApplyRadialDistortionIntrinsics (libs/camera/include/mvg/camera/brown_pinhole_camera.h:14-41);
 *   - Jacobian by forward-mode dual numbers (include/ceres/jet.h, internal/autodiff.h:208-311);
 *   - HuberLoss + Corrector (internal/ceres/loss_function.cc:47-61, corrector.cc:41-155,
 *     residual_block.cc:68-196);
 *   - Schur ordering (parameter_block_ordering.cc:50-154, graph_algorithms.h:173-235,
 *     reorder_program.cc:249-531);
 *   - Jacobi column scaling and LM diagonal (trust_region_minimizer.cc:103-109,
 *     levenberg_marquardt_strategy.cc:65-145);
 *   - Schur elimination / back substitution (schur_eliminator_impl.h:170-366) with a
 *     dense Cholesky of the reduced camera matrix (schur_complement_solver.cc:179-223);
 *   - implicit-Schur CG with the JACOBI preconditioner (implicit_schur_complement.cc:55-222,
 *     iterative_schur_complement_solver.cc:65-179, conjugate_gradients_solver.cc:66-245);
 *   - the trust-region LM loop (trust_region_minimizer.cc:117-712, solver_utils.h:41-56).
 *
 * Parity pinning of each part is listed with the tests (tests/test_oracle_*.py).
 */
#ifndef MBA_ORACLE_H_
#define MBA_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LAYOUT_PACKED9 = 0, ORC_LAYOUT_INTR6_EXT6 = 1 };
enum { ORC_SPARSE_SCHUR = 0, ORC_ITERATIVE_SCHUR = 1 };
enum { ORC_CONVERGENCE = 0, ORC_NO_CONVERGENCE = 1, ORC_FAILURE = 2 };

typedef struct orc_problem {
  int32_t layout;
  int32_t num_cameras;    /* packed cameras, or extrinsic blocks */
  int32_t num_intrinsics; /* intrinsic groups (layout 1) */
  int32_t num_points;
  int64_t num_obs;
  double* cameras;        /* 9*Nc (layout 0) or 6*Nc [aa, t] (layout 1) */
  double* intrinsics;     /* 6*Ni [f, ppx, ppy, k1, k2, k3] (layout 1) */
  double* points;         /* 3*Np */
  const double* obs;      /* 2*Nobs */
  const int32_t* cam_idx;
  const int32_t* intr_idx;
  const int32_t* pt_idx;
  const uint8_t* cam_constant; /* Nc flags or NULL */
} orc_problem;

typedef struct orc_options {
  int32_t max_num_iterations;
  int32_t linear_solver;
  int32_t max_num_consecutive_invalid_steps;
  int32_t max_linear_solver_iterations;
  int32_t min_linear_solver_iterations;
  int32_t jacobi_scaling;
  int32_t num_threads;
  int32_t pad_;
  double huber_a;  /* <= 0: no loss function */
  double function_tolerance;
  double gradient_tolerance;
  double parameter_tolerance;
  double initial_trust_region_radius;
  double max_trust_region_radius;
  double min_trust_region_radius;
  double min_relative_decrease;
  double min_lm_diagonal;
  double max_lm_diagonal;
  double eta;
  double r_tolerance;
} orc_options;

typedef struct orc_iteration {
  int32_t iteration;
  int32_t step_is_valid;
  int32_t step_is_successful;
  int32_t linear_solver_iterations;
  double cost;
  double cost_change;
  double gradient_max_norm;
  double gradient_norm;
  double step_norm;
  double relative_decrease;
  double trust_region_radius;
} orc_iteration;

typedef struct orc_summary {
  double initial_cost;
  double final_cost;
  double fixed_cost;
  double total_time_s;
  int32_t num_successful_steps;
  int32_t num_unsuccessful_steps;
  int32_t num_iterations;   /* entries written to the iteration array */
  int32_t termination_type;
  int32_t num_e_blocks;
  int32_t num_f_blocks;
} orc_summary;

void orc_default_options(orc_options* o);

/* Per-observation CostFunction::Evaluate with Jets (no loss).  J blocks are
 * row-major 2 x k: layout 0 -> jac[0] 2x9 (camera), jac[1] 2x3 (point);
 * layout 1 -> jac[0] 2x6 (intrinsics), jac[1] 2x6 (extrinsics), jac[2] 2x3. */
int orc_evaluate_obs(int32_t layout, const double* intr, const double* cam,
                     const double* pt, const double* obs, double* residual,
                     double* jac0, double* jac1, double* jac2);

/* Whole-problem evaluation in input order: residuals (2*Nobs, uncorrected),
 * Jacobians (row-major blocks, NULL to skip) and the robustified cost. */
int orc_evaluate(const orc_problem* P, double huber_a, double* residuals,
                 double* jac_cam, double* jac_pt, double* cost);

/* Ceres's Schur ordering on the reduced program.
 *   param_order[k]: block id of the k-th column block (cameras 0..Nc-1,
 *     points Nc..Nc+Np-1 for layout 0; intrinsics, extrinsics, points for 1);
 *   returns the number of E blocks (or -1);
 *   residual_order[k]: input index of the k-th residual row block. */
int32_t orc_schur_ordering(const orc_problem* P, int32_t* param_order,
                           int32_t* num_blocks, int64_t* residual_order,
                           int64_t* num_residuals);

/* Upper-triangle block pairs (i <= j, reduced-camera indices) of the reduced
 * camera matrix, in std::set order.  Returns the count, or -1 if `cap` is too
 * small (call with pairs = NULL to query). */
int64_t orc_schur_block_pairs(const orc_problem* P, int32_t* pairs, int64_t cap);

/* Full LM solve; parameters updated in place when the result is usable. */
int orc_solve(orc_problem* P, const orc_options* opt, orc_summary* summary,
              orc_iteration* iterations, int32_t max_records);

/* One Jacobian + Schur evaluation in FP64 (the unit of the obs/s metric):
 * residuals, Jets, Huber correction, Jacobi scaling, per-point V = E'E + D^2
 * and its inverse, g_p, per-camera U = F'F + D^2 and g_c, and the reduced
 * right-hand side rhs_c = F'(b - E V^-1 g_p).  `D` (length 9Nc + 3Np, camera
 * blocks first then points, input ids) may be NULL (no damping).  Outputs are
 * in input ids: U (45 per camera, upper triangle row-major), gc (9), V (6 per
 * point: xx xy xz yy yz zz), gp (3), rhs (9 per camera), cost. Packed layout
 * only.  `scale` (9Nc + 3Np) is the Jacobi column scale to apply (NULL: 1). */
int orc_jacobian_schur(const orc_problem* P, double huber_a, const double* scale,
                       const double* D, double* U, double* gc, double* V,
                       double* gp, double* rhs, double* cost, int32_t num_threads);

#ifdef __cplusplus
}
#endif
#endif
