/*
 * mba.h -- C ABI of the MI355X bundle-adjustment engine.
 *
 * Drop-in for the Ceres call that libs/sfm makes for bundle adjustment
 * (libs/sfm/src/sfm_incremental_engine.cpp:1413-1472,
 *  libs/sfm/src/sfm_global_engine.cpp:1548-1594 and :1758-1802):
 *
 *   ceres::Problem problem;                                  -> mba_problem
 *   problem.AddResidualBlock(cost, HuberLoss(a), cam, point) -> one observation
 *       (include/ceres/problem.h:214-258)
 *   problem.SetParameterBlockConstant(cam)                   -> camera_constant[]
 *       (include/ceres/problem.h:307)
 *   ceres::Solver::Options options                           -> mba_options
 *       (include/ceres/solver.h:59-129)
 *   ceres::Solve(options, &problem, &summary)                -> mba_solve
 *       (include/ceres/solver.h:1020, internal/ceres/solver.cc:540-545)
 *   summary.initial_cost / final_cost / IsSolutionUsable()   -> mba_summary
 *       (include/ceres/solver.h:753-831, internal/ceres/solver_utils.h:41-56)
 *   CostFunction::Evaluate(params, residuals, jacobians)     -> mba_evaluate
 *       (include/ceres/cost_function.h:114-116)
 *
 * Conventions (same as the reference):
 *   - the residual of an observation is predicted - observed (pixels) with
 *     the libs/sfm angle-axis pinhole model and radial distortion
 *     (libs/sfm/include/mvg/sfm/pinhole_brown_rt_ceres_functor.h:60-107);
 *     a camera is packed as 9 doubles [aa(3), t(3), f, k1, k2] (principal
 *     point 0, k3 = 0), a point as 3 doubles;
 *   - cost = 1/2 sum rho(|r|^2) (HuberLoss with threshold huber_a pixels,
 *     include/ceres/loss_function.h:175-184; huber_a <= 0 means no loss);
 *   - all arrays are caller-owned; cameras and points are written back in
 *     place only when the solution is usable (CONVERGENCE or
 *     NO_CONVERGENCE), as Ceres does (internal/ceres/solver.cc:442-445);
 *   - no entry point aborts the process: every failure is a status code,
 *     with a message from mba_last_error() (thread-local).
 *
 * Threading: every call works on its own context and HIP stream; independent
 * problems may be solved concurrently from several host threads (the global
 * engine does this from an OpenMP loop, sfm_global_engine.cpp:985).
 */
#ifndef MBA_H_
#define MBA_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBA_ABI_VERSION 4

typedef enum mba_status {
  MBA_OK = 0,
  MBA_ERR_INVALID_ARGUMENT = -1, /* bad sizes, indices out of range, NULL */
  MBA_ERR_UNSUPPORTED = -2,      /* structure the engine does not handle */
  MBA_ERR_DEVICE = -3,           /* HIP runtime error / no device */
  MBA_ERR_OUT_OF_MEMORY = -4,
  MBA_ERR_NOT_FINITE = -5        /* non-finite input parameters */
} mba_status;

/* Ceres TerminationType (include/ceres/types.h:315-350); USER_* unused. */
typedef enum mba_termination {
  MBA_CONVERGENCE = 0,
  MBA_NO_CONVERGENCE = 1,
  MBA_FAILURE = 2
} mba_termination;

typedef enum mba_layout {
  /* camera [aa(3), t(3), f, k1, k2] + point(3): the camera of the
   * benchmark scenes and of the functor of bundle_adjustment_unittest.cpp
   * (principal point 0, k3 0). */
  MBA_LAYOUT_PACKED9 = 0,
  /* The incremental engine's layout (sfm_incremental_engine.cpp:1273-1446,
   * ErrorFunc_Refine_Camera_3DPoints in
   * libs/sfm/include/mvg/sfm/pinhole_brown_rt_ceres_functor.h:42-110):
   * residual blocks (intrinsics, camera, point) with camera [aa(3), t(3)]
   * and a shared intrinsic group [f, ppx, ppy, k1, k2, k3] per camera. */
  MBA_LAYOUT_GROUPED = 1
} mba_layout;

typedef enum mba_ordering {
  /* ComputeStableSchurOrdering (parameter_block_ordering.cc:50-79): the
   * greedy independent set, which must come out as the point blocks. */
  MBA_ORDER_AUTOMATIC = 0,
  /* A user ordering with every point block in group 0 and every other block
   * in group 1 (examples/bundle_adjuster.cc's "user" ordering).  Ceres keeps
   * a group's blocks in address order (ordered_groups.h), which for the
   * libs/sfm containers ([cameras | intrinsics | points]) is: points by
   * index, then cameras by index, then intrinsic groups by index.  Needed
   * when the automatic set would eliminate cameras (few points, many
   * cameras, e.g. the libmv camera-refinement problems). */
  MBA_ORDER_POINTS_FIRST = 1
} mba_ordering;

/* Linear solver of the LM step.  EXPLICIT_SCHUR forms the reduced camera
 * matrix S and factors it (the SPARSE_SCHUR / DENSE_SCHUR role);
 * IMPLICIT_SCHUR_PCG is ITERATIVE_SCHUR with the block-Jacobi (JACOBI)
 * preconditioner. */
typedef enum mba_linear_solver {
  MBA_EXPLICIT_SCHUR = 0,
  MBA_IMPLICIT_SCHUR_PCG = 1
} mba_linear_solver;

/* Preconditioner of IMPLICIT_SCHUR_PCG (Solver::Options::preconditioner_type,
 * iterative_schur_complement_solver.cc:126-146). */
typedef enum mba_preconditioner {
  /* Block diagonal of F'F + D^2 (implicit_schur_complement.cc:71-93). */
  MBA_PRECOND_JACOBI = 0,
  /* Exact diagonal blocks of S (schur_jacobi_preconditioner.cc:83-103):
   * U + D^2 - sum over the camera's rows of W V^-1 W'.  Packed cameras that
   * observe each point at most once. */
  MBA_PRECOND_SCHUR_JACOBI = 1,
  /* The visibility-based preconditioners need SuiteSparse, which the
   * reference's Ceres build leaves off (3rdparty/ceres/CMakeLists.txt:72):
   * its Solve rejects them for every linear solver (solver.cc:135-149).
   * So does this library: MBA_ERR_UNSUPPORTED, with the reference's
   * message, before any device work. */
  MBA_PRECOND_CLUSTER_JACOBI = 2,
  MBA_PRECOND_CLUSTER_TRIDIAGONAL = 3
} mba_preconditioner;

/* Arithmetic of the residual / Jacobian / Schur / PCG kernels.  The LM cost
 * and gain ratio are always FP64; point-block inverses are always FP64.
 * GROUPED problems with MBA_EXPLICIT_SCHUR always run in FP64: an FP32
 * reduced matrix loses their weakly determined intrinsic directions. */
typedef enum mba_precision {
  MBA_PRECISION_FP64 = 0,
  MBA_PRECISION_MIXED = 1 /* FP32 residuals, Jacobians, matvecs, S */
} mba_precision;

typedef struct mba_problem {
  int32_t layout;       /* mba_layout */
  int32_t num_cameras;
  int32_t num_points;
  int32_t num_intrinsics;       /* GROUPED: intrinsic groups (else ignored) */
  int64_t num_observations;
  double* cameras;              /* 9 (PACKED9) or 6 (GROUPED) per camera, in/out */
  double* points;               /* 3 * num_points, in/out */
  const double* observations;   /* 2 * num_observations, pixels */
  const int32_t* camera_index;  /* per observation */
  const int32_t* point_index;   /* per observation */
  const uint8_t* camera_constant; /* num_cameras flags, or NULL */
  /* GROUPED only: 6 * num_intrinsics [f, ppx, ppy, k1, k2, k3], in/out; the
   * group of each observation (every observation of a camera must use the
   * same group, as the incremental engine builds them); per-group constant
   * flags (SetParameterBlockConstant) or NULL. */
  double* intrinsics;
  const int32_t* intrinsic_index;
  const uint8_t* intrinsic_constant;
  /* GROUPED only: per-group SubsetParameterization (include/ceres/
   * local_parameterization.h:173), bit k set = entry k of
   * [f, ppx, ppy, k1, k2, k3] held constant (the incremental engine's
   * focal-only refinement is 0x3E, sfm_incremental_engine.cpp:1396-1409);
   * NULL or 0: all six free. */
  const uint8_t* intrinsic_subset;
  /* Which entries of every camera vector are parameters (bit k = entry k);
   * the others are fixed data of the residual, not part of the parameter
   * block (they take no step and do not count in |x|).  0 = all (PACKED9
   * 0x1FF, GROUPED 0x3F).  The global engine's functors
   * (libs/sfm/include/mvg/sfm/sfm_bundle_adjustment_helper_tonly.h):
   *   PinholeReprojectionError_Rt <2,6,3>: PACKED9, 0x3F, focal in entry 6,
   *     k1 = k2 = 0 (sfm_global_engine.cpp:1144-1174, 1758-1802);
   *   PinholeReprojectionError_t <2,3,3>: PACKED9, 0x38 (translation only;
   *     rotation and focal are data, sfm_global_engine.cpp:1548-1594). */
  int32_t camera_params;
  /* mba_ordering: the Schur elimination ordering (Solver::Options::
   * linear_solver_ordering, include/ceres/solver.h:432-520). */
  int32_t ordering;
} mba_problem;

typedef struct mba_options {
  /* Solver::Options fields (defaults of include/ceres/solver.h:62-129). */
  int32_t max_num_iterations;                 /* 50 */
  int32_t max_num_consecutive_invalid_steps;  /* 5 */
  int32_t max_linear_solver_iterations;       /* 500 */
  int32_t min_linear_solver_iterations;       /* 0 */
  int32_t jacobi_scaling;                     /* 1 */
  int32_t linear_solver;                      /* mba_linear_solver, default EXPLICIT */
  int32_t precision;                          /* mba_precision, default FP64 */
  int32_t device;                             /* HIP device ordinal, default 0 */
  double function_tolerance;                  /* 1e-6 */
  double gradient_tolerance;                  /* 1e-10 */
  double parameter_tolerance;                 /* 1e-8 */
  double initial_trust_region_radius;         /* 1e4 */
  double max_trust_region_radius;             /* 1e16 */
  double min_trust_region_radius;             /* 1e-32 */
  double min_relative_decrease;               /* 1e-3 */
  double min_lm_diagonal;                     /* 1e-6 */
  double max_lm_diagonal;                     /* 1e32 */
  double eta;                                 /* 0.1: CG q-tolerance */
  /* PCG relative-residual tolerance |r| <= r_tolerance |b|.  The reference
   * LM strategy disables it (-1, levenberg_marquardt_strategy.cc:91-96). */
  double pcg_r_tolerance;                     /* -1 */
  double huber_a;                             /* <= 0: no loss (NULL) */
  /* Explicit Schur, MIXED precision: FP64 iterative-refinement sweeps after
   * the FP32 Cholesky solve of S; 0 disables.  An FP64 factorisation is
   * used once, as the reference's sparse Cholesky solve is. */
  int32_t refinement_iterations;              /* 2 */
  int32_t preconditioner;                     /* mba_preconditioner, JACOBI */
  /* Trust-region problem dumps, TEXTFILE format
   * (Solver::Options::trust_region_minimizer_iterations_to_dump and
   * trust_region_problem_dump_directory; trust_region_minimizer.cc:287-301,
   * linear_least_squares_problems.cc:534-603).  For every listed k, the
   * linear least-squares problem of the step computed after iteration
   * record k is written as <dump_directory>/ceres_solver_iteration_%03d
   * {_A.txt, _D.txt, _b.txt, _x.txt, .m}: A the Jacobi-scaled, robustified
   * Jacobian of the reduced program ("row col value" per entry, row blocks
   * in residual order, E columns first), D the LM diagonal, b the
   * robustified residuals, x the step.  A is evaluated in FP64 at the
   * current parameters.  mba_solve only; a non-empty list needs a non-empty
   * directory (MBA_ERR_INVALID_ARGUMENT, as solver.cc:243-247). */
  const char* dump_directory;                 /* NULL */
  const int32_t* dump_iterations;             /* NULL */
  int32_t num_dump_iterations;                /* 0 */
  /* Launch shapes (no counterpart in the reference; 0 = the library's
   * tuned default for the problem's size, from its autotuning table).
   * point_rows_per_lane: rows each lane of the point-major kernels takes
   * (a point of track length t gets the smallest power-of-two lane count
   * L <= 64 with t <= point_rows_per_lane * L); 1..64.
   * camera_chunk_rows: rows per camera-major chunk (one workgroup's
   * partial sums), rounded up to a multiple of 64; 64..8192. */
  int32_t point_rows_per_lane;                /* 0 */
  int32_t camera_chunk_rows;                  /* 0 */
  /* Implicit-Schur PCG launches (no counterpart in the reference): 0
   * launches each CG iteration's kernels one by one, the host queueing up to
   * four iterations ahead; 1 replays the iteration as a HIP graph captured
   * on first use.  Same kernels, same arguments, bit-identical results; on
   * MI355X (ROCm 7.2) the graph replay costs 8-13 us more per iteration. */
  int32_t pcg_graph_launch;                   /* 0 */
} mba_options;

typedef struct mba_iteration {
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
  double cumulative_time_s;
} mba_iteration;

typedef struct mba_summary {
  double initial_cost;
  double final_cost;
  double fixed_cost;
  double preprocessor_time_s;
  double jacobian_evaluation_time_s;   /* residual + Jacobian + Schur blocks */
  double residual_evaluation_time_s;   /* cost-only evaluations */
  double linear_solver_time_s;
  double total_time_s;
  int32_t num_successful_steps;
  int32_t num_unsuccessful_steps;
  int32_t num_iterations;              /* iteration records produced */
  int32_t termination_type;            /* mba_termination */
  int32_t num_e_blocks;                /* eliminated (point) blocks */
  int32_t num_f_blocks;                /* reduced (camera) blocks */
  int32_t num_linear_solver_iterations;
  int32_t is_solution_usable;
} mba_summary;

/* Options with the reference defaults. */
mba_status mba_default_options(mba_options* options);

/* ceres::Solve for a bundle-adjustment problem.  `iterations` (may be NULL)
 * receives up to `max_iterations` IterationSummary records. */
mba_status mba_solve(const mba_options* options, mba_problem* problem,
                     mba_summary* summary, mba_iteration* iterations,
                     int32_t max_iterations);

/* Many independent problems ("windows", e.g. local-BA windows of uneven
 * size): the role of the global engine's concurrent ceres::Solve calls over
 * image pairs (libs/sfm/src/sfm_global_engine.cpp:985-1001, one Solve per
 * pair at :1168).  With MBA_IMPLICIT_SCHUR_PCG every window is solved in the
 * same device launches: each window keeps its own trust region, PCG state
 * and termination, exactly as its own mba_solve would decide them (sums are
 * accumulated in another order, so values agree to rounding).  With
 * MBA_EXPLICIT_SCHUR (the DENSE_SCHUR role of the 2-view problems,
 * sfm_global_engine.cpp:1168) the windows also share launches: each
 * window's dense S is assembled in its own region and factored (FP64) by
 * one workgroup; windows of more than 64 cameras, and grouped-intrinsics
 * problems, are solved one after another.  summaries[i] is window i's; its
 * time fields are the whole batch's.  Usable windows are written back. */
mba_status mba_solve_batched(const mba_options* options, mba_problem* problems,
                             int32_t num_problems, mba_summary* summaries);

/* CostFunction::Evaluate for every observation, input order, computed on the
 * device in `precision`.  residuals: 2 per observation (no loss applied);
 * jac_camera: row-major 2x9 per observation (2x6 for GROUPED); jac_point:
 * row-major 2x3; each may be NULL.  cost: robustified 1/2 sum rho (FP64),
 * may be NULL.  Same as mba_evaluate_blocks with jac_intrinsics NULL. */
mba_status mba_evaluate(const mba_problem* problem, int32_t precision,
                        double huber_a, double* residuals, double* jac_camera,
                        double* jac_point, double* cost, int32_t device);

/* mba_evaluate for either layout with the intrinsics block too: GROUPED
 * gives jac_intrinsics 2x6 ([f, ppx, ppy, k1, k2, k3]) and jac_camera 2x6
 * ([aa, t]) per observation; PACKED9 gives jac_camera 2x9 and leaves
 * jac_intrinsics untouched.  Any output may be NULL. */
mba_status mba_evaluate_blocks(const mba_problem* problem, int32_t precision,
                               double huber_a, double* residuals,
                               double* jac_intrinsics, double* jac_camera,
                               double* jac_point, double* cost, int32_t device);

/* mba_evaluate for a PACKED9 problem whose arrays are already in device
 * memory, with no allocation of any kind (no counterpart in the reference:
 * for callers that cannot allocate, such as code under HIP graph capture or
 * a tight serving loop).  Every pointer is device memory on `device`:
 * cameras 9 per camera, points 3 per point, observations 2 per
 * observation, camera_index / point_index per observation; residuals (2),
 * jac_camera (row-major 2x9), jac_point (row-major 2x3) per observation,
 * and cost (one double: the FP64 sum of 1/2 rho(|r|^2) with HuberLoss
 * threshold huber_a, <= 0 no loss); each output may be NULL.  Values are
 * those of mba_evaluate in the same precision.  Asynchronous on `stream`
 * (a hipStream_t; NULL is the null stream): the outputs are valid once the
 * stream reaches this point.  Intermediates live in a fixed device scratch
 * array of the library (one per device, held until the process ends), so
 * at most 16384 cameras (more: MBA_ERR_INVALID_ARGUMENT); calls that
 * overlap on one device race on it.  An observation whose camera or point
 * index is out of range gets NaN outputs (and makes the cost NaN); nothing
 * is read out of bounds. */
mba_status mba_evaluate_noalloc(int32_t num_cameras, int32_t num_points, int64_t num_observations,
                                const double* cameras, const double* points, const double* observations,
                                const int32_t* camera_index, const int32_t* point_index, int32_t precision,
                                double huber_a, double* residuals, double* jac_camera, double* jac_point,
                                double* cost, int32_t device, void* stream);

/* The bit-exact structure the reference's preprocessor builds
 * (reorder_program.cc:436-531, schur_complement_solver.cc:246-313):
 *   param_order: column blocks in order (PACKED9: cameras 0..Nc-1, points
 *     Nc..; GROUPED: intrinsic groups 0..Ni-1, cameras Ni.., points Ni+Nc..),
 *     room for every block (Ni + Nc + Np);
 *   num_e_blocks: size of the eliminated group;
 *   residual_order: input observation index of each row block;
 *   pairs: upper-triangle block pairs of S (reduced indices), set order.
 * Any output may be NULL; *num_pairs receives the pair count; returns
 * MBA_ERR_INVALID_ARGUMENT if pairs_capacity is too small. */
mba_status mba_schur_structure(const mba_problem* problem, int32_t* param_order,
                               int32_t* num_param_blocks, int32_t* num_e_blocks,
                               int64_t* residual_order, int32_t* pairs,
                               int64_t pairs_capacity, int64_t* num_pairs);

/* ------------------------------------------------- resident contexts ----
 * A context owns the preprocessed problem on one device (ordering, device
 * copies, stream).  It lets a caller (or a benchmark) run the engine's
 * passes repeatedly on HBM-resident data. */
typedef struct mba_context mba_context;

mba_status mba_context_create(const mba_problem* problem,
                              const mba_options* options, mba_context** out);
mba_status mba_context_destroy(mba_context* ctx);

/* One Jacobian + Schur pass at the context's current parameters: residuals,
 * Jacobians (Huber-corrected, Jacobi-scaled), V_p = E'E + D_p^2 and its FP64
 * inverse, g_p, U_c = F'F + D_c^2, g_c and the reduced right-hand side
 * rhs_c = F'(r - E V^-1 g_p), for the LM diagonal at trust radius `radius`
 * (radius <= 0: no damping).  Asynchronous on the context stream. */
mba_status mba_context_jacobian_schur(mba_context* ctx, double radius);

/* The Jacobian + Schur pass as the LM loop runs it at the start of an
 * iteration: the FP64 cost of the context's current parameters first (into
 * *cost, may be NULL), whose residuals, rounded once to FP32, are the
 * residuals of a MIXED context's pass (gradient, rhs and Huber weights);
 * then the pass of mba_context_jacobian_schur at trust radius `radius`.  An
 * FP64 context computes the same blocks as mba_context_jacobian_schur.
 * Synchronous for the cost; the pass is queued on the context stream. */
mba_status mba_context_lm_pass(mba_context* ctx, double radius, double* cost);

/* Copy the blocks of the last pass to the host, in INPUT ids: U (45 per
 * camera, upper triangle row-major), g_c (9), rhs (9), V (6 per point:
 * xx xy xz yy yz zz), g_p (3).  Any output may be NULL. */
mba_status mba_context_read_schur(mba_context* ctx, double* U, double* gc,
                                  double* rhs, double* V, double* gp);

/* Jacobi column scaling in use (1 / (1 + |column|) of the first Jacobian,
 * trust_region_minimizer.cc:103-109), input ids: 9 per camera, 3 per point
 * (1 for blocks outside the reduced program).  Either output may be NULL. */
mba_status mba_context_read_scale(mba_context* ctx, double* cam_scale, double* pt_scale);

/* Implicit reduced-camera matvec y = S x of the last pass (S = U + D^2 -
 * W V^-1 W'), x and y 9 per camera in input ids (host arrays). */
mba_status mba_context_schur_matvec(mba_context* ctx, const double* x, double* y);

/* Run the full LM solve on the context (parameters stay on the device until
 * mba_context_download). */
mba_status mba_context_solve(mba_context* ctx, mba_summary* summary,
                             mba_iteration* iterations, int32_t max_iterations);

/* Copy the context parameters back to the problem arrays. */
mba_status mba_context_download(mba_context* ctx, mba_problem* problem);

/* Block until the context stream is idle. */
mba_status mba_context_synchronize(mba_context* ctx);

/* The HIP stream (hipStream_t) the context launches on. */
void* mba_context_stream(mba_context* ctx);

/* Timing measured with HIP events on the context stream, enabled with
 * mba_context_set_timing(ctx, 1).  Of the last Jacobian + Schur pass:
 * ms[0] point pass, ms[1] camera pass, ms[2] total; of the last explicit
 * Schur solve (0 if none ran): ms[3] assembly of S, ms[4] Cholesky
 * factorisation, ms[5] triangular solves with refinement.  n <= 6 values. */
mba_status mba_context_set_timing(mba_context* ctx, int32_t enable);
mba_status mba_context_last_timing(mba_context* ctx, double* ms, int32_t n);

/* ----------------------------------------- conversions and file formats --
 * Host-only (no device needed). */

/* RotationMatrixToAngleAxis / AngleAxisToRotationMatrix
 * (include/ceres/rotation.h:357-415, the conversions the incremental engine
 * applies on entry and exit, sfm_incremental_engine.cpp:1306,1505): n
 * column-major 3x3 matrices <-> n angle-axis vectors. */
mba_status mba_rotation_matrix_to_angle_axis(const double* R, double* angle_axis, int64_t n);
mba_status mba_angle_axis_to_rotation_matrix(const double* angle_axis, double* R, int64_t n);

/* A libmv bundle-adjustment problem file (the reader of
 * 3rdparty/ceres/examples/libmv_bundle_adjuster.cc:255-437; the files of
 * 3rdparty/ceres/data/libmv-ba-problems) as a GROUPED problem with one
 * intrinsic group: cameras [aa, t] (RotationMatrixToAngleAxis of the file's
 * R), points, and the markers whose image and track exist, in file order;
 * cameras and points numbered in ascending image / track id
 * (image_of_camera, track_of_point).  intrinsics: 6 values [f, ppx, ppy,
 * k1, k2, k3]; image_space: 1 for pixel markers.  Tangential distortion
 * (p1, p2 != 0) is MBA_ERR_UNSUPPORTED.  Call with NULL arrays for the
 * sizes, then with arrays of those sizes. */
mba_status mba_read_libmv(const char* path, int32_t* num_cameras, int32_t* num_points,
                          int64_t* num_observations, int32_t* image_space, double* intrinsics,
                          double* cameras, double* points, double* observations,
                          int32_t* camera_index, int32_t* point_index,
                          int32_t* image_of_camera, int32_t* track_of_point);

/* BAL text problems (examples/bal_problem.cc:72-175) as PACKED9 problems.
 * A BAL camera looks down -z (Snavely's model); the libs/sfm model predicts
 * the negated image point with the same camera vector, so observations are
 * negated on reading and writing (cost and normal equations unchanged).
 * mba_read_bal: NULL arrays give the sizes.  mba_write_bal: the text of
 * BALProblem::WriteToFile. */
mba_status mba_read_bal(const char* path, int32_t* num_cameras, int32_t* num_points,
                        int64_t* num_observations, double* cameras, double* points,
                        double* observations, int32_t* camera_index, int32_t* point_index);
mba_status mba_write_bal(const char* path, const mba_problem* problem);

/* --------------------------------------------------------- utilities --- */
const char* mba_status_string(mba_status status);
const char* mba_last_error(void);
int32_t mba_abi_version(void);

/* Seeded synthetic scene (no counterpart in the reference; the reference
 * test builds scenes with NRealisticCamerasRing,
 * libs/multiview/src/nview_data_sets.cpp:23-64).  preset 1..5 are the
 * packed-camera benchmark scenes (4 = window `window` of the batched set).  Call with
 * NULL arrays to get the sizes, then with arrays of those sizes. */
mba_status mba_scene_generate(int32_t preset, int32_t window, int32_t* num_cameras,
                              int32_t* num_points, int64_t* num_observations,
                              double* cam_true, double* cam_init, double* pt_true,
                              double* pt_init, double* observations,
                              int32_t* camera_index, int32_t* point_index);

/* Seeded grouped-intrinsics scene (preset 6: 300 cameras in 2 intrinsic
 * groups, 100,000 points, tracks of 2..10, 5% outliers of 20 px; no
 * counterpart in the reference): cameras 6 per camera [aa, t], intrinsics
 * 6 per group.  Call with NULL arrays to get the sizes. */
mba_status mba_scene_generate_grouped(int32_t preset, int32_t* num_cameras, int32_t* num_points,
                                      int32_t* num_intrinsics, int64_t* num_observations,
                                      double* cam_init, double* intr_init, double* pt_init,
                                      double* observations, int32_t* camera_index,
                                      int32_t* intrinsic_index, int32_t* point_index);

#ifdef __cplusplus
}
#endif
#endif /* MBA_H_ */
