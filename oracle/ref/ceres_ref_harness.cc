// Harness around the reference's own Ceres Solver 1.10 (compiled from
// 3rdparty/ceres, 3rdparty/cxsparse and 3rdparty/eigen3 of the reference by
// oracle/ref/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: it is
// used in this container to pin the FP64 oracle and to write the golden
// fixtures under tests/golden/.  Nothing in the product links it.
//
// The libs/sfm residual functor (libs/sfm/include/mvg/sfm/
// pinhole_brown_rt_ceres_functor.h:60-107) cannot be included directly: its
// header chain needs the CMake-generated mvg/config.h.  Its arithmetic
// (ceres::AngleAxisRotatePoint from the reference's Ceres headers + the
// radial polynomial of brown_pinhole_camera.h:14-41) is therefore written
// out below, operation for operation, as a Ceres autodiff functor.  The
// global engine's functors (PinholeReprojectionError_Rt / _t) need only the
// Ceres headers and are included from the reference as they are.

#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "ceres/ceres.h"
#include "ceres/rotation.h"
#include "mvg/sfm/sfm_bundle_adjustment_helper_tonly.h"
#include "bal_problem.h"                 // 3rdparty/ceres/examples
#include "snavely_reprojection_error.h"  // 3rdparty/ceres/examples
#include "ceres/problem_impl.h"
#include "ceres/program.h"
#include "ceres/parameter_block.h"
#include "ceres/residual_block.h"
#include "ceres/reorder_program.h"
#include "ceres/ordered_groups.h"
#include "ceres/evaluator.h"
#include "ceres/block_sparse_matrix.h"
#include "ceres/linear_solver.h"
#define private public
#define protected public
#include "ceres/schur_complement_solver.h"
#include "ceres/block_random_access_sparse_matrix.h"
#undef private
#undef protected
#include "ceres/detect_structure.h"
#include "ceres/schur_eliminator.h"
#include <chrono>

namespace {

// [aa(3), t(3), f, k1, k2] camera; intrinsics [f, 0, 0, k1, k2, 0].
struct Packed9Functor {
  explicit Packed9Functor(const double* o) { obs[0] = o[0]; obs[1] = o[1]; }
  template <typename T>
  bool operator()(const T* const cam, const T* const X, T* r) const {
    T x[3];
    ceres::AngleAxisRotatePoint(cam, X, x);
    x[0] += cam[3];
    x[1] += cam[4];
    x[2] += cam[5];
    const T xe = x[0] / x[2];
    const T ye = x[1] / x[2];
    const T& f = cam[6];
    const T ppx = T(0.0), ppy = T(0.0), k3 = T(0.0);
    const T& k1 = cam[7];
    const T& k2 = cam[8];
    const T r2 = xe * xe + ye * ye;
    const T r4 = r2 * r2;
    const T r6 = r4 * r2;
    const T rc = (T(1) + k1 * r2 + k2 * r4 + k3 * r6);
    const T xd = xe * rc;
    const T yd = ye * rc;
    const T px = ppx + f * xd;
    const T py = ppy + f * yd;
    r[0] = px - T(obs[0]);
    r[1] = py - T(obs[1]);
    return true;
  }
  double obs[2];
};

// Intrinsic block [f, ppx, ppy, k1, k2, k3], extrinsic [aa, t], point.
struct Intr6Ext6Functor {
  explicit Intr6Ext6Functor(const double* o) { obs[0] = o[0]; obs[1] = o[1]; }
  template <typename T>
  bool operator()(const T* const in, const T* const rt, const T* const X, T* r) const {
    T x[3];
    ceres::AngleAxisRotatePoint(rt, X, x);
    x[0] += rt[3];
    x[1] += rt[4];
    x[2] += rt[5];
    const T xe = x[0] / x[2];
    const T ye = x[1] / x[2];
    const T r2 = xe * xe + ye * ye;
    const T r4 = r2 * r2;
    const T r6 = r4 * r2;
    const T rc = (T(1) + in[3] * r2 + in[4] * r4 + in[5] * r6);
    const T xd = xe * rc;
    const T yd = ye * rc;
    const T px = in[1] + in[0] * xd;
    const T py = in[2] + in[0] * yd;
    r[0] = px - T(obs[0]);
    r[1] = py - T(obs[1]);
    return true;
  }
  double obs[2];
};

struct Scene {
  int layout, nc, ni, np;
  long nobs;
  double *cams, *intr, *pts;
  const double* obs;
  const int *cam_idx, *intr_idx, *pt_idx;
  const unsigned char* cam_const;
  // Layout 1: per-group SetParameterBlockConstant flags and
  // SubsetParameterization masks (bit k: entry k constant); may be NULL.
  const unsigned char* intr_const;
  const unsigned char* intr_subset;
};

// Layouts: 0 packed [aa, t, f, k1, k2]; 1 (intrinsics, [aa, t], point);
// 2 PinholeReprojectionError_Rt ([aa, t], focal = intr[0]); 3
// PinholeReprojectionError_t ([t], focal = intr[0], rotation of camera c at
// intr + 1 + 3 c).
int CamStride(int layout) { return layout == 0 ? 9 : layout == 3 ? 3 : 6; }

ceres::CostFunction* MakeCost(const Scene& s, long i) {
  if (s.layout == 0)
    return new ceres::AutoDiffCostFunction<Packed9Functor, 2, 9, 3>(new Packed9Functor(s.obs + 2 * i));
  if (s.layout == 2)
    return new ceres::AutoDiffCostFunction<mvg::sfm::PinholeReprojectionError_Rt, 2, 6, 3>(
        new mvg::sfm::PinholeReprojectionError_Rt(s.obs + 2 * i, s.intr[0]));
  if (s.layout == 3)
    return new ceres::AutoDiffCostFunction<mvg::sfm::PinholeReprojectionError_t, 2, 3, 3>(
        new mvg::sfm::PinholeReprojectionError_t(s.obs + 2 * i, s.intr[0], s.intr + 1 + 3L * s.cam_idx[i]));
  return new ceres::AutoDiffCostFunction<Intr6Ext6Functor, 2, 6, 6, 3>(new Intr6Ext6Functor(s.obs + 2 * i));
}

// Block id of a user-state pointer: layout 0 cameras then points; layout 1
// intrinsics, extrinsics, points.
int BlockId(const Scene& s, const double* p) {
  const int cs = CamStride(s.layout);
  if (s.layout == 1 && p >= s.intr && p < s.intr + 6 * s.ni) return (int)((p - s.intr) / 6);
  const int ni = s.layout == 1 ? s.ni : 0;
  if (p >= s.cams && p < s.cams + (long)cs * s.nc) return ni + (int)((p - s.cams) / cs);
  return ni + s.nc + (int)((p - s.pts) / 3);
}

template <typename P>
void AddAll(const Scene& s, P* problem, ceres::LossFunction* loss,
            std::map<const ceres::CostFunction*, long>* ids) {
  const int cs = CamStride(s.layout);
  for (long i = 0; i < s.nobs; ++i) {
    ceres::CostFunction* cf = MakeCost(s, i);
    if (ids) (*ids)[cf] = i;
    double* cam = s.cams + (long)cs * s.cam_idx[i];
    double* pt = s.pts + 3L * s.pt_idx[i];
    if (s.layout != 1) {
      problem->AddResidualBlock(cf, loss, cam, pt);
    } else {
      problem->AddResidualBlock(cf, loss, s.intr + 6L * s.intr_idx[i], cam, pt);
    }
  }
  if (s.cam_const)
    for (int c = 0; c < s.nc; ++c)
      if (s.cam_const[c]) problem->SetParameterBlockConstant(s.cams + (long)cs * c);
  if (s.layout == 1) {
    // One SubsetParameterization per distinct mask, shared by the groups
    // that use it (the incremental engine shares one, :1396-1436).
    std::map<int, ceres::LocalParameterization*> param;
    for (int g = 0; g < s.ni; ++g) {
      double* blk = s.intr + 6L * g;
      if (s.intr_const && s.intr_const[g]) {
        problem->SetParameterBlockConstant(blk);
      } else if (s.intr_subset && s.intr_subset[g]) {
        const int m = s.intr_subset[g];
        if (!param.count(m)) {
          std::vector<int> held;
          for (int k = 0; k < 6; ++k)
            if ((m >> k) & 1) held.push_back(k);
          param[m] = new ceres::SubsetParameterization(6, held);
        }
        problem->SetParameterization(blk, param[m]);
      }
    }
  }
}

// User ordering "points first" (Solver::Options::linear_solver_ordering
// with every point block in group 0, the rest in group 1); set by
// ref_set_points_first.  Ceres keeps each group in address order, so for
// layout 1 the cameras and intrinsics are moved into one buffer
// [cameras | intrinsics] (the libs/sfm container order) while solving.
int g_points_first = 0;
// Preconditioner of ITERATIVE_SCHUR in ref_solve: 0 JACOBI, 1 SCHUR_JACOBI,
// 2 CLUSTER_JACOBI, 3 CLUSTER_TRIDIAGONAL (the last two need SuiteSparse,
// which the reference's Ceres build leaves off: Solve rejects the options).
int g_precond = 0;
// Solver::Options::trust_region_minimizer_iterations_to_dump and
// trust_region_problem_dump_directory (TEXTFILE) of ref_solve.
std::vector<int> g_dump_iters;
std::string g_dump_dir;
// Summary::message of the last ref_solve.
std::string g_message;
// Solver::Options::eta and ::function_tolerance of ref_solve; a value <= 0
// keeps Ceres's default (eta 1e-1, function tolerance 1e-6).
double g_eta = 0.0;
double g_function_tolerance = 0.0;

ceres::PreconditionerType PrecondType(int k) {
  switch (k) {
    case 1: return ceres::SCHUR_JACOBI;
    case 2: return ceres::CLUSTER_JACOBI;
    case 3: return ceres::CLUSTER_TRIDIAGONAL;
    default: return ceres::JACOBI;
  }
}

void AddPointsFirst(const Scene& s, const std::vector<double*>& blocks, ceres::ParameterBlockOrdering* o) {
  for (double* b : blocks) o->AddElementToGroup(b, (b >= s.pts && b < s.pts + 3L * s.np) ? 0 : 1);
}

struct ContiguousCamsIntr {
  Scene* s;
  std::vector<double> buf;
  double *cams0, *intr0;
  explicit ContiguousCamsIntr(Scene* sc) : s(sc), cams0(sc->cams), intr0(sc->intr) {
    if (!g_points_first || s->layout != 1) return;
    buf.assign(cams0, cams0 + 6L * s->nc);
    buf.insert(buf.end(), intr0, intr0 + 6L * s->ni);
    s->cams = buf.data();
    s->intr = buf.data() + 6L * s->nc;
  }
  ~ContiguousCamsIntr() {
    if (buf.empty()) return;
    std::copy(buf.begin(), buf.begin() + 6L * s->nc, cams0);
    std::copy(buf.begin() + 6L * s->nc, buf.end(), intr0);
    s->cams = cams0;
    s->intr = intr0;
  }
};

}  // namespace

extern "C" {

void ref_set_points_first(int on) { g_points_first = on; }
void ref_set_schur_jacobi(int on) { g_precond = on ? 1 : 0; }
void ref_set_preconditioner(int k) { g_precond = k; }
void ref_set_lm_tolerances(double eta, double function_tolerance) {
  g_eta = eta;
  g_function_tolerance = function_tolerance;
}
void ref_set_dump(const char* dir, const int* iters, int n) {
  g_dump_dir = dir ? dir : "";
  g_dump_iters.assign(iters, iters + (n > 0 ? n : 0));
}
// Copies the last ref_solve's Summary::message; returns its length.
int ref_last_message(char* buf, int cap) {
  if (buf && cap > 0) {
    const int n = std::min<int>(cap - 1, (int)g_message.size());
    std::copy(g_message.begin(), g_message.begin() + n, buf);
    buf[n] = 0;
  }
  return (int)g_message.size();
}

// ceres::RotationMatrixToAngleAxis on n column-major matrices.
void ref_rotation_matrix_to_angle_axis(const double* R, double* aa, long n) {
  for (long i = 0; i < n; ++i) ceres::RotationMatrixToAngleAxis(R + 9 * i, aa + 3 * i);
}

// BALProblem (examples/bal_problem.cc): read `in`, write it back to `out`.
int ref_bal_roundtrip(const char* in, const char* out) {
  ceres::examples::BALProblem bal(in, false);
  bal.WriteToFile(out);
  return 1;
}

// BALProblem's arrays: sizes[3] = {Nc, Np, Nobs}; params 9 Nc + 3 Np.
int ref_bal_arrays(const char* in, int* sizes, double* params, double* obs, int* cam_idx, int* pt_idx) {
  ceres::examples::BALProblem bal(in, false);
  sizes[0] = bal.num_cameras();
  sizes[1] = bal.num_points();
  sizes[2] = bal.num_observations();
  if (params) std::copy(bal.parameters(), bal.parameters() + bal.num_parameters(), params);
  if (obs) std::copy(bal.observations(), bal.observations() + 2 * bal.num_observations(), obs);
  if (cam_idx) std::copy(bal.camera_index(), bal.camera_index() + bal.num_observations(), cam_idx);
  if (pt_idx) std::copy(bal.point_index(), bal.point_index() + bal.num_observations(), pt_idx);
  return 1;
}

// A BAL file solved with the examples' SnavelyReprojectionError (the
// bundle_adjuster example's functor), options as ref_solve; summary and
// iteration records as ref_solve, final parameters into params_out.
int ref_bal_solve(const char* in, double huber_a, int linear_solver, int max_iters, double* summary,
                  double* iters, int max_rec, double* params_out) {
  ceres::examples::BALProblem bal(in, false);
  ceres::Problem problem;
  ceres::LossFunction* loss = huber_a > 0.0 ? new ceres::HuberLoss(huber_a) : NULL;
  const double* o = bal.observations();
  for (int i = 0; i < bal.num_observations(); ++i)
    problem.AddResidualBlock(ceres::examples::SnavelyReprojectionError::Create(o[2 * i], o[2 * i + 1]), loss,
                             bal.mutable_cameras() + 9 * bal.camera_index()[i],
                             bal.mutable_points() + 3 * bal.point_index()[i]);
  ceres::Solver::Options options;
  options.preconditioner_type = ceres::JACOBI;
  options.linear_solver_type = linear_solver == 1 ? ceres::ITERATIVE_SCHUR : ceres::SPARSE_SCHUR;
  options.sparse_linear_algebra_library_type = ceres::CX_SPARSE;
  options.logging_type = ceres::SILENT;
  options.max_num_iterations = max_iters;
  ceres::Solver::Summary sum;
  ceres::Solve(options, &problem, &sum);
  summary[0] = sum.initial_cost;
  summary[1] = sum.final_cost;
  summary[2] = sum.num_successful_steps;
  summary[3] = sum.num_unsuccessful_steps;
  summary[4] = sum.termination_type;
  summary[5] = (double)sum.iterations.size();
  for (int k = 0; k < (int)sum.iterations.size() && k < max_rec; ++k) {
    const ceres::IterationSummary& it = sum.iterations[k];
    double* r = iters + 9 * k;
    r[0] = it.iteration; r[1] = it.cost; r[2] = it.cost_change; r[3] = it.trust_region_radius;
    r[4] = it.step_is_successful; r[5] = it.step_is_valid; r[6] = it.linear_solver_iterations;
    r[7] = it.relative_decrease; r[8] = it.gradient_max_norm;
  }
  if (params_out) std::copy(bal.parameters(), bal.parameters() + bal.num_parameters(), params_out);
  return sum.IsSolutionUsable() ? 1 : 0;
}

// AutoDiffCostFunction::Evaluate for one observation (no loss).
int ref_evaluate_obs(int layout, const double* intr, const double* cam, const double* pt,
                     const double* obs, double* residual, double* j0, double* j1, double* j2) {
  if (layout == 0) {
    ceres::AutoDiffCostFunction<Packed9Functor, 2, 9, 3> cf(new Packed9Functor(obs));
    const double* params[2] = {cam, pt};
    double* jac[2] = {j0, j1};
    return cf.Evaluate(params, residual, (j0 || j1) ? jac : NULL) ? 1 : 0;
  }
  ceres::AutoDiffCostFunction<Intr6Ext6Functor, 2, 6, 6, 3> cf(new Intr6Ext6Functor(obs));
  const double* params[3] = {intr, cam, pt};
  double* jac[3] = {j0, j1, j2};
  return cf.Evaluate(params, residual, (j0 || j1 || j2) ? jac : NULL) ? 1 : 0;
}

// Schur ordering of the reduced program through Ceres's own preprocessing
// calls (program.cc:269, reorder_program.cc:436-531).  Returns num E blocks.
int ref_schur_ordering(int layout, int nc, int ni, int np, long nobs, double* cams, double* intr,
                       double* pts, const double* obs, const int* cam_idx, const int* intr_idx,
                       const int* pt_idx, const unsigned char* cam_const, int* param_order,
                       int* num_blocks, long* residual_order, long* num_residuals,
                       int* pairs, long pairs_cap, long* num_pairs) {
  Scene s = {layout, nc, ni, np, nobs, cams, intr, pts, obs, cam_idx, intr_idx, pt_idx, cam_const};
  ContiguousCamsIntr contiguous(&s);
  ceres::Problem::Options po;
  ceres::internal::ProblemImpl problem(po);
  std::map<const ceres::CostFunction*, long> ids;
  AddAll(s, &problem, NULL, &ids);
  std::vector<double*> removed;
  double fixed_cost = 0.0;
  std::string error;
  ceres::internal::scoped_ptr<ceres::internal::Program> reduced(
      problem.mutable_program()->CreateReducedProgram(&removed, &fixed_cost, &error));
  if (reduced.get() == NULL) return -1;
  ceres::ParameterBlockOrdering ordering;
  std::vector<double*> blocks;
  for (size_t i = 0; i < reduced->parameter_blocks().size(); ++i)
    blocks.push_back(reduced->parameter_blocks()[i]->mutable_user_state());
  if (g_points_first) {
    AddPointsFirst(s, blocks, &ordering);
  } else {
    for (double* b : blocks) ordering.AddElementToGroup(b, 0);
  }
  if (!ceres::internal::ReorderProgramForSchurTypeLinearSolver(
          ceres::SPARSE_SCHUR, ceres::CX_SPARSE, problem.parameter_map(), &ordering,
          reduced.get(), &error))
    return -2;
  const int ne = ordering.group_to_elements().begin()->second.size();
  const std::vector<ceres::internal::ParameterBlock*>& pbs = reduced->parameter_blocks();
  for (size_t i = 0; i < pbs.size(); ++i) param_order[i] = BlockId(s, pbs[i]->user_state());
  *num_blocks = (int)pbs.size();
  const std::vector<ceres::internal::ResidualBlock*>& rbs = reduced->residual_blocks();
  for (size_t i = 0; i < rbs.size(); ++i) residual_order[i] = ids[rbs[i]->cost_function()];
  *num_residuals = (long)rbs.size();

  // S block pairs from SparseSchurComplementSolver::InitStorage itself.
  if (pairs != NULL || num_pairs != NULL) {
    ceres::internal::Evaluator::Options eo;
    eo.linear_solver_type = ceres::SPARSE_SCHUR;
    eo.num_eliminate_blocks = ne;
    eo.num_threads = 1;
    ceres::internal::scoped_ptr<ceres::internal::Evaluator> ev(
        ceres::internal::Evaluator::Create(eo, reduced.get(), &error));
    ceres::internal::scoped_ptr<ceres::internal::SparseMatrix> jac(ev->CreateJacobian());
    ceres::internal::BlockSparseMatrix* bsm =
        static_cast<ceres::internal::BlockSparseMatrix*>(jac.get());
    ceres::internal::LinearSolver::Options lo;
    lo.type = ceres::SPARSE_SCHUR;
    lo.sparse_linear_algebra_library_type = ceres::CX_SPARSE;
    lo.elimination_groups.push_back(ne);
    lo.elimination_groups.push_back((int)pbs.size() - ne);
    ceres::internal::SparseSchurComplementSolver solver(lo);
    solver.InitStorage(bsm->block_structure());
    const ceres::internal::BlockRandomAccessSparseMatrix* m =
        static_cast<const ceres::internal::BlockRandomAccessSparseMatrix*>(solver.lhs());
    const long np_ = (long)m->cell_values_.size();
    if (num_pairs) *num_pairs = np_;
    if (pairs) {
      if (np_ > pairs_cap) return -3;
      for (long i = 0; i < np_; ++i) {
        pairs[2 * i] = m->cell_values_[i].first.first;
        pairs[2 * i + 1] = m->cell_values_[i].first.second;
      }
    }
  }
  return ne;
}

// One "Jacobian + Schur" pass of the reference's SPARSE_SCHUR path, as its
// LM iteration runs it: ProgramEvaluator::Evaluate (residuals, robustified
// Jacobian; program_evaluator.h:138-285), BlockSparseMatrix::ScaleColumns
// (trust_region_minimizer.cc:239-244), then SchurEliminator::Eliminate into
// the block-sparse S and rhs (schur_complement_solver.cc:120-128), on
// num_threads OpenMP threads.  Problem building, ordering and storage set-up
// are outside the timing.  seconds[0] = mean seconds per pass over `reps`
// passes after one untimed pass; seconds[1] = the set-up time.  Returns the
// number of residual blocks, or < 0 on error.
long ref_jacobian_schur_pass(int layout, int nc, int ni, int np, long nobs, double* cams, double* intr,
                             double* pts, const double* obs, const int* cam_idx, const int* intr_idx,
                             const int* pt_idx, const unsigned char* cam_const, double huber_a,
                             int num_threads, int reps, double* seconds) {
  typedef std::chrono::steady_clock Clock;
  const Clock::time_point t0 = Clock::now();
  Scene s = {layout, nc, ni, np, nobs, cams, intr, pts, obs, cam_idx, intr_idx, pt_idx, cam_const};
  ceres::Problem::Options po;
  ceres::internal::ProblemImpl problem(po);
  ceres::LossFunction* loss = huber_a > 0.0 ? new ceres::HuberLoss(huber_a) : NULL;
  AddAll(s, &problem, loss, NULL);
  std::vector<double*> removed;
  double fixed_cost = 0.0;
  std::string error;
  ceres::internal::scoped_ptr<ceres::internal::Program> reduced(
      problem.mutable_program()->CreateReducedProgram(&removed, &fixed_cost, &error));
  if (reduced.get() == NULL) return -1;
  ceres::ParameterBlockOrdering ordering;
  for (size_t i = 0; i < reduced->parameter_blocks().size(); ++i)
    ordering.AddElementToGroup(reduced->parameter_blocks()[i]->mutable_user_state(), 0);
  if (!ceres::internal::ReorderProgramForSchurTypeLinearSolver(
          ceres::SPARSE_SCHUR, ceres::CX_SPARSE, problem.parameter_map(), &ordering, reduced.get(), &error))
    return -2;
  const int ne = ordering.group_to_elements().begin()->second.size();
  ceres::internal::Evaluator::Options eo;
  eo.linear_solver_type = ceres::SPARSE_SCHUR;
  eo.num_eliminate_blocks = ne;
  eo.num_threads = num_threads;
  ceres::internal::scoped_ptr<ceres::internal::Evaluator> ev(
      ceres::internal::Evaluator::Create(eo, reduced.get(), &error));
  if (ev.get() == NULL) return -3;
  ceres::internal::scoped_ptr<ceres::internal::SparseMatrix> jac(ev->CreateJacobian());
  ceres::internal::BlockSparseMatrix* A = static_cast<ceres::internal::BlockSparseMatrix*>(jac.get());
  ceres::internal::LinearSolver::Options lo;
  lo.type = ceres::SPARSE_SCHUR;
  lo.sparse_linear_algebra_library_type = ceres::CX_SPARSE;
  lo.elimination_groups.push_back(ne);
  lo.elimination_groups.push_back((int)reduced->parameter_blocks().size() - ne);
  lo.num_threads = num_threads;
  ceres::internal::DetectStructure(*A->block_structure(), ne, &lo.row_block_size, &lo.e_block_size,
                                   &lo.f_block_size);
  ceres::internal::SparseSchurComplementSolver solver(lo);
  solver.InitStorage(A->block_structure());
  ceres::internal::scoped_ptr<ceres::internal::SchurEliminatorBase> elim(
      ceres::internal::SchurEliminatorBase::Create(lo));
  elim->Init(ne, A->block_structure());
  const int ncols = A->num_cols();
  std::vector<double> state(reduced->NumParameters()), residuals(A->num_rows()), scale(ncols, 0.5),
      D(ncols, 1e-2), rhs(solver.lhs()->num_rows());
  reduced->ParameterBlocksToStateVector(state.data());
  double cost = 0.0;
  const Clock::time_point t1 = Clock::now();
  double total = 0.0;
  for (int r = 0; r <= reps; ++r) {
    const Clock::time_point a = Clock::now();
    if (!ev->Evaluate(state.data(), &cost, residuals.data(), NULL, A)) return -4;
    A->ScaleColumns(scale.data());
    elim->Eliminate(A, residuals.data(), D.data(), solver.lhs_.get(), rhs.data());
    if (r > 0) total += std::chrono::duration<double>(Clock::now() - a).count();
  }
  seconds[0] = reps > 0 ? total / reps : 0.0;
  seconds[1] = std::chrono::duration<double>(t1 - t0).count();
  return (long)reduced->NumResidualBlocks();
}

// ceres::Solve as libs/sfm configures it (sfm_incremental_engine.cpp:1449-1472):
// linear_solver 0 = SPARSE_SCHUR + CX_SPARSE, 1 = ITERATIVE_SCHUR + JACOBI,
// 2 = DENSE_SCHUR.  summary[8] = {initial_cost, final_cost, successful,
// unsuccessful, termination_type, iterations recorded, total_time,
// linear_solver_time}.  iters[9 * k] = {iteration, cost, cost_change,
// radius, successful, valid, linear iterations, relative_decrease,
// gradient_max_norm}.
int ref_solve(int layout, int nc, int ni, int np, long nobs, double* cams, double* intr,
              double* pts, const double* obs, const int* cam_idx, const int* intr_idx,
              const int* pt_idx, const unsigned char* cam_const, const unsigned char* intr_const,
              const unsigned char* intr_subset, double huber_a, int linear_solver, int max_iters,
              int num_threads, double* summary, double* iters, int max_rec) {
  Scene s = {layout, nc, ni, np, nobs, cams, intr, pts, obs, cam_idx, intr_idx, pt_idx, cam_const,
             intr_const, intr_subset};
  ContiguousCamsIntr contiguous(&s);
  ceres::Problem problem;
  ceres::LossFunction* loss = huber_a > 0.0 ? new ceres::HuberLoss(huber_a) : NULL;
  AddAll(s, &problem, loss, NULL);
  ceres::Solver::Options options;
  if (g_points_first) {
    std::vector<double*> blocks;
    problem.GetParameterBlocks(&blocks);
    ceres::ParameterBlockOrdering* o = new ceres::ParameterBlockOrdering;
    AddPointsFirst(s, blocks, o);
    options.linear_solver_ordering.reset(o);
  }
  options.preconditioner_type = PrecondType(g_precond);
  options.linear_solver_type = linear_solver == 1 ? ceres::ITERATIVE_SCHUR
                               : linear_solver == 2 ? ceres::DENSE_SCHUR
                                                    : ceres::SPARSE_SCHUR;
  options.sparse_linear_algebra_library_type = ceres::CX_SPARSE;
  options.minimizer_progress_to_stdout = false;
  options.logging_type = ceres::SILENT;
  options.max_num_iterations = max_iters;
  if (g_eta > 0.0) options.eta = g_eta;
  if (g_function_tolerance > 0.0) options.function_tolerance = g_function_tolerance;
  options.num_threads = num_threads;
  options.num_linear_solver_threads = num_threads;
  options.trust_region_minimizer_iterations_to_dump = g_dump_iters;
  options.trust_region_problem_dump_directory = g_dump_dir;
  options.trust_region_problem_dump_format_type = ceres::TEXTFILE;
  ceres::Solver::Summary sum;
  ceres::Solve(options, &problem, &sum);
  g_message = sum.message;
  summary[0] = sum.initial_cost;
  summary[1] = sum.final_cost;
  summary[2] = sum.num_successful_steps;
  summary[3] = sum.num_unsuccessful_steps;
  summary[4] = sum.termination_type;
  summary[5] = (double)sum.iterations.size();
  summary[6] = sum.total_time_in_seconds;
  summary[7] = sum.linear_solver_time_in_seconds;
  for (int k = 0; k < (int)sum.iterations.size() && k < max_rec; ++k) {
    const ceres::IterationSummary& it = sum.iterations[k];
    double* o = iters + 9 * k;
    o[0] = it.iteration;
    o[1] = it.cost;
    o[2] = it.cost_change;
    o[3] = it.trust_region_radius;
    o[4] = it.step_is_successful;
    o[5] = it.step_is_valid;
    o[6] = it.linear_solver_iterations;
    o[7] = it.relative_decrease;
    o[8] = it.gradient_max_norm;
  }
  return sum.IsSolutionUsable() ? 1 : 0;
}

}  // extern "C"
