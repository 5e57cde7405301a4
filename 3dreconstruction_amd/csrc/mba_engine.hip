// Host runtime of the MI355X bundle-adjustment engine: the reference's
// preprocessing (Schur ordering, residual order), the device layout, the
// trust-region Levenberg-Marquardt controller, the PCG / explicit-Schur
// linear solvers and the C ABI of include/mba.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/mba.h"
#include "mba_device.h"
#include "mba_scene.h"

namespace mba {
namespace {

thread_local std::string g_last_error;

mba_status fail(mba_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      return fail(e_ == hipErrorOutOfMemory ? MBA_ERR_OUT_OF_MEMORY : MBA_ERR_DEVICE, \
                  std::string(#expr) + ": " + hipGetErrorString(e_));              \
    }                                                                              \
  } while (0)

#define MBA_TRY(expr)              \
  do {                             \
    mba_status s_ = (expr);        \
    if (s_ != MBA_OK) return s_;   \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

// ---------------------------------------------------------------- ordering
// The reference's preprocessing for a (camera, point) bundle problem:
//   program order = first appearance in AddResidualBlock argument lists
//     (internal/ceres/problem_impl.cc:143,271), constant blocks removed
//     (program.cc:269-300);
//   ComputeStableSchurOrdering (parameter_block_ordering.cc:50-79): vertices
//     stable-sorted by ascending Hessian-graph degree, greedy independent set
//     (graph_algorithms.h:173-235), E = the set, F = the rest, both in queue
//     order;
//   LexicographicallyOrderResidualBlocks (reorder_program.cc:249-328): rows
//     bucketed by their E block, each bucket filled back to front.
struct Structure {
  int32_t nc = 0, np = 0;
  int64_t nobs = 0;
  std::vector<int32_t> order;      // column blocks: cameras c, points nc + p
  int32_t num_e = 0;
  std::vector<int64_t> res_order;  // row -> input observation
  bool standard = false;           // E == active points, F == active cameras
  // Derived (standard structure only).
  std::vector<int32_t> e_to_pt, f_to_cam, cam_to_f, pt_start;
  int32_t max_track = 0;
};

mba_status validate(const mba_problem* p) {
  if (!p) return fail(MBA_ERR_INVALID_ARGUMENT, "problem is NULL");
  if (p->layout != MBA_LAYOUT_PACKED9) return fail(MBA_ERR_UNSUPPORTED, "only the packed 9-parameter camera layout is supported");
  if (p->num_cameras < 0 || p->num_points < 0 || p->num_observations < 0)
    return fail(MBA_ERR_INVALID_ARGUMENT, "negative sizes");
  if (p->num_observations > 0x7fffffffLL) return fail(MBA_ERR_UNSUPPORTED, "more than 2^31-1 observations");
  if (p->num_observations > 0 && (!p->observations || !p->camera_index || !p->point_index))
    return fail(MBA_ERR_INVALID_ARGUMENT, "observation arrays are NULL");
  if ((p->num_cameras > 0 && !p->cameras) || (p->num_points > 0 && !p->points))
    return fail(MBA_ERR_INVALID_ARGUMENT, "parameter arrays are NULL");
  for (int64_t i = 0; i < p->num_observations; ++i) {
    const int32_t c = p->camera_index[i], q = p->point_index[i];
    if (c < 0 || c >= p->num_cameras || q < 0 || q >= p->num_points)
      return fail(MBA_ERR_INVALID_ARGUMENT, "observation " + std::to_string(i) + " has an index out of range");
  }
  return MBA_OK;
}

mba_status build_structure(const mba_problem* p, Structure* st) {
  MBA_TRY(validate(p));
  const int32_t nc = p->num_cameras, np = p->num_points;
  const int64_t no = p->num_observations;
  st->nc = nc;
  st->np = np;
  st->nobs = no;
  const int32_t nb = nc + np;
  auto is_const = [&](int32_t b) -> bool {
    return b < nc && p->camera_constant != nullptr && p->camera_constant[b] != 0;
  };
  // First appearance.
  std::vector<int32_t> prog;
  prog.reserve(nb);
  std::vector<uint8_t> seen(nb, 0);
  for (int64_t i = 0; i < no; ++i) {
    const int32_t a = p->camera_index[i], b = nc + p->point_index[i];
    if (!seen[a]) { seen[a] = 1; prog.push_back(a); }
    if (!seen[b]) { seen[b] = 1; prog.push_back(b); }
  }
  // Edges of the Hessian graph: distinct (camera, point) pairs between
  // non-constant blocks.  CSR of observations by camera, marker per point.
  std::vector<int64_t> cstart(nc + 1, 0);
  for (int64_t i = 0; i < no; ++i) cstart[p->camera_index[i] + 1]++;
  for (int32_t c = 0; c < nc; ++c) cstart[c + 1] += cstart[c];
  std::vector<int32_t> cobs_pt(no);
  {
    std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
    for (int64_t i = 0; i < no; ++i) cobs_pt[fill[p->camera_index[i]]++] = p->point_index[i];
  }
  std::vector<int32_t> deg(nb, 0);
  std::vector<int32_t> mark(np, -1);
  // Adjacency CSR for both sides (deduplicated).
  std::vector<int64_t> cadj_start(nc + 1, 0);
  std::vector<int32_t> cadj;
  cadj.reserve(no);
  for (int32_t c = 0; c < nc; ++c) {
    cadj_start[c] = (int64_t)cadj.size();
    if (is_const(c)) continue;
    for (int64_t k = cstart[c]; k < cstart[c + 1]; ++k) {
      const int32_t q = cobs_pt[k];
      if (mark[q] == c) continue;
      mark[q] = c;
      cadj.push_back(q);
    }
  }
  cadj_start[nc] = (int64_t)cadj.size();
  std::vector<int64_t> padj_start(np + 1, 0);
  for (int32_t c = 0; c < nc; ++c)
    for (int64_t k = cadj_start[c]; k < cadj_start[c + 1]; ++k) padj_start[cadj[k] + 1]++;
  for (int32_t q = 0; q < np; ++q) padj_start[q + 1] += padj_start[q];
  std::vector<int32_t> padj(cadj.size());
  {
    std::vector<int64_t> fill(padj_start.begin(), padj_start.end() - 1);
    for (int32_t c = 0; c < nc; ++c)
      for (int64_t k = cadj_start[c]; k < cadj_start[c + 1]; ++k) padj[fill[cadj[k]]++] = c;
  }
  for (int32_t c = 0; c < nc; ++c) deg[c] = (int32_t)(cadj_start[c + 1] - cadj_start[c]);
  for (int32_t q = 0; q < np; ++q) deg[nc + q] = (int32_t)(padj_start[q + 1] - padj_start[q]);
  // Active vertices in program order, stable counting sort by degree.
  std::vector<int32_t> active;
  active.reserve(prog.size());
  for (int32_t b : prog)
    if (!is_const(b)) active.push_back(b);
  int32_t maxdeg = 0;
  for (int32_t b : active) maxdeg = std::max(maxdeg, deg[b]);
  std::vector<int64_t> cnt((size_t)maxdeg + 2, 0);
  for (int32_t b : active) cnt[deg[b] + 1]++;
  for (int32_t d = 0; d <= maxdeg; ++d) cnt[d + 1] += cnt[d];
  std::vector<int32_t> queue(active.size());
  for (int32_t b : active) queue[cnt[deg[b]]++] = b;
  // Greedy independent set.
  std::vector<uint8_t> color(nb, 0);  // 0 white, 1 grey, 2 black
  st->order.clear();
  st->order.reserve(active.size());
  for (int32_t b : queue) {
    if (color[b] != 0) continue;
    st->order.push_back(b);
    color[b] = 2;
    if (b < nc) {
      for (int64_t k = cadj_start[b]; k < cadj_start[b + 1]; ++k) color[nc + cadj[k]] = 1;
    } else {
      const int32_t q = b - nc;
      for (int64_t k = padj_start[q]; k < padj_start[q + 1]; ++k) color[padj[k]] = 1;
    }
  }
  st->num_e = (int32_t)st->order.size();
  for (int32_t b : queue)
    if (color[b] != 2) st->order.push_back(b);
  // Column-block index of every block.
  std::vector<int32_t> col(nb, -1);
  for (size_t i = 0; i < st->order.size(); ++i) col[st->order[i]] = (int32_t)i;
  // Residual order: all rows have a non-constant block (points are never
  // constant); bucket = smallest column block index.
  std::vector<int64_t> hist((size_t)st->num_e + 1, 0);
  std::vector<int32_t> bucket(no);
  for (int64_t i = 0; i < no; ++i) {
    const int32_t a = p->camera_index[i], b = nc + p->point_index[i];
    int32_t m = st->num_e;
    if (!is_const(a)) m = std::min(m, col[a]);
    m = std::min(m, col[b]);
    bucket[i] = m;
    hist[m]++;
  }
  for (int32_t e = 1; e <= st->num_e; ++e) hist[e] += hist[e - 1];
  st->res_order.assign(no, 0);
  for (int64_t i = 0; i < no; ++i) st->res_order[--hist[bucket[i]]] = i;

  // Standard BA structure: E = every active point, F = every active camera.
  int32_t n_act_pts = 0, n_act_cams = 0;
  for (int32_t b : active) (b < nc ? n_act_cams : n_act_pts)++;
  bool ok = (st->num_e == n_act_pts);
  for (int32_t i = 0; ok && i < st->num_e; ++i) ok = st->order[i] >= nc;
  st->standard = ok;
  if (!ok) return MBA_OK;
  st->e_to_pt.resize(st->num_e);
  for (int32_t i = 0; i < st->num_e; ++i) st->e_to_pt[i] = st->order[i] - nc;
  const int32_t nf = (int32_t)st->order.size() - st->num_e;
  st->f_to_cam.resize(nf);
  st->cam_to_f.assign(nc, -1);
  for (int32_t i = 0; i < nf; ++i) {
    st->f_to_cam[i] = st->order[st->num_e + i];
    st->cam_to_f[st->f_to_cam[i]] = i;
  }
  st->pt_start.assign((size_t)st->num_e + 1, 0);
  for (int64_t r = 0; r < no; ++r) st->pt_start[bucket[st->res_order[r]] + 1]++;
  for (int32_t e = 0; e < st->num_e; ++e) st->pt_start[e + 1] += st->pt_start[e];
  st->max_track = 0;
  for (int32_t e = 0; e < st->num_e; ++e)
    st->max_track = std::max(st->max_track, st->pt_start[e + 1] - st->pt_start[e]);
  (void)n_act_cams;
  return MBA_OK;
}

// Upper-triangle block pairs of the reduced camera matrix
// (SparseSchurComplementSolver::InitStorage, schur_complement_solver.cc:246-313).
std::vector<int64_t> block_pairs(const mba_problem* p, const Structure& st) {
  const int32_t nc = st.nc;
  const int32_t nf = (int32_t)st.order.size() - st.num_e;
  std::vector<int32_t> col(nc + st.np, -1);
  for (size_t i = 0; i < st.order.size(); ++i) col[st.order[i]] = (int32_t)i;
  std::vector<int64_t> keys;
  keys.reserve(nf);
  for (int32_t i = 0; i < nf; ++i) keys.push_back((int64_t)i * nf + i);
  auto is_const = [&](int32_t b) { return b < nc && p->camera_constant && p->camera_constant[b]; };
  auto cells = [&](int64_t obs, int32_t* out) {
    int n = 0;
    const int32_t a = p->camera_index[obs], b = nc + p->point_index[obs];
    if (!is_const(a)) out[n++] = col[a];
    out[n++] = col[b];
    std::sort(out, out + n);
    return n;
  };
  int64_t r = 0;
  const int64_t nr = (int64_t)st.res_order.size();
  std::vector<int32_t> fb;
  int32_t c2[2];
  while (r < nr) {
    cells(st.res_order[r], c2);
    if (c2[0] >= st.num_e) break;
    const int32_t e = c2[0];
    fb.clear();
    for (; r < nr; ++r) {
      const int n = cells(st.res_order[r], c2);
      if (c2[0] != e) break;
      for (int k = 1; k < n; ++k) fb.push_back(c2[k] - st.num_e);
    }
    std::sort(fb.begin(), fb.end());
    fb.erase(std::unique(fb.begin(), fb.end()), fb.end());
    for (size_t i = 0; i < fb.size(); ++i)
      for (size_t j = i + 1; j < fb.size(); ++j) keys.push_back((int64_t)fb[i] * nf + fb[j]);
  }
  for (; r < nr; ++r) {
    const int n = cells(st.res_order[r], c2);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (c2[i] <= c2[j]) keys.push_back((int64_t)(c2[i] - st.num_e) * nf + (c2[j] - st.num_e));
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  return keys;
}

// ------------------------------------------------------------ device pool
struct Pool {
  std::vector<void*> ptrs;
  template <typename X>
  mba_status alloc(X** out, size_t count) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(X);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? MBA_ERR_OUT_OF_MEMORY : MBA_ERR_DEVICE,
                  "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    ptrs.push_back(p);
    *out = (X*)p;
    return MBA_OK;
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
};

template <typename T>
struct Dev {
  // rows in the reference's residual order
  int32_t* row_cam = nullptr;
  T* row_obs = nullptr;
  int32_t* pt_start = nullptr;
  // camera-major chunks
  int32_t *cm_e = nullptr, *chunk_f = nullptr, *chunk_begin = nullptr, *chunk_end = nullptr;
  int32_t *f_chunk_start = nullptr, *f_to_cam = nullptr, *cam_to_f = nullptr;
  T* cm_obs = nullptr;
  // kernel-precision parameters
  T *ctab = nullptr, *pscale = nullptr, *cscale = nullptr;
  // blocks
  T *V = nullptr, *gp = nullptr, *Vinv = nullptr, *z = nullptr, *pk = nullptr;
  T *U = nullptr, *gc = nullptr, *rhs = nullptr, *Minv = nullptr, *dtab = nullptr, *partial = nullptr;
  // PCG / step
  T *x = nullptr, *r = nullptr, *zc = nullptr, *p = nullptr, *q = nullptr, *tmp = nullptr;
  T *step_p = nullptr, *step_c = nullptr;
  // explicit S
  T *S = nullptr, *L = nullptr, *sv = nullptr;
};

}  // namespace
}  // namespace mba

using namespace mba;

struct mba_context {
  mba_options opt;
  Structure st;
  int device = 0;
  hipStream_t stream = nullptr;
  Pool pool;
  int32_t nc = 0, np = 0, ne = 0, nf = 0, nrows = 0, nchunks = 0, ncm = 0;
  // FP64 state and LM data
  double *cams64 = nullptr, *cams64_t = nullptr, *pts64 = nullptr, *pts64_t = nullptr;
  double *ctab64 = nullptr, *row_obs64 = nullptr, *pscale64 = nullptr, *cscale64 = nullptr;
  double *diag_p = nullptr, *diag_c = nullptr, *red = nullptr, *red2 = nullptr, *stepd = nullptr;
  double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr;
  Scalars* sc = nullptr;      // device
  PcgState* pcg = nullptr;    // device
  int32_t* flag = nullptr;    // device
  int32_t* info = nullptr;    // device
  Dev<float> df;
  Dev<double> dd;
  bool mixed = false;
  bool scale_ready = false;
  double last_radius = 0.0;
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_ms[3] = {0, 0, 0};
  // Device point order (E blocks stably grouped by lane class) and the
  // point-major launch layout.
  std::vector<int32_t> dev_e_to_pt;
  Seg seg;
  int32_t seg_blocks = 0;
};

namespace mba {
namespace {

template <typename T> Dev<T>& dev(mba_context* c);
template <> Dev<float>& dev<float>(mba_context* c) { return c->df; }
template <> Dev<double>& dev<double>(mba_context* c) { return c->dd; }

// Copy a host vector to a new device allocation.
template <typename X>
mba_status upload(mba_context* c, X** out, const std::vector<X>& v) {
  MBA_TRY(c->pool.alloc(out, v.size()));
  if (!v.empty()) HIP_TRY(hipMemcpyAsync(*out, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, c->stream));
  return MBA_OK;
}

template <typename T>
mba_status alloc_precision(mba_context* c, const std::vector<int32_t>& dev_pt_start,
                           const std::vector<int32_t>& row_cam,
                           const std::vector<double>& row_obs, const std::vector<int32_t>& cm_e,
                           const std::vector<double>& cm_obs, const std::vector<int32_t>& chunk_f,
                           const std::vector<int32_t>& chunk_begin, const std::vector<int32_t>& chunk_end,
                           const std::vector<int32_t>& f_chunk_start) {
  Dev<T>& d = dev<T>(c);
  std::vector<T> ro(row_obs.begin(), row_obs.end()), co(cm_obs.begin(), cm_obs.end());
  MBA_TRY(upload(c, &d.row_cam, row_cam));
  MBA_TRY(upload(c, &d.row_obs, ro));
  MBA_TRY(upload(c, &d.pt_start, dev_pt_start));
  MBA_TRY(upload(c, &d.cm_e, cm_e));
  MBA_TRY(upload(c, &d.cm_obs, co));
  MBA_TRY(upload(c, &d.chunk_f, chunk_f));
  MBA_TRY(upload(c, &d.chunk_begin, chunk_begin));
  MBA_TRY(upload(c, &d.chunk_end, chunk_end));
  MBA_TRY(upload(c, &d.f_chunk_start, f_chunk_start));
  MBA_TRY(upload(c, &d.f_to_cam, c->st.f_to_cam));
  MBA_TRY(upload(c, &d.cam_to_f, c->st.cam_to_f));
  const size_t ne = c->ne, nf = c->nf, nc = c->nc;
  MBA_TRY(c->pool.alloc(&d.ctab, kCamTab * nc));
  MBA_TRY(c->pool.alloc(&d.pscale, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.cscale, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.V, 6 * ne));
  MBA_TRY(c->pool.alloc(&d.gp, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.Vinv, 6 * ne));
  MBA_TRY(c->pool.alloc(&d.z, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.pk, kPk * ne));
  MBA_TRY(c->pool.alloc(&d.U, 45 * nf));
  MBA_TRY(c->pool.alloc(&d.gc, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.rhs, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.Minv, 81 * nf));
  MBA_TRY(c->pool.alloc(&d.dtab, kDirTab * nc));
  MBA_TRY(c->pool.alloc(&d.partial, (size_t)kPartial * c->nchunks));
  MBA_TRY(c->pool.alloc(&d.x, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.r, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.zc, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.p, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.q, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.tmp, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.step_p, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.step_c, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.sv, 9 * nf));
  return MBA_OK;
}

// Rows per camera-major chunk: aim for >= ~2048 workgroups on large
// problems, 64..2048 rows each.
int chunk_rows_for(int64_t nrows) {
  int64_t c = nrows / 2048;
  c = std::max<int64_t>(64, std::min<int64_t>(2048, c));
  return (int)((c + 63) / 64 * 64);
}

mba_status context_create(const mba_problem* p, const mba_options* o, mba_context** out) {
  if (!out) return fail(MBA_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  const double t0 = now_s();
  mba_context* c = new mba_context();
  std::unique_ptr<mba_context> guard(c);
  if (o) c->opt = *o; else mba_default_options(&c->opt);
  c->device = c->opt.device;
  c->mixed = (c->opt.precision == MBA_PRECISION_MIXED);
  MBA_TRY(build_structure(p, &c->st));
  if (!c->st.standard)
    return fail(MBA_ERR_UNSUPPORTED, "the Schur ordering did not eliminate exactly the point blocks");
  for (int32_t q = 0; q < p->num_points; ++q)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(p->points[3 * (size_t)q + k])) return fail(MBA_ERR_NOT_FINITE, "non-finite point");
  for (int32_t q = 0; q < p->num_cameras; ++q)
    for (int k = 0; k < 9; ++k)
      if (!std::isfinite(p->cameras[9 * (size_t)q + k])) return fail(MBA_ERR_NOT_FINITE, "non-finite camera");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (c->device < 0 || c->device >= ndev) return fail(MBA_ERR_DEVICE, "no such HIP device");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const Structure& st = c->st;
  c->nc = st.nc;
  c->np = st.np;
  c->ne = st.num_e;
  c->nf = (int32_t)st.order.size() - st.num_e;
  c->nrows = (int32_t)st.nobs;
  // Device point order: the reference's E order stably grouped by lane
  // class (lanes_for), so each point-major launch segment is contiguous.
  std::vector<int32_t> dperm(c->ne), dev_pt_start((size_t)c->ne + 1, 0);
  {
    std::vector<int32_t> cls(c->ne);
    int32_t ncls[8] = {0};
    for (int32_t e = 0; e < c->ne; ++e) {
      const int t = st.pt_start[e + 1] - st.pt_start[e];
      int k = 0;
      while (k < 6 && t > 4 * (1 << k)) ++k;
      cls[e] = k;
      ncls[k]++;
    }
    int32_t start[8] = {0};
    for (int k = 1; k < 8; ++k) start[k] = start[k - 1] + ncls[k - 1];
    int32_t fill[8];
    std::memcpy(fill, start, sizeof(fill));
    for (int32_t e = 0; e < c->ne; ++e) dperm[fill[cls[e]]++] = e;
    c->seg.n = 0;
    int32_t blk = 0;
    for (int k = 0; k < 7; ++k) {
      if (ncls[k] == 0) continue;
      const int j = c->seg.n++;
      c->seg.e0[j] = start[k];
      c->seg.e1[j] = start[k] + ncls[k];
      c->seg.lanes[j] = 1 << k;
      c->seg.blk0[j] = blk;
      blk += (int32_t)(((int64_t)ncls[k] * (1 << k) + kBlock - 1) / kBlock);
    }
    c->seg.blk0[c->seg.n] = blk;
    if (c->seg.n == 0) {  // no points: a harmless empty segment
      c->seg.n = 1;
      c->seg.e0[0] = c->seg.e1[0] = 0;
      c->seg.lanes[0] = 1;
      c->seg.blk0[0] = 0;
      c->seg.blk0[1] = 0;
    }
    c->seg_blocks = blk;
    c->dev_e_to_pt.resize(c->ne);
    for (int32_t de = 0; de < c->ne; ++de) {
      const int32_t e = dperm[de];
      c->dev_e_to_pt[de] = st.e_to_pt[e];
      dev_pt_start[de + 1] = dev_pt_start[de] + (st.pt_start[e + 1] - st.pt_start[e]);
    }
  }
  // Rows in device order (each point's rows keep the reference's order).
  std::vector<int32_t> row_cam(c->nrows), row_e(c->nrows);
  std::vector<double> row_obs(2 * (size_t)c->nrows);
  for (int32_t de = 0; de < c->ne; ++de) {
    const int32_t e = dperm[de];
    int32_t r = dev_pt_start[de];
    for (int32_t rr = st.pt_start[e]; rr < st.pt_start[e + 1]; ++rr, ++r) {
      const int64_t i = st.res_order[rr];
      row_e[r] = de;
      row_cam[r] = p->camera_index[i];
      row_obs[2 * (size_t)r] = p->observations[2 * i];
      row_obs[2 * (size_t)r + 1] = p->observations[2 * i + 1];
    }
  }
  // Camera-major rows (stable by row order), constant cameras excluded.
  std::vector<int64_t> fstart((size_t)c->nf + 1, 0);
  for (int32_t r = 0; r < c->nrows; ++r) {
    const int32_t f = st.cam_to_f[row_cam[r]];
    if (f >= 0) fstart[f + 1]++;
  }
  for (int32_t f = 0; f < c->nf; ++f) fstart[f + 1] += fstart[f];
  c->ncm = (int32_t)fstart[c->nf];
  std::vector<int32_t> cm_e(c->ncm);
  std::vector<double> cm_obs(2 * (size_t)c->ncm);
  {
    std::vector<int64_t> fill(fstart.begin(), fstart.end() - 1);
    for (int32_t r = 0; r < c->nrows; ++r) {
      const int32_t f = st.cam_to_f[row_cam[r]];
      if (f < 0) continue;
      const int64_t k = fill[f]++;
      cm_e[k] = row_e[r];
      cm_obs[2 * k] = row_obs[2 * (size_t)r];
      cm_obs[2 * k + 1] = row_obs[2 * (size_t)r + 1];
    }
  }
  const int cr = chunk_rows_for(c->ncm);
  std::vector<int32_t> chunk_f, chunk_begin, chunk_end, f_chunk_start((size_t)c->nf + 1, 0);
  for (int32_t f = 0; f < c->nf; ++f) {
    f_chunk_start[f] = (int32_t)chunk_f.size();
    for (int64_t b = fstart[f]; b < fstart[f + 1]; b += cr) {
      chunk_f.push_back(f);
      chunk_begin.push_back((int32_t)b);
      chunk_end.push_back((int32_t)std::min<int64_t>(b + cr, fstart[f + 1]));
    }
  }
  f_chunk_start[c->nf] = (int32_t)chunk_f.size();
  c->nchunks = (int32_t)chunk_f.size();
  // FP64 state.
  std::vector<double> pts(3 * (size_t)c->ne);
  for (int32_t e = 0; e < c->ne; ++e)
    for (int k = 0; k < 3; ++k) pts[3 * (size_t)e + k] = p->points[3 * (size_t)c->dev_e_to_pt[e] + k];
  std::vector<double> cams(p->cameras, p->cameras + 9 * (size_t)c->nc);
  MBA_TRY(upload(c, &c->cams64, cams));
  MBA_TRY(upload(c, &c->cams64_t, cams));
  MBA_TRY(upload(c, &c->pts64, pts));
  MBA_TRY(upload(c, &c->pts64_t, pts));
  MBA_TRY(upload(c, &c->row_obs64, row_obs));
  MBA_TRY(c->pool.alloc(&c->ctab64, kCamTab * (size_t)std::max(1, c->nc)));
  MBA_TRY(c->pool.alloc(&c->pscale64, 3 * (size_t)c->ne));
  MBA_TRY(c->pool.alloc(&c->cscale64, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->diag_p, 3 * (size_t)c->ne));
  MBA_TRY(c->pool.alloc(&c->diag_c, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->stepd, 3 * (size_t)c->ne + 9 * (size_t)c->nf));
  const int64_t maxb = nblocks(std::max<int64_t>({(int64_t)c->ne, 9 * (int64_t)c->nf, (int64_t)c->nrows, 1})) + 64;
  // Reduction scratch: block partials (<= 2 * maxb) plus the gradient
  // max/sumsq partials at offset 2048 (<= 1024 blocks each).
  MBA_TRY(c->pool.alloc(&c->red, std::max<size_t>(4 * (size_t)maxb, 4096)));
  MBA_TRY(c->pool.alloc(&c->red2, std::max<size_t>(4 * (size_t)maxb, 4096)));
  MBA_TRY(c->pool.alloc(&c->sc, 4));
  MBA_TRY(c->pool.alloc(&c->pcg, 1));
  MBA_TRY(c->pool.alloc(&c->flag, 4));
  MBA_TRY(c->pool.alloc(&c->info, 4));
  MBA_TRY(c->pool.alloc(&c->b64, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->x64, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->r64, 9 * (size_t)c->nf));
  if (c->mixed)
    MBA_TRY(alloc_precision<float>(c, dev_pt_start, row_cam, row_obs, cm_e, cm_obs, chunk_f, chunk_begin, chunk_end,
                                   f_chunk_start));
  else
    MBA_TRY(alloc_precision<double>(c, dev_pt_start, row_cam, row_obs, cm_e, cm_obs, chunk_f, chunk_begin, chunk_end,
                                    f_chunk_start));
  for (int i = 0; i < 4; ++i) HIP_TRY(hipEventCreate(&c->ev[i]));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)t0;
  *out = guard.release();
  return MBA_OK;
}

void context_free(mba_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 4; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  c->pool.release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// ------------------------------------------------------------ LM passes
template <typename T>
struct Engine {
  mba_context* c;
  Dev<T>& d;
  hipStream_t s;
  T huber;
  explicit Engine(mba_context* ctx) : c(ctx), d(dev<T>(ctx)), s(ctx->stream), huber((T)ctx->opt.huber_a) {}

  mba_status sync() {
    HIP_TRY(hipStreamSynchronize(s));
    return MBA_OK;
  }
  mba_status read_scalar(const double* dptr, double* out) {
    HIP_TRY(hipMemcpyAsync(out, dptr, sizeof(double), hipMemcpyDeviceToHost, s));
    return sync();
  }
  mba_status clear_flag() {
    HIP_TRY(hipMemsetAsync(c->flag, 0, sizeof(int32_t), s));
    return MBA_OK;
  }
  mba_status read_flag(int32_t* f) {
    HIP_TRY(hipMemcpyAsync(f, c->flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    return sync();
  }

  // Kernel-precision tables from the FP64 state.
  void tables(const double* cams, const double* pts) {
    (void)pts;  // points are read in FP64 by the point pass
    hipLaunchKernelGGL(k_cam_table<T>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, cams, d.ctab);
  }

  // FP64 cost at (cams, pts) -> *cost; returns false on a non-finite value.
  mba_status cost64(const double* cams, const double* pts, double* cost, bool* finite) {
    MBA_TRY(clear_flag());
    hipLaunchKernelGGL(k_cam_table<double>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, cams, c->ctab64);
    const int nb = c->seg_blocks;
    HIP_TRY(hipMemsetAsync(c->red, 0, sizeof(double), s));
    if (nb > 0)
      hipLaunchKernelGGL(k_cost64, dim3(nb), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam, c->row_obs64, pts,
                         c->ctab64, c->opt.huber_a, c->red, c->flag);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, c->red, std::max(nb, 1), &c->sc->sum);
    MBA_TRY(read_scalar(&c->sc->sum, cost));
    int32_t f = 0;
    MBA_TRY(read_flag(&f));
    *finite = (f == 0) && std::isfinite(*cost);
    return MBA_OK;
  }

  // Jacobian + Schur pass at the current x (tables must be current).
  void jacobian_pass(double radius, int recompute_diag) {
    const double inv_r = radius > 0.0 ? 1.0 / radius : 0.0;
    if (c->timing) (void)hipEventRecord(c->ev[0], s);
    if (c->seg_blocks > 0)
      hipLaunchKernelGGL(k_point_pass<T>, dim3(c->seg_blocks), dim3(kBlock), 0, s, c->seg, d.pt_start,
                         d.row_cam, d.row_obs, c->pts64, d.pscale, d.ctab, huber, inv_r,
                         c->opt.min_lm_diagonal, c->opt.max_lm_diagonal, c->diag_p, recompute_diag,
                         d.V, d.gp, d.Vinv, d.pk, c->flag);
    if (c->timing) (void)hipEventRecord(c->ev[1], s);
    if (c->nchunks > 0)
      hipLaunchKernelGGL(k_cam_pass<T>, dim3(c->nchunks), dim3(kBlock), 0, s, d.chunk_f, d.chunk_begin,
                         d.chunk_end, d.f_to_cam, d.cm_e, d.cm_obs, d.pk, d.ctab, huber, 1, d.partial);
    if (c->timing) (void)hipEventRecord(c->ev[2], s);
    if (c->nf > 0)
      hipLaunchKernelGGL(k_cam_finalize<T>, dim3(c->nf), dim3(64), 0, s, c->nf,
                         d.f_chunk_start, d.partial, d.cscale, inv_r, c->opt.min_lm_diagonal, c->opt.max_lm_diagonal,
                         c->diag_c, recompute_diag, 1, d.U, d.gc, d.rhs, d.Minv, c->flag);
    if (c->timing) (void)hipEventRecord(c->ev[3], s);
    c->last_radius = radius;
  }

  // Re-damp after a rejected step and rebuild the reduced rhs.
  void damp(double radius) {
    const double inv_r = 1.0 / radius;
    hipLaunchKernelGGL(k_point_damp<T>, dim3(nblocks(c->ne)), dim3(kBlock), 0, s, c->ne, d.V, d.gp,
                       c->diag_p, inv_r, d.pscale, d.Vinv, d.pk, c->flag);
    if (c->nf > 0) {
      hipLaunchKernelGGL(k_cam_damp<T>, dim3(nblocks(c->nf, 64)), dim3(64), 0, s, c->nf, d.U, c->diag_c,
                         inv_r, d.Minv, c->flag);
      if (c->nchunks > 0)
        hipLaunchKernelGGL(k_mv_cam<T>, dim3(c->nchunks), dim3(kBlock), 0, s, d.chunk_f, d.chunk_begin,
                           d.chunk_end, d.f_to_cam, d.cm_e, d.cm_obs, d.pk, d.cscale, d.ctab, huber,
                           d.partial, (const PcgState*)nullptr);
      hipLaunchKernelGGL(k_rhs_finalize<T>, dim3(nblocks(c->nf)), dim3(kBlock), 0, s, c->nf,
                         d.f_chunk_start, d.partial, d.gc, d.rhs);
    }
    c->last_radius = radius;
  }

  // y = S x (implicit), x and y F-order.  `st` may be null.
  void matvec(const T* x, T* y, const PcgState* st) {
    const double inv_r = c->last_radius > 0.0 ? 1.0 / c->last_radius : 0.0;
    hipLaunchKernelGGL(k_cam_dir<T>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f, d.cscale,
                       d.ctab, x, (T)1, d.dtab, st);
    if (c->seg_blocks > 0)
      hipLaunchKernelGGL(k_mv_point<T>, dim3(c->seg_blocks), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam,
                         d.row_obs, d.pk, d.pscale, d.ctab, d.dtab, d.Vinv, huber, (T)0, 1, d.z, d.pk, st);
    if (c->nchunks > 0)
      hipLaunchKernelGGL(k_mv_cam<T>, dim3(c->nchunks), dim3(kBlock), 0, s, d.chunk_f, d.chunk_begin,
                         d.chunk_end, d.f_to_cam, d.cm_e, d.cm_obs, d.pk, d.cscale, d.ctab, huber,
                         d.partial, st);
    hipLaunchKernelGGL(k_mv_finalize<T>, dim3(nblocks(c->nf)), dim3(kBlock), 0, s, c->nf, d.f_chunk_start,
                       d.partial, d.U, c->diag_c, inv_r, x, 1, y, st);
  }

  // PCG on S x = rhs (x in d.x).  Returns iterations and status.
  mba_status pcg(int* iters, int* status) {
    const int nf = c->nf;
    const int nb = nblocks(nf);
    PcgState h;
    std::memset(&h, 0, sizeof(h));
    h.iter = 1;
    h.tol_r = c->opt.pcg_r_tolerance;
    HIP_TRY(hipMemcpyAsync(c->pcg, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const double eta = c->opt.eta;
    const int minit = c->opt.min_linear_solver_iterations, maxit = c->opt.max_linear_solver_iterations;
    hipLaunchKernelGGL(k_pcg_update<T>, dim3(nb), dim3(kBlock), 0, s, nf, c->pcg, 3, d.p, d.q, d.rhs, d.tmp,
                       d.Minv, d.x, d.r, d.zc, c->red);
    hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(1), 0, s, c->red, nb, c->pcg, eta, minit, maxit, 1);
    int check_every = 1;
    for (int it = 1; it <= maxit; ++it) {
      hipLaunchKernelGGL(k_pcg_p<T>, dim3(nb), dim3(kBlock), 0, s, nf, c->pcg, d.zc, d.p);
      matvec(d.p, d.q, c->pcg);
      hipLaunchKernelGGL(k_dot_partial<T>, dim3(nblocks(9 * (int64_t)nf)), dim3(kBlock), 0, s,
                         (int64_t)9 * nf, c->pcg, d.p, d.q, c->red);
      hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(1), 0, s, c->red, nblocks(9 * (int64_t)nf), c->pcg);
      if (it % 10 == 0) {
        hipLaunchKernelGGL(k_pcg_update<T>, dim3(nb), dim3(kBlock), 0, s, nf, c->pcg, 1, d.p, d.q, d.rhs,
                           d.tmp, d.Minv, d.x, d.r, d.zc, c->red);
        matvec(d.x, d.tmp, c->pcg);
        hipLaunchKernelGGL(k_pcg_update<T>, dim3(nb), dim3(kBlock), 0, s, nf, c->pcg, 2, d.p, d.q, d.rhs,
                           d.tmp, d.Minv, d.x, d.r, d.zc, c->red);
      } else {
        hipLaunchKernelGGL(k_pcg_update<T>, dim3(nb), dim3(kBlock), 0, s, nf, c->pcg, 0, d.p, d.q, d.rhs,
                           d.tmp, d.Minv, d.x, d.r, d.zc, c->red);
      }
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(1), 0, s, c->red, nb, c->pcg, eta, minit, maxit, 0);
      if (it % check_every == 0 || it == maxit) {
        HIP_TRY(hipMemcpyAsync(&h, c->pcg, sizeof(h), hipMemcpyDeviceToHost, s));
        MBA_TRY(sync());
        if (h.done) break;
        check_every = std::min(8, check_every * 2);
      }
    }
    HIP_TRY(hipMemcpyAsync(&h, c->pcg, sizeof(h), hipMemcpyDeviceToHost, s));
    MBA_TRY(sync());
    *iters = h.iter;
    *status = h.status;
    return MBA_OK;
  }

  // Explicit Schur: assemble dense S, Cholesky, solve, FP64 refinement.
  mba_status explicit_solve(bool* ok) {
    const int64_t n = 9 * (int64_t)c->nf;
    *ok = true;
    if (n == 0) return MBA_OK;
    if (!d.S) {
      MBA_TRY(c->pool.alloc(&d.S, (size_t)(n * n)));
      MBA_TRY(c->pool.alloc(&d.L, (size_t)(n * n)));
    }
    const double inv_r = c->last_radius > 0.0 ? 1.0 / c->last_radius : 0.0;
    HIP_TRY(hipMemsetAsync(d.S, 0, sizeof(T) * n * n, s));
    hipLaunchKernelGGL(k_s_diag<T>, dim3(nblocks(81 * (int64_t)c->nf)), dim3(kBlock), 0, s, c->nf, d.U,
                       c->diag_c, inv_r, n, d.S);
    hipLaunchKernelGGL(k_s_assemble<T>, dim3(std::max(1, c->ne)), dim3(64), 0, s, c->ne, d.pt_start, d.row_cam,
                       d.cam_to_f, d.row_obs, d.pk, d.pscale, d.cscale, d.ctab, d.Vinv, huber, n, d.S);
    HIP_TRY(hipMemcpyAsync(d.L, d.S, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(c->info, 0, sizeof(int32_t), s));
    for (int64_t k = 0; k < n; k += kNB) {
      const int kb = (int)std::min<int64_t>(kNB, n - k);
      hipLaunchKernelGGL(k_potrf_diag<T>, dim3(1), dim3(256), 0, s, d.L, n, (int)k, kb, c->info);
      const int64_t rest = n - k - kb;
      if (rest > 0) {
        hipLaunchKernelGGL(k_trsm_panel<T>, dim3((unsigned)((rest + 63) / 64)), dim3(64), 0, s, d.L, n, (int)k, kb);
        const int nt = (int)((rest + kNB - 1) / kNB);
        hipLaunchKernelGGL(k_syrk_update<T>, dim3((unsigned)((int64_t)nt * (nt + 1) / 2)), dim3(256), 0, s,
                           d.L, n, (int)k, kb, nt);
      }
    }
    int32_t info = 0;
    HIP_TRY(hipMemcpyAsync(&info, c->info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MBA_TRY(sync());
    if (info != 0) {
      *ok = false;
      return MBA_OK;
    }
    // x = S^-1 rhs with FP64 iterative refinement of the residual.
    HIP_TRY(hipMemcpyAsync(d.x, d.rhs, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_trsv_lower<T>, dim3(1), dim3(256), 0, s, d.L, n, d.x);
    hipLaunchKernelGGL(k_trsv_upper<T>, dim3(1), dim3(256), 0, s, d.L, n, d.x);
    if (c->opt.refinement_iterations > 0) {
      hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, d.rhs, c->b64);
      hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, d.x, c->x64);
      for (int it = 0; it < c->opt.refinement_iterations; ++it) {
        hipLaunchKernelGGL(k_sym_residual64<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, d.S, n, c->b64, c->x64,
                           c->r64);
        hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, c->r64, d.sv);
        hipLaunchKernelGGL(k_trsv_lower<T>, dim3(1), dim3(256), 0, s, d.L, n, d.sv);
        hipLaunchKernelGGL(k_trsv_upper<T>, dim3(1), dim3(256), 0, s, d.L, n, d.sv);
        hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, d.sv, c->r64);
        hipLaunchKernelGGL(k_axpy_inplace, dim3(nblocks(n)), dim3(kBlock), 0, s, n, c->r64, c->x64);
      }
      hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, c->x64, d.x);
    }
    return MBA_OK;
  }

  // Back substitution: y_e = Vinv E'(b - F z) into d.z; step = -y.
  void back_substitute() {
    hipLaunchKernelGGL(k_cam_dir<T>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f, d.cscale,
                       d.ctab, d.x, (T)-1, d.dtab, (const PcgState*)nullptr);
    if (c->seg_blocks > 0)
      hipLaunchKernelGGL(k_mv_point<T>, dim3(c->seg_blocks), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam,
                         d.row_obs, d.pk, d.pscale, d.ctab, d.dtab, d.Vinv, huber, (T)1, 1, d.z, (T*)nullptr,
                         (const PcgState*)nullptr);
    hipLaunchKernelGGL(k_negate<T>, dim3(nblocks(3 * (int64_t)c->ne)), dim3(kBlock), 0, s,
                       (int64_t)3 * c->ne, d.z, d.step_p);
    hipLaunchKernelGGL(k_negate<T>, dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s,
                       (int64_t)9 * c->nf, d.x, d.step_c);
  }

  // Model cost change -J step . (r + J step / 2).
  mba_status model_cost_change(double* mcc) {
    hipLaunchKernelGGL(k_cam_dir<T>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f, d.cscale,
                       d.ctab, d.step_c, (T)1, d.dtab, (const PcgState*)nullptr);
    const int nb = c->seg_blocks;
    HIP_TRY(hipMemsetAsync(c->red, 0, sizeof(double), s));
    if (nb > 0)
      hipLaunchKernelGGL(k_model_cost<T>, dim3(nb), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam, d.row_obs,
                         d.pk, d.pscale, d.ctab, d.dtab, d.step_p, huber, c->red);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, c->red, std::max(nb, 1), &c->sc->sum);
    double v = 0.0;
    MBA_TRY(read_scalar(&c->sc->sum, &v));
    *mcc = -v;
    return MBA_OK;
  }

  // Step finite check + x_trial = x + step .* scale, returns |delta|.
  mba_status plus(double* step_norm, bool* finite) {
    const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
    hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(np3)), dim3(kBlock), 0, s, np3, d.step_p, c->stepd);
    hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(nf9)), dim3(kBlock), 0, s, nf9, d.step_c,
                       c->stepd + np3);
    const int nb1 = nblocks(np3), nb2 = nblocks(nf9);
    hipLaunchKernelGGL(k_plus, dim3(nb1), dim3(kBlock), 0, s, np3, c->pts64, c->stepd, c->pscale64, c->pts64_t,
                       c->red);
    hipLaunchKernelGGL(k_plus_cams, dim3(nb2), dim3(kBlock), 0, s, c->nf, d.f_to_cam, c->cams64,
                       c->stepd + np3, c->cscale64, c->cams64_t, c->red + nb1);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, c->red, nb1 + nb2, &c->sc->sum);
    double v = 0.0;
    MBA_TRY(read_scalar(&c->sc->sum, &v));
    *step_norm = std::sqrt(v);
    *finite = std::isfinite(v);
    return MBA_OK;
  }

  mba_status x_norm(double* out) {
    const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
    hipLaunchKernelGGL(k_gather_cams, dim3(nblocks(nf9)), dim3(kBlock), 0, s, c->nf, d.f_to_cam, c->cams64,
                       c->stepd + np3);
    hipLaunchKernelGGL(k_copy64, dim3(nblocks(np3)), dim3(kBlock), 0, s, np3, c->pts64, c->stepd);
    const int nb = nblocks(np3 + nf9);
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, s, np3 + nf9, c->stepd, c->red);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, c->red, nb, &c->sc->sum);
    double v = 0.0;
    MBA_TRY(read_scalar(&c->sc->sum, &v));
    *out = std::sqrt(v);
    return MBA_OK;
  }

  // Max |g| and |g| of the unscaled gradient.
  mba_status gradient_norms(double* gmax, double* gnorm) {
    double init[2] = {0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(c->red2, init, sizeof(init), hipMemcpyHostToDevice, s));
    const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
    int nb = std::min(1024, nblocks(np3));
    hipLaunchKernelGGL(k_grad_partial<T>, dim3(nb), dim3(kBlock), 0, s, np3, d.gp, c->pscale64, c->red,
                       c->red + 2048);
    hipLaunchKernelGGL(k_grad_final, dim3(1), dim3(1), 0, s, c->red, c->red + 2048, nb, c->red2);
    nb = std::min(1024, nblocks(nf9));
    if (nf9 > 0) {
      hipLaunchKernelGGL(k_grad_partial<T>, dim3(nb), dim3(kBlock), 0, s, nf9, d.gc, c->cscale64, c->red,
                         c->red + 2048);
      hipLaunchKernelGGL(k_grad_final, dim3(1), dim3(1), 0, s, c->red, c->red + 2048, nb, c->red2);
    }
    double v[2];
    HIP_TRY(hipMemcpyAsync(v, c->red2, sizeof(v), hipMemcpyDeviceToHost, s));
    MBA_TRY(sync());
    *gmax = v[0];
    *gnorm = std::sqrt(v[1]);
    return MBA_OK;
  }

  // EstimateScale (trust_region_minimizer.cc:103-109) from an unscaled pass.
  mba_status init_scale(bool jacobi) {
    hipLaunchKernelGGL(k_fill64, dim3(nblocks(3 * (int64_t)c->ne)), dim3(kBlock), 0, s, (int64_t)3 * c->ne,
                       1.0, c->pscale64);
    hipLaunchKernelGGL(k_fill64, dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s, (int64_t)9 * c->nf,
                       1.0, c->cscale64);
    sync_scale_T();
    if (!jacobi) {
      c->scale_ready = true;
      return MBA_OK;
    }
    jacobian_pass(0.0, 1);
    hipLaunchKernelGGL(k_make_scale<T>, dim3(nblocks(std::max(c->ne, c->nf))), dim3(kBlock), 0, s, c->ne,
                       c->nf, d.V, d.U, c->pscale64, c->cscale64);
    sync_scale_T();
    c->scale_ready = true;
    return MBA_OK;
  }
  void sync_scale_T() {
    hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(3 * (int64_t)c->ne)), dim3(kBlock), 0, s,
                       (int64_t)3 * c->ne, c->pscale64, d.pscale);
    hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s,
                       (int64_t)9 * c->nf, c->cscale64, d.cscale);
  }

  // Solve the LM system at the current radius; fills step_p / step_c.
  mba_status compute_step(bool* solved, int* lin_iters) {
    *solved = true;
    *lin_iters = 0;
    if (c->opt.linear_solver == MBA_IMPLICIT_SCHUR_PCG) {
      int status = 0;
      if (c->nf > 0) {
        MBA_TRY(pcg(lin_iters, &status));
      } else {
        HIP_TRY(hipMemsetAsync(d.x, 0, sizeof(T) * 9, s));
      }
      if (status < 0) *solved = false;
    } else {
      bool ok = true;
      MBA_TRY(explicit_solve(&ok));
      if (!ok) *solved = false;
    }
    if (*solved) back_substitute();
    return MBA_OK;
  }

  // TrustRegionMinimizer::Minimize + LevenbergMarquardtStrategy.
  mba_status solve(mba_summary* sum, mba_iteration* its, int32_t max_its) {
    const mba_options& o = c->opt;
    const double t_start = now_s();
    std::memset(sum, 0, sizeof(*sum));
    sum->num_e_blocks = c->ne;
    sum->num_f_blocks = c->nf;
    sum->termination_type = MBA_NO_CONVERGENCE;
    int nrec = 0;
    double rec_min = std::numeric_limits<double>::max();
    auto push = [&](const mba_iteration& it) {
      if (its && nrec < max_its) its[nrec] = it;
      ++nrec;
      rec_min = std::min(rec_min, it.cost);
    };
    double radius = o.initial_trust_region_radius, decrease_factor = 2.0;
    double t_jac = 0.0, t_res = 0.0, t_lin = 0.0;
    mba_iteration it;
    std::memset(&it, 0, sizeof(it));
    it.trust_region_radius = radius;
    double cost = 0.0;
    bool finite = true;
    double t0 = now_s();
    MBA_TRY(cost64(c->cams64, c->pts64, &cost, &finite));
    t_res += now_s() - t0;
    int num_invalid = 0;
    bool have_jac = false;
    if (!finite) {
      sum->termination_type = MBA_FAILURE;
    } else {
      sum->initial_cost = cost;
      it.cost = cost;
      t0 = now_s();
      tables(c->cams64, c->pts64);
      MBA_TRY(init_scale(o.jacobi_scaling != 0));
      MBA_TRY(clear_flag());
      jacobian_pass(radius, 1);
      int32_t fl = 0;
      MBA_TRY(read_flag(&fl));
      t_jac += now_s() - t0;
      have_jac = true;
      MBA_TRY(gradient_norms(&it.gradient_max_norm, &it.gradient_norm));
      if (it.gradient_max_norm <= o.gradient_tolerance) {
        sum->termination_type = MBA_CONVERGENCE;
        it.cumulative_time_s = now_s() - t_start;
        push(it);
        goto done;
      }
      it.cumulative_time_s = now_s() - t_start;
      push(it);
      double x_norm = 0.0;
      MBA_TRY(x_norm_fn(&x_norm));
      for (;;) {
        if (it.iteration >= o.max_num_iterations) {
          sum->termination_type = MBA_NO_CONVERGENCE;
          break;
        }
        bool solved = true;
        int lin = 0;
        t0 = now_s();
        MBA_TRY(clear_flag());
        MBA_TRY(compute_step(&solved, &lin));
        int32_t fl2 = 0;
        MBA_TRY(read_flag(&fl2));
        if (fl2) solved = false;
        t_lin += now_s() - t0;
        sum->num_linear_solver_iterations += lin;
        const int prev = it.iteration;
        const double pgm = it.gradient_max_norm, pgn = it.gradient_norm;
        std::memset(&it, 0, sizeof(it));
        it.iteration = prev + 1;
        it.linear_solver_iterations = lin;
        double mcc = 0.0;
        if (solved) {
          MBA_TRY(model_cost_change(&mcc));
          it.step_is_valid = std::isfinite(mcc) && mcc >= 0.0;
        }
        if (!it.step_is_valid) {
          if (++num_invalid >= o.max_num_consecutive_invalid_steps) {
            sum->termination_type = MBA_FAILURE;
            break;
          }
          it.cost = cost;
          it.gradient_max_norm = pgm;
          it.gradient_norm = pgn;
        } else {
          num_invalid = 0;
          double step_norm = 0.0;
          bool sf = true;
          MBA_TRY(plus(&step_norm, &sf));
          double new_cost = std::numeric_limits<double>::max();
          bool cf = true;
          t0 = now_s();
          MBA_TRY(cost64(c->cams64_t, c->pts64_t, &new_cost, &cf));
          t_res += now_s() - t0;
          if (!cf || !sf) new_cost = std::numeric_limits<double>::max();
          it.step_norm = step_norm;
          if (step_norm <= o.parameter_tolerance * (x_norm + o.parameter_tolerance)) {
            sum->termination_type = MBA_CONVERGENCE;
            break;
          }
          it.cost_change = cost - new_cost;
          if (std::fabs(it.cost_change) <= o.function_tolerance * cost) {
            sum->termination_type = MBA_CONVERGENCE;
            break;
          }
          it.relative_decrease = it.cost_change / mcc;
          it.step_is_successful = it.relative_decrease > o.min_relative_decrease;
          if (it.step_is_successful) {
            // Accept: radius update first (StepAccepted), then evaluate.
            radius = radius / std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * it.relative_decrease - 1.0, 3));
            radius = std::min(o.max_trust_region_radius, radius);
            decrease_factor = 2.0;
            std::swap(c->cams64, c->cams64_t);
            std::swap(c->pts64, c->pts64_t);
            // Keep the constant cameras of the trial copy in sync.
            HIP_TRY(hipMemcpyAsync(c->cams64_t, c->cams64, sizeof(double) * 9 * c->nc, hipMemcpyDeviceToDevice, s));
            cost = new_cost;
            MBA_TRY(x_norm_fn(&x_norm));
            t0 = now_s();
            tables(c->cams64, c->pts64);
            jacobian_pass(radius, 1);
            MBA_TRY(sync());
            t_jac += now_s() - t0;
            MBA_TRY(gradient_norms(&it.gradient_max_norm, &it.gradient_norm));
          }
        }
        if (it.step_is_successful) {
          ++sum->num_successful_steps;
        } else {
          ++sum->num_unsuccessful_steps;
          radius = radius / decrease_factor;  // StepRejected / StepIsInvalid
          decrease_factor *= 2.0;
          damp(radius);
        }
        it.cost = cost;
        it.trust_region_radius = radius;
        it.cumulative_time_s = now_s() - t_start;
        push(it);
        if (it.step_is_successful) {
          if (it.gradient_max_norm <= o.gradient_tolerance) {
            sum->termination_type = MBA_CONVERGENCE;
            break;
          }
        } else if (radius < o.min_trust_region_radius) {
          sum->termination_type = MBA_CONVERGENCE;
          break;
        }
      }
    }
  done:
    (void)have_jac;
    sum->final_cost = std::min(sum->initial_cost, rec_min);
    if (sum->termination_type == MBA_FAILURE && nrec == 0) sum->final_cost = sum->initial_cost;
    sum->num_iterations = nrec;
    sum->is_solution_usable = sum->termination_type != MBA_FAILURE;
    sum->jacobian_evaluation_time_s = t_jac;
    sum->residual_evaluation_time_s = t_res;
    sum->linear_solver_time_s = t_lin;
    sum->total_time_s = now_s() - t_start;
    return sync();
  }

  mba_status x_norm_fn(double* out) { return x_norm(out); }
};

}  // namespace
}  // namespace mba

// ======================================================================= ABI
extern "C" {

mba_status mba_default_options(mba_options* o) {
  if (!o) return fail(MBA_ERR_INVALID_ARGUMENT, "options is NULL");
  std::memset(o, 0, sizeof(*o));
  o->max_num_iterations = 50;
  o->max_num_consecutive_invalid_steps = 5;
  o->max_linear_solver_iterations = 500;
  o->min_linear_solver_iterations = 0;
  o->jacobi_scaling = 1;
  o->linear_solver = MBA_EXPLICIT_SCHUR;
  o->precision = MBA_PRECISION_FP64;
  o->device = 0;
  o->function_tolerance = 1e-6;
  o->gradient_tolerance = 1e-10;
  o->parameter_tolerance = 1e-8;
  o->initial_trust_region_radius = 1e4;
  o->max_trust_region_radius = 1e16;
  o->min_trust_region_radius = 1e-32;
  o->min_relative_decrease = 1e-3;
  o->min_lm_diagonal = 1e-6;
  o->max_lm_diagonal = 1e32;
  o->eta = 1e-1;
  o->pcg_r_tolerance = -1.0;
  o->huber_a = 0.0;
  o->refinement_iterations = 2;
  return MBA_OK;
}

const char* mba_status_string(mba_status s) {
  switch (s) {
    case MBA_OK: return "ok";
    case MBA_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MBA_ERR_UNSUPPORTED: return "unsupported";
    case MBA_ERR_DEVICE: return "device error";
    case MBA_ERR_OUT_OF_MEMORY: return "out of memory";
    case MBA_ERR_NOT_FINITE: return "non-finite input";
  }
  return "unknown";
}

const char* mba_last_error(void) { return g_last_error.c_str(); }
int32_t mba_abi_version(void) { return MBA_ABI_VERSION; }

mba_status mba_context_create(const mba_problem* problem, const mba_options* options, mba_context** out) {
  g_last_error.clear();
  return context_create(problem, options, out);
}

mba_status mba_context_destroy(mba_context* ctx) {
  context_free(ctx);
  return MBA_OK;
}

void* mba_context_stream(mba_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

mba_status mba_context_synchronize(mba_context* ctx) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MBA_OK;
}

mba_status mba_context_set_timing(mba_context* ctx, int32_t enable) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  ctx->timing = enable != 0;
  return MBA_OK;
}

mba_status mba_context_last_timing(mba_context* ctx, double* ms, int32_t n) {
  if (!ctx || !ms) return fail(MBA_ERR_INVALID_ARGUMENT, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  float a = 0.f, b = 0.f, t = 0.f;
  HIP_TRY(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
  HIP_TRY(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
  HIP_TRY(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[3]));
  const double v[3] = {a, b, t};
  for (int i = 0; i < n && i < 3; ++i) ms[i] = v[i];
  return MBA_OK;
}

mba_status mba_context_jacobian_schur(mba_context* ctx, double radius) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  auto run = [&](auto& eng) -> mba_status {
    if (!ctx->scale_ready) {
      eng.tables(ctx->cams64, ctx->pts64);
      MBA_TRY(eng.init_scale(ctx->opt.jacobi_scaling != 0));
    }
    eng.tables(ctx->cams64, ctx->pts64);
    eng.jacobian_pass(radius, 1);
    return MBA_OK;
  };
  if (ctx->mixed) {
    Engine<float> e(ctx);
    return run(e);
  }
  Engine<double> e(ctx);
  return run(e);
}

mba_status mba_context_read_scale(mba_context* ctx, double* cam_scale, double* pt_scale);

mba_status mba_context_read_schur(mba_context* ctx, double* U, double* gc, double* rhs, double* V, double* gp) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  const Structure& st = ctx->st;
  auto rd = [&](auto* src, size_t n, std::vector<double>* out) -> mba_status {
    using X = std::remove_cv_t<std::remove_pointer_t<decltype(src)>>;
    std::vector<X> h(n);
    if (n) HIP_TRY(hipMemcpyAsync(h.data(), src, n * sizeof(X), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    out->assign(h.begin(), h.end());
    return MBA_OK;
  };
  auto go = [&](auto& d) -> mba_status {
    std::vector<double> hU, hg, hr, hV, hgp;
    MBA_TRY(rd(d.U, 45 * (size_t)ctx->nf, &hU));
    MBA_TRY(rd(d.gc, 9 * (size_t)ctx->nf, &hg));
    MBA_TRY(rd(d.rhs, 9 * (size_t)ctx->nf, &hr));
    MBA_TRY(rd(d.V, 6 * (size_t)ctx->ne, &hV));
    MBA_TRY(rd(d.gp, 3 * (size_t)ctx->ne, &hgp));
    for (int32_t f = 0; f < ctx->nf; ++f) {
      const int32_t cam = st.f_to_cam[f];
      if (U) std::memcpy(U + 45 * (size_t)cam, hU.data() + 45 * (size_t)f, 45 * sizeof(double));
      if (gc) std::memcpy(gc + 9 * (size_t)cam, hg.data() + 9 * (size_t)f, 9 * sizeof(double));
      if (rhs) std::memcpy(rhs + 9 * (size_t)cam, hr.data() + 9 * (size_t)f, 9 * sizeof(double));
    }
    for (int32_t e = 0; e < ctx->ne; ++e) {
      const int32_t q = ctx->dev_e_to_pt[e];
      if (V) std::memcpy(V + 6 * (size_t)q, hV.data() + 6 * (size_t)e, 6 * sizeof(double));
      if (gp) std::memcpy(gp + 3 * (size_t)q, hgp.data() + 3 * (size_t)e, 3 * sizeof(double));
    }
    return MBA_OK;
  };
  if (ctx->mixed) return go(ctx->df);
  return go(ctx->dd);
}

mba_status mba_context_read_scale(mba_context* ctx, double* cam_scale, double* pt_scale) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<double> cs(9 * (size_t)ctx->nf), ps(3 * (size_t)ctx->ne);
  if (!cs.empty()) HIP_TRY(hipMemcpyAsync(cs.data(), ctx->cscale64, cs.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (!ps.empty()) HIP_TRY(hipMemcpyAsync(ps.data(), ctx->pscale64, ps.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (cam_scale) {
    for (int64_t i = 0; i < 9 * (int64_t)ctx->nc; ++i) cam_scale[i] = 1.0;
    for (int32_t f = 0; f < ctx->nf; ++f)
      std::memcpy(cam_scale + 9 * (size_t)ctx->st.f_to_cam[f], cs.data() + 9 * (size_t)f, 72);
  }
  if (pt_scale) {
    for (int64_t i = 0; i < 3 * (int64_t)ctx->np; ++i) pt_scale[i] = 1.0;
    for (int32_t e = 0; e < ctx->ne; ++e)
      std::memcpy(pt_scale + 3 * (size_t)ctx->dev_e_to_pt[e], ps.data() + 3 * (size_t)e, 24);
  }
  return MBA_OK;
}

mba_status mba_context_schur_matvec(mba_context* ctx, const double* x, double* y) {
  if (!ctx || !x || !y) return fail(MBA_ERR_INVALID_ARGUMENT, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  auto go = [&](auto& eng, auto* dx, auto* dy) -> mba_status {
    using T = std::remove_pointer_t<decltype(dx)>;
    std::vector<T> hx(9 * (size_t)ctx->nf), hy(9 * (size_t)ctx->nf);
    for (int32_t f = 0; f < ctx->nf; ++f)
      for (int k = 0; k < 9; ++k) hx[9 * (size_t)f + k] = (T)x[9 * (size_t)ctx->st.f_to_cam[f] + k];
    if (!hx.empty()) HIP_TRY(hipMemcpyAsync(dx, hx.data(), hx.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    eng.matvec(dx, dy, nullptr);
    if (!hy.empty()) HIP_TRY(hipMemcpyAsync(hy.data(), dy, hy.size() * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < 9 * (int64_t)ctx->nc; ++i) y[i] = 0.0;
    for (int32_t f = 0; f < ctx->nf; ++f)
      for (int k = 0; k < 9; ++k) y[9 * (size_t)ctx->st.f_to_cam[f] + k] = (double)hy[9 * (size_t)f + k];
    return MBA_OK;
  };
  if (ctx->mixed) {
    Engine<float> e(ctx);
    return go(e, ctx->df.p, ctx->df.q);
  }
  Engine<double> e(ctx);
  return go(e, ctx->dd.p, ctx->dd.q);
}

mba_status mba_context_solve(mba_context* ctx, mba_summary* summary, mba_iteration* its, int32_t max_its) {
  if (!ctx || !summary) return fail(MBA_ERR_INVALID_ARGUMENT, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->opt.linear_solver == MBA_EXPLICIT_SCHUR) {
    const double n = 9.0 * ctx->nf;
    const double bytes = 2.0 * n * n * (ctx->mixed ? 4.0 : 8.0);
    if (bytes > 64.0 * (1ull << 30))
      return fail(MBA_ERR_UNSUPPORTED, "explicit Schur matrix too large; use MBA_IMPLICIT_SCHUR_PCG");
    if (ctx->st.max_track > 128)
      return fail(MBA_ERR_UNSUPPORTED, "explicit Schur assembly supports tracks of at most 128 observations");
  }
  if (ctx->mixed) {
    Engine<float> e(ctx);
    return e.solve(summary, its, max_its);
  }
  Engine<double> e(ctx);
  return e.solve(summary, its, max_its);
}

mba_status mba_context_download(mba_context* ctx, mba_problem* p) {
  if (!ctx || !p) return fail(MBA_ERR_INVALID_ARGUMENT, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<double> cams(9 * (size_t)ctx->nc), pts(3 * (size_t)ctx->ne);
  if (!cams.empty()) HIP_TRY(hipMemcpyAsync(cams.data(), ctx->cams64, cams.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (!pts.empty()) HIP_TRY(hipMemcpyAsync(pts.data(), ctx->pts64, pts.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int32_t f = 0; f < ctx->nf; ++f) {
    const int32_t cam = ctx->st.f_to_cam[f];
    std::memcpy(p->cameras + 9 * (size_t)cam, cams.data() + 9 * (size_t)cam, 72);
  }
  for (int32_t e = 0; e < ctx->ne; ++e)
    std::memcpy(p->points + 3 * (size_t)ctx->dev_e_to_pt[e], pts.data() + 3 * (size_t)e, 24);
  return MBA_OK;
}

mba_status mba_solve(const mba_options* options, mba_problem* problem, mba_summary* summary,
                     mba_iteration* iterations, int32_t max_iterations) {
  g_last_error.clear();
  if (!summary) return fail(MBA_ERR_INVALID_ARGUMENT, "summary is NULL");
  std::memset(summary, 0, sizeof(*summary));
  const double t0 = now_s();
  mba_context* ctx = nullptr;
  mba_status s = context_create(problem, options, &ctx);
  if (s != MBA_OK) return s;
  const double t_pre = now_s() - t0;
  s = mba_context_solve(ctx, summary, iterations, max_iterations);
  if (s == MBA_OK && summary->is_solution_usable) s = mba_context_download(ctx, problem);
  summary->preprocessor_time_s = t_pre;
  summary->total_time_s += t_pre;
  context_free(ctx);
  return s;
}

mba_status mba_schur_structure(const mba_problem* problem, int32_t* param_order, int32_t* num_param_blocks,
                               int32_t* num_e_blocks, int64_t* residual_order, int32_t* pairs,
                               int64_t pairs_capacity, int64_t* num_pairs) {
  g_last_error.clear();
  Structure st;
  MBA_TRY(build_structure(problem, &st));
  if (param_order) std::memcpy(param_order, st.order.data(), st.order.size() * sizeof(int32_t));
  if (num_param_blocks) *num_param_blocks = (int32_t)st.order.size();
  if (num_e_blocks) *num_e_blocks = st.num_e;
  if (residual_order) std::memcpy(residual_order, st.res_order.data(), st.res_order.size() * sizeof(int64_t));
  if (pairs || num_pairs) {
    const std::vector<int64_t> keys = block_pairs(problem, st);
    const int64_t nf = (int64_t)st.order.size() - st.num_e;
    if (num_pairs) *num_pairs = (int64_t)keys.size();
    if (pairs) {
      if (pairs_capacity < (int64_t)keys.size()) return fail(MBA_ERR_INVALID_ARGUMENT, "pairs_capacity too small");
      for (size_t i = 0; i < keys.size(); ++i) {
        pairs[2 * i] = (int32_t)(keys[i] / nf);
        pairs[2 * i + 1] = (int32_t)(keys[i] % nf);
      }
    }
  }
  return MBA_OK;
}

}  // extern "C"

template <typename T>
static mba_status evaluate_impl(const mba_problem* p, double huber_a, double* residuals, double* jac_camera,
                                double* jac_point, double* cost, int32_t device) {
  MBA_TRY(validate(p));
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(MBA_ERR_DEVICE, "no such HIP device");
  HIP_TRY(hipSetDevice(device));
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  Pool pool;
  const int64_t n = p->num_observations;
  const int32_t nc = p->num_cameras, np = p->num_points;
  double* cams = nullptr;
  double *res = nullptr, *jc = nullptr, *jp = nullptr;
  T *tab = nullptr, *pts = nullptr, *obs = nullptr;
  int32_t *ci = nullptr, *pi = nullptr;
  auto run = [&]() -> mba_status {
    MBA_TRY(pool.alloc(&cams, 9 * (size_t)nc));
    MBA_TRY(pool.alloc(&tab, kCamTab * (size_t)nc));
    MBA_TRY(pool.alloc(&pts, 3 * (size_t)np));
    MBA_TRY(pool.alloc(&obs, 2 * (size_t)n));
    MBA_TRY(pool.alloc(&ci, (size_t)n));
    MBA_TRY(pool.alloc(&pi, (size_t)n));
    MBA_TRY(pool.alloc(&res, 2 * (size_t)n));
    if (jac_camera) MBA_TRY(pool.alloc(&jc, 18 * (size_t)n));
    if (jac_point) MBA_TRY(pool.alloc(&jp, 6 * (size_t)n));
    std::vector<T> hp(p->points, p->points + 3 * (size_t)np), ho(p->observations, p->observations + 2 * n);
    if (nc) HIP_TRY(hipMemcpyAsync(cams, p->cameras, 72 * (size_t)nc, hipMemcpyHostToDevice, s));
    if (np) HIP_TRY(hipMemcpyAsync(pts, hp.data(), hp.size() * sizeof(T), hipMemcpyHostToDevice, s));
    if (n) {
      HIP_TRY(hipMemcpyAsync(obs, ho.data(), ho.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(ci, p->camera_index, 4 * (size_t)n, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(pi, p->point_index, 4 * (size_t)n, hipMemcpyHostToDevice, s));
    }
    if (nc) hipLaunchKernelGGL(k_cam_table<T>, dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, cams, tab);
    if (n) hipLaunchKernelGGL(k_eval_obs<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, ci, pi, obs, tab, pts,
                              res, jc, jp);
    std::vector<double> hr(2 * (size_t)n);
    if (n) HIP_TRY(hipMemcpyAsync(hr.data(), res, hr.size() * 8, hipMemcpyDeviceToHost, s));
    if (jc && n) HIP_TRY(hipMemcpyAsync(jac_camera, jc, 18 * 8 * (size_t)n, hipMemcpyDeviceToHost, s));
    if (jp && n) HIP_TRY(hipMemcpyAsync(jac_point, jp, 6 * 8 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    if (residuals) std::memcpy(residuals, hr.data(), hr.size() * 8);
    if (cost) {
      double tot = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        const double sq = hr[2 * i] * hr[2 * i] + hr[2 * i + 1] * hr[2 * i + 1];
        tot += (huber_a > 0.0 && sq > huber_a * huber_a) ? 0.5 * (2.0 * huber_a * std::sqrt(sq) - huber_a * huber_a)
                                                         : 0.5 * sq;
      }
      *cost = tot;
    }
    return MBA_OK;
  };
  const mba_status st = run();
  (void)hipStreamSynchronize(s);
  pool.release();
  (void)hipStreamDestroy(s);
  return st;
}

extern "C" {

mba_status mba_evaluate(const mba_problem* p, int32_t precision, double huber_a, double* residuals,
                        double* jac_camera, double* jac_point, double* cost, int32_t device) {
  g_last_error.clear();
  if (precision == MBA_PRECISION_MIXED)
    return evaluate_impl<float>(p, huber_a, residuals, jac_camera, jac_point, cost, device);
  return evaluate_impl<double>(p, huber_a, residuals, jac_camera, jac_point, cost, device);
}

mba_status mba_scene_generate(int32_t preset, int32_t window, int32_t* num_cameras, int32_t* num_points,
                              int64_t* num_observations, double* cam_true, double* cam_init, double* pt_true,
                              double* pt_init, double* observations, int32_t* camera_index,
                              int32_t* point_index) {
  if (preset < 1 || preset > 5) return fail(MBA_ERR_INVALID_ARGUMENT, "preset must be 1..5");
  const mba_scene_spec sp = mba_scene_preset(preset, window);
  std::vector<int32_t> tl(sp.num_points);
  const int64_t n = mba_scene_count(&sp, tl.data());
  if (num_cameras) *num_cameras = sp.num_cameras;
  if (num_points) *num_points = sp.num_points;
  if (num_observations) *num_observations = n;
  if (!cam_init || !pt_init || !observations || !camera_index || !point_index) return MBA_OK;
  std::vector<int64_t> off((size_t)sp.num_points + 1);
  mba_scene_fill(&sp, tl.data(), cam_true, cam_init, pt_true, pt_init, observations, camera_index,
                 point_index, off.data());
  return MBA_OK;
}

}  // extern "C"
