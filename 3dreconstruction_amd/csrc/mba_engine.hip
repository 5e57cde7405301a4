// Host runtime of the MI355X bundle-adjustment engine: the reference's
// preprocessing (Schur ordering, residual order), the device layout, the
// trust-region Levenberg-Marquardt controller, the PCG / explicit-Schur
// linear solvers and the C ABI of include/mba.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>

#include <sys/mman.h>
#include <unistd.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <type_traits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mba.h"
#include "mba_device.h"
#include "mba_device_grouped.h"
#include "mba_scene.h"
#include "mba_tuning.h"

namespace mba {
namespace {

thread_local std::string g_last_error;

mba_status fail(mba_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

}  // namespace
void set_last_error(const std::string& msg) { g_last_error = msg; }
namespace {

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

// Host buffers that are fully written before they are read: default-
// initialised (no zero fill of tens of MB on the preprocessing path), and
// from 4 MB on backed by transparent huge pages where the kernel offers them
// (2 MB-aligned, MADV_HUGEPAGE): the first touch of a 200 MB row array by
// 16 threads then takes ~100 page faults instead of ~50,000.
constexpr size_t kHugePage = size_t(2) << 20, kHugeMin = size_t(4) << 20;
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U>&) noexcept {}
  T* allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < kHugeMin) return std::allocator<T>::allocate(n);
    const size_t sz = (bytes + kHugePage - 1) / kHugePage * kHugePage;
    void* q = std::aligned_alloc(kHugePage, sz);
    if (!q) throw std::bad_alloc();
    (void)madvise(q, sz, MADV_HUGEPAGE);
    return static_cast<T*>(q);
  }
  void deallocate(T* q, size_t n) {
    if (n * sizeof(T) < kHugeMin) {
      std::allocator<T>::deallocate(q, n);
      return;
    }
    std::free(q);
  }
  template <typename U>
  void construct(U* q) noexcept {
    ::new ((void*)q) U;
  }
  template <typename U, typename... A>
  void construct(U* q, A&&... a) {
    ::new ((void*)q) U(std::forward<A>(a)...);
  }
};
template <typename T>
using Buf = std::vector<T, NoInit<T>>;

// Host threads for n items (at most 16, about 64k items each).
int host_threads(int64_t n, int64_t grain = 1 << 16) {
  const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<int64_t>(1, std::min<int64_t>({std::min<int64_t>(16, hw), (n + grain - 1) / grain}));
}

// Persistent host workers (up to 15, started on first use, never joined):
// run() hands fn(t), t = 1 .. nt-1, to workers, runs fn(0) on the caller
// and returns when all have finished.  One job at a time: a call made while
// the pool is busy (another thread's job, or a nested call), or made in a
// child of fork() (which has no workers), returns false.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* p = new HostPool();
    return *p;
  }
  bool run(int nt, const std::function<void(int)>& fn) {
    if (nt - 1 > (int)nworkers_ || getpid() != pid_) return false;
    std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
    if (!busy.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> g(m_);
      job_ = &fn;
      nparticipants_ = nt - 1;
      remaining_ = nt - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [&] { return remaining_ == 0; });
    job_ = nullptr;
    return true;
  }

 private:
  HostPool() : pid_(getpid()) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    nworkers_ = std::min(15u, hw > 1 ? hw - 1 : 0u);
    for (unsigned i = 0; i < nworkers_; ++i) std::thread([this, i]() { loop((int)i); }).detach();
  }
  void loop(int i) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* job = nullptr;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
        if (i >= nparticipants_) continue;
        job = job_;
      }
      (*job)(i + 1);
      std::lock_guard<std::mutex> g(m_);
      if (--remaining_ == 0) done_.notify_one();
    }
  }
  const pid_t pid_;
  unsigned nworkers_ = 0;
  std::mutex busy_, m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int nparticipants_ = 0, remaining_ = 0;
  uint64_t gen_ = 0;
};

// fn(t) for t in [0, nt) on nt threads (inline when nt == 1): the
// persistent workers when free, else threads of its own.
template <typename F>
void run_threads(int nt, F&& fn) {
  if (nt <= 1) {
    fn(0);
    return;
  }
  const std::function<void(int)> job = [&fn](int t) { fn(t); };
  if (HostPool::get().run(nt, job)) return;
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t) th.emplace_back([&fn, t]() { fn(t); });
  for (std::thread& x : th) x.join();
}

// fn(begin, end) over equal ranges of [0, n).
template <typename F>
void parallel_for(int64_t n, F&& fn, int64_t grain = 1 << 16) {
  if (n <= 0) return;
  const int nt = host_threads(n, grain);
  const int64_t step = (n + nt - 1) / nt;
  run_threads(nt, [&](int t) {
    const int64_t b = t * step, e = std::min(n, b + step);
    if (b < e) fn(b, e);
  });
}

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
  int32_t layout = MBA_LAYOUT_PACKED9;
  int32_t ng = 0, nc = 0, np = 0;  // ng: intrinsic groups (grouped layout)
  int64_t nobs = 0;
  // Column blocks in order.  Packed: cameras c, points nc + p.  Grouped:
  // groups g, cameras ng + c, points ng + nc + p.
  std::vector<int32_t> order;
  int32_t num_e = 0;
  Buf<int64_t> res_order;          // row -> input observation
  bool standard = false;           // E == active points
  // Derived (standard structure only).  f_to_cam: the F block's parameter
  // slot block (packed: its camera; grouped: camera c at c, group g at
  // nc + g); cam_to_f is its inverse (-1: constant or unused).
  std::vector<int32_t> e_to_pt, f_to_cam, cam_to_f, pt_start;
  std::vector<int32_t> cam_grp;    // grouped: the group of each camera (-1: none)
  int32_t max_track = 0;
};

// Phase timer of the host preprocessing: with MBA_TRACE_PREP=1 in the
// environment, context_create prints each phase's wall time to stderr.
struct PrepTrace {
  bool on = false;
  double t = 0.0;
  PrepTrace() {
    const char* e = std::getenv("MBA_TRACE_PREP");
    on = e && e[0] == '1';
    t = now_s();
  }
  void mark(const char* what) {
    if (!on) return;
    const double n = now_s();
    std::fprintf(stderr, "mba prep: %-28s %8.2f ms\n", what, 1e3 * (n - t));
    t = n;
  }
};

mba_status validate(const mba_problem* p) {
  if (!p) return fail(MBA_ERR_INVALID_ARGUMENT, "problem is NULL");
  if (p->layout != MBA_LAYOUT_PACKED9 && p->layout != MBA_LAYOUT_GROUPED)
    return fail(MBA_ERR_INVALID_ARGUMENT, "unknown camera layout " + std::to_string(p->layout));
  if (p->num_cameras < 0 || p->num_points < 0 || p->num_observations < 0)
    return fail(MBA_ERR_INVALID_ARGUMENT, "negative sizes");
  if (p->ordering != MBA_ORDER_AUTOMATIC && p->ordering != MBA_ORDER_POINTS_FIRST)
    return fail(MBA_ERR_INVALID_ARGUMENT, "unknown ordering " + std::to_string(p->ordering));
  {
    const int32_t all = p->layout == MBA_LAYOUT_GROUPED ? 0x3F : 0x1FF;
    if (p->camera_params < 0 || (p->camera_params & ~all) != 0)
      return fail(MBA_ERR_INVALID_ARGUMENT, "camera_params has bits outside the camera vector");
  }
  if (p->layout == MBA_LAYOUT_GROUPED) {
    if (p->num_intrinsics < 0) return fail(MBA_ERR_INVALID_ARGUMENT, "negative num_intrinsics");
    if (p->num_intrinsics > 0 && !p->intrinsics) return fail(MBA_ERR_INVALID_ARGUMENT, "intrinsics is NULL");
    if (p->num_observations > 0 && !p->intrinsic_index)
      return fail(MBA_ERR_INVALID_ARGUMENT, "intrinsic_index is NULL");
    for (int64_t i = 0; i < p->num_observations; ++i)
      if (p->intrinsic_index[i] < 0 || p->intrinsic_index[i] >= p->num_intrinsics)
        return fail(MBA_ERR_INVALID_ARGUMENT,
                    "observation " + std::to_string(i) + " has an intrinsic index out of range");
    // SubsetParameterization needs at least one constant and one free entry
    // (local_parameterization.cc:93-109); a fully held block is constant.
    if (p->intrinsic_subset)
      for (int32_t g = 0; g < p->num_intrinsics; ++g)
        if (p->intrinsic_subset[g] > 0x3E && p->intrinsic_subset[g] != 0)
          return fail(MBA_ERR_INVALID_ARGUMENT, "intrinsic_subset of group " + std::to_string(g) +
                                                    " must hold 1..5 of the 6 entries (use intrinsic_constant)");
  }
  if (p->num_observations > 0x7fffffffLL) return fail(MBA_ERR_UNSUPPORTED, "more than 2^31-1 observations");
  if (p->num_observations > 0 && (!p->observations || !p->camera_index || !p->point_index))
    return fail(MBA_ERR_INVALID_ARGUMENT, "observation arrays are NULL");
  if ((p->num_cameras > 0 && !p->cameras) || (p->num_points > 0 && !p->points))
    return fail(MBA_ERR_INVALID_ARGUMENT, "parameter arrays are NULL");
  // Index ranges, in parallel; the first offending observation is named.
  std::atomic<int64_t> bad{INT64_MAX};
  parallel_for(p->num_observations, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const int32_t c = p->camera_index[i], q = p->point_index[i];
      if (c < 0 || c >= p->num_cameras || q < 0 || q >= p->num_points) {
        int64_t cur = bad.load();
        while (i < cur && !bad.compare_exchange_weak(cur, i)) {
        }
        return;
      }
    }
  }, 1 << 18);
  if (bad.load() != INT64_MAX)
    return fail(MBA_ERR_INVALID_ARGUMENT, "observation " + std::to_string(bad.load()) + " has an index out of range");
  return MBA_OK;
}

mba_status build_structure_grouped(const mba_problem* p, Structure* st);

// MBA_ORDER_POINTS_FIRST: group 0 = every active point block, group 1 = the
// rest, each group in address order of the libs/sfm containers (cameras,
// intrinsics, points): vertex ids in the structure are already in that
// order within each group (points are the ids >= `first_point`).
void user_points_first(const std::vector<int32_t>& active, int32_t first_point, Structure* st) {
  std::vector<int32_t> pts, rest;
  for (int32_t b : active) (b >= first_point ? pts : rest).push_back(b);
  std::sort(pts.begin(), pts.end());
  std::sort(rest.begin(), rest.end(), [&](int32_t a, int32_t b) {
    // Grouped ids: groups 0..ng-1 precede cameras, but in the containers
    // the cameras come first; packed ids are cameras only.
    const bool ga = a < st->ng, gb = b < st->ng;
    if (ga != gb) return gb;
    return a < b;
  });
  st->order = pts;
  st->num_e = (int32_t)pts.size();
  st->order.insert(st->order.end(), rest.begin(), rest.end());
}

// Relaxed atomic add on a plain int (host threads).
inline int32_t atomic_inc(int32_t* v) { return __atomic_fetch_add(v, 1, __ATOMIC_RELAXED); }

// Residual order from each residual's bucket (0 .. ne, ne last): buckets
// in order, each filled back to front (descending residual index), as
// LexicographicallyOrderResidualBlocks (reorder_program.cc:295-309).  Each
// thread owns a range of buckets: it counts them over all residuals, then
// fills them scanning the residuals downwards.  bstart: bucket offsets
// (ne + 2 entries).
template <typename B, typename R>
void order_residuals(const B& bucket, int64_t no, int32_t ne, std::vector<int64_t>* bstart_out, R* res_order) {
  std::vector<int32_t> hist((size_t)ne + 1, 0);
  std::vector<int64_t>& bstart = *bstart_out;
  bstart.assign((size_t)ne + 2, 0);
  res_order->resize(no);
  const int nt = host_threads(no);
  const int64_t span = ((int64_t)ne + 1 + nt - 1) / std::max(1, nt);
  run_threads(nt, [&](int t) {
    const int32_t b0 = (int32_t)std::min<int64_t>(ne + 1, t * span), b1 = (int32_t)std::min<int64_t>(ne + 1, (t + 1) * span);
    // Branch-free: buckets outside the range count into a spare slot.
    const uint32_t w = (uint32_t)(b1 - b0);
    std::vector<int32_t> local((size_t)w + 1, 0);
    for (int64_t i = 0; i < no; ++i) {
      const uint32_t k = (uint32_t)(bucket[i] - b0);
      local[k < w ? k : w]++;
    }
    std::copy(local.begin(), local.begin() + w, hist.begin() + b0);
  });
  for (int32_t e = 0; e <= ne; ++e) bstart[e + 1] = bstart[e] + hist[e];
  run_threads(nt, [&](int t) {
    const int32_t b0 = (int32_t)std::min<int64_t>(ne + 1, t * span), b1 = (int32_t)std::min<int64_t>(ne + 1, (t + 1) * span);
    if (b0 >= b1) return;
    std::vector<int64_t> pos(bstart.begin() + b0, bstart.begin() + b1);
    for (int64_t i = no - 1; i >= 0; --i) {
      const int32_t m = bucket[i];
      if (m >= b0 && m < b1) (*res_order)[pos[m - b0]++] = i;
    }
  });
}

// Program order (problem_impl.cc:143,271): parameter blocks in order of
// first appearance in the residuals' argument lists.  blocks(i, out)
// lists residual i's block ids (at most kmax); a block's first position
// k_max i + k is found by a parallel atomic minimum, and the positions
// (distinct) are scanned in order.
template <typename F>
std::vector<int32_t> first_appearance(int64_t no, int32_t nb, int kmax, F&& blocks) {
  std::vector<int64_t> first(nb, std::numeric_limits<int64_t>::max());
  parallel_for(no, [&](int64_t b, int64_t e) {
    int32_t rb[3];
    for (int64_t i = b; i < e; ++i) {
      const int n = blocks(i, rb);
      for (int k = 0; k < n; ++k) {
        const int64_t key = kmax * i + k;
        int64_t cur = __atomic_load_n(&first[rb[k]], __ATOMIC_RELAXED);
        while (key < cur &&
               !__atomic_compare_exchange_n(&first[rb[k]], &cur, key, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
      }
    }
  });
  Buf<int32_t> slot(kmax * (size_t)no);
  parallel_for((int64_t)slot.size(), [&](int64_t b, int64_t e) { std::fill(slot.begin() + b, slot.begin() + e, -1); });
  parallel_for(nb, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v)
      if (first[v] != std::numeric_limits<int64_t>::max()) slot[first[v]] = (int32_t)v;
  });
  std::vector<int32_t> prog;
  prog.reserve(nb);
  for (int32_t v : slot)
    if (v >= 0) prog.push_back(v);
  return prog;
}

// Exclusive prefix sum out[0..n] of v[0..n) (out[n] = total), in parallel
// ranges: per-range totals, their scan, then each range's own.
template <typename A, typename B>
void exclusive_scan(const A* v, int64_t n, B* out) {
  const int nt = host_threads(n, 1 << 18);
  const int64_t step = (n + nt - 1) / std::max(1, nt);
  std::vector<B> part((size_t)nt + 1, 0);
  run_threads(nt, [&](int t) {
    B acc = 0;
    for (int64_t i = t * step, e = std::min(n, (t + 1) * step); i < e; ++i) acc += (B)v[i];
    part[t + 1] = acc;
  });
  for (int t = 0; t < nt; ++t) part[t + 1] += part[t];
  run_threads(nt, [&](int t) {
    B acc = part[t];
    for (int64_t i = t * step, e = std::min(n, (t + 1) * step); i < e; ++i) {
      out[i] = acc;
      acc += (B)v[i];
    }
  });
  out[n] = part[nt];
}

// The ordering of the common bundle-adjustment structure, without the
// Hessian adjacency: one counting sort of the observations by point (each
// point's list in descending observation index) gives every point its rows,
// its first appearance (the list's last entry), its degree (distinct
// non-constant cameras) and every camera's degree (distinct points).  When
// every point's degree is below every active camera's, the stable degree sort
// queues all points before all cameras, the greedy independent set is
// exactly the points, and each row's bucket is its point's E position -- the
// general path's result (point-first branch of build_structure), built here
// directly: E = points by (degree, first appearance), F = cameras by
// (degree, first appearance), rows = the points' lists in E order.  Returns
// false, leaving st untouched, when that condition fails.
bool build_structure_points_first(const mba_problem* p, Structure* st, PrepTrace& tr) {
  const int32_t nc = p->num_cameras, np = p->num_points;
  const int64_t no = p->num_observations;
  if (no == 0 || np == 0) return false;
  const int32_t* cam = p->camera_index;
  const int32_t* pt = p->point_index;
  auto is_const = [&](int32_t c) { return p->camera_constant != nullptr && p->camera_constant[c] != 0; };
  // Rows per point, then each point's list filled from its end (atomic
  // cursors: any order) and sorted descending below.
  Buf<int32_t> cnt(np);
  parallel_for(np, [&](int64_t b, int64_t e) { std::fill(cnt.begin() + b, cnt.begin() + e, 0); }, 1 << 18);
  // One host thread (a small problem, or a window of a batch): plain
  // increments; several: relaxed atomics.
  const bool serial = host_threads(no) == 1;
  if (serial) {
    for (int64_t i = 0; i < no; ++i) cnt[pt[i]]++;
  } else {
    parallel_for(no, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) __atomic_fetch_add(&cnt[pt[i]], 1, __ATOMIC_RELAXED);
    });
  }
  Buf<int32_t> pstart((size_t)np + 1);
  exclusive_scan(cnt.data(), np, pstart.data());
  Buf<int32_t> cursor(np);
  parallel_for(np, [&](int64_t b, int64_t e) {
    for (int64_t q = b; q < e; ++q) cursor[q] = pstart[q + 1];
  }, 1 << 18);
  Buf<int32_t> rows(no);
  if (serial) {
    for (int64_t i = 0; i < no; ++i) rows[--cursor[pt[i]]] = (int32_t)i;
  } else {
    parallel_for(no, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) rows[__atomic_sub_fetch(&cursor[pt[i]], 1, __ATOMIC_RELAXED)] = (int32_t)i;
    });
  }
  tr.mark("  rows by point");
  // Per point: sort its list descending, its degree; per camera (thread
  // histograms): distinct points and first appearance.
  const int nt = host_threads(np, 1 << 14);
  const int64_t step = ((int64_t)np + nt - 1) / nt;
  std::vector<std::vector<int32_t>> cdeg(nt, std::vector<int32_t>((size_t)nc, 0));
  std::vector<std::vector<int32_t>> cfirst(nt, std::vector<int32_t>((size_t)nc, INT32_MAX));
  std::vector<int32_t> tmax(nt, 0);
  Buf<int32_t> pdeg(np);
  run_threads(nt, [&](int t) {
    int32_t* cd = cdeg[t].data();
    int32_t* cf = cfirst[t].data();
    std::vector<int32_t> seen;
    int32_t mx = 0;
    for (int64_t q = t * step, qe = std::min<int64_t>(np, (t + 1) * step); q < qe; ++q) {
      int32_t* r = rows.data() + pstart[q];
      const int32_t n = pstart[q + 1] - pstart[q];
      if (n == 0) {
        pdeg[q] = -1;  // not in any residual: not a program block
        continue;
      }
      std::sort(r, r + n, [](int32_t a, int32_t b) { return a > b; });
      seen.clear();
      for (int32_t k = 0; k < n; ++k) {
        const int32_t c = cam[r[k]];
        cf[c] = std::min(cf[c], r[k]);
        if (is_const(c) || std::find(seen.begin(), seen.end(), c) != seen.end()) continue;
        seen.push_back(c);
        cd[c]++;
      }
      pdeg[q] = (int32_t)seen.size();
      mx = std::max(mx, pdeg[q]);
    }
    tmax[t] = mx;
  });
  const int32_t max_pt = *std::max_element(tmax.begin(), tmax.end());
  std::vector<int32_t> deg((size_t)nc, 0), first((size_t)nc, INT32_MAX);
  for (int t = 0; t < nt; ++t)
    for (int32_t c = 0; c < nc; ++c) {
      deg[c] += cdeg[t][c];
      first[c] = std::min(first[c], cfirst[t][c]);
    }
  int32_t min_cam = INT32_MAX;
  std::vector<int32_t> cams;
  for (int32_t c = 0; c < nc; ++c)
    if (first[c] != INT32_MAX && !is_const(c)) {
      cams.push_back(c);
      min_cam = std::min(min_cam, deg[c]);
    }
  tr.mark("  degrees");
  if (!(max_pt < min_cam)) return false;
  // E: points in first-appearance order (a point's first appearance is its
  // list's last entry; distinct across points), then stably by degree.
  Buf<int32_t> slot(no);
  parallel_for(no, [&](int64_t b, int64_t e) { std::fill(slot.begin() + b, slot.begin() + e, -1); });
  parallel_for(np, [&](int64_t b, int64_t e) {
    for (int64_t q = b; q < e; ++q)
      if (pdeg[q] >= 0) slot[rows[pstart[q + 1] - 1]] = (int32_t)q;
  }, 1 << 16);
  // Stable counting sort by degree: per-range histograms, then each range
  // fills its slots in index order.
  const int nt2 = host_threads(no);
  const int64_t step2 = (no + nt2 - 1) / nt2;
  const int32_t nd = max_pt + 1;
  std::vector<std::vector<int64_t>> hist(nt2, std::vector<int64_t>((size_t)nd, 0));
  run_threads(nt2, [&](int t) {
    int64_t* h = hist[t].data();
    for (int64_t i = t * step2, e = std::min(no, (t + 1) * step2); i < e; ++i)
      if (slot[i] >= 0) h[pdeg[slot[i]]]++;
  });
  int64_t run = 0;
  for (int32_t d = 0; d < nd; ++d)
    for (int t = 0; t < nt2; ++t) {
      const int64_t k = hist[t][d];
      hist[t][d] = run;
      run += k;
    }
  const int32_t ne = (int32_t)run;
  std::vector<int32_t> epts((size_t)ne);
  run_threads(nt2, [&](int t) {
    int64_t* h = hist[t].data();
    for (int64_t i = t * step2, e = std::min(no, (t + 1) * step2); i < e; ++i)
      if (slot[i] >= 0) epts[h[pdeg[slot[i]]]++] = slot[i];
  });
  // F: cameras by (degree, first appearance); first appearances are distinct.
  std::sort(cams.begin(), cams.end(), [&](int32_t a, int32_t b) {
    return deg[a] != deg[b] ? deg[a] < deg[b] : first[a] < first[b];
  });
  tr.mark("  E/F order");
  // Rows in E order.
  st->num_e = ne;
  st->order.resize((size_t)ne + cams.size());
  st->e_to_pt.assign(epts.begin(), epts.end());
  st->pt_start.resize((size_t)ne + 1);
  {
    Buf<int32_t> len(ne);
    std::vector<int32_t> lmax(host_threads(ne, 1 << 14), 0);
    const int64_t lstep = ((int64_t)ne + (int64_t)lmax.size() - 1) / (int64_t)lmax.size();
    run_threads((int)lmax.size(), [&](int t) {
      for (int64_t e = t * lstep, ee = std::min<int64_t>(ne, (t + 1) * lstep); e < ee; ++e) {
        const int32_t q = epts[e];
        len[e] = pstart[q + 1] - pstart[q];
        st->order[e] = nc + q;
        lmax[t] = std::max(lmax[t], len[e]);
      }
    });
    exclusive_scan(len.data(), ne, st->pt_start.data());
    st->max_track = *std::max_element(lmax.begin(), lmax.end());
  }
  st->res_order.resize(no);
  parallel_for(ne, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) {
      const int32_t q = epts[k];
      int64_t o = st->pt_start[k];
      for (int32_t r = pstart[q]; r < pstart[q + 1]; ++r) st->res_order[o++] = rows[r];
    }
  }, 1 << 14);
  const int32_t nf = (int32_t)cams.size();
  st->f_to_cam.assign(cams.begin(), cams.end());
  st->cam_to_f.assign(nc, -1);
  for (int32_t f = 0; f < nf; ++f) {
    st->order[(size_t)ne + f] = cams[f];
    st->cam_to_f[cams[f]] = f;
  }
  st->standard = true;
  tr.mark("  residual order");
  return true;
}

mba_status build_structure(const mba_problem* p, Structure* st) {
  MBA_TRY(validate(p));
  if (p->layout == MBA_LAYOUT_GROUPED) return build_structure_grouped(p, st);
  const int32_t nc = p->num_cameras, np = p->num_points;
  const int64_t no = p->num_observations;
  st->nc = nc;
  st->np = np;
  st->nobs = no;
  const int32_t nb = nc + np;
  auto is_const = [&](int32_t b) -> bool {
    return b < nc && p->camera_constant != nullptr && p->camera_constant[b] != 0;
  };
  PrepTrace tr;
  if (p->ordering == MBA_ORDER_AUTOMATIC && build_structure_points_first(p, st, tr)) return MBA_OK;
  // First appearance (program order).
  const std::vector<int32_t> prog = first_appearance(no, nb, 2, [&](int64_t i, int32_t* rb) {
    rb[0] = p->camera_index[i];
    rb[1] = nc + p->point_index[i];
    return 2;
  });
  tr.mark("  first appearance");
  // CSR of observations by camera: per-thread histograms over observation
  // ranges (the cameras are few), then each thread fills its own slots.
  std::vector<int64_t> cstart((size_t)nc + 1, 0);
  Buf<int32_t> cobs_pt(no);
  {
    const int nt = host_threads(no);
    const int64_t step = (no + nt - 1) / std::max(1, nt);
    std::vector<std::vector<int64_t>> h(nt, std::vector<int64_t>((size_t)nc, 0));
    run_threads(nt, [&](int t) {
      const int64_t e = std::min<int64_t>(no, (t + 1) * step);
      for (int64_t i = t * step; i < e; ++i) h[t][p->camera_index[i]]++;
    });
    int64_t run = 0;
    for (int32_t c = 0; c < nc; ++c) {
      cstart[c] = run;
      for (int t = 0; t < nt; ++t) {
        const int64_t k = h[t][c];
        h[t][c] = run;
        run += k;
      }
    }
    cstart[nc] = run;
    run_threads(nt, [&](int t) {
      const int64_t e = std::min<int64_t>(no, (t + 1) * step);
      for (int64_t i = t * step; i < e; ++i) cobs_pt[h[t][p->camera_index[i]]++] = p->point_index[i];
    });
  }
  tr.mark("  camera CSR");
  // Hessian-graph edges: the distinct points of every non-constant camera
  // (cameras in parallel, a point marker per thread), compacted into cadj;
  // a point's degree counts the cameras that list it.
  // cadj: camera c's distinct points at [cstart[c], cstart[c] + deg[c]).
  std::vector<int32_t> deg(nb, 0);
  Buf<int32_t> cadj(no);
  {
    const int nt = host_threads(no);
    std::atomic<int32_t> next{0};
    run_threads(nt, [&](int) {
      std::vector<int32_t> mark(np, -1);
      for (int32_t c = next.fetch_add(1); c < nc; c = next.fetch_add(1)) {
        if (is_const(c)) continue;
        int64_t o = cstart[c];
        for (int64_t k = cstart[c]; k < cstart[c + 1]; ++k) {
          const int32_t q = cobs_pt[k];
          if (mark[q] == c) continue;
          mark[q] = c;
          cadj[o++] = q;
          atomic_inc(&deg[nc + q]);
        }
        deg[c] = (int32_t)(o - cstart[c]);
      }
    });
  }
  tr.mark("  adjacency");
  // Active vertices in program order, stable counting sort by degree.
  std::vector<int32_t> active;
  active.reserve(prog.size());
  for (int32_t b : prog)
    if (!is_const(b)) active.push_back(b);
  int32_t maxdeg = 0, max_pt = 0, min_cam = std::numeric_limits<int32_t>::max();
  for (int32_t b : active) {
    maxdeg = std::max(maxdeg, deg[b]);
    if (b < nc) min_cam = std::min(min_cam, deg[b]);
    else max_pt = std::max(max_pt, deg[b]);
  }
  std::vector<int64_t> cnt((size_t)maxdeg + 2, 0);
  for (int32_t b : active) cnt[deg[b] + 1]++;
  for (int32_t d = 0; d <= maxdeg; ++d) cnt[d + 1] += cnt[d];
  std::vector<int32_t> queue(active.size());
  for (int32_t b : active) queue[cnt[deg[b]]++] = b;
  tr.mark("  degree sort");
  st->order.clear();
  st->order.reserve(active.size());
  if (p->ordering == MBA_ORDER_POINTS_FIRST) {
    user_points_first(active, nc, st);
  } else if (max_pt < min_cam) {
    // Every point precedes every camera in the queue: each point is still
    // white when reached (only cameras neighbour it) and every camera is
    // then grey, so the greedy set is the points in queue order.
    for (int32_t b : queue)
      if (b >= nc) st->order.push_back(b);
    st->num_e = (int32_t)st->order.size();
    for (int32_t b : queue)
      if (b < nc) st->order.push_back(b);
  } else {
    // Greedy independent set over both sides of the adjacency.
    std::vector<int64_t> padj_start((size_t)np + 1, 0);
    for (int32_t q = 0; q < np; ++q) padj_start[q + 1] = padj_start[q] + deg[nc + q];
    std::vector<int32_t> padj((size_t)padj_start[np]);
    {
      std::vector<int64_t> fill(padj_start.begin(), padj_start.end() - 1);
      for (int32_t c = 0; c < nc; ++c)
        for (int64_t k = cstart[c]; k < cstart[c] + deg[c]; ++k) padj[fill[cadj[k]]++] = c;
    }
    std::vector<uint8_t> color(nb, 0);  // 0 white, 1 grey, 2 black
    for (int32_t b : queue) {
      if (color[b] != 0) continue;
      st->order.push_back(b);
      color[b] = 2;
      if (b < nc) {
        for (int64_t k = cstart[b]; k < cstart[b] + deg[b]; ++k) color[nc + cadj[k]] = 1;
      } else {
        const int32_t q = b - nc;
        for (int64_t k = padj_start[q]; k < padj_start[q + 1]; ++k) color[padj[k]] = 1;
      }
    }
    st->num_e = (int32_t)st->order.size();
    for (int32_t b : queue)
      if (color[b] != 2) st->order.push_back(b);
  }
  tr.mark("  independent set");
  // Column-block index of every block.
  std::vector<int32_t> col(nb, -1);
  for (size_t i = 0; i < st->order.size(); ++i) col[st->order[i]] = (int32_t)i;
  // Residual order: all rows have a non-constant block (points are never
  // constant); bucket = smallest column block index; buckets in column
  // order, each filled back to front (descending observation index).  The
  // buckets are counted and filled in parallel, then each bucket is sorted.
  const int32_t ne = st->num_e;
  Buf<int32_t> bucket(no);
  parallel_for(no, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const int32_t a = p->camera_index[i], q = nc + p->point_index[i];
      int32_t m = ne;
      if (!is_const(a)) m = std::min(m, col[a]);
      m = std::min(m, col[q]);
      bucket[i] = m;
    }
  });
  std::vector<int64_t> bstart;
  order_residuals(bucket, no, ne, &bstart, &st->res_order);
  tr.mark("  residual order");
  // Standard BA structure: E = every active point, F = every active camera.
  int32_t n_act_pts = 0;
  for (int32_t b : active) n_act_pts += b >= nc;
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
  // Every row's bucket is its point's E block, so the point-major row
  // ranges are the buckets.
  st->pt_start.assign((size_t)st->num_e + 1, 0);
  for (int32_t e = 0; e <= st->num_e; ++e) st->pt_start[e] = (int32_t)bstart[e];
  st->max_track = 0;
  for (int32_t e = 0; e < st->num_e; ++e)
    st->max_track = std::max(st->max_track, (int32_t)(bstart[e + 1] - bstart[e]));
  tr.mark("  E/F maps");
  return MBA_OK;
}

// Parameter blocks of observation i in AddResidualBlock argument order, as
// vertex ids of the structure; returns their count.
inline int residual_blocks(const mba_problem* p, int64_t i, int32_t out[3]) {
  if (p->layout == MBA_LAYOUT_GROUPED) {
    const int32_t ng = p->num_intrinsics, nc = p->num_cameras;
    out[0] = p->intrinsic_index[i];
    out[1] = ng + p->camera_index[i];
    out[2] = ng + nc + p->point_index[i];
    return 3;
  }
  out[0] = p->camera_index[i];
  out[1] = p->num_cameras + p->point_index[i];
  return 2;
}

// SetParameterBlockConstant of a vertex (points are never constant).
inline bool block_constant(const mba_problem* p, int32_t b) {
  if (p->layout == MBA_LAYOUT_GROUPED) {
    const int32_t ng = p->num_intrinsics, nc = p->num_cameras;
    if (b < ng) return p->intrinsic_constant && p->intrinsic_constant[b];
    if (b < ng + nc) return p->camera_constant && p->camera_constant[b - ng];
    return false;
  }
  return b < p->num_cameras && p->camera_constant && p->camera_constant[b];
}

// The same preprocessing for the grouped layout's residual blocks
// (intrinsics, camera, point): Hessian-graph edges join every pair of
// non-constant blocks of a residual (parameter_block_ordering.cc:132-151), so
// a shared group is adjacent to its cameras and to every point they see.
// Every observation of a camera must use one group (as the incremental
// engine builds them, sfm_incremental_engine.cpp:1273-1393).
mba_status build_structure_grouped(const mba_problem* p, Structure* st) {
  const int32_t ng = p->num_intrinsics, nc = p->num_cameras, np = p->num_points;
  const int64_t no = p->num_observations;
  st->layout = MBA_LAYOUT_GROUPED;
  st->ng = ng;
  st->nc = nc;
  st->np = np;
  st->nobs = no;
  if ((int64_t)ng + nc + np > 0x7fffffffLL) return fail(MBA_ERR_UNSUPPORTED, "more than 2^31-1 parameter blocks");
  const int32_t nb = ng + nc + np;
  PrepTrace tr;
  st->cam_grp.assign(nc, -1);
  for (int64_t i = 0; i < no; ++i) {
    const int32_t c = p->camera_index[i], g = p->intrinsic_index[i];
    if (st->cam_grp[c] < 0) st->cam_grp[c] = g;
    else if (st->cam_grp[c] != g)
      return fail(MBA_ERR_UNSUPPORTED, "camera " + std::to_string(c) + " is observed with two intrinsic groups");
  }
  auto is_const = [&](int32_t b) { return block_constant(p, b); };
  const std::vector<int32_t> prog =
      first_appearance(no, nb, 3, [&](int64_t i, int32_t* out) { return residual_blocks(p, i, out); });
  int32_t rb[3];
  std::vector<int32_t> active;
  active.reserve(prog.size());
  for (int32_t b : prog)
    if (!is_const(b)) active.push_back(b);
  tr.mark("  first appearance");
  st->order.clear();
  st->order.reserve(active.size());
  if (p->ordering == MBA_ORDER_POINTS_FIRST) {
    user_points_first(active, ng + nc, st);
  } else {
    // Vertex degrees of the Hessian graph without listing its edges: per
    // point (in parallel, over its observations) the distinct non-constant
    // cameras and groups it meets (camera-point and group-point edges),
    // plus one camera-group edge per non-constant camera in a non-constant
    // group.
    std::vector<int32_t> deg(nb, 0);
    std::vector<int64_t> pstart((size_t)np + 1, 0);
    Buf<int32_t> pobs(no);
    {
      // Per-thread histograms over observation ranges, then each thread
      // fills its own slots (observations of a point in index order).
      const int nt = host_threads(no);
      const int64_t step = (no + nt - 1) / std::max(1, nt);
      std::vector<std::vector<int64_t>> h(nt, std::vector<int64_t>((size_t)np, 0));
      run_threads(nt, [&](int t) {
        const int64_t e = std::min<int64_t>(no, (t + 1) * step);
        for (int64_t i = t * step; i < e; ++i) h[t][p->point_index[i]]++;
      });
      int64_t run = 0;
      for (int32_t q = 0; q < np; ++q) {
        pstart[q] = run;
        for (int t = 0; t < nt; ++t) {
          const int64_t k = h[t][q];
          h[t][q] = run;
          run += k;
        }
      }
      pstart[np] = run;
      run_threads(nt, [&](int t) {
        const int64_t e = std::min<int64_t>(no, (t + 1) * step);
        for (int64_t i = t * step; i < e; ++i) pobs[h[t][p->point_index[i]]++] = (int32_t)i;
      });
    }
    // Camera and group degrees are counted per thread, then added.
    parallel_for(np, [&](int64_t b, int64_t e) {
      std::vector<int32_t> cams, grps, dloc((size_t)ng + nc, 0);
      for (int64_t q = b; q < e; ++q) {
        cams.clear();
        grps.clear();
        for (int64_t k = pstart[q]; k < pstart[q + 1]; ++k) {
          const int64_t i = pobs[k];
          const int32_t cv = ng + p->camera_index[i], gv = p->intrinsic_index[i];
          if (!is_const(cv) && std::find(cams.begin(), cams.end(), cv) == cams.end()) cams.push_back(cv);
          if (!is_const(gv) && std::find(grps.begin(), grps.end(), gv) == grps.end()) grps.push_back(gv);
        }
        deg[ng + nc + q] = (int32_t)(cams.size() + grps.size());
        for (int32_t v : cams) dloc[v]++;
        for (int32_t v : grps) dloc[v]++;
      }
      for (int32_t v = 0; v < ng + nc; ++v)
        if (dloc[v]) __atomic_fetch_add(&deg[v], dloc[v], __ATOMIC_RELAXED);
    }, 1 << 12);
    for (int32_t c = 0; c < nc; ++c) {
      const int32_t g = st->cam_grp[c];
      if (g >= 0 && !is_const(ng + c) && !is_const(g)) {
        deg[ng + c]++;
        deg[g]++;
      }
    }
    tr.mark("  degrees");
    int32_t maxdeg = 0, max_pt = 0, min_other = std::numeric_limits<int32_t>::max();
    for (int32_t b : active) {
      maxdeg = std::max(maxdeg, deg[b]);
      if (b >= ng + nc) max_pt = std::max(max_pt, deg[b]);
      else min_other = std::min(min_other, deg[b]);
    }
    std::vector<int64_t> cnt((size_t)maxdeg + 2, 0);
    for (int32_t b : active) cnt[deg[b] + 1]++;
    for (int32_t d = 0; d <= maxdeg; ++d) cnt[d + 1] += cnt[d];
    std::vector<int32_t> queue(active.size());
    for (int32_t b : active) queue[cnt[deg[b]]++] = b;
    if (max_pt < min_other) {
      // Every point precedes every camera and group in the queue, and only
      // cameras and groups neighbour a point: the greedy set is the points
      // in queue order.
      for (int32_t b : queue)
        if (b >= ng + nc) st->order.push_back(b);
      st->num_e = (int32_t)st->order.size();
      for (int32_t b : queue)
        if (b < ng + nc) st->order.push_back(b);
    } else {
      // Greedy independent set over the explicit adjacency.
      std::vector<int64_t> keys;
      keys.reserve(3 * (size_t)no);
      for (int64_t i = 0; i < no; ++i) {
        const int n = residual_blocks(p, i, rb);
        for (int a = 0; a < n; ++a) {
          if (is_const(rb[a])) continue;
          for (int bb = a + 1; bb < n; ++bb) {
            if (is_const(rb[bb]) || rb[a] == rb[bb]) continue;
            const int32_t lo = std::min(rb[a], rb[bb]), hi = std::max(rb[a], rb[bb]);
            keys.push_back((int64_t)lo * nb + hi);
          }
        }
      }
      std::sort(keys.begin(), keys.end());
      keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
      std::vector<int64_t> adj_start((size_t)nb + 1, 0);
      for (int64_t k : keys) {
        adj_start[k / nb + 1]++;
        adj_start[k % nb + 1]++;
      }
      for (int32_t v = 0; v < nb; ++v) adj_start[v + 1] += adj_start[v];
      std::vector<int32_t> adj((size_t)adj_start[nb]);
      {
        std::vector<int64_t> fill(adj_start.begin(), adj_start.end() - 1);
        for (int64_t k : keys) {
          const int32_t a = (int32_t)(k / nb), bb = (int32_t)(k % nb);
          adj[fill[a]++] = bb;
          adj[fill[bb]++] = a;
        }
      }
      std::vector<uint8_t> color(nb, 0);
      for (int32_t b : queue) {
        if (color[b] != 0) continue;
        st->order.push_back(b);
        color[b] = 2;
        for (int64_t k = adj_start[b]; k < adj_start[b + 1]; ++k) color[adj[k]] = 1;
      }
      st->num_e = (int32_t)st->order.size();
      for (int32_t b : queue)
        if (color[b] != 2) st->order.push_back(b);
    }
  }
  tr.mark("  independent set");
  std::vector<int32_t> col(nb, -1);
  for (size_t i = 0; i < st->order.size(); ++i) col[st->order[i]] = (int32_t)i;
  // Residual order: bucket = the smallest column block among the residual's
  // non-constant blocks, buckets in column order, each filled back to
  // front (as the packed layout, in parallel).
  const int32_t ne = st->num_e;
  Buf<int32_t> bucket(no);
  parallel_for(no, [&](int64_t b, int64_t e) {
    int32_t r3[3];
    for (int64_t i = b; i < e; ++i) {
      const int n = residual_blocks(p, i, r3);
      int32_t m = ne;
      for (int k = 0; k < n; ++k)
        if (!is_const(r3[k])) m = std::min(m, col[r3[k]]);
      bucket[i] = m;
    }
  });
  std::vector<int64_t> bstart;
  order_residuals(bucket, no, ne, &bstart, &st->res_order);
  tr.mark("  residual order");
  int32_t n_act_pts = 0;
  for (int32_t b : active) n_act_pts += b >= ng + nc;
  bool ok = st->num_e == n_act_pts;
  for (int32_t i = 0; ok && i < st->num_e; ++i) ok = st->order[i] >= ng + nc;
  st->standard = ok;
  if (!ok) return MBA_OK;
  st->e_to_pt.resize(st->num_e);
  for (int32_t i = 0; i < st->num_e; ++i) st->e_to_pt[i] = st->order[i] - ng - nc;
  const int32_t nf = (int32_t)st->order.size() - st->num_e;
  st->f_to_cam.resize(nf);
  st->cam_to_f.assign((size_t)nc + ng, -1);
  for (int32_t i = 0; i < nf; ++i) {
    const int32_t v = st->order[st->num_e + i];
    st->f_to_cam[i] = v < ng ? nc + v : v - ng;
    st->cam_to_f[st->f_to_cam[i]] = i;
  }
  // Every row's bucket is its point's E block (points are never constant).
  st->pt_start.assign((size_t)st->num_e + 1, 0);
  st->max_track = 0;
  for (int32_t e = 0; e <= st->num_e; ++e) st->pt_start[e] = (int32_t)bstart[e];
  for (int32_t e = 0; e < st->num_e; ++e)
    st->max_track = std::max(st->max_track, st->pt_start[e + 1] - st->pt_start[e]);
  return MBA_OK;
}

// Upper-triangle block pairs of the reduced camera matrix
// (SparseSchurComplementSolver::InitStorage, schur_complement_solver.cc:246-313).
std::vector<int64_t> block_pairs(const mba_problem* p, const Structure& st) {
  const int32_t nf = (int32_t)st.order.size() - st.num_e;
  std::vector<int32_t> col((size_t)st.ng + st.nc + st.np, -1);
  for (size_t i = 0; i < st.order.size(); ++i) col[st.order[i]] = (int32_t)i;
  std::vector<int64_t> keys;
  keys.reserve(nf);
  for (int32_t i = 0; i < nf; ++i) keys.push_back((int64_t)i * nf + i);
  auto cells = [&](int64_t obs, int32_t* out) {
    int32_t rb[3];
    const int nrb = residual_blocks(p, obs, rb);
    int n = 0;
    for (int k = 0; k < nrb; ++k)
      if (!block_constant(p, rb[k])) out[n++] = col[rb[k]];
    std::sort(out, out + n);
    return n;
  };
  int64_t r = 0;
  const int64_t nr = (int64_t)st.res_order.size();
  std::vector<int32_t> fb;
  int32_t c2[3];
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
  // camera-major chunks; chunk_f is the chunk's F block (packed) or camera
  // (grouped); an F block's partial rows are [f_lo, f_hi)
  int32_t *cm_e = nullptr, *chunk_f = nullptr, *chunk_begin = nullptr, *chunk_end = nullptr;
  int32_t *f_lo = nullptr, *f_hi = nullptr, *f_to_cam = nullptr, *cam_to_f = nullptr;
  T* cm_obs = nullptr;
  // kernel-precision parameters
  T *ctab = nullptr, *pscale = nullptr, *cscale = nullptr;
  // blocks (V in FP64 in either precision; Vinv64, the FP64 inverse of the
  // damped V, for FP32 kernels: the FP64 kernels' Vinv is already FP64)
  double *V = nullptr, *Vinv64 = nullptr;
  T *gp = nullptr, *Vinv = nullptr, *z = nullptr, *pk = nullptr;
  T *U = nullptr, *gc = nullptr, *rhs = nullptr, *Minv = nullptr, *dtab = nullptr, *partial = nullptr;
  // PCG / step
  T *x = nullptr, *r = nullptr, *zc = nullptr, *p = nullptr, *q = nullptr, *tmp = nullptr;
  T *step_p = nullptr, *step_c = nullptr;
  // explicit S
  T *S = nullptr, *L = nullptr, *sv = nullptr, *WY = nullptr, *Linv = nullptr, *cross = nullptr, *slab = nullptr;
  // Inverses of the factor's 64 x 64 diagonal blocks and their transposes
  // (one per block), and the triangular solves' two work vectors.
  T *LinvT = nullptr, *ty = nullptr;
  // Triangular solves: n values as epoch words (two per FP64 value), then
  // the failed-epoch word; epoch of the last launch (host side).
  uint64_t* flow = nullptr;
  unsigned flow_epoch = 0;
  T* s_scratch = nullptr;
  // SCHUR_JACOBI preconditioner terms (IMPLICIT_SCHUR_PCG only)
  T* partial_sj = nullptr;
  double* Tsj = nullptr;
  // Packed layout with a loss function: each row's loss weight sqrt(rho')
  // as the last Jacobian pass used it, in row order (point pass) and
  // camera-major order (camera pass); the matvecs and the SCHUR_JACOBI pass
  // read them instead of re-deriving them from kernel-precision residuals.
  // Mixed single problems keep a second slot (wrow1, wcm1), written with
  // the FP64 residuals of a trial state, as rres / cres are.
  T *wrow = nullptr, *wcm = nullptr, *wrow1 = nullptr, *wcm1 = nullptr;
  // FP32 packed layout: per camera the left Jacobian G = J_l(aa) (9) that
  // takes the camera-major kernels' local rotation basis to the
  // parameters.
  T* gtab = nullptr;
};

constexpr size_t kBigChunk = (size_t)8 << 20;  // pinned staging buffers for large uploads (two)
constexpr size_t kStageMin = (size_t)1 << 20;  // smaller uploads go directly
constexpr int kSumSlots = 8;  // per-window sums read back per LM iteration
// Sum slots: model cost change, |step|^2, trial cost, |x|^2, gradient max
// and |g|^2 (two slots), a slot for sums read at once.
constexpr int kSlotModel = 0, kSlotPlus = 1, kSlotCost = 2, kSlotXnorm = 3, kSlotGrad = 4, kSlotSync = 7;
constexpr int kStage = 8;     // pinned upload slots in rotation
constexpr int kLook = 2;      // PCG iterations the host may run ahead (each queued
                              // past the last costs ~22 us of early-exit launches)
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace
}  // namespace mba

using namespace mba;

// Grid and dynamic LDS bytes of a persistent LDS-staged launch (grid 0: not
// used for this context).
struct LdsLaunch {
  int grid = 0;
  size_t bytes = 0;
};

struct mba_context {
  mba_options opt;
  // Independent problems ("windows") held together; a single problem is
  // W = 1.  Windows are concatenated: cameras, points, E and F blocks of
  // window w start at the offsets below (W + 1 prefix sums).
  int32_t W = 1;
  std::vector<Structure> wst;
  std::vector<int32_t> cam_off, pt_off, e_off, f_off;
  std::vector<int32_t> f_to_cam, cam_to_f;  // F block <-> global camera (slot block)
  int32_t max_track = 0;
  // Grouped layout (single problem): intrinsic groups, the FP64 slot blocks
  // (cameras, then groups), each camera's group and group F block.
  bool grouped = false;
  int32_t ng = 0, nslot = 0;
  std::vector<int32_t> cam_grp, cam_gf;
  int32_t *d_cam_grp = nullptr, *d_cam_gf = nullptr;
  // Launch slot -> camera-major chunk of the packed kernels (null: slot =
  // chunk); nslots launch slots (nchunks without the map).
  int32_t* d_slot_chunk = nullptr;
  int32_t nslots = 0;
  // Mixed precision, one packed problem: residuals of the FP64 cost at the
  // current state (slot rcur) and at the trial state (1 - rcur), rounded to
  // FP32, in row order (rres) and camera-major order (cres, through row_cm;
  // -1 for rows of constant cameras).  Inside the LM loop (rvalid) the FP32
  // passes use them instead of their own pred - obs, whose rounding (~1e-5
  // relative) would put noise into the gradient.
  bool rfix = false, rvalid = false;
  int rcur = 0;
  float *rres[2] = {}, *cres[2] = {};
  int32_t* row_cm = nullptr;
  // 9 per F block (NULL when every slot is a free parameter): fmask = 1 for
  // a slot that takes steps, xmask = 1 for a slot of the parameter block
  // (counted in |x|; a SubsetParameterization entry is held but counted).
  uint8_t *fmask = nullptr, *xmask = nullptr;
  // Group F blocks: their chunk partial rows [g_lo, g_hi) are summed into
  // row g_out (after the 2 nchunks chunk rows) before any finalize.
  int32_t ngf = 0;
  int32_t *g_lo = nullptr, *g_hi = nullptr, *g_out = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  Pool pool;
  int32_t nc = 0, np = 0, ne = 0, nf = 0, nrows = 0, nchunks = 0, ncm = 0;
  // Launch shapes in use (options or the tuned defaults): rows per lane of
  // the point-major kernels, rows per camera-major chunk (0: by size).
  int32_t rows_per_lane = 4, chunk_rows = 0;
  // FP64 state and LM data
  double *cams64 = nullptr, *cams64_t = nullptr, *pts64 = nullptr, *pts64_t = nullptr;
  double *ctab64 = nullptr, *row_obs64 = nullptr, *pscale64 = nullptr, *cscale64 = nullptr;
  double *diag_c = nullptr, *red = nullptr, *red2 = nullptr, *stepd = nullptr;
  double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr, *gemv_part = nullptr;
  double* dtab64 = nullptr;  // FP64 direction table of the step (mixed: the model cost change)
  Scalars* sc = nullptr;        // device
  PcgState* pcg = nullptr;      // device, one per window
  int32_t* flag = nullptr;      // device, one per window
  int32_t* info = nullptr;      // device
  int32_t* all_done = nullptr;  // device: every window's PCG has stopped
  unsigned int* counter = nullptr;  // device: last-workgroup election (2)
  // Window descriptors (device): window of each device E / F block (W > 1
  // only), 1 / radius per window, per-window accumulators, accept flags.
  int32_t *of_e = nullptr, *of_f = nullptr;
  double *inv_r = nullptr, *wacc = nullptr, *wacc2 = nullptr;
  uint8_t* accept = nullptr;
  uint8_t* wactive = nullptr;  // windows still iterating
  // Explicit Schur (single problem): the blocks of S that points couple and
  // their observation-row pairs, sorted by block.
  // Grouped layout: rows are virtual rows (camera part r, group part
  // nrows + r); s_blk_cross marks camera / group block pairs.
  int32_t *s_blk_fa = nullptr, *s_blk_fb = nullptr, *s_pair_a = nullptr, *s_pair_b = nullptr;
  int32_t* s_blk_cross = nullptr;
  // Work items of the block assembly (a block, or a slice of a block with
  // more than kSplitPairs pairs whose sums go to scratch slots).
  int32_t *s_item_blk = nullptr, *s_item_slot = nullptr;
  int64_t *s_item_p0 = nullptr, *s_item_p1 = nullptr;
  int32_t *s_split_blk = nullptr, *s_split_lo = nullptr, *s_split_hi = nullptr;
  int64_t s_nitems = 0, s_nsplit = 0, s_nslots = 0;
  int64_t s_nblk = 0, s_npairs = 0;
  // Grouped layout: merged group rows (virtual rows 2 nrows + m), each the
  // sum of one point's group parts (virtual rows nrows + r) whose cameras
  // share an intrinsic group: rows s_m_row[s_m_start[m] .. s_m_start[m+1]).
  int32_t *s_m_start = nullptr, *s_m_row = nullptr;
  int64_t s_nmerged = 0;
  // Batched explicit solve (W > 1): window w's dense S at offset d_woff[w]
  // (s_wtotal entries in all), window F-block offsets on the device.
  int64_t* d_woff = nullptr;
  int32_t* d_f_off = nullptr;
  // Batched PCG in one persistent launch (k_pcg_window): each window's
  // device E blocks (CSR over windows), its camera range, and each F
  // block's camera-major rows.
  bool pcg_window = false;
  int32_t *d_win_e_lo = nullptr, *d_win_e = nullptr, *d_cam_off = nullptr, *d_win_order = nullptr;
  int32_t *d_fcm_lo = nullptr, *d_fcm_hi = nullptr;
  int64_t s_wtotal = 0;
  // Host copies of what was last uploaded.
  std::vector<double> h_inv_r;
  std::vector<uint8_t> h_active;
  Dev<float> df;
  Dev<double> dd;
  bool mixed = false;
  bool scale_ready = false;
  bool timing = false;
  hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool timed_explicit = false;
  // Pinned host memory (hipHostMalloc) for the LM loop's readbacks and
  // small uploads: kSumSlots x W sums, 2 x W flags, W PCG states, kStage
  // upload slots of 8 W bytes.  An upload slot is rewritten only after the
  // event recorded behind its copy has completed.  The PCG "all stopped"
  // word is written by the device into coherent pinned memory (d_done is
  // its device address).
  char* pin = nullptr;
  double* h_sums = nullptr;
  int32_t* h_flags = nullptr;
  PcgState* h_pcgs = nullptr;
  int32_t* pin_done = nullptr;
  volatile int32_t* h_done = nullptr;
  int32_t* d_done = nullptr;
  char* h_stage = nullptr;
  size_t stage_bytes = 0;
  int stage_next = 0;
  hipEvent_t stage_ev[kStage] = {};
  // Captured PCG iterations, [row-weight slot][residual reset]; made on
  // first use.
  hipGraphExec_t pcg_graph[2][2] = {};
  // FP32 point-major pass and matvec with LDS-staged camera records (grid
  // > 0): grids and dynamic LDS bytes.
  LdsLaunch lds_pass, lds_mv;
  // Polls per cross-block wait of the triangular solves; MBA_FLOW_SPIN_LIMIT
  // in the environment at context creation overrides it (0: every wait
  // gives up at once, which tests the failure path without a hang).
  unsigned flow_spin = kFlowSpinMax;
  hipEvent_t tev[3] = {};          // LM phase timing
  // Dense Cholesky look-ahead: the trailing tiles past the next block column
  // run on stream2 beside the next diagonal factor; chol_ev[0] orders them
  // after the panel, chol_ev[1] the next panel after them (rings of 4).
  hipStream_t stream2 = nullptr;
  hipEvent_t chol_ev[2][4] = {};
  // Large host -> device uploads: two pinned kBigChunk buffers in turn.
  char* big = nullptr;
  hipEvent_t big_ev[2] = {};
  int big_next = 0;
  // Trust-region problem dumps (single problem): iterations and directory
  // copied from the options; `src` is the problem mba_solve was given.
  std::vector<int32_t> dump_iters;
  std::string dump_dir;
  const mba_problem* src = nullptr;
  // Device point order (E blocks of all windows stably grouped by lane
  // class) and the point-major launch layout.
  std::vector<int32_t> dev_e_to_pt;  // device E block -> global point
  Seg seg;
  int32_t seg_blocks = 0;
};

// LM state at a dumped step (device order, FP64): parameters, Jacobi
// scales, clamped squared column norms, the step.
struct LsqpState {
  std::vector<double> cams, pts, pscale, cscale, diag_p, diag_c, step_p, step_c;
  double radius = 0.0;
};
mba_status write_lsqp(mba_context* c, int k, const LsqpState& S);

namespace mba {
namespace {

template <typename T> Dev<T>& dev(mba_context* c);
template <> Dev<float>& dev<float>(mba_context* c) { return c->df; }
template <> Dev<double>& dev<double>(mba_context* c) { return c->dd; }

// Host -> device copy on the context stream.  A large copy goes through the
// two pinned staging buffers in turn: filled by a parallel memcpy, sent by
// an async DMA, refilled once the event behind its last copy has passed.
// The source may be released on return.
mba_status copy_to_device(mba_context* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return MBA_OK;
  if (bytes < kStageMin || !c->big) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return MBA_OK;
  }
  for (size_t off = 0; off < bytes; off += kBigChunk) {
    const size_t n = std::min(kBigChunk, bytes - off);
    const int i = c->big_next;
    c->big_next ^= 1;
    HIP_TRY(hipEventSynchronize(c->big_ev[i]));
    char* h = c->big + (size_t)i * kBigChunk;
    const char* from = (const char*)src + off;
    parallel_for((int64_t)n, [&](int64_t b, int64_t e) { std::memcpy(h + b, from + b, (size_t)(e - b)); }, 1 << 20);
    HIP_TRY(hipMemcpyAsync((char*)dst + off, h, n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipEventRecord(c->big_ev[i], c->stream));
  }
  return MBA_OK;
}

// Copy a host vector to a new device allocation.
template <typename X, typename A>
mba_status upload(mba_context* c, X** out, const std::vector<X, A>& v) {
  MBA_TRY(c->pool.alloc(out, v.size()));
  return copy_to_device(c, *out, v.data(), v.size() * sizeof(X));
}

template <typename T>
mba_status alloc_precision(mba_context* c, const std::vector<int32_t>& dev_pt_start,
                           const Buf<int32_t>& row_cam, const Buf<int32_t>& cm_src,
                           const std::vector<int32_t>& chunk_f, const std::vector<int32_t>& chunk_begin,
                           const std::vector<int32_t>& chunk_end, const std::vector<int32_t>& f_lo,
                           const std::vector<int32_t>& f_hi) {
  Dev<T>& d = dev<T>(c);
  const hipStream_t s = c->stream;
  MBA_TRY(upload(c, &d.row_cam, row_cam));
  MBA_TRY(upload(c, &d.pt_start, dev_pt_start));
  // Kernel-precision observations and the camera-major copies are built
  // on the device from the FP64 point-major rows.
  if constexpr (std::is_same<T, double>::value) {
    d.row_obs = c->row_obs64;
  } else {
    MBA_TRY(c->pool.alloc(&d.row_obs, 2 * (size_t)c->nrows));
    hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(2 * (int64_t)c->nrows)), dim3(kBlock), 0, s,
                       2 * (int64_t)c->nrows, (const double*)c->row_obs64, d.row_obs);
  }
  int32_t *src = nullptr, *row_e = nullptr;
  MBA_TRY(upload(c, &src, cm_src));
  if (c->row_cm) {
    HIP_TRY(hipMemsetAsync(c->row_cm, 0xff, sizeof(int32_t) * (size_t)c->nrows, s));
    if (c->ncm > 0)
      hipLaunchKernelGGL(k_slot_of_row, dim3(nblocks(c->ncm)), dim3(kBlock), 0, s, c->ncm, (const int32_t*)src,
                         c->row_cm);
  }
  MBA_TRY(c->pool.alloc(&row_e, (size_t)c->nrows));
  MBA_TRY(c->pool.alloc(&d.cm_e, (size_t)c->ncm));
  MBA_TRY(c->pool.alloc(&d.cm_obs, 2 * (size_t)c->ncm));
  hipLaunchKernelGGL(k_row_e, dim3(nblocks(c->ne)), dim3(kBlock), 0, s, c->ne, (const int32_t*)d.pt_start, row_e);
  hipLaunchKernelGGL(k_gather_cm<T>, dim3(nblocks(c->ncm)), dim3(kBlock), 0, s, c->ncm, (const int32_t*)src,
                     (const int32_t*)row_e, (const double*)c->row_obs64, d.cm_e, d.cm_obs);
  MBA_TRY(upload(c, &d.chunk_f, chunk_f));
  MBA_TRY(upload(c, &d.chunk_begin, chunk_begin));
  MBA_TRY(upload(c, &d.chunk_end, chunk_end));
  MBA_TRY(upload(c, &d.f_lo, f_lo));
  MBA_TRY(upload(c, &d.f_hi, f_hi));
  MBA_TRY(upload(c, &d.f_to_cam, c->f_to_cam));
  MBA_TRY(upload(c, &d.cam_to_f, c->cam_to_f));
  const size_t ne = c->ne, nf = c->nf, nc = c->nc;
  MBA_TRY(c->pool.alloc(&d.ctab, kCamTab * nc));
  MBA_TRY(c->pool.alloc(&d.pscale, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.cscale, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.V, 6 * ne));
  MBA_TRY(c->pool.alloc(&d.gp, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.Vinv, 6 * ne));
  if constexpr (std::is_same<T, float>::value) MBA_TRY(c->pool.alloc(&d.Vinv64, 6 * std::max<size_t>(1, ne)));
  MBA_TRY(c->pool.alloc(&d.z, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.pk, kPk * ne));
  MBA_TRY(c->pool.alloc(&d.U, 45 * nf));
  MBA_TRY(c->pool.alloc(&d.gc, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.rhs, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.Minv, 81 * nf));
  MBA_TRY(c->pool.alloc(&d.dtab, (c->grouped ? kDirTabG : kDirTab) * nc));
  MBA_TRY(c->pool.alloc(&d.partial, (size_t)kPartial * (c->grouped ? 2 * (size_t)c->nchunks + c->ngf
                                                                    : (size_t)c->nchunks)));
  MBA_TRY(c->pool.alloc(&d.x, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.r, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.zc, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.p, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.q, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.tmp, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.step_p, 3 * ne));
  MBA_TRY(c->pool.alloc(&d.step_c, 9 * nf));
  MBA_TRY(c->pool.alloc(&d.sv, 9 * nf));
  if (c->opt.preconditioner == MBA_PRECOND_SCHUR_JACOBI && c->opt.linear_solver == MBA_IMPLICIT_SCHUR_PCG) {
    MBA_TRY(c->pool.alloc(&d.partial_sj, (size_t)kPartial * std::max(1, c->nchunks)));
    MBA_TRY(c->pool.alloc(&d.Tsj, 45 * std::max<size_t>(1, nf)));
  }
  if (c->opt.huber_a > 0.0 && !c->grouped && c->nrows > 0) {
    MBA_TRY(c->pool.alloc(&d.wrow, (size_t)c->nrows));
    MBA_TRY(c->pool.alloc(&d.wcm, (size_t)std::max(1, c->ncm)));
    if (c->rfix) {
      MBA_TRY(c->pool.alloc(&d.wrow1, (size_t)c->nrows));
      MBA_TRY(c->pool.alloc(&d.wcm1, (size_t)std::max(1, c->ncm)));
    }
  }
  if constexpr (std::is_same<T, float>::value)
    if (!c->grouped) MBA_TRY(c->pool.alloc(&d.gtab, 9 * std::max<size_t>(1, nc)));
  return MBA_OK;
}

// Rows per camera-major chunk: aim for >= ~2048 workgroups on large
// problems, 64..2048 rows each.
// XCDs of the part (workgroups are dealt to them round robin) and the
// camera-major row count from which the chunks are split by point range.
constexpr int kXcd = 8;
constexpr int64_t kXcdSplitMin = 1 << 18;
constexpr int kXcdChunkRows = 768;

int chunk_rows_for(int64_t nrows) {
  int64_t c = nrows / 2048;
  c = std::max<int64_t>(64, std::min<int64_t>(2048, c));
  return (int)((c + 63) / 64 * 64);
}

void context_free(mba_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 8; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  for (hipEvent_t e : c->stage_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& row : c->pcg_graph)
    for (hipGraphExec_t g : row)
      if (g) (void)hipGraphExecDestroy(g);
  for (hipEvent_t e : c->tev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto& ring : c->chol_ev)
    for (hipEvent_t e : ring)
      if (e) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->pin) (void)hipHostFree(c->pin);
  if (c->pin_done) (void)hipHostFree(c->pin_done);
  for (hipEvent_t e : c->big_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->big) (void)hipHostFree(c->big);
  c->pool.release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

struct ContextGuard {
  mba_context* c;
  ~ContextGuard() { context_free(c); }
};

std::string window_prefix(int32_t W, int32_t w) {
  return W > 1 ? "window " + std::to_string(w) + ": " : std::string();
}

// Blocks (fa, fb), fa >= fb, of the reduced camera matrix that points couple,
// and for each the pairs (a, b) of point-major rows of a common point with
// F blocks fa and fb (both orders when fa == fb), sorted by block and, in a
// block, by point and row.  This is the block structure the reference sizes
// its reduced matrix with (schur_complement_solver.cc:246-313); here it also
// fixes who computes each block.
mba_status build_s_pairs(mba_context* c, const std::vector<int32_t>& pt_start,
                         const Buf<int32_t>& row_cam) {
  const int32_t nf = c->nf, ne = c->ne;
  // Virtual rows: one per observation (packed).  Grouped layout: the camera
  // part of row r at r, its group part at nrows + r, and one merged row per
  // (point, group F block) at 2 nrows + m summing the point's group parts
  // in that group.  W and Y = W V^-1 are linear in the rows, so the group
  // blocks of S pair merged rows: one pair per point for a group-group
  // block and one per camera row for a camera-group block, instead of one
  // per pair of the point's rows.
  const int32_t nrows = c->nrows;
  const bool grouped = c->grouped;
  if ((grouped ? 3 : 1) * (int64_t)nrows > 0x7fffffffLL) return fail(MBA_ERR_UNSUPPORTED, "explicit Schur: too many rows");
  std::vector<int32_t> row_f((size_t)(grouped ? 2 : 1) * nrows);
  parallel_for(c->nrows, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) {
      row_f[r] = c->cam_to_f[row_cam[r]];
      if (grouped) row_f[nrows + r] = c->cam_gf[row_cam[r]];
    }
  });
  std::vector<int32_t> e_m0, m_start, m_row;
  if (grouped) {
    e_m0.assign((size_t)ne + 1, 0);
    m_start.push_back(0);
    std::vector<int32_t> fm;  // F block of each merged row
    for (int32_t e = 0; e < ne; ++e) {
      const int32_t m0 = (int32_t)fm.size();
      e_m0[e] = m0;
      for (int32_t r = pt_start[e]; r < pt_start[e + 1]; ++r) {
        const int32_t gf = row_f[(size_t)nrows + r];
        if (gf < 0) continue;
        bool seen = false;
        for (int32_t m = m0; m < (int32_t)fm.size(); ++m) seen = seen || fm[m] == gf;
        if (!seen) fm.push_back(gf);
      }
      for (int32_t m = m0; m < (int32_t)fm.size(); ++m) {
        for (int32_t r = pt_start[e]; r < pt_start[e + 1]; ++r)
          if (row_f[(size_t)nrows + r] == fm[m]) m_row.push_back(r);
        m_start.push_back((int32_t)m_row.size());
      }
    }
    e_m0[ne] = (int32_t)fm.size();
    c->s_nmerged = (int64_t)fm.size();
    row_f.insert(row_f.end(), fm.begin(), fm.end());
    MBA_TRY(upload(c, &c->s_m_start, m_start));
    MBA_TRY(upload(c, &c->s_m_row, m_row));
  }
  const int nt = host_threads(ne, 4096);
  const int64_t step = ((int64_t)ne + nt - 1) / nt;
  // Pairs per fa, per thread (point ranges), then bucket offsets.
  std::vector<std::vector<int64_t>> cnt(nt, std::vector<int64_t>((size_t)nf, 0));
  auto for_pairs = [&](int t, auto&& fn) {
    const int64_t end = std::min<int64_t>(ne, (t + 1) * step);
    std::vector<int32_t> vr;
    for (int64_t e = t * step; e < end; ++e) {
      vr.clear();
      for (int32_t r = pt_start[e]; r < pt_start[e + 1]; ++r)
        if (row_f[r] >= 0) vr.push_back(r);
      if (grouped)
        for (int32_t m = e_m0[e]; m < e_m0[e + 1]; ++m) vr.push_back(2 * nrows + m);
      for (const int32_t a : vr) {
        const int32_t fa = row_f[a];
        for (const int32_t b : vr) {
          const int32_t fb = row_f[b];
          if (fb <= fa) fn(fa, a, b);
        }
      }
    }
  };
  run_threads(nt, [&](int t) { for_pairs(t, [&](int32_t fa, int32_t, int32_t) { cnt[t][fa]++; }); });
  std::vector<int64_t> fa_start((size_t)nf + 1, 0);
  int64_t run = 0;
  for (int32_t fa = 0; fa < nf; ++fa) {
    fa_start[fa] = run;
    for (int t = 0; t < nt; ++t) {
      const int64_t k = cnt[t][fa];
      cnt[t][fa] = run;
      run += k;
    }
  }
  fa_start[nf] = run;
  if (run > 0x7fffffffLL) return fail(MBA_ERR_UNSUPPORTED, "explicit Schur: more than 2^31-1 coupled row pairs");
  c->s_npairs = run;
  std::vector<int32_t> pa((size_t)run), pb((size_t)run);
  run_threads(nt, [&](int t) {
    for_pairs(t, [&](int32_t fa, int32_t a, int32_t b) {
      const int64_t k = cnt[t][fa]++;
      pa[k] = a;
      pb[k] = b;
    });
  });
  // Within each fa bucket: stable counting sort by fb, block list.
  std::vector<int32_t> sa((size_t)run), sb((size_t)run);
  std::vector<std::vector<int32_t>> bfb(nf);
  std::vector<std::vector<int64_t>> bcnt(nf);
  {
    std::atomic<int32_t> next{0};
    const int ntb = (int)std::min<int64_t>(std::max(1, host_threads(run, 1 << 14)), std::max(1, nf));
    run_threads(ntb, [&](int) {
      std::vector<int64_t> h((size_t)nf + 1, 0);
      for (int32_t fa = next.fetch_add(1); fa < nf; fa = next.fetch_add(1)) {
        const int64_t b0 = fa_start[fa], b1 = fa_start[fa + 1];
        for (int64_t k = b0; k < b1; ++k) h[row_f[pb[k]]]++;
        int64_t o = b0;
        for (int32_t fb = 0; fb <= fa; ++fb) {
          const int64_t m = h[fb];
          if (m == 0) continue;
          bfb[fa].push_back(fb);
          bcnt[fa].push_back(m);
          h[fb] = o;
          o += m;
        }
        for (int64_t k = b0; k < b1; ++k) {
          const int64_t dst = h[row_f[pb[k]]]++;
          sa[dst] = pa[k];
          sb[dst] = pb[k];
        }
        for (int32_t fb = 0; fb <= fa; ++fb) h[fb] = 0;
      }
    });
  }
  std::vector<int64_t> blk_start;
  std::vector<int32_t> blk_fa, blk_fb;
  int64_t o = 0;
  for (int32_t fa = 0; fa < nf; ++fa)
    for (size_t i = 0; i < bfb[fa].size(); ++i) {
      blk_start.push_back(o);
      blk_fa.push_back(fa);
      blk_fb.push_back(bfb[fa][i]);
      o += bcnt[fa][i];
    }
  blk_start.push_back(o);
  c->s_nblk = (int64_t)blk_fa.size();
  {
    std::vector<int32_t> iblk, islot, sblk, slo, shi;
    std::vector<int64_t> ip0, ip1;
    int32_t nslots = 0;
    for (int64_t b = 0; b < c->s_nblk; ++b) {
      const int64_t p0 = blk_start[b], p1 = blk_start[b + 1];
      if (p1 - p0 <= kSplitPairs) {
        iblk.push_back((int32_t)b);
        ip0.push_back(p0);
        ip1.push_back(p1);
        islot.push_back(-1);
        continue;
      }
      sblk.push_back((int32_t)b);
      slo.push_back(nslots);
      for (int64_t q = p0; q < p1; q += kSplitPairs) {
        iblk.push_back((int32_t)b);
        ip0.push_back(q);
        ip1.push_back(std::min(p1, q + kSplitPairs));
        islot.push_back(nslots++);
      }
      shi.push_back(nslots);
    }
    c->s_nitems = (int64_t)iblk.size();
    c->s_nsplit = (int64_t)sblk.size();
    c->s_nslots = nslots;
    MBA_TRY(upload(c, &c->s_item_blk, iblk));
    MBA_TRY(upload(c, &c->s_item_p0, ip0));
    MBA_TRY(upload(c, &c->s_item_p1, ip1));
    MBA_TRY(upload(c, &c->s_item_slot, islot));
    MBA_TRY(upload(c, &c->s_split_blk, sblk));
    MBA_TRY(upload(c, &c->s_split_lo, slo));
    MBA_TRY(upload(c, &c->s_split_hi, shi));
  }
  if (c->grouped) {
    // Camera / group pairs: 2 c (+1 when the group block is fa).
    std::vector<int32_t> cross(blk_fa.size(), -1);
    for (size_t k = 0; k < blk_fa.size(); ++k) {
      const int32_t ba = c->f_to_cam[blk_fa[k]], bb = c->f_to_cam[blk_fb[k]];
      if (ba < c->nc && bb >= c->nc && c->cam_gf[ba] == blk_fb[k]) cross[k] = 2 * ba;
      else if (bb < c->nc && ba >= c->nc && c->cam_gf[bb] == blk_fa[k]) cross[k] = 2 * bb + 1;
    }
    MBA_TRY(upload(c, &c->s_blk_cross, cross));
  }
  MBA_TRY(upload(c, &c->s_blk_fa, blk_fa));
  MBA_TRY(upload(c, &c->s_blk_fb, blk_fb));
  MBA_TRY(upload(c, &c->s_pair_a, sa));
  MBA_TRY(upload(c, &c->s_pair_b, sb));
  return hipStreamSynchronize(c->stream) == hipSuccess ? MBA_OK : fail(MBA_ERR_DEVICE, "pair upload failed");
}

// Preprocess W independent problems into one device layout.
mba_status context_create(const mba_problem* probs, int32_t W, const mba_options* o, mba_context** out) {
  if (!out) return fail(MBA_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!probs || W < 1) return fail(MBA_ERR_INVALID_ARGUMENT, "no problem given");
  mba_context* c = new mba_context();
  ContextGuard guard{c};
  if (o) c->opt = *o; else mba_default_options(&c->opt);
  c->W = W;
  c->device = c->opt.device;
  c->mixed = (c->opt.precision == MBA_PRECISION_MIXED);
  c->wst.resize(W);
  c->cam_off.assign((size_t)W + 1, 0);
  c->pt_off.assign((size_t)W + 1, 0);
  c->e_off.assign((size_t)W + 1, 0);
  c->f_off.assign((size_t)W + 1, 0);
  // Per-window preprocessing (the reference's ordering), windows in
  // parallel on host threads.
  std::vector<mba_status> wstat(W, MBA_OK);
  std::vector<std::string> werr(W);
  for (int32_t w = 0; w < W; ++w)
    if (probs[w].layout == MBA_LAYOUT_GROUPED && W > 1)
      return fail(MBA_ERR_UNSUPPORTED, "grouped-intrinsics problems are solved one context each");
  if (c->opt.preconditioner == MBA_PRECOND_CLUSTER_JACOBI || c->opt.preconditioner == MBA_PRECOND_CLUSTER_TRIDIAGONAL)
    return fail(MBA_ERR_UNSUPPORTED, std::string(c->opt.preconditioner == MBA_PRECOND_CLUSTER_JACOBI
                                                     ? "CLUSTER_JACOBI"
                                                     : "CLUSTER_TRIDIAGONAL") +
                                         " requires Solver::Options::sparse_linear_algebra_library_type to be "
                                         "SUITE_SPARSE");
  if (c->opt.preconditioner != MBA_PRECOND_JACOBI && c->opt.preconditioner != MBA_PRECOND_SCHUR_JACOBI)
    return fail(MBA_ERR_INVALID_ARGUMENT, "unknown preconditioner " + std::to_string(c->opt.preconditioner));
  if (c->opt.num_dump_iterations < 0 || (c->opt.num_dump_iterations > 0 && !c->opt.dump_iterations))
    return fail(MBA_ERR_INVALID_ARGUMENT, "bad dump_iterations");
  if (c->opt.point_rows_per_lane < 0 || c->opt.point_rows_per_lane > 64)
    return fail(MBA_ERR_INVALID_ARGUMENT, "point_rows_per_lane must be 0 (tuned) or 1..64");
  if (c->opt.camera_chunk_rows != 0 && (c->opt.camera_chunk_rows < 64 || c->opt.camera_chunk_rows > 8192))
    return fail(MBA_ERR_INVALID_ARGUMENT, "camera_chunk_rows must be 0 (tuned) or 64..8192");
  if (c->opt.num_dump_iterations > 0) {
    if (!c->opt.dump_directory || !c->opt.dump_directory[0])
      return fail(MBA_ERR_INVALID_ARGUMENT, "Solver::Options::trust_region_problem_dump_directory is empty.");
    if (W > 1) return fail(MBA_ERR_UNSUPPORTED, "trust-region problem dumps are written for a single problem");
    c->dump_iters.assign(c->opt.dump_iterations, c->opt.dump_iterations + c->opt.num_dump_iterations);
    c->dump_dir = c->opt.dump_directory;
  }
  c->opt.dump_iterations = nullptr;
  c->opt.dump_directory = nullptr;
  c->opt.num_dump_iterations = 0;
  const bool schur_jacobi =
      c->opt.preconditioner == MBA_PRECOND_SCHUR_JACOBI && c->opt.linear_solver == MBA_IMPLICIT_SCHUR_PCG;
  if (schur_jacobi && probs[0].layout == MBA_LAYOUT_GROUPED)
    return fail(MBA_ERR_UNSUPPORTED, "SCHUR_JACOBI is implemented for packed cameras");
  PrepTrace trace;
  auto prep = [&](int32_t w) {
    const mba_problem* p = &probs[w];
    Structure& st = c->wst[w];
    mba_status s = build_structure(p, &st);
    if (s != MBA_OK) {
      wstat[w] = s;
      werr[w] = g_last_error;  // thread-local to this worker
      return;
    }
    if (!st.standard) {
      wstat[w] = MBA_ERR_UNSUPPORTED;
      werr[w] = "the Schur ordering did not eliminate exactly the point blocks";
      return;
    }
    for (int64_t i = 0; i < 3 * (int64_t)p->num_points; ++i)
      if (!std::isfinite(p->points[i])) {
        wstat[w] = MBA_ERR_NOT_FINITE;
        werr[w] = "non-finite point";
        return;
      }
    const int cw = p->layout == MBA_LAYOUT_GROUPED ? 6 : 9;
    for (int64_t i = 0; i < cw * (int64_t)p->num_cameras; ++i)
      if (!std::isfinite(p->cameras[i])) {
        wstat[w] = MBA_ERR_NOT_FINITE;
        werr[w] = "non-finite camera";
        return;
      }
    if (p->layout == MBA_LAYOUT_GROUPED)
      for (int64_t i = 0; i < 6 * (int64_t)p->num_intrinsics; ++i)
        if (!std::isfinite(p->intrinsics[i])) {
          wstat[w] = MBA_ERR_NOT_FINITE;
          werr[w] = "non-finite intrinsics";
          return;
        }
  };
  if (W == 1) {
    prep(0);
  } else {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<unsigned>((unsigned)W, std::min(16u, hw));
    std::atomic<int32_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&]() {
        for (int32_t w = next.fetch_add(1); w < W; w = next.fetch_add(1)) prep(w);
      });
    for (std::thread& t : pool) t.join();
  }
  trace.mark("ordering (per window)");
  int64_t nrows = 0;
  for (int32_t w = 0; w < W; ++w) {
    if (wstat[w] != MBA_OK) return fail(wstat[w], window_prefix(W, w) + werr[w]);
    const Structure& st = c->wst[w];
    const int64_t nc_tot = (int64_t)c->cam_off[w] + st.nc, np_tot = (int64_t)c->pt_off[w] + st.np;
    nrows += st.nobs;
    if (nc_tot > 0x7fffffffLL || np_tot > 0x7fffffffLL || nrows > 0x7fffffffLL)
      return fail(MBA_ERR_UNSUPPORTED, "more than 2^31-1 cameras, points or observations in one call");
    c->cam_off[w + 1] = (int32_t)nc_tot;
    c->pt_off[w + 1] = (int32_t)np_tot;
    c->e_off[w + 1] = c->e_off[w] + st.num_e;
    c->f_off[w + 1] = c->f_off[w] + ((int32_t)st.order.size() - st.num_e);
    c->max_track = std::max(c->max_track, st.max_track);
  }
  c->nc = c->cam_off[W];
  c->np = c->pt_off[W];
  c->ne = c->e_off[W];
  c->nf = c->f_off[W];
  c->nrows = (int32_t)nrows;
  c->grouped = probs[0].layout == MBA_LAYOUT_GROUPED;
  {
    const TunedShape ts = tuned_shape(nrows);
    c->rows_per_lane = c->opt.point_rows_per_lane > 0 ? c->opt.point_rows_per_lane : ts.point_rows_per_lane;
    c->chunk_rows = c->opt.camera_chunk_rows > 0 ? (c->opt.camera_chunk_rows + 63) / 64 * 64 : ts.camera_chunk_rows;
  }
  // An explicit S assembled in FP32 loses the weakly determined directions
  // of a grouped problem (shared intrinsics with nearly collinear k1..k3
  // columns) to cancellation in U - W Vinv W', which FP64 refinement against
  // the same S cannot restore: grouped explicit solves run in FP64.
  if (c->grouped && c->opt.linear_solver == MBA_EXPLICIT_SCHUR) c->mixed = false;
  // Batched explicit solves (several windows) factor each window's S in
  // FP64 by column steps; windows of more than kWinMaxF F blocks are
  // solved one context each by the caller.
  if (W > 1 && c->opt.linear_solver == MBA_EXPLICIT_SCHUR) {
    c->mixed = false;
    for (int32_t w = 0; w < W; ++w)
      if (c->f_off[w + 1] - c->f_off[w] > kWinMaxF)
        return fail(MBA_ERR_UNSUPPORTED, window_prefix(W, w) + "more than " + std::to_string(kWinMaxF) +
                                             " camera blocks for a batched explicit solve");
  }
  c->ng = c->grouped ? c->wst[0].ng : 0;
  c->nslot = c->nc + c->ng;
  // Global E blocks in window order, each window in its reference E order.
  Buf<int32_t> ge_win(c->ne), ge_local(c->ne);
  parallel_for(W, [&](int64_t b, int64_t end) {
    for (int64_t w = b; w < end; ++w)
      for (int32_t e = 0; e < c->wst[w].num_e; ++e) {
        ge_win[c->e_off[w] + e] = (int32_t)w;
        ge_local[c->e_off[w] + e] = e;
      }
  }, 16);
  auto track_len = [&](int32_t ge) {
    const Structure& st = c->wst[ge_win[ge]];
    const int32_t e = ge_local[ge];
    return st.pt_start[e + 1] - st.pt_start[e];
  };
  // Device point order: global E order stably sorted by track length, so
  // each lane class (a point-major launch segment) is contiguous and the
  // points sharing a wave share their rows' trip count.
  std::vector<int32_t> dperm(c->ne), dev_pt_start((size_t)c->ne + 1, 0);
  {
    // Stable counting sort by track length, per-thread counts over E ranges.
    Buf<int32_t> tlv(c->ne);
    const int nt = host_threads(c->ne, 1 << 14);
    const int64_t step = ((int64_t)c->ne + nt - 1) / nt;
    std::vector<int32_t> tmax(nt, 0);
    run_threads(nt, [&](int t) {
      const int64_t end = std::min<int64_t>(c->ne, (t + 1) * step);
      for (int64_t e = t * step; e < end; ++e) {
        tlv[e] = track_len((int32_t)e);
        tmax[t] = std::max(tmax[t], tlv[e]);
      }
    });
    const int32_t nkey = *std::max_element(tmax.begin(), tmax.end()) + 1;
    std::vector<int32_t> tc((size_t)nt * nkey, 0);
    run_threads(nt, [&](int t) {
      const int64_t end = std::min<int64_t>(c->ne, (t + 1) * step);
      int32_t* h = tc.data() + (size_t)t * nkey;
      for (int64_t e = t * step; e < end; ++e) h[tlv[e]]++;
    });
    // Lane class of a track length: 1 lane up to 4 rows, 2 up to 8, ...,
    // 64 beyond 128.
    const int32_t rpl = c->rows_per_lane;
    auto cls_of = [rpl](int32_t tl) {
      int k = 0;
      while (k < 6 && tl > rpl * (1 << k)) ++k;
      return k;
    };
    int32_t ncls[8] = {0}, start[8] = {0};
    {
      int32_t run = 0;
      for (int32_t key = 0; key < nkey; ++key)
        for (int t = 0; t < nt; ++t) {
          int32_t& v = tc[(size_t)t * nkey + key];
          const int32_t n = v;
          v = run;
          run += n;
          ncls[cls_of(key)] += n;
        }
    }
    for (int k = 1; k < 8; ++k) start[k] = start[k - 1] + ncls[k - 1];
    run_threads(nt, [&](int t) {
      const int64_t end = std::min<int64_t>(c->ne, (t + 1) * step);
      int32_t* h = tc.data() + (size_t)t * nkey;
      for (int64_t e = t * step; e < end; ++e) dperm[h[tlv[e]]++] = (int32_t)e;
    });
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
    Buf<int32_t> dlen(c->ne);
    parallel_for(c->ne, [&](int64_t b, int64_t end) {
      for (int64_t de = b; de < end; ++de) {
        const int32_t ge = dperm[de], w = ge_win[ge];
        c->dev_e_to_pt[de] = c->pt_off[w] + c->wst[w].e_to_pt[ge_local[ge]];
        dlen[de] = track_len(ge);
      }
    }, 1 << 14);
    exclusive_scan(dlen.data(), c->ne, dev_pt_start.data());
  }
  trace.mark("device point order");
  // F blocks: window-contiguous, each window in its reference F order.
  c->f_to_cam.resize(c->nf);
  c->cam_to_f.assign(c->nslot, -1);
  for (int32_t w = 0; w < W; ++w)
    for (size_t f = 0; f < c->wst[w].f_to_cam.size(); ++f) {
      const int32_t gf = c->f_off[w] + (int32_t)f, cam = c->cam_off[w] + c->wst[w].f_to_cam[f];
      c->f_to_cam[gf] = cam;
      c->cam_to_f[cam] = gf;
    }
  if (c->grouped) {
    c->cam_grp = c->wst[0].cam_grp;
    c->cam_gf.assign(c->nc, -1);
    for (int32_t cam = 0; cam < c->nc; ++cam)
      if (c->cam_grp[cam] >= 0) c->cam_gf[cam] = c->cam_to_f[c->nc + c->cam_grp[cam]];
  }
  // Rows in device order (each point's rows keep its window's reference
  // order); camera indices are global.  Points are independent.
  // A single problem's observations are uploaded in the caller's order and
  // put in row order on the device (row_src: each row's observation), so
  // the host reads only the camera index at random.
  const bool obs_on_device = W == 1;
  Buf<int32_t> row_cam(c->nrows), row_src(obs_on_device ? (size_t)c->nrows : 0);
  Buf<double> row_obs(obs_on_device ? 0 : 2 * (size_t)c->nrows);
  parallel_for(c->ne, [&](int64_t b, int64_t end) {
    for (int64_t de = b; de < end; ++de) {
      const int32_t ge = dperm[de], w = ge_win[ge], e = ge_local[ge];
      const Structure& st = c->wst[w];
      const mba_problem* p = &probs[w];
      int32_t r = dev_pt_start[de];
      for (int32_t rr = st.pt_start[e]; rr < st.pt_start[e + 1]; ++rr, ++r) {
        const int64_t i = st.res_order[rr];
        row_cam[r] = c->cam_off[w] + p->camera_index[i];
        if (obs_on_device) {
          row_src[r] = (int32_t)i;
        } else {
          row_obs[2 * (size_t)r] = p->observations[2 * i];
          row_obs[2 * (size_t)r + 1] = p->observations[2 * i + 1];
        }
      }
    }
  }, 1 << 14);
  trace.mark("rows");
  if (schur_jacobi) {
    // The per-row W Vinv W' terms are the camera's diagonal block only if
    // no camera observes a point twice.
    std::atomic<bool> dup{false};
    parallel_for(c->ne, [&](int64_t b, int64_t end) {
      std::vector<int32_t> cams;
      for (int64_t de = b; de < end && !dup; ++de) {
        cams.assign(row_cam.begin() + dev_pt_start[de], row_cam.begin() + dev_pt_start[de + 1]);
        std::sort(cams.begin(), cams.end());
        if (std::adjacent_find(cams.begin(), cams.end()) != cams.end()) dup = true;
      }
    });
    if (dup) return fail(MBA_ERR_UNSUPPORTED, "SCHUR_JACOBI: a camera observes a point more than once");
  }
  // Camera-major rows: a stable counting sort of the rows by key with
  // per-thread histograms over row ranges; cm_src[k] is the point-major row
  // of camera-major slot k.  Packed: key = F block (constant cameras
  // excluded).  Grouped: key = the camera's rank in (group, camera) order,
  // over cameras whose block or group block is free, so each group's
  // cameras are contiguous.
  std::vector<int32_t> key_of_cam, key_cam;
  int32_t nkeys = c->nf;
  if (c->grouped) {
    key_of_cam.assign(c->nc, -1);
    std::vector<int32_t> cams_sorted;
    for (int32_t cam = 0; cam < c->nc; ++cam)
      if (c->cam_to_f[cam] >= 0 || c->cam_gf[cam] >= 0) cams_sorted.push_back(cam);
    std::stable_sort(cams_sorted.begin(), cams_sorted.end(),
                     [&](int32_t a, int32_t b) { return c->cam_grp[a] < c->cam_grp[b]; });
    for (size_t k = 0; k < cams_sorted.size(); ++k) key_of_cam[cams_sorted[k]] = (int32_t)k;
    key_cam = cams_sorted;
    nkeys = (int32_t)cams_sorted.size();
  }
  auto row_key = [&](int32_t cam) { return c->grouped ? key_of_cam[cam] : c->cam_to_f[cam]; };
  std::vector<int64_t> fstart((size_t)nkeys + 1, 0);
  Buf<int32_t> cm_src;
  {
    const int nt = host_threads(c->nrows);
    const int64_t step = ((int64_t)c->nrows + nt - 1) / nt;
    std::vector<std::vector<int64_t>> cnt(nt, std::vector<int64_t>((size_t)nkeys, 0));
    run_threads(nt, [&](int t) {
      std::vector<int64_t>& h = cnt[t];
      const int64_t end = std::min<int64_t>(c->nrows, (t + 1) * step);
      for (int64_t r = t * step; r < end; ++r) {
        const int32_t f = row_key(row_cam[r]);
        if (f >= 0) h[f]++;
      }
    });
    int64_t run = 0;
    for (int32_t f = 0; f < nkeys; ++f) {
      fstart[f] = run;
      for (int t = 0; t < nt; ++t) {
        const int64_t n = cnt[t][f];
        cnt[t][f] = run;
        run += n;
      }
    }
    fstart[nkeys] = run;
    c->ncm = (int32_t)run;
    cm_src.resize((size_t)run);
    run_threads(nt, [&](int t) {
      std::vector<int64_t>& o = cnt[t];
      const int64_t end = std::min<int64_t>(c->nrows, (t + 1) * step);
      for (int64_t r = t * step; r < end; ++r) {
        const int32_t f = row_key(row_cam[r]);
        if (f >= 0) cm_src[o[f]++] = (int32_t)r;
      }
    });
  }
  trace.mark("camera-major sort");
  const int cr = chunk_rows_for(c->ncm);
  // Packed layout, large problems: each camera's rows (in row order, i.e.
  // by device point) are split at the bounds of kXcd point ranges of equal
  // row count, and range q's chunks get the launch slots congruent to q
  // mod kXcd, so that each XCD gathers from one range of point records.
  const bool xcd_split = !c->grouped && c->ncm >= kXcdSplitMin;
  std::vector<int32_t> range_row(kXcd + 1, 0);
  if (xcd_split) {
    for (int q = 1; q < kXcd; ++q) {
      const int64_t target = (int64_t)c->nrows * q / kXcd;
      const size_t de = (size_t)(std::lower_bound(dev_pt_start.begin(), dev_pt_start.end(), target) -
                                 dev_pt_start.begin());
      range_row[q] = dev_pt_start[std::min(de, (size_t)c->ne)];
    }
    range_row[kXcd] = c->nrows;
  }
  const int nq = xcd_split ? kXcd : 1;
  // Split chunks run as one-wave workgroups of (by default) kXcdChunkRows
  // rows; an explicit or tuned chunk size applies to either layout.
  const int crq = c->chunk_rows > 0 ? c->chunk_rows : (xcd_split ? kXcdChunkRows : cr);
  // Two parallel passes over the keys (the binary searches in cm_src are
  // cache misses: serially they cost ~10 ms at 6.8M rows): the range split
  // points and chunk counts of each key, then the chunks themselves.
  std::vector<int64_t> split((size_t)nkeys * (nq + 1));
  std::vector<int32_t> key_chunk((size_t)nkeys + 1, 0);
  auto chunk_len = [&](int64_t rows) -> int64_t {  // ceil(rows / crq) chunks of equal size
    const int64_t nck = (rows + crq - 1) / crq;
    return nck > 0 ? (rows + nck - 1) / nck : 0;
  };
  parallel_for(nkeys, [&](int64_t b, int64_t e) {
    for (int64_t f = b; f < e; ++f) {
      int64_t* sp = &split[(size_t)f * (nq + 1)];
      sp[0] = fstart[f];
      int32_t n = 0;
      for (int q = 0; q < nq; ++q) {
        int64_t jend = fstart[f + 1];
        if (q + 1 < nq)
          jend = std::lower_bound(cm_src.begin() + sp[q], cm_src.begin() + jend, range_row[q + 1]) - cm_src.begin();
        sp[q + 1] = jend;
        const int64_t rows = jend - sp[q];
        if (rows > 0) n += (int32_t)((rows + chunk_len(rows) - 1) / chunk_len(rows));
      }
      key_chunk[f + 1] = n;
    }
  }, 1 << 10);
  for (int32_t f = 0; f < nkeys; ++f) key_chunk[f + 1] += key_chunk[f];
  const size_t total_chunks = (size_t)key_chunk[nkeys];
  std::vector<int32_t> chunk_f(total_chunks), chunk_begin(total_chunks), chunk_end(total_chunks),
      chunk_q(total_chunks);
  parallel_for(nkeys, [&](int64_t b, int64_t e) {
    for (int64_t f = b; f < e; ++f) {
      const int64_t* sp = &split[(size_t)f * (nq + 1)];
      size_t k = (size_t)key_chunk[f];
      for (int q = 0; q < nq; ++q) {
        const int64_t len = chunk_len(sp[q + 1] - sp[q]);
        for (int64_t r = sp[q]; r < sp[q + 1]; r += len, ++k) {
          chunk_f[k] = c->grouped ? key_cam[f] : (int32_t)f;
          chunk_begin[k] = (int32_t)r;
          chunk_end[k] = (int32_t)std::min<int64_t>(r + len, sp[q + 1]);
          chunk_q[k] = q;
        }
      }
    }
  }, 1 << 10);
  key_chunk[nkeys] = (int32_t)chunk_f.size();
  c->nchunks = (int32_t)chunk_f.size();
  std::vector<int32_t> slot_chunk;
  if (xcd_split) {
    std::vector<std::vector<int32_t>> lists(kXcd);
    for (int32_t ch = 0; ch < c->nchunks; ++ch) lists[chunk_q[ch]].push_back(ch);
    size_t most = 0;
    for (const auto& l : lists) most = std::max(most, l.size());
    slot_chunk.assign(most * kXcd, -1);
    for (int q = 0; q < kXcd; ++q)
      for (size_t i = 0; i < lists[q].size(); ++i) slot_chunk[i * kXcd + q] = lists[q][i];
  }
  c->nslots = xcd_split ? (int32_t)slot_chunk.size() : c->nchunks;
  trace.mark("  chunks");
  // Host-side checks are done: the device from here on.
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (c->device < 0 || c->device >= ndev) return fail(MBA_ERR_DEVICE, "no such HIP device");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  trace.mark("  stream");
  HIP_TRY(hipHostMalloc((void**)&c->big, 2 * kBigChunk, hipHostMallocDefault));
  trace.mark("  pinned staging");
  for (hipEvent_t& e : c->big_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!slot_chunk.empty()) MBA_TRY(upload(c, &c->d_slot_chunk, slot_chunk));
  c->rfix = c->mixed && W == 1 && !c->grouped && c->nrows > 0;
  if (c->rfix) {
    // row -> camera-major slot (-1: a constant camera's row), filled on the
    // device from the camera-major order (alloc_precision).
    MBA_TRY(c->pool.alloc(&c->row_cm, (size_t)c->nrows));
    for (int q = 0; q < 2; ++q) {
      MBA_TRY(c->pool.alloc(&c->rres[q], 2 * (size_t)c->nrows));
      MBA_TRY(c->pool.alloc(&c->cres[q], 2 * (size_t)std::max(1, c->ncm)));
    }
  }
  // Partial rows of each F block: its chunks (packed), or its camera's
  // chunks / its cameras' group rows after the nchunks camera rows.
  std::vector<int32_t> f_lo(c->nf), f_hi(c->nf);
  if (!c->grouped) {
    for (int32_t f = 0; f < c->nf; ++f) {
      f_lo[f] = key_chunk[f];
      f_hi[f] = key_chunk[f + 1];
    }
  } else {
    for (int32_t f = 0; f < c->nf; ++f) {
      f_lo[f] = 0;
      f_hi[f] = 0;
    }
    for (int32_t k = 0; k < nkeys; ++k) {
      const int32_t cam = key_cam[k], fc = c->cam_to_f[cam], fg = c->cam_gf[cam];
      if (fc >= 0) {
        f_lo[fc] = key_chunk[k];
        f_hi[fc] = key_chunk[k + 1];
      }
      if (fg >= 0) {
        if (f_hi[fg] == 0) f_lo[fg] = c->nchunks + key_chunk[k];
        f_hi[fg] = c->nchunks + key_chunk[k + 1];
      }
    }
    std::vector<int32_t> glo, ghi, gout;
    for (int32_t f = 0; f < c->nf; ++f) {
      if (c->f_to_cam[f] < c->nc) continue;
      glo.push_back(f_lo[f]);
      ghi.push_back(f_hi[f]);
      gout.push_back(2 * c->nchunks + (int32_t)gout.size());
      f_lo[f] = gout.back();
      f_hi[f] = gout.back() + 1;
    }
    c->ngf = (int32_t)gout.size();
    MBA_TRY(upload(c, &c->g_lo, glo));
    MBA_TRY(upload(c, &c->g_hi, ghi));
    MBA_TRY(upload(c, &c->g_out, gout));
  }
  if (const char* e = std::getenv("MBA_FLOW_SPIN_LIMIT")) c->flow_spin = (unsigned)std::strtoul(e, nullptr, 10);
  // FP32 point-major kernels with the camera records staged in LDS (the
  // pass: projection quads; the matvec: projection and direction quads),
  // one or two workgroups per CU, as many as the table and the kernel's
  // registers leave room for.
  if (c->mixed && !c->grouped && c->nc > 0 && c->seg_blocks > 0) {
    int ncu = 0, lds_max = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device));
    auto plan = [&](const void* kernel, int rec_floats, LdsLaunch* l) -> mba_status {
      const size_t bytes = sizeof(float) * rec_floats * (size_t)c->nc;
      if (ncu <= 0 || bytes > (size_t)lds_max) return MBA_OK;
      HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
      // Workgroups resident per CU: the table's LDS and the kernel's
      // registers both limit it.
      int resident = 0;
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, kStageThreads, bytes));
      const int per_cu = (int)std::max<size_t>(1, std::min<size_t>({2, (size_t)lds_max / bytes, (size_t)resident}));
      l->bytes = bytes;
      l->grid = (int)std::min<int64_t>(((int64_t)c->seg_blocks + 3) / 4, (int64_t)ncu * per_cu);
      return MBA_OK;
    };
    MBA_TRY(plan((const void*)k_point_pass_lds, stage_rec<4>(), &c->lds_pass));
    MBA_TRY(plan((const void*)k_mv_point_lds, stage_rec<8>(), &c->lds_mv));
  }
  trace.mark("chunks + device setup");
  // FP64 state: cameras (slot blocks), then intrinsic groups (grouped).
  Buf<double> pts(3 * (size_t)c->ne);
  std::vector<double> cams(9 * (size_t)c->nslot, 0.0);
  parallel_for(c->ne, [&](int64_t b, int64_t end) {
    for (int64_t de = b; de < end; ++de) {
      const int32_t w = ge_win[dperm[de]];
      const int32_t q = c->dev_e_to_pt[de] - c->pt_off[w];
      for (int k = 0; k < 3; ++k) pts[3 * (size_t)de + k] = probs[w].points[3 * (size_t)q + k];
    }
  });
  if (c->grouped) {
    for (int32_t cam = 0; cam < c->nc; ++cam)
      for (int k = 0; k < 6; ++k) cams[9 * (size_t)cam + k] = probs[0].cameras[6 * (size_t)cam + k];
    for (int32_t g = 0; g < c->ng; ++g)
      for (int k = 0; k < 6; ++k) cams[9 * ((size_t)c->nc + g) + k] = probs[0].intrinsics[6 * (size_t)g + k];
    MBA_TRY(upload(c, &c->d_cam_grp, c->cam_grp));
    MBA_TRY(upload(c, &c->d_cam_gf, c->cam_gf));
  } else {
    for (int32_t w = 0; w < W; ++w)
      if (probs[w].num_cameras > 0)
        std::memcpy(cams.data() + 9 * (size_t)c->cam_off[w], probs[w].cameras,
                    72 * (size_t)probs[w].num_cameras);
  }
  {
    // Slot masks: a camera's parameter entries (camera_params); a group's
    // entries outside its SubsetParameterization constants (still counted
    // in |x|); unused slots of the 9 neither.
    std::vector<uint8_t> fm(9 * (size_t)c->nf, 0), xm(9 * (size_t)c->nf, 0);
    bool trivial = true;
    int32_t w = 0;
    for (int32_t f = 0; f < c->nf; ++f) {
      while (f >= c->f_off[w + 1]) ++w;
      const mba_problem& pw = probs[w];
      const int32_t b = c->f_to_cam[f];
      int32_t free_bits, in_bits;
      if (c->grouped && b >= c->nc) {
        const int32_t sub = pw.intrinsic_subset ? pw.intrinsic_subset[b - c->nc] : 0;
        in_bits = 0x3F;
        free_bits = 0x3F & ~sub;
      } else {
        const int32_t all = c->grouped ? 0x3F : 0x1FF;
        in_bits = free_bits = pw.camera_params ? pw.camera_params : all;
      }
      for (int k = 0; k < 9; ++k) {
        fm[9 * (size_t)f + k] = (free_bits >> k) & 1;
        xm[9 * (size_t)f + k] = (in_bits >> k) & 1;
        trivial = trivial && fm[9 * (size_t)f + k] && xm[9 * (size_t)f + k];
      }
    }
    if (!trivial) {
      MBA_TRY(upload(c, &c->fmask, fm));
      MBA_TRY(upload(c, &c->xmask, xm));
    }
  }
  trace.mark("state gather");
  MBA_TRY(upload(c, &c->cams64, cams));
  MBA_TRY(c->pool.alloc(&c->cams64_t, cams.size()));
  HIP_TRY(hipMemcpyAsync(c->cams64_t, c->cams64, 8 * cams.size(), hipMemcpyDeviceToDevice, c->stream));
  MBA_TRY(upload(c, &c->pts64, pts));
  MBA_TRY(c->pool.alloc(&c->pts64_t, pts.size()));
  if (!pts.empty())
    HIP_TRY(hipMemcpyAsync(c->pts64_t, c->pts64, 8 * pts.size(), hipMemcpyDeviceToDevice, c->stream));
  if (obs_on_device) {
    double* obs_in = nullptr;
    int32_t* src = nullptr;
    MBA_TRY(c->pool.alloc(&obs_in, 2 * (size_t)c->nrows));
    MBA_TRY(copy_to_device(c, obs_in, probs[0].observations, sizeof(double) * 2 * (size_t)c->nrows));
    MBA_TRY(upload(c, &src, row_src));
    MBA_TRY(c->pool.alloc(&c->row_obs64, 2 * (size_t)c->nrows));
    if (c->nrows > 0)
      hipLaunchKernelGGL(k_gather_obs, dim3(nblocks((int64_t)c->nrows)), dim3(kBlock), 0, c->stream,
                         (int64_t)c->nrows, (const int32_t*)src, (const double2*)obs_in, (double2*)c->row_obs64);
    HIP_TRY(hipGetLastError());
  } else {
    MBA_TRY(upload(c, &c->row_obs64, row_obs));
  }
  trace.mark("state uploads");
  MBA_TRY(c->pool.alloc(&c->ctab64, kCamTab * (size_t)std::max(1, c->nc)));
  MBA_TRY(c->pool.alloc(&c->dtab64, (size_t)std::max(kDirTab, kDirTabG) * std::max(1, c->nc)));
  MBA_TRY(c->pool.alloc(&c->pscale64, 3 * (size_t)c->ne));
  MBA_TRY(c->pool.alloc(&c->cscale64, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->diag_c, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->stepd, 3 * (size_t)c->ne + 9 * (size_t)c->nf));
  const int64_t maxb = nblocks(std::max<int64_t>({(int64_t)c->ne, 9 * (int64_t)c->nf, (int64_t)c->nrows, 1})) + 64;
  // Reduction scratch: block partials (<= 2 * maxb) plus the gradient
  // max/sumsq partials at offset 2048 (<= 1024 blocks each).
  const size_t nred = std::max<size_t>({4 * (size_t)maxb, 4096, 4 * (size_t)nblocks((int64_t)kGrp * c->nf)});
  MBA_TRY(c->pool.alloc(&c->red, nred));
  MBA_TRY(c->pool.alloc(&c->red2, std::max<size_t>(4 * (size_t)maxb, 4096)));
  MBA_TRY(c->pool.alloc(&c->sc, 4));
  MBA_TRY(c->pool.alloc(&c->pcg, (size_t)W));
  MBA_TRY(c->pool.alloc(&c->flag, std::max<size_t>(4, (size_t)W)));
  MBA_TRY(c->pool.alloc(&c->info, 4));
  MBA_TRY(c->pool.alloc(&c->all_done, 4));
  MBA_TRY(upload(c, &c->counter, std::vector<unsigned int>(4, 0u)));
  MBA_TRY(c->pool.alloc(&c->wacc, 4 * (size_t)W));
  MBA_TRY(c->pool.alloc(&c->wacc2, 2 * (size_t)W));
  MBA_TRY(c->pool.alloc(&c->accept, (size_t)W));
  c->h_inv_r.assign(W, 0.0);
  MBA_TRY(upload(c, &c->inv_r, c->h_inv_r));
  c->h_active.assign(W, 1);
  MBA_TRY(upload(c, &c->wactive, c->h_active));
  if (W > 1) {
    Buf<int32_t> of_e(c->ne);
    std::vector<int32_t> of_f(c->nf);
    parallel_for(c->ne, [&](int64_t b, int64_t e) {
      for (int64_t de = b; de < e; ++de) of_e[de] = ge_win[dperm[de]];
    });
    for (int32_t w = 0; w < W; ++w)
      for (int32_t f = c->f_off[w]; f < c->f_off[w + 1]; ++f) of_f[f] = w;
    MBA_TRY(upload(c, &c->of_e, of_e));
    MBA_TRY(upload(c, &c->of_f, of_f));
  }
  MBA_TRY(c->pool.alloc(&c->b64, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->x64, 9 * (size_t)c->nf));
  MBA_TRY(c->pool.alloc(&c->r64, 9 * (size_t)c->nf));
  trace.mark("scratch + window maps");
  if (c->mixed)
    MBA_TRY(alloc_precision<float>(c, dev_pt_start, row_cam, cm_src, chunk_f, chunk_begin, chunk_end, f_lo, f_hi));
  else
    MBA_TRY(alloc_precision<double>(c, dev_pt_start, row_cam, cm_src, chunk_f, chunk_begin, chunk_end, f_lo, f_hi));
  // Batched PCG, packed windows of at most kPcgWinMaxCams cameras: one
  // workgroup per window runs the window's whole PCG in one launch.
  if (W > 1 && !c->grouped && c->opt.linear_solver == MBA_IMPLICIT_SCHUR_PCG && c->nf > 0) {
    bool fits = true;
    for (int32_t w = 0; w < W && fits; ++w) fits = c->cam_off[w + 1] - c->cam_off[w] <= kPcgWinMaxCams;
    if (fits) {
      // Device points by window (each window's in device order): a stable
      // counting sort with per-range histograms.
      std::vector<int32_t> cnt((size_t)W + 1, 0), list(c->ne);
      {
        const int nt = host_threads(c->ne, 1 << 15);
        const int64_t step = ((int64_t)c->ne + nt - 1) / nt;
        std::vector<std::vector<int32_t>> h(nt, std::vector<int32_t>((size_t)W, 0));
        run_threads(nt, [&](int t) {
          for (int64_t de = t * step, e = std::min<int64_t>(c->ne, (t + 1) * step); de < e; ++de)
            h[t][ge_win[dperm[de]]]++;
        });
        int32_t run = 0;
        for (int32_t w = 0; w < W; ++w) {
          cnt[w] = run;
          for (int t = 0; t < nt; ++t) {
            const int32_t k = h[t][w];
            h[t][w] = run;
            run += k;
          }
        }
        cnt[W] = run;
        run_threads(nt, [&](int t) {
          for (int64_t de = t * step, e = std::min<int64_t>(c->ne, (t + 1) * step); de < e; ++de)
            list[h[t][ge_win[dperm[de]]]++] = (int32_t)de;
        });
      }
      std::vector<int32_t> lo(c->nf, 0), hi(c->nf, 0);
      for (int32_t f = 0; f < c->nf; ++f)
        if (f_hi[f] > f_lo[f]) {
          lo[f] = chunk_begin[f_lo[f]];
          hi[f] = chunk_end[f_hi[f] - 1];
        }
      // Workgroups take the windows with the most rows first.
      std::vector<int32_t> order(W);
      std::vector<int64_t> wrows(W, 0);
      for (int32_t w = 0; w < W; ++w) {
        order[w] = w;
        for (int32_t i = cnt[w]; i < cnt[w + 1]; ++i) wrows[w] += dev_pt_start[list[i] + 1] - dev_pt_start[list[i]];
      }
      std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return wrows[a] > wrows[b]; });
      MBA_TRY(upload(c, &c->d_win_order, order));
      MBA_TRY(upload(c, &c->d_win_e_lo, cnt));
      MBA_TRY(upload(c, &c->d_win_e, list));
      MBA_TRY(upload(c, &c->d_cam_off, c->cam_off));
      MBA_TRY(upload(c, &c->d_fcm_lo, lo));
      MBA_TRY(upload(c, &c->d_fcm_hi, hi));
      if (!c->d_f_off) MBA_TRY(upload(c, &c->d_f_off, c->f_off));
      c->pcg_window = true;
    }
  }
  if (c->opt.linear_solver == MBA_EXPLICIT_SCHUR) {
    MBA_TRY(build_s_pairs(c, dev_pt_start, row_cam));
    if (W > 1) {
      // One dense S per window: window w's of order 9 nf_w at woff[w].
      std::vector<int64_t> woff((size_t)W + 1, 0);
      for (int32_t w = 0; w < W; ++w) {
        const int64_t nw = 9 * (int64_t)(c->f_off[w + 1] - c->f_off[w]);
        woff[w + 1] = woff[w] + nw * nw;
      }
      c->s_wtotal = woff[W];
      MBA_TRY(upload(c, &c->d_woff, woff));
      if (!c->d_f_off) MBA_TRY(upload(c, &c->d_f_off, c->f_off));
    }
  }
  trace.mark("precision layout + S pairs");
  for (int i = 0; i < 8; ++i) HIP_TRY(hipEventCreate(&c->ev[i]));
  for (hipEvent_t& e : c->stage_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (hipEvent_t& e : c->tev) HIP_TRY(hipEventCreate(&e));
  if (c->opt.linear_solver == MBA_EXPLICIT_SCHUR) {
    HIP_TRY(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    for (auto& ring : c->chol_ev)
      for (hipEvent_t& e : ring) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  {
    const size_t Wz = (size_t)W;
    const size_t o_flags = sizeof(double) * kSumSlots * Wz;
    const size_t o_pcg = align_up(o_flags + sizeof(int32_t) * 2 * Wz, 64);
    const size_t o_stage = align_up(o_pcg + sizeof(PcgState) * Wz, 64);
    c->stage_bytes = align_up(8 * Wz, 64);
    HIP_TRY(hipHostMalloc((void**)&c->pin, o_stage + kStage * c->stage_bytes, hipHostMallocDefault));
    std::memset(c->pin, 0, o_stage + kStage * c->stage_bytes);
    c->h_sums = (double*)c->pin;
    c->h_flags = (int32_t*)(c->pin + o_flags);
    c->h_pcgs = (PcgState*)(c->pin + o_pcg);
    c->h_stage = c->pin + o_stage;
    HIP_TRY(hipHostMalloc((void**)&c->pin_done, 64, hipHostMallocCoherent));
    c->pin_done[0] = 0;
    c->h_done = c->pin_done;
    HIP_TRY(hipHostGetDevicePointer((void**)&c->d_done, c->pin_done, 0));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  trace.mark("drain");
  *out = c;
  guard.c = nullptr;
  return MBA_OK;
}

// The window descriptor handed to kernels; `acc` is the per-window
// accumulator array of a reduction (unused for a single problem).  Windows
// that have stopped iterating (and, inside PCG, whose solve has stopped)
// are skipped by the kernels.
Win win_desc(mba_context* c, double* acc, bool in_pcg) {
  Win w;
  const bool multi = c->W > 1;
  w.of_e = multi ? c->of_e : nullptr;
  w.of_f = multi ? c->of_f : nullptr;
  w.inv_r = c->inv_r;
  w.acc = multi ? acc : nullptr;
  w.flag = c->flag;
  w.active = multi ? c->wactive : nullptr;
  w.pcg = multi && in_pcg ? c->pcg : nullptr;
  return w;
}

// ------------------------------------------------------------ LM passes
template <typename T>
struct Engine {
  mba_context* c;
  Dev<T>& d;
  hipStream_t s;
  T huber;
  int W;
  explicit Engine(mba_context* ctx)
      : c(ctx), d(dev<T>(ctx)), s(ctx->stream), huber((T)ctx->opt.huber_a), W(ctx->W) {}

  mba_status sync() {
    HIP_TRY(hipStreamSynchronize(s));
    return MBA_OK;
  }
  // Seconds between two completed events.
  static double elapsed_s(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? 1e-3 * ms : 0.0;
  }
  Win win(double* acc = nullptr, bool in_pcg = false) { return win_desc(c, acc, in_pcg); }

  // Host -> device upload of n <= 8 W bytes through the next pinned slot,
  // in stream order and without a sync: the slot is rewritten only after
  // the event recorded behind its last copy has completed.
  mba_status stage_upload(void* dst, const void* src, size_t n) {
    if (n > c->stage_bytes) return fail(MBA_ERR_INVALID_ARGUMENT, "staged upload larger than its slot");
    const int k = c->stage_next;
    c->stage_next = (k + 1) % kStage;
    HIP_TRY(hipEventSynchronize(c->stage_ev[k]));
    char* h = c->h_stage + (size_t)k * c->stage_bytes;
    std::memcpy(h, src, n);
    HIP_TRY(hipMemcpyAsync(dst, h, n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(c->stage_ev[k], s));
    return MBA_OK;
  }

  // Windows still iterating; uploaded when it changes.
  mba_status set_active(const std::vector<uint8_t>& act) {
    if (W == 1 || act == c->h_active) return MBA_OK;
    c->h_active = act;
    return stage_upload(c->wactive, act.data(), (size_t)W);
  }

  mba_status clear_flags() {
    HIP_TRY(hipMemsetAsync(c->flag, 0, sizeof(int32_t) * W, s));
    return MBA_OK;
  }
  // Flags into pinned slot `slot` (0 or 1), valid after the next sync.
  mba_status flags_async(int slot) {
    HIP_TRY(hipMemcpyAsync(c->h_flags + (size_t)slot * W, c->flag, sizeof(int32_t) * W, hipMemcpyDeviceToHost, s));
    return MBA_OK;
  }
  int32_t flag_at(int slot, int w) const { return c->h_flags[(size_t)slot * W + w]; }

  // Per-window sums.  A single problem reduces block partials in c->red
  // with one deterministic k_reduce; several windows add into c->wacc.
  // sum_async leaves the result in pinned sum slot `slot`, valid after
  // the next sync; end_sum waits for it.
  mba_status begin_sum() {
    if (W > 1) HIP_TRY(hipMemsetAsync(c->wacc, 0, sizeof(double) * W, s));
    else HIP_TRY(hipMemsetAsync(c->red, 0, sizeof(double), s));
    return MBA_OK;
  }
  mba_status sum_async(int npartials, int slot) {
    double* h = c->h_sums + (size_t)slot * W;
    if (W == 1) {
      hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, c->red, std::max(npartials, 1), &c->sc->sum);
      HIP_TRY(hipMemcpyAsync(h, &c->sc->sum, sizeof(double), hipMemcpyDeviceToHost, s));
    } else {
      HIP_TRY(hipMemcpyAsync(h, c->wacc, sizeof(double) * W, hipMemcpyDeviceToHost, s));
    }
    return MBA_OK;
  }
  double sum_at(int slot, int w) const { return c->h_sums[(size_t)slot * W + w]; }
  mba_status end_sum(int npartials, std::vector<double>* out) {
    MBA_TRY(sum_async(npartials, kSlotSync));
    MBA_TRY(sync());
    out->assign(W, 0.0);
    for (int w = 0; w < W; ++w) (*out)[w] = sum_at(kSlotSync, w);
    return MBA_OK;
  }

  // 1 / radius per window (0 for radius 0: no damping), uploaded when it
  // changes: kernels read it from device memory, so a captured PCG
  // iteration replays under every radius.
  mba_status set_radii(const std::vector<double>& radius) {
    std::vector<double> inv(W);
    for (int w = 0; w < W; ++w) inv[w] = radius[w] > 0.0 ? 1.0 / radius[w] : 0.0;
    if (inv == c->h_inv_r) return MBA_OK;
    c->h_inv_r = inv;
    return stage_upload(c->inv_r, c->h_inv_r.data(), sizeof(double) * W);
  }
  mba_status set_radius(double r) { return set_radii(std::vector<double>(W, r)); }

  // Kernel-precision camera table from the FP64 state (and, with gtab, the
  // local rotation basis in the same launch).
  template <typename X>
  void table_into(const double* cams, X* tab, X* gtab = nullptr) {
    if (c->grouped)
      hipLaunchKernelGGL(k_cam_table_g<X>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, cams,
                         (const int32_t*)c->d_cam_grp, tab);
    else
      hipLaunchKernelGGL(k_cam_table<X>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, cams, tab, gtab);
  }
  void tables(const double* cams) { table_into<T>(cams, d.ctab, c->grouped ? nullptr : d.gtab); }
  // Row weights of the current state (slot rcur for mixed single problems).
  int wslot() const { return c->rfix ? c->rcur : 0; }
  T* wrow_cur() const { return wslot() ? d.wrow1 : d.wrow; }
  T* wcm_cur() const { return wslot() ? d.wcm1 : d.wcm; }

  // FP64 cost per window at (cams, pts) into sum slot kSlotCost and its
  // non-finite flags into flag slot 1; cost_result reads them after a sync.
  // slot >= 0 (residual-fix contexts): the residuals also go to slot `slot`
  // of rres / cres.  The flags are clear again when this returns (in stream
  // order), so what the next Jacobian pass or re-damping flags (a point or
  // camera block that is not positive definite) reaches the step's readback.
  // cm = false (a trial state): the camera-major copies are made only if the
  // state is accepted (cm_copies).
  mba_status cost64_async(const double* cams, const double* pts, int slot = -1, bool cm = true) {
    MBA_TRY(clear_flags());
    table_into<double>(cams, c->ctab64);
    const int nb = c->seg_blocks;
    MBA_TRY(begin_sum());
    const bool wr = c->rfix && slot >= 0, wc = wr && cm;
    if (nb > 0)
      hipLaunchKernelGGL((c->grouped ? k_cost64<1> : k_cost64<0>), dim3(nb), dim3(kBlock), 0, s, c->seg, d.pt_start,
                         d.row_cam, c->row_obs64, pts, c->ctab64, c->opt.huber_a, c->red, win(c->wacc),
                         wr ? c->rres[slot] : nullptr, wc ? c->cres[slot] : nullptr,
                         wc ? (const int32_t*)c->row_cm : nullptr, wr ? (slot ? c->df.wrow1 : c->df.wrow) : nullptr,
                         wc ? (slot ? c->df.wcm1 : c->df.wcm) : nullptr);
    MBA_TRY(sum_async(nb, kSlotCost));
    MBA_TRY(flags_async(1));
    return clear_flags();
  }
  // Camera-major copies of slot `slot`'s row-order residuals and weights.
  void cm_copies(int slot) {
    if (!c->rfix || c->nrows == 0) return;
    hipLaunchKernelGGL(k_rows_to_cm, dim3(nblocks((int64_t)c->nrows)), dim3(kBlock), 0, s, (int64_t)c->nrows,
                       (const int32_t*)c->row_cm, (const float*)c->rres[slot],
                       (const float*)(slot ? c->df.wrow1 : c->df.wrow), c->cres[slot], slot ? c->df.wcm1 : c->df.wcm);
  }
  void cost_result(std::vector<double>* cost, std::vector<uint8_t>* finite) const {
    cost->assign(W, 0.0);
    finite->assign(W, 1);
    for (int w = 0; w < W; ++w) {
      (*cost)[w] = sum_at(kSlotCost, w);
      (*finite)[w] = flag_at(1, w) == 0 && std::isfinite((*cost)[w]);
    }
  }
  mba_status cost64(const double* cams, const double* pts, std::vector<double>* cost, std::vector<uint8_t>* finite,
                    int slot = -1) {
    MBA_TRY(cost64_async(cams, pts, slot));
    MBA_TRY(sync());
    cost_result(cost, finite);
    return MBA_OK;
  }

  // Jacobian + Schur pass at the current x (tables and radii current).
  // FP64-evaluated residuals of the current state for the FP32 passes
  // (null: the passes form pred - obs themselves).
  const T* fixed_rows() const {
    if constexpr (std::is_same<T, float>::value)
      return c->rvalid ? c->rres[c->rcur] : nullptr;
    else
      return nullptr;
  }
  // The kernel-precision inverses of the point blocks, written only where
  // they are read: FP64 kernels (they are the FP64 inverses) and the
  // explicit solver's S assembly.
  T* vinv_out() const {
    if constexpr (std::is_same<T, float>::value)
      return c->opt.linear_solver == MBA_EXPLICIT_SCHUR ? d.Vinv : nullptr;
    else
      return d.Vinv;
  }
  // FP64 inverses of the damped point blocks.
  const double* vinv64() const {
    if constexpr (std::is_same<T, double>::value)
      return d.Vinv;
    else
      return d.Vinv64;
  }
  const T* fixed_cm() const {
    if constexpr (std::is_same<T, float>::value)
      return c->rvalid ? c->cres[c->rcur] : nullptr;
    else
      return nullptr;
  }

  void jacobian_pass(int recompute_diag) {
    if (c->timing) (void)hipEventRecord(c->ev[0], s);
    bool staged = false;
    if constexpr (std::is_same<T, float>::value) {
      if (c->lds_pass.grid > 0) {
        hipLaunchKernelGGL(k_point_pass_lds, dim3(c->lds_pass.grid), dim3(kStageThreads), c->lds_pass.bytes, s, c->seg,
                           c->nc, d.pt_start, d.row_cam, d.row_obs, c->pts64, d.pscale, d.ctab, huber, win(),
                           c->opt.min_lm_diagonal, c->opt.max_lm_diagonal, d.V, d.gp, vinv_out(), d.pk,
                           fixed_rows(), wrow_cur(), d.Vinv64);
        staged = true;
      }
    }
    if (c->seg_blocks > 0 && !staged)
      hipLaunchKernelGGL((c->grouped ? k_point_pass<T, 1> : k_point_pass<T, 0>), dim3(c->seg_blocks), dim3(kBlock), 0,
                         s, c->seg, d.pt_start, d.row_cam, d.row_obs, c->pts64, d.pscale, d.ctab, huber, win(),
                         c->opt.min_lm_diagonal, c->opt.max_lm_diagonal, d.V, d.gp, vinv_out(), d.pk, fixed_rows(),
                         c->grouped ? nullptr : wrow_cur(), d.Vinv64);
    if (c->timing) (void)hipEventRecord(c->ev[1], s);
    if (c->nchunks > 0) {
      if (c->grouped) {
        hipLaunchKernelGGL(k_cam_pass_g<T>, dim3(c->nchunks), dim3(kBlock), 0, s, c->nchunks, d.chunk_f,
                           d.chunk_begin, d.chunk_end, d.cm_e, d.cm_obs, d.pk, d.ctab, huber, d.partial);
        group_rows(63, nullptr);
      } else
        hipLaunchKernelGGL((c->d_slot_chunk ? k_cam_pass<T, 64> : k_cam_pass<T, kBlock>), dim3(c->nslots),
                           dim3(c->d_slot_chunk ? 64 : kBlock), 0, s, d.chunk_f, d.chunk_begin,
                           d.chunk_end, d.f_to_cam, d.cm_e, d.cm_obs, d.pk, d.ctab, huber, 1, d.partial, win(),
                           (const int32_t*)c->d_slot_chunk, fixed_cm(), wcm_cur(), (const T*)d.gtab);
    }
    sj_terms();
    if (c->timing) (void)hipEventRecord(c->ev[2], s);
    if (c->nf > 0)
      hipLaunchKernelGGL(k_cam_finalize<T>, dim3(c->nf), dim3(64), 0, s, c->nf, d.f_lo, d.f_hi, d.partial,
                         d.cscale, win(), c->opt.min_lm_diagonal, c->opt.max_lm_diagonal, c->diag_c,
                         recompute_diag, 1, d.U, d.gc, d.rhs, d.Minv, (const double*)d.Tsj);
    if (c->timing) (void)hipEventRecord(c->ev[3], s);
  }

  // SCHUR_JACOBI: the W Vinv W' terms of the cameras' diagonal blocks of S
  // at the current point damping (d.Tsj; NULL with JACOBI).
  void sj_terms() {
    if (!d.Tsj || c->nf == 0) return;
    if (c->nchunks > 0)
      hipLaunchKernelGGL((c->d_slot_chunk ? k_sj_pass<T, 64> : k_sj_pass<T, kBlock>), dim3(c->nslots),
                         dim3(c->d_slot_chunk ? 64 : kBlock), 0, s, d.chunk_f, d.chunk_begin, d.chunk_end,
                         d.f_to_cam, d.cm_e, d.cm_obs, d.pk, d.pscale, vinv64(), d.ctab, huber,
                         d.partial_sj, win(),
                         (const int32_t*)c->d_slot_chunk, (const T*)wcm_cur());
    hipLaunchKernelGGL(k_sj_finalize<T>, dim3(nblocks(45 * (int64_t)c->nf)), dim3(kBlock), 0, s, c->nf, d.f_lo,
                       d.f_hi, (const T*)d.partial_sj, (const T*)d.cscale, win(), d.Tsj);
  }

  // Re-damp after a rejected step (radii current) and rebuild the rhs.
  void damp() {
    hipLaunchKernelGGL(k_point_damp<T>, dim3(nblocks(c->ne)), dim3(kBlock), 0, s, c->ne, (const double*)d.V, d.gp,
                       c->opt.min_lm_diagonal, c->opt.max_lm_diagonal, win(), d.pscale, vinv_out(), d.pk, d.Vinv64);
    if (c->nf > 0) {
      sj_terms();
      hipLaunchKernelGGL(k_cam_damp<T>, dim3(nblocks(c->nf, 64)), dim3(64), 0, s, c->nf, d.U, c->diag_c, win(),
                         d.Minv, (const double*)d.Tsj);
      mv_cam(nullptr, nullptr);
      hipLaunchKernelGGL((c->grouped ? k_rhs_finalize<T, 1> : k_rhs_finalize<T, 0>), dim3(nblocks(c->nf)),
                         dim3(kBlock), 0, s, c->nf, d.f_lo, d.f_hi, d.partial, d.gc, d.rhs, win());
    }
  }

  // Direction table of an F-space vector x (sign * scale .* x) for the
  // point-major kernels.
  void cam_dir(const T* x, T sign, const int32_t* done, const Win& wv) {
    if (c->grouped)
      hipLaunchKernelGGL(k_cam_dir_g<T>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f,
                         (const int32_t*)c->d_cam_gf, d.cscale, d.ctab, x, sign, d.dtab, done);
    else
      hipLaunchKernelGGL(k_cam_dir<T>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f, d.cscale,
                         d.ctab, x, sign, d.dtab, done, wv);
  }
  // Point-major: z_e = Vinv E'(beta r + F x) (x through d.dtab); the rows'
  // loss weights from the last pass and, for beta != 0 inside the LM loop,
  // the FP64 cost's residuals (the ones the pass's g_p was formed from).
  void mv_point(T beta, T* pk_out, const int32_t* done, const Win& wv) {
    if (c->seg_blocks == 0) return;
    if constexpr (std::is_same<T, float>::value) {
      if (c->lds_mv.grid > 0) {
        hipLaunchKernelGGL(k_mv_point_lds, dim3(c->lds_mv.grid), dim3(kStageThreads), c->lds_mv.bytes, s, c->seg, c->nc,
                           d.pt_start, d.row_cam, d.row_obs, d.pk, d.pscale, d.ctab, d.dtab, d.Vinv, huber, beta, 1,
                           d.z, pk_out, done, wv, (const T*)wrow_cur(), beta != T(0) ? fixed_rows() : nullptr,
                           vinv64());
        return;
      }
    }
    hipLaunchKernelGGL((c->grouped ? k_mv_point<T, 1> : k_mv_point<T, 0>), dim3(c->seg_blocks), dim3(kBlock), 0, s,
                       c->seg, d.pt_start, d.row_cam, d.row_obs, d.pk, d.pscale, d.ctab, d.dtab, d.Vinv, huber,
                       beta, 1, d.z, pk_out, done, wv, (const T*)wrow_cur(), beta != T(0) ? fixed_rows() : nullptr,
                       vinv64());
  }
  // Camera-major partials: packed, sum Jc'(Jp z); grouped, sum F'(F x - E z)
  // (dtab: the direction table of x, or NULL for x = 0).
  // Grouped layout: sum each group block's chunk partials into its row.
  void group_rows(int width, const int32_t* done) {
    if (c->ngf > 0)
      hipLaunchKernelGGL(k_group_rows<T>, dim3(c->ngf), dim3(kGroupReduceThreads), 0, s, (const int32_t*)c->g_lo,
                         (const int32_t*)c->g_hi, (const int32_t*)c->g_out, width, d.partial, done);
  }
  void mv_cam(const T* dtab, const int32_t* done, const Win* wv = nullptr) {
    if (c->nchunks == 0) return;
    if (c->grouped) {
      hipLaunchKernelGGL(k_mv_cam_g<T>, dim3(c->nchunks), dim3(kBlock), 0, s, c->nchunks, d.chunk_f, d.chunk_begin,
                         d.chunk_end, d.cam_to_f, (const int32_t*)c->d_cam_gf, d.cm_e, d.cm_obs, d.pk, d.cscale,
                         d.ctab, dtab, huber, d.partial, done);
      group_rows(9, done);
    } else
      hipLaunchKernelGGL((c->d_slot_chunk ? k_mv_cam<T, 64> : k_mv_cam<T, kBlock>), dim3(c->nslots),
                         dim3(c->d_slot_chunk ? 64 : kBlock), 0, s, d.chunk_f, d.chunk_begin,
                         d.chunk_end, d.f_to_cam, d.cm_e, d.cm_obs, d.pk, d.cscale, d.ctab, huber, d.partial,
                         done, wv ? *wv : win(), (const int32_t*)c->d_slot_chunk, (const T*)wcm_cur(),
                         (const T*)d.gtab);
  }

  // y = S x (implicit), x and y F-order.  Inside PCG (`done` set) the
  // launches stop once every window's solve has stopped.
  void matvec(const T* x, T* y, const int32_t* done) {
    const Win wv = win(nullptr, done != nullptr);
    cam_dir(x, (T)1, done, wv);
    mv_point((T)0, d.pk, done, wv);
    mv_cam(d.dtab, done, &wv);
    hipLaunchKernelGGL((c->grouped ? k_mv_finalize<T, 1> : k_mv_finalize<T, 0>), dim3(nblocks(c->nf)), dim3(kBlock),
                       0, s, c->nf, d.f_lo, d.f_hi, d.partial, d.U, c->diag_c, wv, x, 1, y, done);
  }

  // Spin until the device's count of completed PCG updates (pinned
  // h_done[1]) reaches `target` or it reports all stopped; a stream that
  // has drained ends the wait too.
  mba_status wait_progress(int target) {
    for (int spin = 1; !*c->h_done && c->h_done[1] < target; ++spin) {
      if (spin % 256 == 0) {
        const hipError_t q = hipStreamQuery(s);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) HIP_TRY(q);
      }
      __builtin_ia32_pause();
    }
    return MBA_OK;
  }

  // PCG on S x = rhs (x in d.x) for every window still iterating (the
  // device's active flags); each window stops on its own tests, the
  // launches stop when all have.  The host runs at most kLook iterations
  // ahead of the device and stops launching once the device has written
  // "all stopped" to pinned memory; launches already queued return at
  // once.  The final states go to pinned memory: pcg_result reads them
  // after the next sync.
  mba_status pcg_async() {
    const int nf = c->nf;
    const int nbg = nblocks((int64_t)kGrp * nf);  // 16 lanes per F block
    if (c->pcg_window) {
      PcgWinArgs<T> a;
      a.order = c->d_win_order;
      a.win_e_lo = c->d_win_e_lo;
      a.win_e = c->d_win_e;
      a.f_off = c->d_f_off;
      a.cam_off = c->d_cam_off;
      a.f_to_cam = d.f_to_cam;
      a.cam_to_f = d.cam_to_f;
      a.fcm_lo = c->d_fcm_lo;
      a.fcm_hi = c->d_fcm_hi;
      a.pt_start = d.pt_start;
      a.row_cam = d.row_cam;
      a.row_obs = d.row_obs;
      a.cm_e = d.cm_e;
      a.cm_obs = d.cm_obs;
      a.pk = d.pk;
      a.pscale = d.pscale;
      a.cscale = d.cscale;
      a.ctab = d.ctab;
      a.Vinv64 = vinv64();
      a.wrow = wrow_cur();
      a.wcm = wcm_cur();
      a.U = d.U;
      a.Minv = d.Minv;
      a.diag_c = c->diag_c;
      a.inv_r = c->inv_r;
      a.b = d.rhs;
      a.x = d.x;
      a.st = c->pcg;
      a.active = c->wactive;
      a.huber_a = huber;
      a.eta = c->opt.eta;
      a.tol_r = c->opt.pcg_r_tolerance;
      a.min_iters = c->opt.min_linear_solver_iterations;
      a.max_iters = c->opt.max_linear_solver_iterations;
      hipLaunchKernelGGL(k_pcg_window<T>, dim3(W), dim3(pcg_win_threads<T>()), 0, s, a);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(c->h_pcgs, c->pcg, sizeof(PcgState) * W, hipMemcpyDeviceToHost, s));
      return MBA_OK;
    }
    *c->h_done = 0;
    c->h_done[1] = 0;
    hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(kBlock), 0, s, W, c->pcg, (const uint8_t*)(W > 1 ? c->wactive : nullptr),
                       c->opt.pcg_r_tolerance, c->all_done, c->counter + 3);
    // Windowed accumulators: dot at wacc[0, W), update sums at wacc[W, 4W);
    // the scalar kernels clear what they read.
    double* dot_acc = c->W > 1 ? c->wacc : nullptr;
    double* upd_acc = c->W > 1 ? c->wacc + W : nullptr;
    if (c->W > 1) HIP_TRY(hipMemsetAsync(c->wacc, 0, sizeof(double) * 4 * W, s));
    const Win wd = win(dot_acc, true), wu = win(upd_acc, true);
    const double eta = c->opt.eta;
    const int minit = c->opt.min_linear_solver_iterations, maxit = c->opt.max_linear_solver_iterations;
    auto update = [&](int mode) {
      hipLaunchKernelGGL(k_pcg_update_scalar<T>, dim3(nbg), dim3(kBlock), 0, s, nf, c->pcg, wu, mode, d.p, d.q, d.rhs,
                         d.tmp, d.Minv, d.x, d.r, d.zc, c->red, c->all_done, W, eta, minit, maxit, c->counter + 1,
                         c->d_done, mode == 1 ? nullptr : c->counter + 3);
    };
    const Win wv = win(nullptr, true);
    // One CG iteration: direction, implicit matvec, alpha, and the x / r
    // updates, or every tenth iteration the residual reset r = b - S x.
    // Every argument is fixed for the context but the row weights' slot, so
    // with pcg_graph_launch the iteration is captured once per (slot, reset)
    // as a HIP graph and replayed; no memset is recorded (the scalar kernels
    // clear what they read).
    auto iteration = [&](int reset) {
      if (c->grouped) {
        // A group's direction feeds all its cameras' rows: update p first.
        hipLaunchKernelGGL(k_pcg_p<T>, dim3(nblocks(9 * (int64_t)nf)), dim3(kBlock), 0, s, nf, (const PcgState*)c->pcg,
                           (const T*)d.zc, d.p, (const int32_t*)c->all_done);
        cam_dir(d.p, (T)1, c->all_done, wv);
      } else {
        hipLaunchKernelGGL(k_pcg_p_dir<T>, dim3(nblocks((int64_t)kGrp * c->nc)), dim3(kBlock), 0, s, c->nc,
                           d.cam_to_f, c->pcg, wv, d.zc, d.p, d.cscale, d.ctab, d.dtab, c->all_done);
      }
      mv_point((T)0, d.pk, c->all_done, wv);
      mv_cam(d.dtab, c->all_done, &wv);
      hipLaunchKernelGGL((c->grouped ? k_mv_finalize_alpha<T, 1> : k_mv_finalize_alpha<T, 0>), dim3(nbg), dim3(kBlock),
                         0, s, nf, d.f_lo, d.f_hi, d.partial, d.U, c->diag_c, wd, d.p, d.q, c->all_done, c->pcg, W,
                         c->red, c->counter);
      if (reset) {
        update(1);
        matvec(d.x, d.tmp, c->all_done);
        update(2);
      } else {
        update(0);
      }
    };
    hipGraphExec_t* graphs = c->opt.pcg_graph_launch ? c->pcg_graph[wslot()] : nullptr;
    auto launch = [&](int reset) -> mba_status {
      if (!graphs) {
        iteration(reset);
        HIP_TRY(hipGetLastError());
        return MBA_OK;
      }
      hipGraphExec_t& ex = graphs[reset];
      if (!ex) {
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        iteration(reset);
        const hipError_t le = hipGetLastError();
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(s, &g);
        HIP_TRY(le);
        HIP_TRY(ce);
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) ex = nullptr;
        HIP_TRY(ie);
      }
      HIP_TRY(hipGraphLaunch(ex, s));
      return MBA_OK;
    };
    update(3);
    for (int it = 1; it <= maxit; ++it) {
      // Wait until the device has finished iteration it - kLook (its update
      // kernel counts completed updates into pinned memory, the first one
      // before the loop), then stop if it has finished altogether.
      MBA_TRY(wait_progress(it + 1 - kLook));
      if (*c->h_done) break;
      MBA_TRY(launch(it % 10 == 0 ? 1 : 0));
    }
    HIP_TRY(hipMemcpyAsync(c->h_pcgs, c->pcg, sizeof(PcgState) * W, hipMemcpyDeviceToHost, s));
    return MBA_OK;
  }
  void pcg_result(const std::vector<uint8_t>& active, std::vector<int>* iters, std::vector<int>* status) const {
    iters->assign(W, 0);
    status->assign(W, 0);
    for (int w = 0; w < W; ++w) {
      if (!active[w]) continue;
      (*iters)[w] = c->h_pcgs[w].iter;
      (*status)[w] = c->h_pcgs[w].status;
    }
  }

  // In-place solve with the factor: L y = v, then L' x = y (work vector
  // d.ty; one launch per direction, blocks handed on through d.flow under a
  // new epoch per launch).
  void trsv(T* v, int64_t n) {
    const int nb = (int)((n + kNB - 1) / kNB);
    uint64_t* err = d.flow + (size_t)flow_words<T>() * n;
    const unsigned e1 = ++d.flow_epoch, e2 = ++d.flow_epoch;
    // Two 64-row blocks per workgroup.
    hipLaunchKernelGGL(k_trsv_lower_flow<T>, dim3((nb + 1) / 2), dim3(kFlowThreads), 0, s, (const T*)d.L,
                       (const T*)d.Linv, n, nb, (const T*)v, d.ty, d.flow, e1, err, c->flow_spin);
    hipLaunchKernelGGL(k_trsv_upper_flow<T>, dim3((nb + 1) / 2), dim3(kFlowThreads), 0, s, (const T*)d.L,
                       (const T*)d.LinvT, n, nb, (const T*)d.ty, v, d.flow, e2, err, c->flow_spin);
  }

  // Explicit Schur (single problem): assemble dense S, Cholesky, solve,
  // FP64 refinement.
  // The S blocks points couple: W_r V^-1 W_r' products summed per block
  // (all windows; idle ones skipped), written with U + D^2 into `sd`.
  void assemble_s(const SDest& sd) {
    const Win wv = win();
    if (c->s_nitems > 0)
      hipLaunchKernelGGL(k_s_blocks<T>, dim3((unsigned)c->s_nitems), dim3(64), 0, s, (const int32_t*)c->s_item_blk,
                         (const int64_t*)c->s_item_p0, (const int64_t*)c->s_item_p1, (const int32_t*)c->s_item_slot,
                         (const int32_t*)c->s_blk_fa, (const int32_t*)c->s_blk_fb, (const int32_t*)c->s_pair_a,
                         (const int32_t*)c->s_pair_b, (const T*)d.WY, (const T*)d.U, (const double*)c->diag_c, wv,
                         sd, d.S, (const int32_t*)c->s_blk_cross, (const T*)d.cross, d.s_scratch);
    if (c->s_nsplit > 0)
      hipLaunchKernelGGL(k_s_split_final<T>, dim3((unsigned)c->s_nsplit), dim3(64), 0, s,
                         (const int32_t*)c->s_split_blk, (const int32_t*)c->s_split_lo,
                         (const int32_t*)c->s_split_hi, (const int32_t*)c->s_blk_fa, (const int32_t*)c->s_blk_fb,
                         (const T*)d.U, (const double*)c->diag_c, wv, sd, d.S, (const int32_t*)c->s_blk_cross,
                         (const T*)d.cross, (const T*)d.s_scratch);
  }

  // Batched explicit solve (several windows, FP64): every window's dense S
  // assembled in its own region, factored and solved by one workgroup per
  // window; a failed factorisation flags its window.
  mba_status batched_explicit_solve() {
    if (c->nf == 0) return MBA_OK;
    if (!d.S) {
      MBA_TRY(c->pool.alloc(&d.S, (size_t)std::max<int64_t>(1, c->s_wtotal)));
      MBA_TRY(c->pool.alloc(&d.WY, (size_t)kWY * c->nrows));
      MBA_TRY(c->pool.alloc(&d.s_scratch, 81 * (size_t)std::max<int64_t>(1, c->s_nslots)));
    }
    HIP_TRY(hipMemsetAsync(d.S, 0, sizeof(T) * c->s_wtotal, s));
    if (c->seg_blocks > 0)
      hipLaunchKernelGGL(k_s_rows<T>, dim3(c->seg_blocks), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam,
                         d.cam_to_f, d.row_obs, d.pk, d.pscale, d.cscale, d.ctab, d.Vinv, huber, d.WY, win());
    assemble_s(SDest{0, (const int64_t*)c->d_woff, (const int32_t*)c->d_f_off, (const int32_t*)c->of_f});
    hipLaunchKernelGGL(k_win_chol_solve<T>, dim3(W), dim3(kWinThreads), 0, s, (const int64_t*)c->d_woff,
                       (const int32_t*)c->d_f_off, d.S, (const T*)d.rhs, d.x, win());
    return MBA_OK;
  }

  mba_status explicit_solve(bool* ok) {
    const int64_t n = 9 * (int64_t)c->nf;
    *ok = true;
    if (W > 1) return batched_explicit_solve();
    if (n == 0) return MBA_OK;
    if (!d.S) {
      MBA_TRY(c->pool.alloc(&d.S, (size_t)(n * n)));
      MBA_TRY(c->pool.alloc(&d.L, (size_t)(n * n)));
      MBA_TRY(c->pool.alloc(&d.WY, (size_t)kWY * ((size_t)c->nrows * (c->grouped ? 2 : 1) + c->s_nmerged)));
      if (c->grouped) MBA_TRY(c->pool.alloc(&d.cross, 36 * (size_t)std::max(1, c->nc)));
      MBA_TRY(c->pool.alloc(&d.s_scratch, 81 * (size_t)std::max<int64_t>(1, c->s_nslots)));
      const size_t nbk = (size_t)((n + kNB - 1) / kNB);
      MBA_TRY(c->pool.alloc(&d.Linv, nbk * kNB * kNB));
      MBA_TRY(c->pool.alloc(&d.slab, (size_t)kNB * kNB));
      MBA_TRY(c->pool.alloc(&d.LinvT, nbk * kNB * kNB));
      MBA_TRY(c->pool.alloc(&d.ty, (size_t)n));
      MBA_TRY(c->pool.alloc(&d.flow, (size_t)flow_words<T>() * n + 1));
      HIP_TRY(hipMemsetAsync(d.flow, 0, sizeof(uint64_t) * ((size_t)flow_words<T>() * n + 1), s));
      d.flow_epoch = 0;
      if (!c->gemv_part) MBA_TRY(c->pool.alloc(&c->gemv_part, (size_t)kGemvSplit * n));
    }
    if (c->timing) (void)hipEventRecord(c->ev[4], s);
    // Blocks no point couples stay zero.
    HIP_TRY(hipMemsetAsync(d.S, 0, sizeof(T) * n * n, s));
    if (c->grouped) {
      if (c->seg_blocks > 0)
        hipLaunchKernelGGL(k_s_rows_g<T>, dim3(c->seg_blocks), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam,
                           d.cam_to_f, (const int32_t*)c->d_cam_gf, d.row_obs, d.pk, d.pscale, d.cscale, d.ctab,
                           d.Vinv, huber, c->nrows, d.WY);
      if (c->s_nmerged > 0)
        hipLaunchKernelGGL(k_merge_rows<T>, dim3((unsigned)((c->s_nmerged + kBlock / 64 - 1) / (kBlock / 64))),
                           dim3(kBlock), 0, s, c->s_nmerged, (const int32_t*)c->s_m_start,
                           (const int32_t*)c->s_m_row, c->nrows, d.WY);
      // The camera-group blocks of F'F (partials reuse the pass's scratch).
      if (c->nchunks > 0) {
        hipLaunchKernelGGL(k_cross_g<T>, dim3(c->nchunks), dim3(kBlock), 0, s, d.chunk_f, d.chunk_begin, d.chunk_end,
                           d.cm_e, d.cm_obs, d.pk, d.ctab, huber, d.partial);
        hipLaunchKernelGGL(k_cross_final<T>, dim3(nblocks(36 * (int64_t)c->nc)), dim3(kBlock), 0, s, c->nc,
                           d.cam_to_f, (const int32_t*)c->d_cam_gf, d.f_lo, d.f_hi, d.partial, d.cscale, d.cross);
      }
    } else if (c->seg_blocks > 0) {
      hipLaunchKernelGGL(k_s_rows<T>, dim3(c->seg_blocks), dim3(kBlock), 0, s, c->seg, d.pt_start, d.row_cam,
                         d.cam_to_f, d.row_obs, d.pk, d.pscale, d.cscale, d.ctab, d.Vinv, huber, d.WY, win());
    }
    assemble_s(SDest{n, nullptr, nullptr, nullptr});
    if (c->timing) (void)hipEventRecord(c->ev[5], s);
    HIP_TRY(hipMemcpyAsync(d.L, d.S, sizeof(T) * n * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(c->info, 0, sizeof(int32_t), s));
    // Outer steps of two 64-column blocks c0 and k2 = c0 + 64 on s: diagonal
    // factor of c0, its panel with the update of block column k2, diagonal
    // factor of k2, its panel alone; then one trailing update with both
    // panels (K = 128): the next outer step's two block columns on s, the
    // other tiles on stream2, where they overlap the next outer step's
    // factors and panels.  Every element still receives all its updates
    // before it is factored or used as a panel.
    int ring = 0, pending = -1;
    for (int64_t c0 = 0; c0 < n; c0 += 2 * kNB) {
      const int kb1 = (int)std::min<int64_t>(kNB, n - c0);
      hipLaunchKernelGGL(k_potrf_diag<T>, dim3(1), dim3(256), 0, s, d.L, n, (int)c0, kb1,
                         d.Linv + (size_t)(c0 / kNB) * kNB * kNB, d.LinvT + (size_t)(c0 / kNB) * kNB * kNB, c->info,
                         d.slab);
      const int64_t k2 = c0 + kb1;
      if (k2 >= n) break;
      const int kb2 = (int)std::min<int64_t>(kNB, n - k2);
      // The panel's workgroup 0 also factors block k2 (it writes that
      // block's last update).
      T* li2 = d.Linv + (size_t)(k2 / kNB) * kNB * kNB;
      hipLaunchKernelGGL(k_panel_col0<T>, dim3((unsigned)((n - k2 + kNB - 1) / kNB)), dim3(256), 0, s, d.L, n,
                         (int)c0, kb1, (const T*)(d.Linv + (size_t)(c0 / kNB) * kNB * kNB), 1, (const T*)d.slab,
                         li2, d.LinvT + (size_t)(k2 / kNB) * kNB * kNB, c->info);
      const int64_t rest = n - k2 - kb2;
      if (rest <= 0) break;
      const int nt = (int)((rest + kNB - 1) / kNB);
      hipLaunchKernelGGL(k_panel_col0<T>, dim3(nt), dim3(256), 0, s, d.L, n, (int)k2, kb2, (const T*)li2, 0,
                         (const T*)nullptr, (T*)nullptr, (T*)nullptr, (int32_t*)nullptr);
      // The next outer step's block columns were also updated by the
      // previous step's stream2 tiles.
      if (pending >= 0) HIP_TRY(hipStreamWaitEvent(s, c->chol_ev[1][pending], 0));
      pending = -1;
      hipLaunchKernelGGL(k_syrk2_mfma<T>, dim3((unsigned)(2 * nt - 1)), dim3(256), 0, s, d.L, n, (int)c0, nt, 1);
      if (nt > 2) {
        HIP_TRY(hipEventRecord(c->chol_ev[0][ring], s));
        HIP_TRY(hipStreamWaitEvent(c->stream2, c->chol_ev[0][ring], 0));
        hipLaunchKernelGGL(k_syrk2_mfma<T>, dim3((unsigned)((int64_t)(nt - 2) * (nt - 1) / 2)), dim3(256), 0,
                           c->stream2, d.L, n, (int)c0, nt, 0);
        HIP_TRY(hipEventRecord(c->chol_ev[1][ring], c->stream2));
        pending = ring;
        ring = (ring + 1) & 3;
      }
    }
    if (pending >= 0) HIP_TRY(hipStreamWaitEvent(s, c->chol_ev[1][pending], 0));
    if (c->timing) (void)hipEventRecord(c->ev[6], s);
    int32_t info = 0;
    HIP_TRY(hipMemcpyAsync(&info, c->info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MBA_TRY(sync());
    if (info != 0) {
      *ok = false;
      return MBA_OK;
    }
    // x = S^-1 rhs with FP64 iterative refinement of the residual.
    HIP_TRY(hipMemcpyAsync(d.x, d.rhs, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    trsv(d.x, n);
    // FP64 refinement of an FP32 (mixed) factorisation; an FP64 factor is
    // used once, as the reference's sparse Cholesky solve is.
    if (c->mixed && c->opt.refinement_iterations > 0) {
      hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, d.rhs, c->b64);
      hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, d.x, c->x64);
      for (int it = 0; it < c->opt.refinement_iterations; ++it) {
        hipLaunchKernelGGL(k_gemv_partial<T>, dim3(nblocks(n), kGemvSplit), dim3(kBlock), 0, s, d.S, n,
                           (const double*)c->x64, c->gemv_part);
        hipLaunchKernelGGL(k_gemv_final, dim3(nblocks(n)), dim3(kBlock), 0, s, n, (const double*)c->gemv_part,
                           (const double*)c->b64, c->r64);
        hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, c->r64, d.sv);
        trsv(d.sv, n);
        hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, d.sv, c->r64);
        hipLaunchKernelGGL(k_axpy_inplace, dim3(nblocks(n)), dim3(kBlock), 0, s, n, c->r64, c->x64);
      }
      hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, c->x64, d.x);
    }
    if (c->timing) {
      (void)hipEventRecord(c->ev[7], s);
      c->timed_explicit = true;
    }
    return MBA_OK;
  }

  // Back substitution: y_e = Vinv E'(b - F z) into d.z; step = -y.
  void back_substitute() {
    cam_dir(d.x, (T)-1, nullptr, win());
    mv_point((T)1, nullptr, nullptr, win());
    hipLaunchKernelGGL(k_negate<T>, dim3(nblocks(3 * (int64_t)c->ne)), dim3(kBlock), 0, s,
                       (int64_t)3 * c->ne, d.z, d.step_p);
    hipLaunchKernelGGL(k_negate<T>, dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s,
                       (int64_t)9 * c->nf, d.x, d.step_c);
  }

  // Model cost change -J step . (r + J step / 2) per window, negated, into
  // sum slot kSlotModel.  Always formed in FP64 (as the reference forms it):
  // with FP32 kernels the step is converted and the Jacobian re-evaluated
  // from the FP64 state, since FP32 residuals (pred - obs at ~500 px) carry
  // ~3e-5 px of cancellation error that would make the gain ratio of the
  // small steps near convergence noisy.
  mba_status model_cost_async() {
    const int nb = c->seg_blocks;
    if constexpr (std::is_same<T, float>::value) {
      const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
      hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(np3)), dim3(kBlock), 0, s, np3, d.step_p, c->stepd);
      hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(nf9)), dim3(kBlock), 0, s, nf9, d.step_c,
                         c->stepd + np3);
      table_into<double>(c->cams64, c->ctab64);
      const Win wv = win();
      if (c->grouped)
        hipLaunchKernelGGL(k_cam_dir_g<double>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f,
                           (const int32_t*)c->d_cam_gf, (const double*)c->cscale64, (const double*)c->ctab64,
                           (const double*)(c->stepd + np3), 1.0, c->dtab64, (const int32_t*)nullptr);
      else
        hipLaunchKernelGGL(k_cam_dir<double>, dim3(nblocks(c->nc)), dim3(kBlock), 0, s, c->nc, d.cam_to_f,
                           (const double*)c->cscale64, (const double*)c->ctab64, (const double*)(c->stepd + np3),
                           1.0, c->dtab64, (const int32_t*)nullptr, wv);
      MBA_TRY(begin_sum());
      if (nb > 0)
        hipLaunchKernelGGL((c->grouped ? k_model_cost<double, 1> : k_model_cost<double, 0>), dim3(nb), dim3(kBlock),
                           0, s, c->seg, d.pt_start, d.row_cam, (const double*)c->row_obs64,
                           (const double*)c->pts64, 3, (const double*)c->pscale64, (const double*)c->ctab64,
                           (const double*)c->dtab64, (const double*)c->stepd, c->opt.huber_a, c->red,
                           win(c->wacc));
    } else {
      cam_dir(d.step_c, (T)1, nullptr, win());
      MBA_TRY(begin_sum());
      if (nb > 0)
        hipLaunchKernelGGL((c->grouped ? k_model_cost<T, 1> : k_model_cost<T, 0>), dim3(nb), dim3(kBlock), 0, s,
                           c->seg, d.pt_start, d.row_cam, d.row_obs, d.pk, kPk, d.pscale, d.ctab, d.dtab, d.step_p,
                           huber, c->red, win(c->wacc));
    }
    return sum_async(nb, kSlotModel);
  }

  // x_trial = x + step .* scale; per window |delta|^2 into sum slot
  // kSlotPlus (non-finite when the trial point is).
  mba_status plus_async() {
    const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
    hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(np3)), dim3(kBlock), 0, s, np3, d.step_p, c->stepd);
    hipLaunchKernelGGL((k_convert<T, double>), dim3(nblocks(nf9)), dim3(kBlock), 0, s, nf9, d.step_c,
                       c->stepd + np3);
    const int nb1 = nblocks(np3), nb2 = nblocks(nf9);
    MBA_TRY(begin_sum());
    hipLaunchKernelGGL(k_plus, dim3(nb1), dim3(kBlock), 0, s, np3, c->pts64, c->stepd, c->pscale64, c->pts64_t,
                       c->red, win(c->wacc));
    hipLaunchKernelGGL(k_plus_cams, dim3(nb2), dim3(kBlock), 0, s, c->nf, d.f_to_cam, c->cams64,
                       c->stepd + np3, c->cscale64, c->cams64_t, c->red + nb1, win(c->wacc));
    return sum_async(nb1 + nb2, kSlotPlus);
  }

  // |x| per window (the parameters of the local parameterisation) into
  // sum slot kSlotXnorm as |x|^2.
  mba_status x_norm_async() {
    const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
    hipLaunchKernelGGL(k_gather_cams, dim3(nblocks(nf9)), dim3(kBlock), 0, s, c->nf, d.f_to_cam, c->cams64,
                       c->stepd + np3, (const uint8_t*)c->xmask);
    hipLaunchKernelGGL(k_copy64, dim3(nblocks(np3)), dim3(kBlock), 0, s, np3, c->pts64, c->stepd);
    const int nb = nblocks(np3 + nf9);
    MBA_TRY(begin_sum());
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, s, np3 + nf9, np3, c->stepd, c->red, win(c->wacc));
    return sum_async(nb, kSlotXnorm);
  }

  // Max |g| and |g|^2 of the unscaled gradient, per window, interleaved in
  // sum slots kSlotGrad, kSlotGrad + 1; gradient_result reads them.
  mba_status gradient_async() {
    const int64_t np3 = 3 * (int64_t)c->ne, nf9 = 9 * (int64_t)c->nf;
    double* v = c->h_sums + (size_t)kSlotGrad * W;
    if (W == 1) {
      HIP_TRY(hipMemsetAsync(c->red2, 0, 2 * sizeof(double), s));
      int nb = std::min(1024, nblocks(np3));
      hipLaunchKernelGGL(k_grad_partial<T>, dim3(nb), dim3(kBlock), 0, s, np3, d.gp, c->pscale64, c->red,
                         c->red + 2048);
      hipLaunchKernelGGL(k_grad_final, dim3(1), dim3(kBlock), 0, s, c->red, c->red + 2048, nb, c->red2);
      nb = std::min(1024, nblocks(nf9));
      if (nf9 > 0) {
        hipLaunchKernelGGL(k_grad_partial<T>, dim3(nb), dim3(kBlock), 0, s, nf9, d.gc, c->cscale64, c->red,
                           c->red + 2048);
        hipLaunchKernelGGL(k_grad_final, dim3(1), dim3(kBlock), 0, s, c->red, c->red + 2048, nb, c->red2);
      }
      HIP_TRY(hipMemcpyAsync(v, c->red2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    } else {
      HIP_TRY(hipMemsetAsync(c->wacc2, 0, 2 * sizeof(double) * W, s));
      hipLaunchKernelGGL(k_grad_win<T>, dim3(nblocks(np3)), dim3(kBlock), 0, s, np3, 3, 0, d.gp, c->pscale64, win(),
                         c->wacc2);
      hipLaunchKernelGGL(k_grad_win<T>, dim3(nblocks(nf9)), dim3(kBlock), 0, s, nf9, 9, 1, d.gc, c->cscale64, win(),
                         c->wacc2);
      HIP_TRY(hipMemcpyAsync(v, c->wacc2, 2 * sizeof(double) * W, hipMemcpyDeviceToHost, s));
    }
    return MBA_OK;
  }
  void gradient_result(std::vector<double>* gmax, std::vector<double>* gnorm) const {
    const double* v = c->h_sums + (size_t)kSlotGrad * W;
    gmax->assign(W, 0.0);
    gnorm->assign(W, 0.0);
    for (int w = 0; w < W; ++w) {
      (*gmax)[w] = v[2 * (size_t)w];
      (*gnorm)[w] = std::sqrt(v[2 * (size_t)w + 1]);
    }
  }

  // EstimateScale (trust_region_minimizer.cc:103-109) from an unscaled,
  // undamped pass.
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
    MBA_TRY(set_radius(0.0));
    jacobian_pass(1);
    hipLaunchKernelGGL(k_make_scale<T>, dim3(nblocks(std::max(c->ne, c->nf))), dim3(kBlock), 0, s, c->ne,
                       c->nf, (const double*)d.V, (const T*)d.U, c->pscale64, c->cscale64);
    sync_scale_T();
    c->scale_ready = true;
    return MBA_OK;
  }
  void sync_scale_T() {
    hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(3 * (int64_t)c->ne)), dim3(kBlock), 0, s,
                       (int64_t)3 * c->ne, c->pscale64, d.pscale);
    hipLaunchKernelGGL((k_convert<double, T>), dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s,
                       (int64_t)9 * c->nf, c->cscale64, d.cscale);
    if (c->fmask && c->nf > 0)
      hipLaunchKernelGGL(k_scale_mask<T>, dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s, (int64_t)9 * c->nf,
                         (const uint8_t*)c->fmask, d.cscale);
  }

  // Solve the LM systems at the current radii and fill step_p / step_c
  // (for every window still iterating: a window whose solve failed gets a
  // step that is never used).  `factored` is false when the explicit
  // factorisation failed; step_result reads the PCG outcome after a sync.
  mba_status compute_step_async(bool* factored) {
    *factored = true;
    if (c->opt.linear_solver == MBA_IMPLICIT_SCHUR_PCG) {
      if (c->nf > 0) MBA_TRY(pcg_async());
      else HIP_TRY(hipMemsetAsync(d.x, 0, sizeof(T) * 9, s));
    } else {
      MBA_TRY(explicit_solve(factored));
    }
    back_substitute();
    return MBA_OK;
  }
  void step_result(const std::vector<uint8_t>& active, bool factored, std::vector<uint8_t>* solved,
                   std::vector<int>* lin) const {
    solved->assign(W, 1);
    lin->assign(W, 0);
    if (c->opt.linear_solver == MBA_IMPLICIT_SCHUR_PCG) {
      if (c->nf > 0) {
        std::vector<int> status;
        pcg_result(active, lin, &status);
        for (int w = 0; w < W; ++w)
          if (status[w] < 0) (*solved)[w] = 0;
      }
    } else if (!factored) {
      (*solved)[0] = 0;
    }
  }

  // Make the trial point current for the windows with accept[w].
  mba_status accept(const std::vector<uint8_t>& acc) {
    if (W == 1) {
      std::swap(c->cams64, c->cams64_t);
      std::swap(c->pts64, c->pts64_t);
      c->rcur = 1 - c->rcur;  // the trial state's residuals are now current
      cm_copies(c->rcur);
      // Keep the constant blocks of the trial copy in sync.
      HIP_TRY(hipMemcpyAsync(c->cams64_t, c->cams64, sizeof(double) * 9 * c->nslot, hipMemcpyDeviceToDevice, s));
      return MBA_OK;
    }
    MBA_TRY(stage_upload(c->accept, acc.data(), (size_t)W));
    hipLaunchKernelGGL(k_accept_points, dim3(nblocks(3 * (int64_t)c->ne)), dim3(kBlock), 0, s, c->ne, win(),
                       c->accept, c->pts64_t, c->pts64);
    hipLaunchKernelGGL(k_accept_cams, dim3(nblocks(9 * (int64_t)c->nf)), dim3(kBlock), 0, s, c->nf, win(),
                       d.f_to_cam, c->accept, c->cams64_t, c->cams64);
    return MBA_OK;
  }

  // Trust-region problem dump of the step just computed (single problem):
  // the device state is read back and written out by write_lsqp.
  mba_status dump_step(int k, double radius) {
    if (!c->src) return fail(MBA_ERR_UNSUPPORTED, "trust-region problem dumps are written by mba_solve");
    LsqpState S;
    const size_t np3 = 3 * (size_t)c->ne, nf9 = 9 * (size_t)c->nf;
    S.cams.resize(9 * (size_t)c->nslot);
    S.pts.resize(np3);
    S.pscale.resize(np3);
    S.diag_p.resize(np3);
    S.cscale.resize(nf9);
    S.diag_c.resize(nf9);
    std::vector<T> sp(np3), scm(nf9);
    auto get = [&](void* dst, const void* src, size_t bytes) -> mba_status {
      if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
      return MBA_OK;
    };
    MBA_TRY(get(S.cams.data(), c->cams64, 8 * S.cams.size()));
    MBA_TRY(get(S.pts.data(), c->pts64, 8 * np3));
    MBA_TRY(get(S.pscale.data(), c->pscale64, 8 * np3));
    // The points' LM diagonal: diag(V) clamped, as the pass forms it.
    std::vector<double> vblk(6 * (size_t)c->ne);
    MBA_TRY(get(vblk.data(), d.V, 8 * vblk.size()));
    MBA_TRY(get(S.cscale.data(), c->cscale64, 8 * nf9));
    MBA_TRY(get(S.diag_c.data(), c->diag_c, 8 * nf9));
    MBA_TRY(get(sp.data(), d.step_p, sizeof(T) * np3));
    MBA_TRY(get(scm.data(), d.step_c, sizeof(T) * nf9));
    MBA_TRY(sync());
    const int di[3] = {0, 3, 5};
    for (size_t e = 0; e < (size_t)c->ne; ++e)
      for (int i = 0; i < 3; ++i)
        S.diag_p[3 * e + i] = std::min(std::max(vblk[6 * e + di[i]], c->opt.min_lm_diagonal), c->opt.max_lm_diagonal);
    S.step_p.assign(sp.begin(), sp.end());
    S.step_c.assign(scm.begin(), scm.end());
    S.radius = radius;
    return write_lsqp(c, k, S);
  }

  // Per-window trust-region state.
  struct WinLM {
    double radius = 0.0, decrease_factor = 2.0, cost = 0.0, x_norm = 0.0;
    int num_invalid = 0;
    bool active = true;
    int nrec = 0;
    double rec_min = std::numeric_limits<double>::max();
    mba_iteration it;
  };

  // TrustRegionMinimizer::Minimize + LevenbergMarquardtStrategy, run for
  // every window at once: each window keeps its own radius, decrease
  // factor, cost, records and termination; the passes cover all windows.
  // `its` (W == 1 only) receives the iteration records.
  mba_status solve(mba_summary* sums, mba_iteration* its, int32_t max_its) {
    const mba_options& o = c->opt;
    const double t_start = now_s();
    // The fixed residuals track the state only inside this loop.
    struct ResidualGuard {
      mba_context* c;
      ~ResidualGuard() { c->rvalid = false; }
    } guard{c};
    c->rvalid = false;
    std::vector<WinLM> L(W);
    for (int w = 0; w < W; ++w) {
      std::memset(&sums[w], 0, sizeof(mba_summary));
      sums[w].num_e_blocks = c->e_off[w + 1] - c->e_off[w];
      sums[w].num_f_blocks = c->f_off[w + 1] - c->f_off[w];
      sums[w].termination_type = MBA_NO_CONVERGENCE;
      std::memset(&L[w].it, 0, sizeof(mba_iteration));
      L[w].radius = o.initial_trust_region_radius;
      L[w].it.trust_region_radius = L[w].radius;
    }
    auto push = [&](int w) {
      WinLM& l = L[w];
      if (W == 1 && its && l.nrec < max_its) its[l.nrec] = l.it;
      ++l.nrec;
      l.rec_min = std::min(l.rec_min, l.it.cost);
    };
    auto finish = [&](int w, int32_t term) {
      sums[w].termination_type = term;
      L[w].active = false;
    };
    auto any_active = [&]() {
      for (const WinLM& l : L)
        if (l.active) return true;
      return false;
    };
    auto radii = [&]() {
      std::vector<double> r(W);
      for (int w = 0; w < W; ++w) r[w] = L[w].radius;
      return r;
    };
    double t_jac = 0.0, t_res = 0.0, t_lin = 0.0;
    std::vector<double> cost, gmax, gnorm;
    std::vector<uint8_t> fin;
    double t0 = now_s();
    MBA_TRY(cost64(c->cams64, c->pts64, &cost, &fin, c->rcur));
    c->rvalid = c->rfix;
    t_res += now_s() - t0;
    for (int w = 0; w < W; ++w) {
      if (!fin[w]) {
        finish(w, MBA_FAILURE);
        continue;
      }
      sums[w].initial_cost = cost[w];
      L[w].cost = cost[w];
      L[w].it.cost = cost[w];
    }
    auto activity = [&]() {
      std::vector<uint8_t> a(W);
      for (int w = 0; w < W; ++w) a[w] = L[w].active;
      return a;
    };
    if (any_active()) {
      MBA_TRY(set_active(activity()));
      t0 = now_s();
      tables(c->cams64);
      MBA_TRY(init_scale(o.jacobi_scaling != 0));
      MBA_TRY(set_radii(radii()));
      MBA_TRY(clear_flags());
      jacobian_pass(1);
      MBA_TRY(gradient_async());
      MBA_TRY(x_norm_async());
      MBA_TRY(sync());
      t_jac += now_s() - t0;
      gradient_result(&gmax, &gnorm);
      for (int w = 0; w < W; ++w) {
        if (!L[w].active) continue;
        L[w].it.gradient_max_norm = gmax[w];
        L[w].it.gradient_norm = gnorm[w];
        L[w].it.cumulative_time_s = now_s() - t_start;
        push(w);
        if (gmax[w] <= o.gradient_tolerance) finish(w, MBA_CONVERGENCE);
      }
      for (int w = 0; w < W; ++w) L[w].x_norm = std::sqrt(sum_at(kSlotXnorm, w));
    }
    std::vector<uint8_t> act(W), solved, stepping(W), acc(W), sf(W), cf;
    std::vector<int> lin;
    std::vector<double> mcc(W, 0.0), pgm(W), pgn(W), sn(W), new_cost;
    while (any_active()) {
      for (int w = 0; w < W; ++w)
        if (L[w].active && L[w].it.iteration >= o.max_num_iterations) finish(w, MBA_NO_CONVERGENCE);
      if (!any_active()) break;
      for (int w = 0; w < W; ++w) act[w] = L[w].active;
      MBA_TRY(set_active(act));
      // The step, its model cost change, the trial point and its cost are
      // queued together and read after one sync; what a window whose step
      // turns out invalid computed is not used.
      // Flag slot 0 = failures of the pass or re-damping that built this
      // system (kept since the last cost evaluation) and of its solve: a
      // flagged window's step is invalid, as a failed linear solve is.
      HIP_TRY(hipEventRecord(c->tev[0], s));
      bool factored = true;
      MBA_TRY(compute_step_async(&factored));
      MBA_TRY(flags_async(0));
      HIP_TRY(hipEventRecord(c->tev[1], s));
      MBA_TRY(model_cost_async());
      MBA_TRY(plus_async());
      MBA_TRY(cost64_async(c->cams64_t, c->pts64_t, 1 - c->rcur, false));
      HIP_TRY(hipEventRecord(c->tev[2], s));
      MBA_TRY(sync());
      t_lin += elapsed_s(c->tev[0], c->tev[1]);
      t_res += elapsed_s(c->tev[1], c->tev[2]);
      step_result(act, factored, &solved, &lin);
      cost_result(&new_cost, &cf);
      // TrustRegionMinimizer numbers the dump by the last iteration record.
      if (!c->dump_iters.empty() &&
          std::find(c->dump_iters.begin(), c->dump_iters.end(), L[0].it.iteration) != c->dump_iters.end())
        MBA_TRY(dump_step(L[0].it.iteration, L[0].radius));
      for (int w = 0; w < W; ++w) {
        const double v = sum_at(kSlotPlus, w);
        mcc[w] = -sum_at(kSlotModel, w);
        sn[w] = std::sqrt(v);
        sf[w] = std::isfinite(v);
      }
      for (int w = 0; w < W; ++w) {
        stepping[w] = 0;
        acc[w] = 0;
        if (!act[w]) continue;
        if (flag_at(0, w)) solved[w] = 0;
        sums[w].num_linear_solver_iterations += lin[w];
        mba_iteration& it = L[w].it;
        const int prev = it.iteration;
        pgm[w] = it.gradient_max_norm;
        pgn[w] = it.gradient_norm;
        std::memset(&it, 0, sizeof(it));
        it.iteration = prev + 1;
        it.linear_solver_iterations = lin[w];
      }
      for (int w = 0; w < W; ++w) {
        if (!act[w]) continue;
        mba_iteration& it = L[w].it;
        it.step_is_valid = solved[w] && std::isfinite(mcc[w]) && mcc[w] >= 0.0;
        stepping[w] = 1;
        if (!it.step_is_valid) {
          if (++L[w].num_invalid >= o.max_num_consecutive_invalid_steps) {
            finish(w, MBA_FAILURE);
            stepping[w] = 0;
            continue;
          }
          it.cost = L[w].cost;
          it.gradient_max_norm = pgm[w];
          it.gradient_norm = pgn[w];
        } else {
          L[w].num_invalid = 0;
        }
      }
      bool any_acc = false, any_step = false;
      for (int w = 0; w < W; ++w) {
        if (!stepping[w]) continue;
        WinLM& l = L[w];
        mba_iteration& it = l.it;
        if (it.step_is_valid) {
          const double nc_w = (cf[w] && sf[w]) ? new_cost[w] : std::numeric_limits<double>::max();
          it.step_norm = sn[w];
          if (sn[w] <= o.parameter_tolerance * (l.x_norm + o.parameter_tolerance)) {
            finish(w, MBA_CONVERGENCE);
            stepping[w] = 0;
            continue;
          }
          it.cost_change = l.cost - nc_w;
          if (std::fabs(it.cost_change) <= o.function_tolerance * l.cost) {
            finish(w, MBA_CONVERGENCE);
            stepping[w] = 0;
            continue;
          }
          it.relative_decrease = it.cost_change / mcc[w];
          it.step_is_successful = it.relative_decrease > o.min_relative_decrease;
          if (it.step_is_successful) {
            // StepAccepted: radius update, then the new point is current.
            l.radius = l.radius / std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * it.relative_decrease - 1.0, 3));
            l.radius = std::min(o.max_trust_region_radius, l.radius);
            l.decrease_factor = 2.0;
            l.cost = nc_w;
            acc[w] = 1;
            any_acc = true;
          }
        }
        if (it.step_is_successful) {
          ++sums[w].num_successful_steps;
        } else {
          ++sums[w].num_unsuccessful_steps;
          l.radius = l.radius / l.decrease_factor;  // StepRejected / StepIsInvalid
          l.decrease_factor *= 2.0;
        }
        any_step = true;
      }
      MBA_TRY(set_active(activity()));
      if (any_acc) {
        MBA_TRY(accept(acc));
        MBA_TRY(x_norm_async());
        HIP_TRY(hipEventRecord(c->tev[0], s));
        tables(c->cams64);
        MBA_TRY(set_radii(radii()));
        // A full pass also re-damps the windows whose step was rejected
        // (their point is unchanged).
        jacobian_pass(1);
        HIP_TRY(hipEventRecord(c->tev[1], s));
        MBA_TRY(gradient_async());
        MBA_TRY(sync());
        t_jac += elapsed_s(c->tev[0], c->tev[1]);
        gradient_result(&gmax, &gnorm);
        for (int w = 0; w < W; ++w)
          if (acc[w]) {
            L[w].x_norm = std::sqrt(sum_at(kSlotXnorm, w));
            L[w].it.gradient_max_norm = gmax[w];
            L[w].it.gradient_norm = gnorm[w];
          }
      } else if (any_step) {
        MBA_TRY(set_radii(radii()));
        damp();
      }
      for (int w = 0; w < W; ++w) {
        if (!stepping[w]) continue;
        WinLM& l = L[w];
        l.it.cost = l.cost;
        l.it.trust_region_radius = l.radius;
        l.it.cumulative_time_s = now_s() - t_start;
        push(w);
        if (l.it.step_is_successful) {
          if (l.it.gradient_max_norm <= o.gradient_tolerance) finish(w, MBA_CONVERGENCE);
        } else if (l.radius < o.min_trust_region_radius) {
          finish(w, MBA_CONVERGENCE);
        }
      }
    }
    const double total = now_s() - t_start;
    for (int w = 0; w < W; ++w) {
      mba_summary& sm = sums[w];
      sm.final_cost = std::min(sm.initial_cost, L[w].rec_min);
      if (sm.termination_type == MBA_FAILURE && L[w].nrec == 0) sm.final_cost = sm.initial_cost;
      sm.num_iterations = L[w].nrec;
      sm.is_solution_usable = sm.termination_type != MBA_FAILURE;
      sm.jacobian_evaluation_time_s = t_jac;
      sm.residual_evaluation_time_s = t_res;
      sm.linear_solver_time_s = t_lin;
      sm.total_time_s = total;
    }
    return sync();
  }
};

// Copy the current parameters of every window back to its problem (all
// windows, or those whose summary says the solution is usable).
mba_status download(mba_context* c, mba_problem* probs, const mba_summary* sums) {
  std::vector<double> cams(9 * (size_t)c->nslot), pts(3 * (size_t)c->ne);
  if (!cams.empty()) HIP_TRY(hipMemcpyAsync(cams.data(), c->cams64, cams.size() * 8, hipMemcpyDeviceToHost, c->stream));
  if (!pts.empty()) HIP_TRY(hipMemcpyAsync(pts.data(), c->pts64, pts.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  auto wanted = [&](int32_t w) { return !sums || sums[w].is_solution_usable; };
  if (c->grouped) {
    if (!wanted(0)) return MBA_OK;
    for (int32_t f = 0; f < c->nf; ++f) {
      const int32_t b = c->f_to_cam[f];
      double* dst = b < c->nc ? probs[0].cameras + 6 * (size_t)b : probs[0].intrinsics + 6 * (size_t)(b - c->nc);
      std::memcpy(dst, cams.data() + 9 * (size_t)b, 48);
    }
    for (int32_t e = 0; e < c->ne; ++e)
      std::memcpy(probs[0].points + 3 * (size_t)c->dev_e_to_pt[e], pts.data() + 3 * (size_t)e, 24);
    return MBA_OK;
  }
  int32_t w = 0;
  for (int32_t f = 0; f < c->nf; ++f) {
    while (f >= c->f_off[w + 1]) ++w;
    if (!wanted(w)) continue;
    const int32_t cam = c->f_to_cam[f];
    std::memcpy(probs[w].cameras + 9 * (size_t)(cam - c->cam_off[w]), cams.data() + 9 * (size_t)cam, 72);
  }
  // Window of each global point by its offset.
  for (int32_t e = 0; e < c->ne; ++e) {
    const int32_t q = c->dev_e_to_pt[e];
    const int32_t wq = (int32_t)(std::upper_bound(c->pt_off.begin(), c->pt_off.end(), q) - c->pt_off.begin()) - 1;
    if (!wanted(wq)) continue;
    std::memcpy(probs[wq].points + 3 * (size_t)(q - c->pt_off[wq]), pts.data() + 3 * (size_t)e, 24);
  }
  return MBA_OK;
}

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
  return context_create(problem, 1, options, out);
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
  float e[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  HIP_TRY(hipEventElapsedTime(&e[0], ctx->ev[0], ctx->ev[1]));
  HIP_TRY(hipEventElapsedTime(&e[1], ctx->ev[1], ctx->ev[2]));
  HIP_TRY(hipEventElapsedTime(&e[2], ctx->ev[0], ctx->ev[3]));
  if (ctx->timed_explicit) {
    HIP_TRY(hipEventElapsedTime(&e[3], ctx->ev[4], ctx->ev[5]));
    HIP_TRY(hipEventElapsedTime(&e[4], ctx->ev[5], ctx->ev[6]));
    HIP_TRY(hipEventElapsedTime(&e[5], ctx->ev[6], ctx->ev[7]));
  }
  for (int i = 0; i < n && i < 6; ++i) ms[i] = e[i];
  return MBA_OK;
}

mba_status mba_context_jacobian_schur(mba_context* ctx, double radius) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  auto run = [&](auto& eng) -> mba_status {
    if (!ctx->scale_ready) {
      eng.tables(ctx->cams64);
      MBA_TRY(eng.init_scale(ctx->opt.jacobi_scaling != 0));
    }
    eng.tables(ctx->cams64);
    MBA_TRY(eng.set_radius(radius));
    eng.jacobian_pass(1);
    return MBA_OK;
  };
  if (ctx->mixed) {
    Engine<float> e(ctx);
    return run(e);
  }
  Engine<double> e(ctx);
  return run(e);
}

mba_status mba_context_lm_pass(mba_context* ctx, double radius, double* cost) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  auto run = [&](auto& eng) -> mba_status {
    std::vector<double> c;
    std::vector<uint8_t> fin;
    // As Engine::solve starts an iteration: the FP64 cost of the state (its
    // residuals, rounded once, kept for the FP32 kernels), then the pass.
    MBA_TRY(eng.cost64(ctx->cams64, ctx->pts64, &c, &fin, ctx->rcur));
    ctx->rvalid = ctx->rfix;
    eng.tables(ctx->cams64);
    if (!ctx->scale_ready) MBA_TRY(eng.init_scale(ctx->opt.jacobi_scaling != 0));
    MBA_TRY(eng.set_radius(radius));
    eng.jacobian_pass(1);
    ctx->rvalid = false;
    if (cost) *cost = c.empty() ? 0.0 : c[0];
    return MBA_OK;
  };
  mba_status st;
  if (ctx->mixed) {
    Engine<float> e(ctx);
    st = run(e);
  } else {
    Engine<double> e(ctx);
    st = run(e);
  }
  ctx->rvalid = false;
  return st;
}

mba_status mba_context_read_scale(mba_context* ctx, double* cam_scale, double* pt_scale);

mba_status mba_context_read_schur(mba_context* ctx, double* U, double* gc, double* rhs, double* V, double* gp) {
  if (!ctx) return fail(MBA_ERR_INVALID_ARGUMENT, "ctx is NULL");
  if (ctx->grouped) return fail(MBA_ERR_UNSUPPORTED, "read_schur reads packed-camera blocks");
  HIP_TRY(hipSetDevice(ctx->device));
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
      const int32_t cam = ctx->f_to_cam[f];
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
  if (ctx->grouped) return fail(MBA_ERR_UNSUPPORTED, "read_scale reads packed-camera blocks");
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<double> cs(9 * (size_t)ctx->nf), ps(3 * (size_t)ctx->ne);
  if (!cs.empty()) HIP_TRY(hipMemcpyAsync(cs.data(), ctx->cscale64, cs.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (!ps.empty()) HIP_TRY(hipMemcpyAsync(ps.data(), ctx->pscale64, ps.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (cam_scale) {
    for (int64_t i = 0; i < 9 * (int64_t)ctx->nc; ++i) cam_scale[i] = 1.0;
    for (int32_t f = 0; f < ctx->nf; ++f)
      std::memcpy(cam_scale + 9 * (size_t)ctx->f_to_cam[f], cs.data() + 9 * (size_t)f, 72);
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
  if (ctx->grouped) return fail(MBA_ERR_UNSUPPORTED, "schur_matvec takes packed-camera vectors");
  HIP_TRY(hipSetDevice(ctx->device));
  auto go = [&](auto& eng, auto* dx, auto* dy) -> mba_status {
    using T = std::remove_pointer_t<decltype(dx)>;
    std::vector<T> hx(9 * (size_t)ctx->nf), hy(9 * (size_t)ctx->nf);
    for (int32_t f = 0; f < ctx->nf; ++f)
      for (int k = 0; k < 9; ++k) hx[9 * (size_t)f + k] = (T)x[9 * (size_t)ctx->f_to_cam[f] + k];
    if (!hx.empty()) HIP_TRY(hipMemcpyAsync(dx, hx.data(), hx.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    eng.matvec(dx, dy, nullptr);
    if (!hy.empty()) HIP_TRY(hipMemcpyAsync(hy.data(), dy, hy.size() * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < 9 * (int64_t)ctx->nc; ++i) y[i] = 0.0;
    for (int32_t f = 0; f < ctx->nf; ++f)
      for (int k = 0; k < 9; ++k) y[9 * (size_t)ctx->f_to_cam[f] + k] = (double)hy[9 * (size_t)f + k];
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
  return download(ctx, p, nullptr);
}

mba_status mba_solve(const mba_options* options, mba_problem* problem, mba_summary* summary,
                     mba_iteration* iterations, int32_t max_iterations) {
  g_last_error.clear();
  if (!summary) return fail(MBA_ERR_INVALID_ARGUMENT, "summary is NULL");
  std::memset(summary, 0, sizeof(*summary));
  const double t0 = now_s();
  mba_context* ctx = nullptr;
  mba_status s = context_create(problem, 1, options, &ctx);
  if (s != MBA_OK) return s;
  ctx->src = problem;
  const double t_pre = now_s() - t0;
  s = mba_context_solve(ctx, summary, iterations, max_iterations);
  if (s == MBA_OK && summary->is_solution_usable) s = mba_context_download(ctx, problem);
  summary->preprocessor_time_s = t_pre;
  summary->total_time_s += t_pre;
  context_free(ctx);
  return s;
}

mba_status mba_solve_batched(const mba_options* options, mba_problem* problems, int32_t num_problems,
                             mba_summary* summaries) {
  g_last_error.clear();
  if (num_problems < 0) return fail(MBA_ERR_INVALID_ARGUMENT, "num_problems < 0");
  if (num_problems == 0) return MBA_OK;
  if (!problems || !summaries) return fail(MBA_ERR_INVALID_ARGUMENT, "problems or summaries is NULL");
  std::memset(summaries, 0, sizeof(mba_summary) * (size_t)num_problems);
  mba_options o;
  if (options) o = *options; else mba_default_options(&o);
  if (o.num_dump_iterations > 0 && num_problems > 1)
    return fail(MBA_ERR_UNSUPPORTED, "trust-region problem dumps are written for a single problem");
  bool any_grouped = false, any_large = false;
  for (int32_t w = 0; w < num_problems; ++w) {
    any_grouped = any_grouped || problems[w].layout == MBA_LAYOUT_GROUPED;
    any_large = any_large || problems[w].num_cameras > kWinMaxF;
  }
  if (any_grouped || (o.linear_solver == MBA_EXPLICIT_SCHUR && any_large)) {
    // Grouped-intrinsics problems, and dense reduced systems larger than a
    // batched explicit solve factors, are solved one window at a time.
    for (int32_t w = 0; w < num_problems; ++w) {
      const mba_status s = mba_solve(&o, &problems[w], &summaries[w], nullptr, 0);
      if (s != MBA_OK) return fail(s, "window " + std::to_string(w) + ": " + g_last_error);
    }
    return MBA_OK;
  }
  const double t0 = now_s();
  mba_context* ctx = nullptr;
  mba_status s = context_create(problems, num_problems, &o, &ctx);
  if (s != MBA_OK) return s;
  const double t_pre = now_s() - t0;
  s = hipSetDevice(ctx->device) == hipSuccess ? MBA_OK : fail(MBA_ERR_DEVICE, "hipSetDevice failed");
  if (s == MBA_OK) {
    if (ctx->mixed) {
      Engine<float> e(ctx);
      s = e.solve(summaries, nullptr, 0);
    } else {
      Engine<double> e(ctx);
      s = e.solve(summaries, nullptr, 0);
    }
  }
  if (s == MBA_OK) s = download(ctx, problems, summaries);
  for (int32_t w = 0; w < num_problems; ++w) {
    summaries[w].preprocessor_time_s = t_pre;
    summaries[w].total_time_s += t_pre;
  }
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
static mba_status evaluate_impl(const mba_problem* p, double huber_a, double* residuals, double* jac_intr,
                                double* jac_camera, double* jac_point, double* cost, int32_t device) {
  MBA_TRY(validate(p));
  const bool grouped = p->layout == MBA_LAYOUT_GROUPED;
  const int64_t n = p->num_observations;
  const int32_t nc = p->num_cameras, np = p->num_points, ng = grouped ? p->num_intrinsics : 0;
  // FP64 slot blocks as the solver holds them: cameras, then groups.
  std::vector<double> slots(9 * ((size_t)nc + ng), 0.0);
  std::vector<int32_t> cam_grp;
  if (grouped) {
    cam_grp.assign(nc, -1);
    for (int64_t i = 0; i < n; ++i) {
      const int32_t cc = p->camera_index[i], g = p->intrinsic_index[i];
      if (cam_grp[cc] < 0) cam_grp[cc] = g;
      else if (cam_grp[cc] != g)
        return fail(MBA_ERR_UNSUPPORTED, "camera " + std::to_string(cc) + " is observed with two intrinsic groups");
    }
    for (int32_t cc = 0; cc < nc; ++cc)
      for (int k = 0; k < 6; ++k) slots[9 * (size_t)cc + k] = p->cameras[6 * (size_t)cc + k];
    for (int32_t g = 0; g < ng; ++g)
      for (int k = 0; k < 6; ++k) slots[9 * ((size_t)nc + g) + k] = p->intrinsics[6 * (size_t)g + k];
  } else if (nc) {
    std::memcpy(slots.data(), p->cameras, 72 * (size_t)nc);
  }
  const int jcw = grouped ? 12 : 18;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(MBA_ERR_DEVICE, "no such HIP device");
  HIP_TRY(hipSetDevice(device));
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  Pool pool;
  double* cams = nullptr;
  double *res = nullptr, *ji = nullptr, *jc = nullptr, *jp = nullptr;
  T *tab = nullptr, *pts = nullptr, *obs = nullptr;
  int32_t *ci = nullptr, *pi = nullptr, *cg = nullptr;
  auto run = [&]() -> mba_status {
    MBA_TRY(pool.alloc(&cams, slots.size()));
    MBA_TRY(pool.alloc(&tab, kCamTab * (size_t)nc));
    MBA_TRY(pool.alloc(&pts, 3 * (size_t)np));
    MBA_TRY(pool.alloc(&obs, 2 * (size_t)n));
    MBA_TRY(pool.alloc(&ci, (size_t)n));
    MBA_TRY(pool.alloc(&pi, (size_t)n));
    MBA_TRY(pool.alloc(&res, 2 * (size_t)n));
    if (grouped) MBA_TRY(pool.alloc(&cg, (size_t)nc));
    if (jac_intr && grouped) MBA_TRY(pool.alloc(&ji, 12 * (size_t)n));
    if (jac_camera) MBA_TRY(pool.alloc(&jc, jcw * (size_t)n));
    if (jac_point) MBA_TRY(pool.alloc(&jp, 6 * (size_t)n));
    std::vector<T> hp(p->points, p->points + 3 * (size_t)np), ho(p->observations, p->observations + 2 * n);
    if (!slots.empty()) HIP_TRY(hipMemcpyAsync(cams, slots.data(), slots.size() * 8, hipMemcpyHostToDevice, s));
    if (grouped && nc) HIP_TRY(hipMemcpyAsync(cg, cam_grp.data(), 4 * (size_t)nc, hipMemcpyHostToDevice, s));
    if (np) HIP_TRY(hipMemcpyAsync(pts, hp.data(), hp.size() * sizeof(T), hipMemcpyHostToDevice, s));
    if (n) {
      HIP_TRY(hipMemcpyAsync(obs, ho.data(), ho.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(ci, p->camera_index, 4 * (size_t)n, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(pi, p->point_index, 4 * (size_t)n, hipMemcpyHostToDevice, s));
    }
    if (nc) {
      if (grouped)
        hipLaunchKernelGGL(k_cam_table_g<T>, dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, (const double*)cams,
                           (const int32_t*)cg, tab);
      else
        hipLaunchKernelGGL(k_cam_table<T>, dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, cams, tab, (T*)nullptr);
    }
    if (n) {
      if (grouped)
        hipLaunchKernelGGL(k_eval_obs_g<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, ci, pi, obs, tab, pts, res, ji,
                           jc, jp);
      else
        hipLaunchKernelGGL(k_eval_obs<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, 0, 0, (const int32_t*)ci,
                           (const int32_t*)pi, (const T*)obs, (const T*)tab, (const T*)pts, res, jc, jp, 0.0,
                           (double*)nullptr);
    }
    std::vector<double> hr(2 * (size_t)n);
    if (n) HIP_TRY(hipMemcpyAsync(hr.data(), res, hr.size() * 8, hipMemcpyDeviceToHost, s));
    if (ji && n) HIP_TRY(hipMemcpyAsync(jac_intr, ji, 12 * 8 * (size_t)n, hipMemcpyDeviceToHost, s));
    if (jc && n) HIP_TRY(hipMemcpyAsync(jac_camera, jc, jcw * 8 * (size_t)n, hipMemcpyDeviceToHost, s));
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

// TEXTFILE dump of the linear least-squares problem of a step
// (DumpLinearLeastSquaresProblemToTextFile, linear_least_squares_problems.cc:
// 534-603; BlockSparseMatrix::ToTextFile, block_sparse_matrix.cc:222-243):
// A = the reduced program's Jacobian, robustified (Corrector with rho'' <= 0:
// rows times sqrt(rho'), corrector.cc:41-85) and Jacobi-scaled, row blocks
// in residual order, each row's cells by column block (E blocks first),
// local columns of each block; D = sqrt(clamped column norm^2 / radius);
// b = robustified residuals; x = the step.
mba_status write_lsqp(mba_context* c, int k, const LsqpState& S) {
  const mba_problem& p0 = *c->src;
  const Structure& st = c->wst[0];
  const bool grouped = c->grouped;
  const int32_t nc = p0.num_cameras, np = p0.num_points, ng = grouped ? p0.num_intrinsics : 0;
  const int64_t n = p0.num_observations;
  const int cw = grouped ? 6 : 9;
  // Current parameters in the problem's layout.
  std::vector<double> cams((size_t)cw * nc), intr(6 * (size_t)ng), pts(p0.points, p0.points + 3 * (size_t)np);
  for (int32_t cc = 0; cc < nc; ++cc)
    std::memcpy(cams.data() + (size_t)cw * cc, S.cams.data() + 9 * (size_t)cc, 8 * (size_t)cw);
  for (int32_t g = 0; g < ng; ++g)
    std::memcpy(intr.data() + 6 * (size_t)g, S.cams.data() + 9 * ((size_t)nc + g), 48);
  for (int32_t e = 0; e < c->ne; ++e)
    std::memcpy(pts.data() + 3 * (size_t)c->dev_e_to_pt[e], S.pts.data() + 3 * (size_t)e, 24);
  mba_problem q = p0;
  q.cameras = cams.data();
  q.points = pts.data();
  if (grouped) q.intrinsics = intr.data();
  std::vector<double> r(2 * (size_t)n), jc((size_t)2 * cw * n), jp(6 * (size_t)n), ji(grouped ? 12 * (size_t)n : 0);
  MBA_TRY(evaluate_impl<double>(&q, 0.0, r.data(), grouped ? ji.data() : nullptr, jc.data(), jp.data(), nullptr,
                                c->device));
  // Columns: E blocks (3 each) in elimination order, then the F blocks'
  // local columns (the free entries of each slot block) in F order.
  const int32_t ne = st.num_e, nslot = c->nslot;
  std::vector<int32_t> pt_de(np, -1), pt_rank(np, -1), e_pt(ne);
  for (int32_t e = 0; e < c->ne; ++e) pt_de[c->dev_e_to_pt[e]] = e;
  for (int32_t i = 0; i < ne; ++i) {
    const int32_t pt = st.order[i] - (grouped ? ng + nc : nc);
    pt_rank[pt] = i;
    e_pt[i] = pt;
  }
  std::vector<int32_t> slot_rank(nslot, -1), slot_col(nslot, 0), slot_bits(nslot, 0), f_slot;
  int64_t ncols = 3 * (int64_t)ne;
  for (size_t j = ne; j < st.order.size(); ++j) {
    const int32_t id = st.order[j];
    const int32_t b = grouped ? (id < ng ? nc + id : id - ng) : id;
    int32_t bits;
    if (grouped && b >= nc) bits = 0x3F & ~(p0.intrinsic_subset ? (int32_t)p0.intrinsic_subset[b - nc] : 0);
    else bits = p0.camera_params ? p0.camera_params : (grouped ? 0x3F : 0x1FF);
    slot_rank[b] = (int32_t)(j - ne);
    slot_col[b] = (int32_t)ncols;
    slot_bits[b] = bits;
    f_slot.push_back(b);
    ncols += __builtin_popcount((unsigned)bits);
  }
  const int64_t nres = (int64_t)st.res_order.size();
  const int64_t nrows = 2 * nres;
  std::string base = c->dump_dir;
  char name[64];
  std::snprintf(name, sizeof(name), "ceres_solver_iteration_%03d", k);
  if (!base.empty() && base.back() != '/') base += '/';
  base += name;
  auto open = [&](const std::string& f, FILE** fp) -> mba_status {
    *fp = std::fopen(f.c_str(), "w");
    return *fp ? MBA_OK : fail(MBA_ERR_INVALID_ARGUMENT, "cannot write " + f);
  };
  const double a = c->opt.huber_a;
  std::vector<double> w(nres);
  {
    FILE* fa = nullptr;
    MBA_TRY(open(base + "_A.txt", &fa));
    struct Cell {
      int32_t id, col, size;
      double v[18];
    };
    for (int64_t row = 0; row < nres; ++row) {
      const int64_t i = st.res_order[row];
      const double r0 = r[2 * i], r1 = r[2 * i + 1], sq = r0 * r0 + r1 * r1;
      w[row] = (a > 0.0 && sq > a * a) ? std::sqrt(std::max(std::numeric_limits<double>::min(), a / std::sqrt(sq)))
                                       : 1.0;
      Cell cells[3];
      int nce = 0;
      const int32_t pt = p0.point_index[i];
      if (pt_rank[pt] >= 0) {
        Cell& ce = cells[nce++];
        ce.id = pt_rank[pt];
        ce.col = 3 * pt_rank[pt];
        ce.size = 3;
        for (int rr = 0; rr < 2; ++rr)
          for (int q3 = 0; q3 < 3; ++q3)
            ce.v[3 * rr + q3] = jp[6 * i + 3 * rr + q3] * S.pscale[3 * (size_t)pt_de[pt] + q3] * w[row];
      }
      auto add_f = [&](int32_t b, const double* J, int width) {
        if (slot_rank[b] < 0) return;
        const int32_t f = c->cam_to_f[b];
        Cell& ce = cells[nce++];
        ce.id = ne + slot_rank[b];
        ce.col = slot_col[b];
        ce.size = __builtin_popcount((unsigned)slot_bits[b]);
        for (int rr = 0, m = 0; rr < 2; ++rr)
          for (int kk = 0; kk < 9; ++kk)
            if ((slot_bits[b] >> kk) & 1) ce.v[m++] = J[width * rr + kk] * S.cscale[9 * (size_t)f + kk] * w[row];
      };
      add_f(p0.camera_index[i], jc.data() + (size_t)2 * cw * i, cw);
      if (grouped) add_f(nc + p0.intrinsic_index[i], ji.data() + 12 * (size_t)i, 6);
      std::sort(cells, cells + nce, [](const Cell& x, const Cell& y) { return x.id < y.id; });
      for (int t = 0; t < nce; ++t)
        for (int rr = 0; rr < 2; ++rr)
          for (int q = 0; q < cells[t].size; ++q)
            std::fprintf(fa, "% 10d % 10d %17f\n", (int)(2 * row + rr), cells[t].col + q,
                         cells[t].v[cells[t].size * rr + q]);
    }
    std::fclose(fa);
  }
  auto write_array = [&](const std::string& f, const std::vector<double>& v) -> mba_status {
    FILE* fp = nullptr;
    MBA_TRY(open(f, &fp));
    for (double x : v) std::fprintf(fp, "%17f\n", x);
    std::fclose(fp);
    return MBA_OK;
  };
  std::vector<double> D, b, x;
  D.reserve(ncols);
  x.reserve(ncols);
  for (int32_t i = 0; i < ne; ++i)
    for (int q3 = 0; q3 < 3; ++q3) {
      const size_t at = 3 * (size_t)pt_de[e_pt[i]] + q3;
      D.push_back(std::sqrt(S.diag_p[at] / S.radius));
      x.push_back(S.step_p[at]);
    }
  for (int32_t bs : f_slot) {
    const int32_t f = c->cam_to_f[bs];
    for (int kk = 0; kk < 9; ++kk)
      if ((slot_bits[bs] >> kk) & 1) {
        D.push_back(std::sqrt(S.diag_c[9 * (size_t)f + kk] / S.radius));
        x.push_back(S.step_c[9 * (size_t)f + kk]);
      }
  }
  b.reserve(nrows);
  for (int64_t row = 0; row < nres; ++row) {
    const int64_t i = st.res_order[row];
    b.push_back(r[2 * i] * w[row]);
    b.push_back(r[2 * i + 1] * w[row]);
  }
  MBA_TRY(write_array(base + "_D.txt", D));
  MBA_TRY(write_array(base + "_b.txt", b));
  MBA_TRY(write_array(base + "_x.txt", x));
  FILE* fm = nullptr;
  MBA_TRY(open(base + ".m", &fm));
  std::fprintf(fm, "function lsqp = load_trust_region_problem()\n");
  std::fprintf(fm, "lsqp.num_rows = %d;\n", (int)nrows);
  std::fprintf(fm, "lsqp.num_cols = %d;\n", (int)ncols);
  std::fprintf(fm, "tmp = load('%s_A.txt', '-ascii');\n", base.c_str());
  std::fprintf(fm, "lsqp.A = sparse(tmp(:, 1) + 1, tmp(:, 2) + 1, tmp(:, 3), %d, %d);\n", (int)nrows, (int)ncols);
  std::fprintf(fm, "lsqp.D = load('%s_D.txt', '-ascii');\n", base.c_str());
  std::fprintf(fm, "lsqp.b = load('%s_b.txt', '-ascii');\n", base.c_str());
  std::fprintf(fm, "lsqp.x = load('%s_x.txt', '-ascii');\n", base.c_str());
  std::fclose(fm);
  return MBA_OK;
}

// ------------------------------------------ evaluation without allocation
// Device scratch of mba_evaluate_noalloc: the camera table (kCamTab values
// of the kernel precision per camera) followed by the block partials of the
// cost.  Each device holds its own copy until the process ends; every call
// on a device shares it, so calls that overlap on one device race.
constexpr int kNoAllocMaxCameras = 16384;
constexpr int kNoAllocBlocks = 2048;  // grid of the evaluation kernel
constexpr size_t kNoAllocTable = (size_t)kCamTab * kNoAllocMaxCameras;  // doubles
static __device__ __align__(256) double scratch[kNoAllocTable + kNoAllocBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device `dev` (looked up
// once per device).  Host code uses this address, never the array's name.
mba_status scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(MBA_ERR_DEVICE, "no device address for the evaluation scratch");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return MBA_OK;
}

template <typename T>
mba_status evaluate_noalloc_launch(double* scr, int32_t nc, int32_t np, int64_t n, const double* cams,
                                   const double* pts, const double* obs, const int32_t* ci, const int32_t* pi,
                                   double huber_a, double* res, double* jc, double* jp, double* cost,
                                   hipStream_t s) {
  T* tab = reinterpret_cast<T*>(scr);
  double* partial = scr + kNoAllocTable;
  if (nc > 0) hipLaunchKernelGGL(k_cam_table<T>, dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, cams, tab, (T*)nullptr);
  const int nb = (int)std::min<int64_t>(kNoAllocBlocks, std::max<int64_t>(1, (n + kBlock - 1) / kBlock));
  if (n > 0 && (res || jc || jp || cost))
    hipLaunchKernelGGL((k_eval_obs<T, double>), dim3(nb), dim3(kBlock), 0, s, n, (int)nc, (int)np, ci, pi, obs,
                       (const T*)tab, pts, res, jc, jp, huber_a, cost ? partial : nullptr);
  if (cost) {
    if (n > 0) hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, (const double*)partial, nb, cost);
    else hipLaunchKernelGGL(k_fill64, dim3(1), dim3(kBlock), 0, s, (int64_t)1, 0.0, cost);
  }
  HIP_TRY(hipGetLastError());
  return MBA_OK;
}
}  // namespace

extern "C" {

mba_status mba_evaluate_noalloc(int32_t num_cameras, int32_t num_points, int64_t num_observations,
                                const double* cameras, const double* points, const double* observations,
                                const int32_t* camera_index, const int32_t* point_index, int32_t precision,
                                double huber_a, double* residuals, double* jac_camera, double* jac_point,
                                double* cost, int32_t device, void* stream) {
  g_last_error.clear();
  if (num_cameras < 0 || num_points < 0 || num_observations < 0)
    return fail(MBA_ERR_INVALID_ARGUMENT, "negative problem size");
  if (num_cameras > kNoAllocMaxCameras)
    return fail(MBA_ERR_INVALID_ARGUMENT, "mba_evaluate_noalloc: " + std::to_string(num_cameras) +
                                              " cameras exceed its scratch limit of " +
                                              std::to_string(kNoAllocMaxCameras));
  if (precision != MBA_PRECISION_FP64 && precision != MBA_PRECISION_MIXED)
    return fail(MBA_ERR_INVALID_ARGUMENT, "unknown precision");
  if ((num_cameras > 0 && !cameras) || (num_points > 0 && !points))
    return fail(MBA_ERR_INVALID_ARGUMENT, "NULL camera or point array");
  if (num_observations > 0 && (!observations || !camera_index || !point_index))
    return fail(MBA_ERR_INVALID_ARGUMENT, "NULL observation or index array");
  if (num_observations > 0 && (num_cameras == 0 || num_points == 0))
    return fail(MBA_ERR_INVALID_ARGUMENT, "observations without cameras or points");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev || device >= kMaxDevices) return fail(MBA_ERR_DEVICE, "no such HIP device");
  HIP_TRY(hipSetDevice(device));
  double* scr = nullptr;
  MBA_TRY(scratch_address(device, &scr));
  const hipStream_t s = (hipStream_t)stream;
  if (precision == MBA_PRECISION_MIXED)
    return evaluate_noalloc_launch<float>(scr, num_cameras, num_points, num_observations, cameras, points,
                                          observations, camera_index, point_index, huber_a, residuals, jac_camera,
                                          jac_point, cost, s);
  return evaluate_noalloc_launch<double>(scr, num_cameras, num_points, num_observations, cameras, points,
                                         observations, camera_index, point_index, huber_a, residuals, jac_camera,
                                         jac_point, cost, s);
}

mba_status mba_evaluate(const mba_problem* p, int32_t precision, double huber_a, double* residuals,
                        double* jac_camera, double* jac_point, double* cost, int32_t device) {
  return mba_evaluate_blocks(p, precision, huber_a, residuals, nullptr, jac_camera, jac_point, cost, device);
}

mba_status mba_evaluate_blocks(const mba_problem* p, int32_t precision, double huber_a, double* residuals,
                               double* jac_intrinsics, double* jac_camera, double* jac_point, double* cost,
                               int32_t device) {
  g_last_error.clear();
  if (precision == MBA_PRECISION_MIXED)
    return evaluate_impl<float>(p, huber_a, residuals, jac_intrinsics, jac_camera, jac_point, cost, device);
  return evaluate_impl<double>(p, huber_a, residuals, jac_intrinsics, jac_camera, jac_point, cost, device);
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

mba_status mba_scene_generate_grouped(int32_t preset, int32_t* num_cameras, int32_t* num_points,
                                      int32_t* num_intrinsics, int64_t* num_observations, double* cam_init,
                                      double* intr_init, double* pt_init, double* observations,
                                      int32_t* camera_index, int32_t* intrinsic_index, int32_t* point_index) {
  if (preset != 6) return fail(MBA_ERR_INVALID_ARGUMENT, "the grouped scene preset is 6");
  const mba_scene_spec sp = mba_scene_preset(preset, 0);
  std::vector<int32_t> tl(sp.num_points);
  const int64_t n = mba_scene_count(&sp, tl.data());
  if (num_cameras) *num_cameras = sp.num_cameras;
  if (num_points) *num_points = sp.num_points;
  if (num_intrinsics) *num_intrinsics = MBA_SCENE_GROUPS;
  if (num_observations) *num_observations = n;
  if (!cam_init || !intr_init || !pt_init || !observations || !camera_index || !intrinsic_index || !point_index)
    return MBA_OK;
  std::vector<int64_t> off((size_t)sp.num_points + 1);
  mba_scene_fill_groups(&sp, MBA_SCENE_GROUPS, tl.data(), nullptr, cam_init, nullptr, intr_init, nullptr, pt_init,
                        observations, camera_index, intrinsic_index, point_index, off.data());
  return MBA_OK;
}

}  // extern "C"
