// HIP kernels of the bundle-adjustment engine (gfx950, wave64).
//
// Precision: every per-observation kernel is templated on T (float for the
// mixed path, double for the FP64 path).  Point-block inverses, the LM cost
// and the model-cost change are always FP64.
//
// Data layout (all in HBM, built once per problem by the host runtime):
//   rows      -- observations in the reference's residual order (point-major:
//                E blocks in ordering order, each bucket reversed), so a
//                point's rows are contiguous;
//   cm rows   -- the same observations sorted by camera (F block) and cut
//                into chunks of <= chunk_rows rows; one workgroup reduces one
//                chunk, so per-camera 9x9 sums need no global atomics;
//   camera table -- per camera R, dR/daa_k, t and the intrinsics,
//                recomputed from the FP64 parameters whenever they change;
//                the Jacobian is recomputed from it in every pass instead of
//                being stored.
//
// Layouts (template parameter L of the per-observation kernels): L = 0 is
// the packed camera [aa, t, f, k1, k2] (principal point 0, k3 = 0); L = 1 is
// the grouped layout, camera [aa, t] plus a shared intrinsic group [f, ppx,
// ppy, k1, k2, k3].  Every F block (a camera, or a group in L = 1) keeps 9
// slots in F-space vectors; a grouped camera uses slots [aa, t] and a group
// [f, ppx, ppy, k1, k2, k3], the rest have zero Jacobian columns.
#ifndef MBA_DEVICE_H_
#define MBA_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>

namespace mba {

// Camera table row: the packed projection part first, so a row-wise gather
// of it touches one 64-byte (FP32) / 128-byte (FP64) line:
//   R 9 | t 3 | f k1 k2 k3 | dR 27 | ppx ppy | pad 3
// (k3, ppx, ppy are zero for the packed layout, whose kernels never read them).
constexpr int kCamTab = 48;
constexpr int kTabT = 9, kTabF = 12, kTabK3 = 15, kTabDR = 16, kTabPP = 43;
constexpr int kDirTab = 16;   // M 9 | dt 3 | df dk1 dk2 | pad
constexpr int kDirTabG = 24;  // M 9 | dt 3 | df dppx dppy dk1 dk2 dk3 | pad
template <int L>
__host__ __device__ constexpr int dir_stride() { return L == 0 ? kDirTab : kDirTabG; }
constexpr int kPartial = 64; // U 45 | g 9 | h 9 | pad
constexpr int kBlock = 256;

struct Scalars {
  double sum;
  double aux;
  int32_t flag;
  int32_t pad;
};

struct PcgState {
  double rho, rho_last, Q0, alpha, beta, pq, norm_b, tol_r, zeta, norm_r;
  int32_t iter, done, status, pad;  // status: 0 ok, 1 no convergence, -1 failure
};

// Point-major launch layout: points (E blocks, device order) are grouped in
// segments by track-length class; a segment gives each point `lanes` lanes
// (1..64, a power of two), and its blocks follow the previous segment's.
constexpr int kMaxSeg = 8;
struct Seg {
  int32_t n;
  int32_t e0[kMaxSeg], e1[kMaxSeg], lanes[kMaxSeg], blk0[kMaxSeg + 1];
};

struct SegPos {
  int e, sub, lanes;
  bool valid;
};

// Position of thread `tid` of (virtual) block b of `bdim` threads.
__device__ __forceinline__ SegPos seg_locate_at(const Seg& sg, int b, int tid, int bdim) {
  int k = 0;
  while (k + 1 < sg.n && b >= sg.blk0[k + 1]) ++k;
  SegPos p;
  p.lanes = sg.lanes[k];
  const int t = (b - sg.blk0[k]) * bdim + tid;
  p.e = sg.e0[k] + t / p.lanes;
  p.sub = t % p.lanes;
  p.valid = p.e < sg.e1[k];
  return p;
}
__device__ __forceinline__ SegPos seg_locate(const Seg& sg) {
  return seg_locate_at(sg, blockIdx.x, threadIdx.x, blockDim.x);
}

// acc += a b + c d as two chained FMAs (the sum form would round a b + c d
// first and add it separately).
template <typename T>
__device__ __forceinline__ void fma2(T& acc, T a, T b, T c, T d) {
  acc = fma(a, b, acc);
  acc = fma(c, d, acc);
}

// Independent problems ("windows") solved together: the window of every
// device point block and camera block, per-window 1 / radius, per-window
// accumulators.  A single problem has of_e = of_f = nullptr (window 0) and
// acc = nullptr (deterministic block partials instead of atomics).
struct Win {
  const int32_t* of_e;
  const int32_t* of_f;
  const double* inv_r;      // 1 / radius per window (device memory, so captured launches stay valid)
  double* acc;
  int32_t* flag;            // per-window "non-finite / failed factorisation" flags
  const uint8_t* active;    // windows still iterating (nullptr: all)
  const PcgState* pcg;      // inside PCG: windows whose solve has stopped are idle
};
__device__ __forceinline__ double win_inv_r(const Win& w, int wi) { return w.inv_r[wi]; }
__device__ __forceinline__ int win_of_e(const Win& w, int e) { return w.of_e ? w.of_e[e] : 0; }
__device__ __forceinline__ int win_of_f(const Win& w, int f) { return w.of_f ? w.of_f[f] : 0; }
__device__ __forceinline__ bool win_idle(const Win& w, int wi) {
  return (w.active && !w.active[wi]) || (w.pcg && w.pcg[wi].done);
}
// Work of an idle window is skipped (a single problem is never idle).
__device__ __forceinline__ bool win_idle_e(const Win& w, int e) { return w.of_e && win_idle(w, w.of_e[e]); }
__device__ __forceinline__ bool win_idle_f(const Win& w, int f) { return w.of_f && win_idle(w, w.of_f[f]); }

// Max of a non-negative double via its ordered bit pattern.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// acc[stride * window + k] = max(., v >= 0) for every lane (wi < 0: none),
// with a wave pre-reduction like win_atomic_add.
__device__ __forceinline__ void win_atomic_max(double* acc, int stride, int k, int wi, double v) {
  const unsigned long long valid = __ballot(wi >= 0);
  if (valid == 0ull) return;
  const int leader = __ffsll((long long)valid) - 1;
  const int wl = __shfl(wi, leader, 64);
  const int wv = wi < 0 ? wl : wi;
  if (__all(wv == wl)) {
    double m = wi >= 0 ? v : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((int)(threadIdx.x & 63) == leader) atomic_max_nonneg(acc + (size_t)stride * wl + k, m);
  } else if (wi >= 0) {
    atomic_max_nonneg(acc + (size_t)stride * wi + k, v);
  }
}

template <typename T> __device__ __forceinline__ T tiny();
template <> __device__ __forceinline__ float tiny<float>() { return FLT_MIN; }
template <> __device__ __forceinline__ double tiny<double>() { return DBL_MIN; }

// Value of lane `l` (uniform) broadcast to the wave (v_readlane).
__device__ __forceinline__ float lane_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v of another lane by a DPP pattern (row-local, every source lane valid).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes, result in every lane (all lanes active): DPP adds
// within each row of 16 lanes (quad_perm xor 1, xor 2, half-row mirror,
// row mirror), then the four row sums read from lanes 0, 16, 32, 48 and
// added in a fixed order.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return (lane_bcast(v, 0) + lane_bcast(v, 16)) + (lane_bcast(v, 32) + lane_bcast(v, 48));
}

// acc[stride * window + k] += v for every lane (wi < 0: no contribution).
// All lanes of the wave must call it.
__device__ __forceinline__ void win_atomic_add(double* acc, int stride, int k, int wi, double v) {
  const unsigned long long valid = __ballot(wi >= 0);
  if (valid == 0ull) return;
  // Windows are laid out in order, so a wave usually holds one or two: the
  // first and last valid lanes' windows get a DPP wave sum and one atomic
  // each; lanes of any window in between add their own value.
  const int lane = (int)(threadIdx.x & 63);
  const int lo = __ffsll((long long)valid) - 1, hi = 63 - __clzll((long long)valid);
  const int w1 = __shfl(wi, lo, 64), w2 = __shfl(wi, hi, 64);
  const double s1 = wave_sum(wi == w1 ? v : 0.0);
  if (lane == lo) atomicAdd(acc + (size_t)stride * w1 + k, s1);
  if (w2 != w1) {
    const double s2 = wave_sum(wi == w2 ? v : 0.0);
    if (lane == hi) atomicAdd(acc + (size_t)stride * w2 + k, s2);
    if (wi >= 0 && wi != w1 && wi != w2 && v != 0.0) atomicAdd(acc + (size_t)stride * wi + k, v);
  }
}


// Sum over the `lanes` lanes of a point group (aligned, power of two,
// uniform in the wave; all lanes active), result in every lane of the
// group: DPP adds over 2, 4, 8 and 16 lanes (quad_perm xor 1, xor 2,
// half-row mirror, row mirror), shuffles beyond.
template <typename T>
__device__ __forceinline__ T group_sum(T v, int lanes) {
  if (lanes > 1) v += dpp_mov<0xB1>(v);
  if (lanes > 2) v += dpp_mov<0x4E>(v);
  if (lanes > 4) v += dpp_mov<0x141>(v);
  if (lanes > 8) v += dpp_mov<0x140>(v);
  if (lanes > 16) v += __shfl_xor(v, 16, 64);
  if (lanes > 32) v += __shfl_xor(v, 32, 64);
  return v;
}

// Reduce-scatter across a wave.  Stage D exchanges with the lane that
// differs in lane bit D: given lo and hi (values i and i + n/2 of the n
// still held), each lane keeps the pair sum of the half its bit selects (lo
// for bit clear, hi for bit set), halving n.  D = 32 and 16 use gfx950's
// permlane swaps (v_permlane32_swap: the upper half of the first operand
// with the lower half of the second; v_permlane16_swap: odd rows of the
// first with even rows of the second), which place both partners' values
// of the kept half in the two results; D <= 8 use DPP (row mirror, half-row
// mirror, quad_perm xor 2, xor 1: each pairs lanes differing in bit D).
// With one value left a stage is a plain butterfly add (lo = hi).
template <int D>
__device__ __forceinline__ float rs_pair(float lo, float hi) {
  if constexpr (D == 32 || D == 16) {
    const auto r = D == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(lo), __float_as_uint(hi), false, false)
                           : __builtin_amdgcn_permlane16_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    constexpr int ctrl = D == 8 ? 0x140 : D == 4 ? 0x141 : D == 2 ? 0x4E : 0xB1;
    const bool bit = (threadIdx.x & D) != 0;
    const float keep = bit ? hi : lo, send = bit ? lo : hi;
    return keep + dpp_mov<ctrl>(send);
  }
}
template <int D>
__device__ __forceinline__ double rs_pair(double lo, double hi) {
  if constexpr (D == 32 || D == 16) {
    const unsigned long long l = (unsigned long long)__double_as_longlong(lo);
    const unsigned long long h = (unsigned long long)__double_as_longlong(hi);
    const auto r0 = D == 32 ? __builtin_amdgcn_permlane32_swap((unsigned)l, (unsigned)h, false, false)
                            : __builtin_amdgcn_permlane16_swap((unsigned)l, (unsigned)h, false, false);
    const auto r1 = D == 32 ? __builtin_amdgcn_permlane32_swap((unsigned)(l >> 32), (unsigned)(h >> 32), false, false)
                            : __builtin_amdgcn_permlane16_swap((unsigned)(l >> 32), (unsigned)(h >> 32), false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)r1[0] << 32) | r0[0]));
    const double b = __longlong_as_double((long long)(((unsigned long long)r1[1] << 32) | r0[1]));
    return a + b;
  } else {
    constexpr int ctrl = D == 8 ? 0x140 : D == 4 ? 0x141 : D == 2 ? 0x4E : 0xB1;
    const bool bit = (threadIdx.x & D) != 0;
    const double keep = bit ? hi : lo, send = bit ? lo : hi;
    return keep + dpp_mov<ctrl>(send);
  }
}
template <typename T, int n, int D>
__device__ __forceinline__ void rs_stages(T* a) {
  if constexpr (D >= 1) {
    if constexpr (n > 1) {
#pragma unroll
      for (int i = 0; i < n / 2; ++i) a[i] = rs_pair<D>(a[i], a[i + n / 2]);
      rs_stages<T, n / 2, D / 2>(a);
    } else {
      a[0] = rs_pair<D>(a[0], a[0]);
      rs_stages<T, 1, D / 2>(a);
    }
  }
}
// log2 of the power of two >= n (n <= 64).
constexpr int rs_bits(int n) { return n <= 1 ? 0 : 1 + rs_bits((n + 1) / 2); }

// Wave sums of N <= 64 accumulators (all lanes active): lane l returns the
// sum of accumulator l >> (6 - rs_bits(N)) over the 64 lanes.
template <typename T, int N>
__device__ __forceinline__ T wave_reduce_scatter(const T (&v)[N]) {
  constexpr int P = 1 << rs_bits(N);
  T a[P];
#pragma unroll
  for (int i = 0; i < P; ++i) a[i] = i < N ? v[i] : T(0);
  rs_stages<T, P, 32>(a);
  return a[0];
}

// Block sums of N <= 64 accumulators over B threads: thread t < N returns
// the sum of accumulator t (one wave: straight from the reduce-scatter when
// N > 32; otherwise through lds[N * B / 64]).
template <typename T, int N, int B>
__device__ __forceinline__ T block_reduce(const T (&acc)[N], T* lds) {
  constexpr int M = rs_bits(N);
  const T v = wave_reduce_scatter<T, N>(acc);
  if constexpr (B == 64 && M == 6) {
    return v;
  } else {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, idx = lane >> (6 - M);
    if ((lane & ((1 << (6 - M)) - 1)) == 0 && idx < N) lds[wave * N + idx] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x < N)
      for (int w = 0; w < B / 64; ++w) s += lds[w * N + threadIdx.x];
    return s;
  }
}

// Block-wide sum of N accumulators (blockDim.x == kBlock): wave sums land
// in lds[wave * N + k] (lane k % 64 of each wave keeps sum k, then writes
// it; no branch per accumulator).
template <typename T, int N>
__device__ __forceinline__ void block_sum(T (&acc)[N], T* lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int M = (N + 63) / 64;
  T mine[M];
#pragma unroll
  for (int m = 0; m < M; ++m) mine[m] = T(0);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const T v = wave_sum(acc[k]);
    mine[k >> 6] = lane == (k & 63) ? v : mine[k >> 6];
  }
#pragma unroll
  for (int m = 0; m < M; ++m)
    if (64 * m + lane < N) lds[wave * N + 64 * m + lane] = mine[m];
  __syncthreads();
}

// ------------------------------------------------------------ camera model
// Rodrigues with the reference's branch point (rotation.h:565-625): first
// order I + [w]x when theta^2 <= DBL_EPSILON.  Writes R and dR/dw_k.
__device__ __forceinline__ void rodrigues_with_jacobian(const double w[3], double R[9],
                                                        double dR[27]) {
  const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (th2 > DBL_EPSILON) {
    const double th = sqrt(th2), c = cos(th), s = sin(th), it = 1.0 / th;
    const double u[3] = {w[0] * it, w[1] * it, w[2] * it};
    const double oc = 1.0 - c;
    // R = c I + s [u]x + (1 - c) u u'
    R[0] = c + oc * u[0] * u[0];        R[1] = -s * u[2] + oc * u[0] * u[1]; R[2] = s * u[1] + oc * u[0] * u[2];
    R[3] = s * u[2] + oc * u[1] * u[0]; R[4] = c + oc * u[1] * u[1];        R[5] = -s * u[0] + oc * u[1] * u[2];
    R[6] = -s * u[1] + oc * u[2] * u[0]; R[7] = s * u[0] + oc * u[2] * u[1]; R[8] = c + oc * u[2] * u[2];
    for (int k = 0; k < 3; ++k) {
      // du_i/dw_k = (delta_ik - u_i u_k) / theta ; dtheta/dw_k = u_k
      double du[3];
      for (int i = 0; i < 3; ++i) du[i] = ((i == k ? 1.0 : 0.0) - u[i] * u[k]) * it;
      const double dc = -s * u[k], ds = c * u[k], doc = s * u[k];
      double* D = dR + 9 * k;
      // d[u]x
      const double X[9] = {0.0, -du[2], du[1], du[2], 0.0, -du[0], -du[1], du[0], 0.0};
      const double Ux[9] = {0.0, -u[2], u[1], u[2], 0.0, -u[0], -u[1], u[0], 0.0};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const int ij = 3 * i + j;
          D[ij] = (i == j ? dc : 0.0) + ds * Ux[ij] + s * X[ij] + doc * u[i] * u[j] +
                  oc * (du[i] * u[j] + u[i] * du[j]);
        }
    }
  } else {
    R[0] = 1.0;   R[1] = -w[2]; R[2] = w[1];
    R[3] = w[2];  R[4] = 1.0;   R[5] = -w[0];
    R[6] = -w[1]; R[7] = w[0];  R[8] = 1.0;
    for (int i = 0; i < 27; ++i) dR[i] = 0.0;
    // d/dw_k [w]x = [e_k]x
    dR[0 * 9 + 5] = -1.0; dR[0 * 9 + 7] = 1.0;
    dR[1 * 9 + 2] = 1.0;  dR[1 * 9 + 6] = -1.0;
    dR[2 * 9 + 1] = -1.0; dR[2 * 9 + 3] = 1.0;
  }
}

// gtab (optional): also the local rotation basis of each camera (cam_lj,
// below), in the same launch.
template <typename T>
__device__ __forceinline__ void cam_lj(const double* p, T* __restrict__ o);
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cam_table(int nc, const double* __restrict__ cams,
                                                      T* __restrict__ tab, T* __restrict__ gtab) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double* p = cams + 9 * (size_t)c;
  double R[9], dR[27];
  const double w[3] = {p[0], p[1], p[2]};
  rodrigues_with_jacobian(w, R, dR);
  T* o = tab + (size_t)kCamTab * c;
  for (int i = 0; i < 9; ++i) o[i] = (T)R[i];
  for (int i = 0; i < 6; ++i) o[kTabT + i] = (T)p[3 + i];
  o[kTabK3] = (T)0;
  for (int i = 0; i < 27; ++i) o[kTabDR + i] = (T)dR[i];
  for (int i = kTabDR + 27; i < kCamTab; ++i) o[i] = (T)0;
  if (gtab) cam_lj(p, gtab + 9 * (size_t)c);
}

template <typename T>
struct CamProj {
  T R[9], t[3], f, k1, k2, k3, ppx, ppy;
};

template <typename T, int L = 0>
__device__ __forceinline__ void load_proj(const T* __restrict__ tab, CamProj<T>& c) {
#pragma unroll
  for (int i = 0; i < 9; ++i) c.R[i] = tab[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) c.t[i] = tab[kTabT + i];
  c.f = tab[kTabF];
  c.k1 = tab[kTabF + 1];
  c.k2 = tab[kTabF + 2];
  if constexpr (L == 1) {
    c.k3 = tab[kTabK3];
    c.ppx = tab[kTabPP];
    c.ppy = tab[kTabPP + 1];
  }
}

template <typename T>
struct Proj {
  T rx, ry;          // residual, predicted - observed
  T P[6];            // d r / d Xc (2 x 3)
  T jf[2], jk1[2], jk2[2], jk3[2];
};

// Residual of libs/sfm's functor (pinhole_brown_rt_ceres_functor.h:60-107,
// brown_pinhole_camera.h:14-41) and its analytic derivative with respect to
// the camera-frame point: pred = pp + f d(r^2) (x/z, y/z) with
// d = 1 + k1 r^2 + k2 r^4 + k3 r^6 (L = 0: principal point 0, k3 = 0).
template <typename T, int L = 0>
__device__ __forceinline__ void project(const CamProj<T>& c, const T X[3], T ox, T oy, Proj<T>& o) {
  const T xc = c.R[0] * X[0] + c.R[1] * X[1] + c.R[2] * X[2] + c.t[0];
  const T yc = c.R[3] * X[0] + c.R[4] * X[1] + c.R[5] * X[2] + c.t[1];
  const T zc = c.R[6] * X[0] + c.R[7] * X[1] + c.R[8] * X[2] + c.t[2];
  const T iz = T(1) / zc;
  const T xe = xc * iz, ye = yc * iz;
  const T r2 = xe * xe + ye * ye;
  T d, fdd;
  if constexpr (L == 0) {
    d = T(1) + r2 * (c.k1 + c.k2 * r2);
    fdd = c.f * T(2) * (c.k1 + T(2) * c.k2 * r2);
  } else {
    d = T(1) + r2 * (c.k1 + r2 * (c.k2 + r2 * c.k3));
    fdd = c.f * T(2) * (c.k1 + r2 * (T(2) * c.k2 + T(3) * c.k3 * r2));
  }
  const T fd = c.f * d;
  if constexpr (L == 0) {
    o.rx = fd * xe - ox;
    o.ry = fd * ye - oy;
  } else {
    o.rx = fd * xe + c.ppx - ox;
    o.ry = fd * ye + c.ppy - oy;
  }
  const T a11 = fd + fdd * xe * xe, a12 = fdd * xe * ye, a22 = fd + fdd * ye * ye;
  o.P[0] = a11 * iz;
  o.P[1] = a12 * iz;
  o.P[2] = -(a11 * xe + a12 * ye) * iz;
  o.P[3] = a12 * iz;
  o.P[4] = a22 * iz;
  o.P[5] = -(a12 * xe + a22 * ye) * iz;
  o.jf[0] = d * xe;
  o.jf[1] = d * ye;
  o.jk1[0] = c.f * xe * r2;
  o.jk1[1] = c.f * ye * r2;
  o.jk2[0] = o.jk1[0] * r2;
  o.jk2[1] = o.jk1[1] * r2;
  if constexpr (L == 1) {
    o.jk3[0] = o.jk2[0] * r2;
    o.jk3[1] = o.jk2[1] * r2;
  }
}

// project() that also returns y = R X, the rotated point before the
// translation (the FP32 camera-major kernels' local rotation Jacobian).
template <typename T>
__device__ __forceinline__ void project_y(const CamProj<T>& c, const T X[3], T ox, T oy, Proj<T>& o, T y[3]) {
  y[0] = c.R[0] * X[0] + c.R[1] * X[1] + c.R[2] * X[2];
  y[1] = c.R[3] * X[0] + c.R[4] * X[1] + c.R[5] * X[2];
  y[2] = c.R[6] * X[0] + c.R[7] * X[1] + c.R[8] * X[2];
  const T zc = y[2] + c.t[2];
  const T iz = T(1) / zc;
  const T xe = (y[0] + c.t[0]) * iz, ye = (y[1] + c.t[1]) * iz;
  const T r2 = xe * xe + ye * ye;
  const T d = T(1) + r2 * (c.k1 + c.k2 * r2);
  const T fdd = c.f * T(2) * (c.k1 + T(2) * c.k2 * r2);
  const T fd = c.f * d;
  o.rx = fd * xe - ox;
  o.ry = fd * ye - oy;
  const T a11 = fd + fdd * xe * xe, a12 = fdd * xe * ye, a22 = fd + fdd * ye * ye;
  o.P[0] = a11 * iz;
  o.P[1] = a12 * iz;
  o.P[2] = -(a11 * xe + a12 * ye) * iz;
  o.P[3] = a12 * iz;
  o.P[4] = a22 * iz;
  o.P[5] = -(a12 * xe + a22 * ye) * iz;
  o.jf[0] = d * xe;
  o.jf[1] = d * ye;
  o.jk1[0] = c.f * xe * r2;
  o.jk1[1] = c.f * ye * r2;
  o.jk2[0] = o.jk1[0] * r2;
  o.jk2[1] = o.jk1[1] * r2;
}

// Local rotation basis of the FP32 camera-major kernels.  With y = R X,
// d(R X)/d aa = -[y]x G where G = J_l(aa) is the left Jacobian of SO(3), so
// a row's camera Jacobian is J_c = Jt B, B = blockdiag(G, I, I), with
// Jt = w [y x P_r | P_r | jf_r, jk1_r, jk2_r] per residual row r (the aa
// columns are cross products instead of P dR_k X).  Sums over rows are
// formed with Jt and taken to the parameter basis once per chunk.
template <typename T>
__device__ __forceinline__ void jac_cam_local(const Proj<T>& o, const T y[3], T wgt, T Jt[18]) {
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const T p0 = wgt * o.P[3 * r], p1 = wgt * o.P[3 * r + 1], p2 = wgt * o.P[3 * r + 2];
    Jt[9 * r + 0] = y[1] * p2 - y[2] * p1;
    Jt[9 * r + 1] = y[2] * p0 - y[0] * p2;
    Jt[9 * r + 2] = y[0] * p1 - y[1] * p0;
    Jt[9 * r + 3] = p0;
    Jt[9 * r + 4] = p1;
    Jt[9 * r + 5] = p2;
  }
  Jt[6] = wgt * o.jf[0];
  Jt[15] = wgt * o.jf[1];
  Jt[7] = wgt * o.jk1[0];
  Jt[16] = wgt * o.jk1[1];
  Jt[8] = wgt * o.jk2[0];
  Jt[17] = wgt * o.jk2[1];
}

// G = J_l(w) = I + a [w]x + b [w]x^2, a = (1 - cos t) / t^2,
// b = (t - sin t) / t^3 (series below t^2 = 1e-6), for the camera at p,
// row-major G[3 i + j], rounded to T.
template <typename T>
__device__ __forceinline__ void cam_lj(const double* p, T* __restrict__ o) {
  const double w0 = p[0], w1 = p[1], w2 = p[2];
  const double t2 = w0 * w0 + w1 * w1 + w2 * w2;
  double a, b;
  if (t2 < 1e-6) {
    a = 0.5 - t2 / 24.0;
    b = 1.0 / 6.0 - t2 / 120.0;
  } else {
    const double t = sqrt(t2);
    a = (1.0 - cos(t)) / t2;
    b = (t - sin(t)) / (t2 * t);
  }
  const double K[9] = {0.0, -w2, w1, w2, 0.0, -w0, -w1, w0, 0.0};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double kk = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
      o[3 * i + j] = (T)((i == j ? 1.0 : 0.0) + a * K[3 * i + j] + b * kk);
    }
}

// Chunk sums in the local basis -> parameter basis, for the 63 camera-pass
// sums held one per lane of a wave (lane k: U upper entry k < 45, g 45..53,
// h 54..62): U = B' Ut B, g = B' gt, h = B' ht.  Every lane of the wave
// calls it.
template <typename T>
__device__ __forceinline__ T chunk_to_param_basis63(T v, int k, const T* __restrict__ G) {
  // Upper-triangle index of (i, j), i <= j, in a row-major 9 x 9.
  auto ut9 = [](int i, int j) { return i * 9 - i * (i - 1) / 2 + (j - i); };
  const T u00 = __shfl(v, 0, 64), u01 = __shfl(v, 1, 64), u02 = __shfl(v, 2, 64);
  const T u11 = __shfl(v, 9, 64), u12 = __shfl(v, 10, 64), u22 = __shfl(v, 17, 64);
  int a, b;  // entry (a, b); b = 9 for g, 10 for h
  if (k < 45) {
    a = 0;
    while (a < 8 && k >= ut9(a + 1, a + 1)) ++a;
    b = a + (k - ut9(a, a));
  } else if (k < 54) {
    a = k - 45;
    b = 9;
  } else {
    a = k < 63 ? k - 54 : 9;
    b = 10;
  }
  auto src = [&](int i) { return b < 9 ? ut9(i, b) : (b == 9 ? 45 + i : 54 + i); };
  const int bs = b < 3 ? 3 : b;  // aa-block columns use the u table below
  const T c0 = __shfl(v, src(0) < 64 && bs == b ? src(0) : 0, 64);
  const T c1 = __shfl(v, src(1) < 64 && bs == b ? src(1) : 0, 64);
  const T c2 = __shfl(v, src(2) < 64 && bs == b ? src(2) : 0, 64);
  if (a >= 3) return v;
  const T g0 = G[a], g1 = G[3 + a], g2 = G[6 + a];  // column a of G
  if (b >= 3) return g0 * c0 + g1 * c1 + g2 * c2;
  const T h0 = G[b], h1 = G[3 + b], h2 = G[6 + b];  // column b of G
  // Ut_aa h, then g' (Ut_aa h)
  const T m0 = u00 * h0 + u01 * h1 + u02 * h2;
  const T m1 = u01 * h0 + u11 * h1 + u12 * h2;
  const T m2 = u02 * h0 + u12 * h1 + u22 * h2;
  return g0 * m0 + g1 * m1 + g2 * m2;
}

// Chunk sums of the camera matvec (lanes 0..8) -> parameter basis and
// camera column scale: p = diag(cs) B' pt.
template <typename T>
__device__ __forceinline__ T chunk_to_param_basis9(T v, int k, const T* __restrict__ G, const T* __restrict__ cs) {
  const T p0 = __shfl(v, 0, 64), p1 = __shfl(v, 1, 64), p2 = __shfl(v, 2, 64);
  if (k >= 9) return v;
  if (k >= 3) return cs[k] * v;
  return cs[k] * (G[k] * p0 + G[3 + k] * p1 + G[6 + k] * p2);
}

// HuberLoss + Corrector for rho'' <= 0 (loss_function.cc:47-61,
// corrector.cc:41-85): residual and Jacobian scale by sqrt(rho').
template <typename T>
__device__ __forceinline__ T huber_weight(T s, T a) {
  if (a > T(0) && s > a * a) {
    T rho1 = a / sqrt(s);
    rho1 = rho1 > tiny<T>() ? rho1 : tiny<T>();
    return sqrt(rho1);
  }
  return T(1);
}

template <typename T>
__device__ __forceinline__ T huber_cost(T s, T a) {
  if (a > T(0) && s > a * a) return T(0.5) * (T(2) * a * sqrt(s) - a * a);
  return T(0.5) * s;
}

// Point Jacobian J_p = w * (P R) .* sp  (2 x 3, row-major).
template <typename T>
__device__ __forceinline__ void jac_point(const Proj<T>& o, const T R[9], T wgt, const T sp[3],
                                          T Jp[6]) {
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Jp[3 * r + j] = wgt * sp[j] *
                      (o.P[3 * r] * R[j] + o.P[3 * r + 1] * R[3 + j] + o.P[3 * r + 2] * R[6 + j]);
}

// Camera Jacobian J_c (2 x 9) from the table's dR: columns aa (P dR_k X),
// t (P), f, k1, k2; scaled by w and the camera scale.
template <typename T>
__device__ __forceinline__ void jac_cam(const Proj<T>& o, const T* dR, const T X[3], T wgt,
                                        const T cs[9], T Jc[18]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T* D = dR + 9 * k;
    const T v0 = D[0] * X[0] + D[1] * X[1] + D[2] * X[2];
    const T v1 = D[3] * X[0] + D[4] * X[1] + D[5] * X[2];
    const T v2 = D[6] * X[0] + D[7] * X[1] + D[8] * X[2];
    Jc[k] = wgt * cs[k] * (o.P[0] * v0 + o.P[1] * v1 + o.P[2] * v2);
    Jc[9 + k] = wgt * cs[k] * (o.P[3] * v0 + o.P[4] * v1 + o.P[5] * v2);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Jc[3 + k] = wgt * cs[3 + k] * o.P[k];
    Jc[12 + k] = wgt * cs[3 + k] * o.P[3 + k];
  }
  Jc[6] = wgt * cs[6] * o.jf[0];
  Jc[15] = wgt * cs[6] * o.jf[1];
  Jc[7] = wgt * cs[7] * o.jk1[0];
  Jc[16] = wgt * cs[7] * o.jk1[1];
  Jc[8] = wgt * cs[8] * o.jk2[0];
  Jc[17] = wgt * cs[8] * o.jk2[1];
}

// J_c x through a direction table (M X + dt, intrinsics), weight applied.
// L = 1: the intrinsic part is the row's group direction
// [df, dppx, dppy, dk1, dk2, dk3].
template <typename T, int L = 0>
__device__ __forceinline__ void jac_cam_dir(const Proj<T>& o, const T* __restrict__ dt, const T X[3],
                                            T wgt, T out[2]) {
  const T v0 = dt[0] * X[0] + dt[1] * X[1] + dt[2] * X[2] + dt[9];
  const T v1 = dt[3] * X[0] + dt[4] * X[1] + dt[5] * X[2] + dt[10];
  const T v2 = dt[6] * X[0] + dt[7] * X[1] + dt[8] * X[2] + dt[11];
  if constexpr (L == 0) {
    out[0] = wgt * (o.P[0] * v0 + o.P[1] * v1 + o.P[2] * v2 + o.jf[0] * dt[12] + o.jk1[0] * dt[13] +
                    o.jk2[0] * dt[14]);
    out[1] = wgt * (o.P[3] * v0 + o.P[4] * v1 + o.P[5] * v2 + o.jf[1] * dt[12] + o.jk1[1] * dt[13] +
                    o.jk2[1] * dt[14]);
  } else {
    out[0] = wgt * (o.P[0] * v0 + o.P[1] * v1 + o.P[2] * v2 + o.jf[0] * dt[12] + dt[13] + o.jk1[0] * dt[15] +
                    o.jk2[0] * dt[16] + o.jk3[0] * dt[17]);
    out[1] = wgt * (o.P[3] * v0 + o.P[4] * v1 + o.P[5] * v2 + o.jf[1] * dt[12] + dt[14] + o.jk1[1] * dt[15] +
                    o.jk2[1] * dt[16] + o.jk3[1] * dt[17]);
  }
}

// Packed point record read by the camera-major kernels: [X(3), v(3), 0, 0]
// with v = s_p .* (V^-1 g_p) or s_p .* z; one aligned 8-element gather per
// observation instead of three scattered ones.
constexpr int kPk = 8;

__device__ __forceinline__ void load_pk(const float* __restrict__ pk, int e, float X[3], float v[3]) {
  const float4* q = reinterpret_cast<const float4*>(pk + (size_t)kPk * e);
  const float4 a = q[0], b = q[1];
  X[0] = a.x; X[1] = a.y; X[2] = a.z; v[0] = a.w; v[1] = b.x; v[2] = b.y;
}
__device__ __forceinline__ void load_pk(const double* __restrict__ pk, int e, double X[3], double v[3]) {
  const double2* q = reinterpret_cast<const double2*>(pk + (size_t)kPk * e);
  const double2 a = q[0], b = q[1], c = q[2];
  X[0] = a.x; X[1] = a.y; X[2] = b.x; v[0] = b.y; v[1] = c.x; v[2] = c.y;
}
template <typename T>
__device__ __forceinline__ void store_pk(T* __restrict__ pk, int e, const T X[3], const T v[3]) {
  T* o = pk + (size_t)kPk * e;
  o[0] = X[0]; o[1] = X[1]; o[2] = X[2]; o[3] = v[0]; o[4] = v[1]; o[5] = v[2]; o[6] = T(0); o[7] = T(0);
}
template <typename T>
__device__ __forceinline__ void store_pk_v(T* __restrict__ pk, int e, const T v[3]) {
  T* o = pk + (size_t)kPk * e;
  o[3] = v[0]; o[4] = v[1]; o[5] = v[2];
}

// u = P R v (2-vector): the point Jacobian (before weight) applied to v.
template <typename T>
__device__ __forceinline__ void jac_point_apply(const Proj<T>& o, const T R[9], const T v[3], T u[2]) {
  const T a0 = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  const T a1 = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  const T a2 = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
  u[0] = o.P[0] * a0 + o.P[1] * a1 + o.P[2] * a2;
  u[1] = o.P[3] * a0 + o.P[4] * a1 + o.P[5] * a2;
}

// FP64 inverse of an SPD 3x3 (xx xy xz yy yz zz): adjugate over the
// determinant, one division.  Positive definiteness by Sylvester's
// criterion (a00 > 0, leading 2x2 minor > 0, det > 0); false otherwise.
__device__ __forceinline__ bool inv3_spd(const double a[6], double o[6]) {
  const double c00 = a[3] * a[5] - a[4] * a[4], c01 = a[2] * a[4] - a[1] * a[5], c02 = a[1] * a[4] - a[2] * a[3];
  const double c11 = a[0] * a[5] - a[2] * a[2], c12 = a[1] * a[2] - a[0] * a[4], c22 = a[0] * a[3] - a[1] * a[1];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(a[0] > 0.0) || !(c22 > 0.0) || !(det > 0.0)) return false;
  const double id = 1.0 / det;
  o[0] = c00 * id;
  o[1] = c01 * id;
  o[2] = c02 * id;
  o[3] = c11 * id;
  o[4] = c12 * id;
  o[5] = c22 * id;
  return true;
}

// Packed upper-triangle index of (i, j), i <= j, in a 9 x 9 block.
__device__ __forceinline__ int ut9(int i, int j) { return i * 9 - i * (i - 1) / 2 + (j - i); }
__device__ __forceinline__ int lt9(int i, int j) { return i * (i + 1) / 2 + j; }  // packed lower, j <= i

// FP64 Cholesky of a 9 x 9 SPD matrix given by its upper triangle (ut9
// order) into packed lower l[45]; false if not positive definite.  Fully
// unrolled: every index is a constant, everything stays in registers.
__device__ __forceinline__ bool chol9(const double a[45], double l[45]) {
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    double d = a[ut9(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= l[lt9(j, k)] * l[lt9(j, k)];
    if (!(d > 0.0)) return false;
    d = sqrt(d);
    l[lt9(j, j)] = d;
#pragma unroll
    for (int i = j + 1; i < 9; ++i) {
      double s = a[ut9(j, i)];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= l[lt9(i, k)] * l[lt9(j, k)];
      l[lt9(i, j)] = s / d;
    }
  }
  return true;
}

// Column c of (L L')^-1 from the packed factor.
__device__ __forceinline__ void chol9_inv_col(const double l[45], int c, double e[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) s -= l[lt9(i, k)] * e[k];
    e[i] = s / l[lt9(i, i)];
  }
#pragma unroll
  for (int i = 8; i >= 0; --i) {
    double s = e[i];
#pragma unroll
    for (int k = i + 1; k < 9; ++k) s -= l[lt9(k, i)] * e[k];
    e[i] = s / l[lt9(i, i)];
  }
}

// --------------------------------------------------------------- FP64 cost
// One thread per E block; block partial sums of 1/2 rho(|r|^2).
template <int L>
__global__ __launch_bounds__(kBlock) void k_cost64(Seg sg, const int32_t* __restrict__ pt_start,
                                                   const int32_t* __restrict__ row_cam,
                                                   const double* __restrict__ row_obs,
                                                   const double* __restrict__ pts,
                                                   const double* __restrict__ ctab, double huber_a,
                                                   double* __restrict__ partial, Win win,
                                                   float* __restrict__ rrow, float* __restrict__ rcm,
                                                   const int32_t* __restrict__ row_cm,
                                                   float* __restrict__ wrow, float* __restrict__ wcm) {
  __shared__ double lds[kBlock / 64];
  SegPos sp = seg_locate(sg);
  if (sp.valid && win_idle_e(win, sp.e)) sp.valid = false;
  double acc[1] = {0.0};
  if (sp.valid) {
    const int e = sp.e;
    const double X[3] = {pts[3 * (size_t)e], pts[3 * (size_t)e + 1], pts[3 * (size_t)e + 2]};
    int bad = 0;
    for (int r = pt_start[e] + sp.sub; r < pt_start[e + 1]; r += sp.lanes) {
      CamProj<double> c;
      load_proj<double, L>(ctab + (size_t)kCamTab * row_cam[r], c);
      Proj<double> o;
      project<double, L>(c, X, row_obs[2 * (size_t)r], row_obs[2 * (size_t)r + 1], o);
      const double s = o.rx * o.rx + o.ry * o.ry;
      if (!isfinite(s)) bad = 1;
      acc[0] += huber_cost(s, huber_a);
      if (rrow) {
        // The FP64 residual, rounded once, for the FP32 passes at this state
        // (row order and, through row_cm, camera-major order).
        rrow[2 * (size_t)r] = (float)o.rx;
        rrow[2 * (size_t)r + 1] = (float)o.ry;
        // With a loss, its weight sqrt(rho') too, evaluated in FP64.
        const float w = wrow ? (float)huber_weight(s, huber_a) : 1.0f;
        if (wrow) wrow[r] = w;
        const int k = rcm ? row_cm[r] : -1;  // rcm null: camera-major copies later (k_rows_to_cm)
        if (k >= 0) {
          rcm[2 * (size_t)k] = (float)o.rx;
          rcm[2 * (size_t)k + 1] = (float)o.ry;
          if (wcm) wcm[k] = w;
        }
      }
    }
    if (bad) win.flag[win_of_e(win, e)] = 1;
  }
  if (win.acc) {
    win_atomic_add(win.acc, 1, 0, sp.valid ? win_of_e(win, sp.e) : -1, acc[0]);
    return;
  }
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

// The camera-major copies of k_cost64's row-order FP32 residuals and weights
// (rcm[2 k], rcm[2 k + 1], wcm[k] for k = row_cm[r] >= 0), made when a trial
// state becomes current: a rejected trial state's copies are never read.
__global__ __launch_bounds__(kBlock) void k_rows_to_cm(int64_t nrows, const int32_t* __restrict__ row_cm,
                                                       const float* __restrict__ rrow,
                                                       const float* __restrict__ wrow, float* __restrict__ rcm,
                                                       float* __restrict__ wcm) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int k = row_cm[r];
  if (k < 0) return;
  const float2 v = reinterpret_cast<const float2*>(rrow)[r];
  reinterpret_cast<float2*>(rcm)[k] = v;
  if (wcm) wcm[k] = wrow ? wrow[r] : 1.0f;
}

// Deterministic single-block sum of partials.
__global__ __launch_bounds__(kBlock) void k_reduce(const double* __restrict__ partial, int n,
                                                   double* __restrict__ out) {
  __shared__ double lds[kBlock / 64];
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += kBlock) acc[0] += partial[i];
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    *out = s;
  }
}

// Max-reduce of |g / scale| (unscaled gradient max norm) and sum of squares.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_grad_partial(int64_t n, const T* __restrict__ g,
                                                         const double* __restrict__ scale,
                                                         double* __restrict__ pmax,
                                                         double* __restrict__ psq) {
  __shared__ double lds[2 * (kBlock / 64)];
  double mx = 0.0, sq = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double v = (double)g[i] / scale[i];
    mx = fmax(mx, fabs(v));
    sq += v * v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, o, 64));
    sq += __shfl_xor(sq, o, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { lds[wave] = mx; lds[4 + wave] = sq; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) { a = fmax(a, lds[w]); b += lds[4 + w]; }
    pmax[blockIdx.x] = a;
    psq[blockIdx.x] = b;
  }
}

// Per-window max |g / scale| and sum of squares into out[2 * window + {0, 1}].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_grad_win(int64_t n, int per, int is_cam, const T* __restrict__ g,
                                                     const double* __restrict__ scale, Win win,
                                                     double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  double v = 0.0;
  int wi = -1;
  if (i < n) {
    const int blk = (int)(i / per);
    if (!(is_cam ? win_idle_f(win, blk) : win_idle_e(win, blk))) {
      v = (double)g[i] / scale[i];
      wi = is_cam ? win_of_f(win, blk) : win_of_e(win, blk);
    }
  }
  win_atomic_max(out, 2, 0, wi, fabs(v));
  win_atomic_add(out, 2, 1, wi, v * v);
}

__global__ __launch_bounds__(kBlock) void k_grad_final(const double* __restrict__ pmax,
                                                       const double* __restrict__ psq, int n,
                                                       double* __restrict__ out /* [max, sumsq] accumulate */) {
  __shared__ double lds[2 * (kBlock / 64)];
  double mx = 0.0, sq = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) {
    mx = fmax(mx, pmax[i]);
    sq += psq[i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, o, 64));
    sq += __shfl_xor(sq, o, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    lds[wave] = mx;
    lds[kBlock / 64 + wave] = sq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = out[0], b = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
      a = fmax(a, lds[w]);
      b += lds[kBlock / 64 + w];
    }
    out[0] = a;
    out[1] += b;
  }
}

// Camera data of a point-major row: projection parameters and the
// direction table entry, from the global tables ...
template <typename T, int L>
struct GlobalCams {
  const T* __restrict__ ctab;
  const T* __restrict__ dtab;
  __device__ __forceinline__ void proj(int cam, CamProj<T>& c) const {
    load_proj<T, L>(ctab + (size_t)kCamTab * cam, c);
  }
  __device__ __forceinline__ const T* dir(int cam) const { return dtab + (size_t)dir_stride<L>() * cam; }
};

// ... or from records staged in LDS by a persistent launch (FP32, packed
// cameras): Q 16-byte quads per camera, [R(9) t(3) | f k1 k2 -] and, for
// Q = 8, the direction (16), 4 Q + 4 floats apart (an odd number of bank
// quads, so that the 16-byte reads of random cameras spread over the 64
// banks).  A camera-table gather per row from L2 limits the point-major
// kernels; an LDS read does not.
constexpr int kStageThreads = 4 * kBlock;  // four virtual blocks per workgroup
template <int Q>
__host__ __device__ constexpr int stage_rec() { return 4 * Q + 4; }
static_assert(kTabT == 9 && kTabF == 12 && kCamTab % 4 == 0 && kDirTab == 16,
              "staged records copy R, t, f, k1, k2 and the direction as 16-byte quads");

// Fill the LDS records of cameras [0, nc): 16-byte copies, eight in flight
// per thread.  Quad k is table quad k (k < 4) or direction quad k - 4.
typedef float stage_quad __attribute__((ext_vector_type(4)));
template <int Q>
__device__ __forceinline__ void stage_cams(float4* __restrict__ lds4, int nc, const float* __restrict__ ctab,
                                           const float* __restrict__ dtab) {
  const stage_quad* ct4 = reinterpret_cast<const stage_quad*>(ctab);
  const stage_quad* dt4 = reinterpret_cast<const stage_quad*>(dtab);
  stage_quad* lds = reinterpret_cast<stage_quad*>(lds4);
  const int nq = Q * nc;
  for (int i0 = 0; i0 < nq; i0 += 8 * (int)blockDim.x) {
    stage_quad v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = min(i0 + u * (int)blockDim.x + (int)threadIdx.x, nq - 1);
      const int cam = i / Q, k = i % Q;
      v[u] = k < 4 ? ct4[(size_t)(kCamTab / 4) * cam + k] : dt4[(size_t)(kDirTab / 4) * cam + (k - 4)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * (int)blockDim.x + (int)threadIdx.x;
      if (i < nq) lds[(stage_rec<Q>() / 4) * (i / Q) + (i % Q)] = v[u];
    }
  }
  __syncthreads();
}

template <int Q>
struct LdsCams {
  const float* rec;
  __device__ __forceinline__ void proj(int cam, CamProj<float>& c) const {
    const float4* q = reinterpret_cast<const float4*>(rec + stage_rec<Q>() * cam);
    const float4 a = q[0], b = q[1], d = q[2], e = q[3];
    c.R[0] = a.x; c.R[1] = a.y; c.R[2] = a.z; c.R[3] = a.w;
    c.R[4] = b.x; c.R[5] = b.y; c.R[6] = b.z; c.R[7] = b.w;
    c.R[8] = d.x; c.t[0] = d.y; c.t[1] = d.z; c.t[2] = d.w;
    c.f = e.x; c.k1 = e.y; c.k2 = e.z;
  }
  __device__ __forceinline__ const float* dir(int cam) const { return rec + stage_rec<Q>() * cam + 16; }
};

// ------------------------------------------------------------- point pass
// V = sum Jp'Jp, g_p = sum Jp' r; LM diagonal of points; Vinv = (V + D^2)^-1
// in FP64; w = Vinv g_p.
// V accumulates in FP64 from the kernel-precision Jacobian rows: V is then
// (up to FP64 rounding) exactly the sum of the rows' J_p'J_p, so V - J_p'J_p
// stays positive semidefinite for every row and the FP64 inverse never meets
// a block that FP32 accumulation rounded to indefinite.
template <typename T, int L, typename Cams>
__device__ __forceinline__ T point_row_accumulate(const Cams& cams, int cam, T ox, T oy,
                                                  const T* __restrict__ rf, const T X[3], const T sp[3],
                                                  T huber_a, double v[6], T g[3], const T* wg = nullptr) {
  CamProj<T> c;
  cams.proj(cam, c);
  Proj<T> o;
  project<T, L>(c, X, ox, oy, o);
  if (rf) {
    o.rx = rf[0];
    o.ry = rf[1];
  }
  // wg (optional): the row's weight, given with the residual rf.
  const T wgt = wg ? *wg : huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
  T Jp[6];
  jac_point(o, c.R, wgt, sp, Jp);
  const T rx = wgt * o.rx, ry = wgt * o.ry;
  double J[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) J[i] = (double)Jp[i];
  fma2(v[0], J[0], J[0], J[3], J[3]);
  fma2(v[1], J[0], J[1], J[3], J[4]);
  fma2(v[2], J[0], J[2], J[3], J[5]);
  fma2(v[3], J[1], J[1], J[4], J[4]);
  fma2(v[4], J[1], J[2], J[4], J[5]);
  fma2(v[5], J[2], J[2], J[5], J[5]);
  fma2(g[0], Jp[0], rx, Jp[3], ry);
  fma2(g[1], Jp[1], rx, Jp[4], ry);
  fma2(g[2], Jp[2], rx, Jp[5], ry);
  return wgt;
}

template <typename T, int L, typename Cams>
__device__ __forceinline__ void point_pass_block(
    SegPos q, const Cams& cams, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const T* __restrict__ row_obs, const double* __restrict__ pts64, const T* __restrict__ pscale, T huber_a,
    const Win& win, double min_diag, double max_diag,
    double* __restrict__ V, T* __restrict__ gp, T* __restrict__ Vinv, T* __restrict__ pk,
    const T* __restrict__ rres, T* __restrict__ wrow, double* __restrict__ Vinv64, int e_fallback) {
  if (q.valid && win_idle_e(win, q.e)) q.valid = false;
  const int e = q.valid ? q.e : e_fallback;
  const T X[3] = {(T)pts64[3 * (size_t)e], (T)pts64[3 * (size_t)e + 1], (T)pts64[3 * (size_t)e + 2]};
  const T sp[3] = {pscale[3 * (size_t)e], pscale[3 * (size_t)e + 1], pscale[3 * (size_t)e + 2]};
  double v[6] = {0, 0, 0, 0, 0, 0};
  T g[3] = {0, 0, 0};
  if (q.valid) {
    const int r1 = pt_start[e + 1];
    const int ln = q.lanes;
    int r = pt_start[e] + q.sub;
    // Two rows per iteration so both rows' loads are in flight together.
    // rres (optional): residuals evaluated in FP64 at this state, used in
    // place of the kernel-precision pred - obs.
    const T* ob = rres ? rres : row_obs;
    // With rres, wrow holds their weights (from the same FP64 evaluation);
    // otherwise the weights are computed here and written to wrow (if any).
    const T* wgi = rres ? wrow : nullptr;
    T* wout = rres ? nullptr : wrow;
    for (; r + ln < r1; r += 2 * ln) {
      const int c0 = row_cam[r], c1 = row_cam[r + ln];
      const T o0x = ob[2 * (size_t)r], o0y = ob[2 * (size_t)r + 1];
      const T o1x = ob[2 * (size_t)(r + ln)], o1y = ob[2 * (size_t)(r + ln) + 1];
      T w0, w1;
      if (rres) {
        const T f0[2] = {o0x, o0y}, f1[2] = {o1x, o1y};
        w0 = point_row_accumulate<T, L>(cams, c0, T(0), T(0), f0, X, sp, huber_a, v, g, wgi ? wgi + r : nullptr);
        w1 = point_row_accumulate<T, L>(cams, c1, T(0), T(0), f1, X, sp, huber_a, v, g,
                                        wgi ? wgi + r + ln : nullptr);
      } else {
        w0 = point_row_accumulate<T, L>(cams, c0, o0x, o0y, nullptr, X, sp, huber_a, v, g);
        w1 = point_row_accumulate<T, L>(cams, c1, o1x, o1y, nullptr, X, sp, huber_a, v, g);
      }
      if (wout) {
        wout[r] = w0;
        wout[r + ln] = w1;
      }
    }
    if (r < r1) {
      const T ox = ob[2 * (size_t)r], oy = ob[2 * (size_t)r + 1];
      const T f0[2] = {ox, oy};
      const T w0 = point_row_accumulate<T, L>(cams, row_cam[r], rres ? T(0) : ox, rres ? T(0) : oy,
                                              rres ? f0 : nullptr, X, sp, huber_a, v, g, wgi ? wgi + r : nullptr);
      if (wout) wout[r] = w0;
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) v[i] = group_sum(v[i], q.lanes);
#pragma unroll
  for (int i = 0; i < 3; ++i) g[i] = group_sum(g[i], q.lanes);
  if (!q.valid || q.sub != 0) return;
#pragma unroll
  for (int i = 0; i < 6; ++i) V[6 * (size_t)e + i] = v[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) gp[3 * (size_t)e + i] = g[i];
  // The LM diagonal, diag(V) clamped (levenberg_marquardt_strategy.cc:
  // 104-111); not stored: V is, and re-damping derives it again.
  double dg[3];
  const int di[3] = {0, 3, 5};
#pragma unroll
  for (int i = 0; i < 3; ++i) dg[i] = fmin(fmax(v[di[i]], min_diag), max_diag);
  const int wi = win_of_e(win, e);
  const double inv_radius = win_inv_r(win, wi);
  double a[6] = {v[0], v[1], v[2], v[3], v[4], v[5]};
  a[0] += dg[0] * inv_radius;
  a[3] += dg[1] * inv_radius;
  a[5] += dg[2] * inv_radius;
  double iv[6];
  if (!inv3_spd(a, iv)) {
    win.flag[wi] = 1;
    for (int i = 0; i < 6; ++i) iv[i] = 0.0;
  }
  // Vinv (kernel precision) may be NULL when only Vinv64 is read.
  if (Vinv)
#pragma unroll
    for (int i = 0; i < 6; ++i) Vinv[6 * (size_t)e + i] = (T)iv[i];
  if (Vinv64)
#pragma unroll
    for (int i = 0; i < 6; ++i) Vinv64[6 * (size_t)e + i] = iv[i];
  const double g0 = g[0], g1 = g[1], g2 = g[2];
  const T sw[3] = {(T)((double)sp[0] * (iv[0] * g0 + iv[1] * g1 + iv[2] * g2)),
                   (T)((double)sp[1] * (iv[1] * g0 + iv[3] * g1 + iv[4] * g2)),
                   (T)((double)sp[2] * (iv[2] * g0 + iv[4] * g1 + iv[5] * g2))};
  store_pk(pk, e, X, sw);
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_point_pass(
    Seg sg, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const T* __restrict__ row_obs, const double* __restrict__ pts64, const T* __restrict__ pscale,
    const T* __restrict__ ctab, T huber_a, Win win, double min_diag, double max_diag,
    double* __restrict__ V, T* __restrict__ gp,
    T* __restrict__ Vinv, T* __restrict__ pk, const T* __restrict__ rres, T* __restrict__ wrow,
    double* __restrict__ Vinv64) {
  point_pass_block<T, L>(seg_locate(sg), GlobalCams<T, L>{ctab, nullptr}, pt_start, row_cam, row_obs, pts64, pscale,
                         huber_a, win, min_diag, max_diag, V, gp, Vinv, pk, rres, wrow, Vinv64, sg.e0[0]);
}

// The same with the cameras' projection records staged in LDS (FP32,
// packed): persistent, each workgroup fills the table once and its four
// groups of kBlock threads take the virtual blocks blockIdx.x * 4 + j,
// stepping by 4 * gridDim.x.
__global__ __launch_bounds__(kStageThreads) void k_point_pass_lds(
    Seg sg, int nc, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const float* __restrict__ row_obs, const double* __restrict__ pts64, const float* __restrict__ pscale,
    const float* __restrict__ ctab, float huber_a, Win win, double min_diag, double max_diag,
    double* __restrict__ V, float* __restrict__ gp,
    float* __restrict__ Vinv, float* __restrict__ pk, const float* __restrict__ rres, float* __restrict__ wrow,
    double* __restrict__ Vinv64) {
  extern __shared__ float4 stage4[];
  stage_cams<4>(stage4, nc, ctab, nullptr);
  const LdsCams<4> cams{reinterpret_cast<const float*>(stage4)};
  const int j = threadIdx.x / kBlock, tid = threadIdx.x % kBlock;
  const int nblk = sg.blk0[sg.n];
  for (int vb = blockIdx.x * 4 + j; vb < nblk; vb += 4 * gridDim.x)
    point_pass_block<float, 0>(seg_locate_at(sg, vb, tid, kBlock), cams, pt_start, row_cam, row_obs, pts64, pscale,
                               huber_a, win, min_diag, max_diag, V, gp, Vinv, pk, rres, wrow, Vinv64, sg.e0[0]);
}

// Re-damp after a rejected step: Vinv, w from stored V and g_p, with the
// LM diagonal (diag(V) clamped, as the pass formed it) under the new radius.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_point_damp(int ne, const double* __restrict__ V,
                                                       const T* __restrict__ gp, double min_diag, double max_diag,
                                                       Win win, const T* __restrict__ pscale,
                                                       T* __restrict__ Vinv, T* __restrict__ pk,
                                                       double* __restrict__ Vinv64) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne || win_idle_e(win, e)) return;
  const int wi = win_of_e(win, e);
  const double inv_radius = win_inv_r(win, wi);
  double a[6];
  for (int i = 0; i < 6; ++i) a[i] = V[6 * (size_t)e + i];
  a[0] += fmin(fmax(a[0], min_diag), max_diag) * inv_radius;
  a[3] += fmin(fmax(a[3], min_diag), max_diag) * inv_radius;
  a[5] += fmin(fmax(a[5], min_diag), max_diag) * inv_radius;
  double iv[6];
  if (!inv3_spd(a, iv)) {
    win.flag[wi] = 1;
    for (int i = 0; i < 6; ++i) iv[i] = 0.0;
  }
  if (Vinv)
    for (int i = 0; i < 6; ++i) Vinv[6 * (size_t)e + i] = (T)iv[i];
  if (Vinv64)
    for (int i = 0; i < 6; ++i) Vinv64[6 * (size_t)e + i] = iv[i];
  const double g0 = gp[3 * (size_t)e], g1 = gp[3 * (size_t)e + 1], g2 = gp[3 * (size_t)e + 2];
  const T w[3] = {(T)(iv[0] * g0 + iv[1] * g1 + iv[2] * g2), (T)(iv[1] * g0 + iv[3] * g1 + iv[4] * g2),
                  (T)(iv[2] * g0 + iv[4] * g1 + iv[5] * g2)};
  const T sw[3] = {pscale[3 * (size_t)e] * w[0], pscale[3 * (size_t)e + 1] * w[1], pscale[3 * (size_t)e + 2] * w[2]};
  store_pk_v(pk, e, sw);
}

// ------------------------------------------------------------ camera pass
// One workgroup per camera-major chunk (all rows of one camera): partial
// U = sum Jc'Jc (45), g_c = sum Jc' r (9) and, with_h, h = sum Jc'(r - Jp w_e).
// Accumulate one camera-major row into U (45), g (9), h (9), with the
// camera Jacobi scale left out (applied per camera in k_cam_finalize).
template <typename T>
__device__ __forceinline__ T cam_row_accumulate(const CamProj<T>& c, const T dR[27], const T X[3],
                                                const T sw[3], T ox, T oy, bool given, const T* wg, T huber_a,
                                                int with_h, T acc[63]) {
  // FP32: Jacobian in the local rotation basis (jac_cam_local); FP64: in the
  // parameter basis through the table's dR.
  constexpr bool kLocal = std::is_same<T, float>::value;
  const T one[9] = {T(1), T(1), T(1), T(1), T(1), T(1), T(1), T(1), T(1)};
  Proj<T> o;
  T y[3];
  if constexpr (kLocal)
    project_y(c, X, given ? T(0) : ox, given ? T(0) : oy, o, y);
  else
    project(c, X, given ? T(0) : ox, given ? T(0) : oy, o);
  if (given) {  // (ox, oy) is the residual itself
    o.rx = ox;
    o.ry = oy;
  }
  // wg (optional): the row's weight, given with the residual.
  const T wgt = wg ? *wg : huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
  T Jc[18];
  if constexpr (kLocal)
    jac_cam_local(o, y, wgt, Jc);
  else
    jac_cam(o, dR, X, wgt, one, Jc);
  const T rx = wgt * o.rx, ry = wgt * o.ry;
  T tx = rx, ty = ry;
  if (with_h) {
    T u[2];
    jac_point_apply(o, c.R, sw, u);
    tx -= wgt * u[0];
    ty -= wgt * u[1];
  }
  int k = 0;
#pragma unroll
  for (int a = 0; a < 9; ++a)
#pragma unroll
    for (int b = a; b < 9; ++b) fma2(acc[k++], Jc[a], Jc[b], Jc[9 + a], Jc[9 + b]);
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    fma2(acc[45 + a], Jc[a], rx, Jc[9 + a], ry);
    fma2(acc[54 + a], Jc[a], tx, Jc[9 + a], ty);
  }
  return wgt;
}

// Camera-major launches: workgroup slot -> chunk (-1: empty slot), or the
// identity without a slot map.  Workgroups are dealt to the XCDs round
// robin, so the host places the chunks of point range q in the slots
// congruent to q mod 8: each XCD's L2 then serves gathers from one eighth
// of the point records.
__device__ __forceinline__ int chunk_of_slot(const int32_t* __restrict__ slot_chunk) {
  return slot_chunk ? slot_chunk[blockIdx.x] : (int)blockIdx.x;
}

template <typename T, int B>
__global__ __launch_bounds__(B) void k_cam_pass(
    const int32_t* __restrict__ chunk_f, const int32_t* __restrict__ chunk_begin,
    const int32_t* __restrict__ chunk_end, const int32_t* __restrict__ f_to_cam,
    const int32_t* __restrict__ cm_e, const T* __restrict__ cm_obs, const T* __restrict__ pk,
    const T* __restrict__ ctab, T huber_a, int with_h, T* __restrict__ partial, Win win,
    const int32_t* __restrict__ slot_chunk, const T* __restrict__ cres, T* __restrict__ wcm,
    const T* __restrict__ gtab) {
  constexpr bool kLocal = std::is_same<T, float>::value;
  __shared__ T lds[63 * (B / 64)];
  const int ch = chunk_of_slot(slot_chunk);
  if (ch < 0) return;
  const int f = chunk_f[ch];
  if (win_idle_f(win, f)) return;
  const int cam = f_to_cam[f];
  const T* tab = ctab + (size_t)kCamTab * cam;
  CamProj<T> c;
  load_proj(tab, c);
  T dR[27];
  if constexpr (!kLocal) {
#pragma unroll
    for (int i = 0; i < 27; ++i) dR[i] = tab[kTabDR + i];
  }
  T acc[63];
#pragma unroll
  for (int i = 0; i < 63; ++i) acc[i] = T(0);
  const int end = chunk_end[ch];
  int i = chunk_begin[ch] + threadIdx.x;
  // cres (optional): residuals evaluated in FP64 at this state, camera-major,
  // used in place of the kernel-precision pred - obs, and then wcm holds
  // their weights; otherwise the weights are computed here and written to
  // wcm (if any) for the matvecs.
  const bool given = cres != nullptr;
  const T* ob = given ? cres : cm_obs;
  const T* wgi = given ? wcm : nullptr;
  T* wout = given ? nullptr : wcm;
  // One row per iteration: with the 63 accumulators live, a second row in
  // flight costs more in registers and scheduling than its loads save
  // (FP32 at 4.0M observations: 47 vs 52 us per pass).
  for (; i < end; i += B) {
    T X0[3], s0[3];
    load_pk(pk, cm_e[i], X0, s0);
    const T w0 = cam_row_accumulate(c, dR, X0, s0, ob[2 * (size_t)i], ob[2 * (size_t)i + 1], given,
                                    wgi ? wgi + i : nullptr, huber_a, with_h, acc);
    if (wout) wout[i] = w0;
  }
  T s = block_reduce<T, 63, B>(acc, lds);
  if constexpr (kLocal) {
    if (threadIdx.x < 64) s = chunk_to_param_basis63(s, (int)threadIdx.x, gtab + 9 * (size_t)cam);
  }
  if (threadIdx.x < 63) partial[(size_t)kPartial * ch + threadIdx.x] = s;
}

// Sum of entry k over rows [lo, hi) of the chunk partials (kPartial wide),
// accumulated in A: four independent sums, so four rows' loads are in
// flight together instead of one dependent load per row.
template <typename A, typename T>
__device__ __forceinline__ A sum_partials(const T* __restrict__ partial, int lo, int hi, int k) {
  A s0 = A(0), s1 = A(0), s2 = A(0), s3 = A(0);
  int ch = lo;
  for (; ch + 4 <= hi; ch += 4) {
    const T p0 = partial[(size_t)kPartial * ch + k], p1 = partial[(size_t)kPartial * (ch + 1) + k];
    const T p2 = partial[(size_t)kPartial * (ch + 2) + k], p3 = partial[(size_t)kPartial * (ch + 3) + k];
    s0 += (A)p0;
    s1 += (A)p1;
    s2 += (A)p2;
    s3 += (A)p3;
  }
  for (; ch < hi; ++ch) s0 += (A)partial[(size_t)kPartial * ch + k];
  return (s0 + s1) + (s2 + s3);
}

// One wave per F block: lanes sum the chunk partials (one value each), lane
// 0 forms the LM diagonal and the FP64 inverse of U + D^2 (block-Jacobi
// preconditioner); rhs = h.
// The partials of F block f are rows [f_lo[f], f_hi[f]) of `partial`.
template <typename T>
__global__ __launch_bounds__(64) void k_cam_finalize(
    int nf, const int32_t* __restrict__ f_lo, const int32_t* __restrict__ f_hi, const T* __restrict__ partial,
    const T* __restrict__ cscale, Win win, double min_diag, double max_diag,
    double* __restrict__ diag_c, int recompute_diag, int with_h, T* __restrict__ U, T* __restrict__ gc,
    T* __restrict__ rhs, T* __restrict__ Minv, const double* __restrict__ Tsj) {
  __shared__ double sum[64];
  __shared__ int ij[64];
  __shared__ double lf[9][9];
  const int f = blockIdx.x;
  const int k = threadIdx.x;
  if (win_idle_f(win, f)) return;
  if (k == 0) {
    int q = 0;
    for (int a = 0; a < 9; ++a)
      for (int b = a; b < 9; ++b) ij[q++] = a * 16 + b;
    for (int a = 0; a < 9; ++a) ij[45 + a] = a * 16 + 15;
    for (int a = 0; a < 9; ++a) ij[54 + a] = a * 16 + 15;
    ij[63] = 0;
  }
  __syncthreads();
  double v = 0.0;
  if (k < 63) {
    v = sum_partials<double>(partial, f_lo[f], f_hi[f], k);
    // Partials carry the unscaled camera columns: scale entry (a, b) by
    // s_a s_b, vectors by s_a.
    const int a = ij[k] >> 4, b = ij[k] & 15;
    v *= (double)cscale[9 * (size_t)f + a] * (b < 9 ? (double)cscale[9 * (size_t)f + b] : 1.0);
  }
  const T vt = (T)v;
  if (k < 45) U[45 * (size_t)f + k] = vt;
  else if (k < 54) gc[9 * (size_t)f + (k - 45)] = vt;
  else if (k < 63 && with_h) rhs[9 * (size_t)f + (k - 54)] = vt;
  sum[k] = (double)vt;
  __syncthreads();
  // Lanes 0..8 factor U + D^2 (FP64) together, lane i holding row i of the
  // factor: the same operations in the same order as chol9, the rows'
  // dot products and divisions in parallel, each finished column published
  // in LDS.  Then lane c writes column c of the inverse (chol9_inv_col).
  // Lanes 9..63 mirror row 8 without writing, so every lane reaches the
  // barriers.
  const int row = k < 9 ? k : 8;
  const int wi = win_of_f(win, f);
  const double inv_radius = win_inv_r(win, wi);
  // The block to invert: U, or with SCHUR_JACOBI the diagonal block of S
  // (Tsj); plus the LM diagonal, taken from the Jacobian's columns (U's
  // diagonal) in either case.
  // Row `row` of that block (stored as its upper triangle, packed).
  double arow[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int idx = j < row ? ut9(j, row) : ut9(row, j);
    arow[j] = Tsj ? Tsj[45 * (size_t)f + idx] : sum[idx];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double dg;
    if (recompute_diag) {
      dg = fmin(fmax(sum[ut9(i, i)], min_diag), max_diag);
      if (k == 0) diag_c[9 * (size_t)f + i] = dg;
    } else {
      dg = diag_c[9 * (size_t)f + i];
    }
    if (i == row) arow[i] += dg * inv_radius;
  }
  double lr[9];  // row `row` of the factor, columns < j so far
#pragma unroll
  for (int q = 0; q < 9; ++q) lr[q] = 0.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    // Row `row`'s entry in column j before the division (lane j: the pivot).
    double sj = arow[j];
#pragma unroll
    for (int q = 0; q < j; ++q) sj -= lr[q] * lf[j][q];
    double d = lane_bcast(sj, j);
    ok = ok && d > 0.0;
    d = ok ? sqrt(d) : 1.0;
    lr[j] = row == j ? d : sj / d;
    if (k < 9 && row >= j) lf[row][j] = lr[j];
    __syncthreads();
  }
  if (k >= 9) return;
  double e[9];
  if (ok) {
    double l[45];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int q = 0; q <= i; ++q) l[lt9(i, q)] = lf[i][q];
    chol9_inv_col(l, k, e);
  } else {
    if (k == 0) win.flag[wi] = 1;
#pragma unroll
    for (int i = 0; i < 9; ++i) e[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) Minv[81 * (size_t)f + 9 * i + k] = (T)e[i];
}

// SCHUR_JACOBI (schur_jacobi_preconditioner.cc:83-103): per camera-major
// chunk, the sum over its rows of Jc' (I - Jp Vinv Jp') Jc (45 values;
// camera columns unscaled like the camera pass, point columns scaled, loss
// weight applied) -- each row's share of U - W Vinv W', the camera's
// diagonal block of S without its LM diagonal, when a camera observes a
// point once.
template <typename T, int B>
__global__ __launch_bounds__(B) void k_sj_pass(
    const int32_t* __restrict__ chunk_f, const int32_t* __restrict__ chunk_begin,
    const int32_t* __restrict__ chunk_end, const int32_t* __restrict__ f_to_cam,
    const int32_t* __restrict__ cm_e, const T* __restrict__ cm_obs, const T* __restrict__ pk,
    const T* __restrict__ pscale, const double* __restrict__ Vinv64, const T* __restrict__ ctab,
    T huber_a, T* __restrict__ partial, Win win, const int32_t* __restrict__ slot_chunk,
    const T* __restrict__ wcm) {
  __shared__ T lds[45 * (B / 64)];
  const int ch = chunk_of_slot(slot_chunk);
  if (ch < 0) return;
  const int f = chunk_f[ch];
  if (win_idle_f(win, f)) return;
  const T* tab = ctab + (size_t)kCamTab * f_to_cam[f];
  CamProj<T> c;
  load_proj(tab, c);
  T dR[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) dR[i] = tab[kTabDR + i];
  const T one[9] = {T(1), T(1), T(1), T(1), T(1), T(1), T(1), T(1), T(1)};
  T acc[45];
#pragma unroll
  for (int i = 0; i < 45; ++i) acc[i] = T(0);
  for (int i = chunk_begin[ch] + threadIdx.x; i < chunk_end[ch]; i += B) {
    const int e = cm_e[i];
    T X[3], unused[3];
    load_pk(pk, e, X, unused);
    const T sp[3] = {pscale[3 * (size_t)e], pscale[3 * (size_t)e + 1], pscale[3 * (size_t)e + 2]};
    const double* iv = Vinv64 + 6 * (size_t)e;
    const double Vi[9] = {iv[0], iv[1], iv[2], iv[1], iv[3], iv[4], iv[2], iv[4], iv[5]};
    Proj<T> o;
    project(c, X, cm_obs[2 * (size_t)i], cm_obs[2 * (size_t)i + 1], o);
    const T wgt = wcm ? wcm[i] : huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T Jc[18], Jp[6];
    jac_cam(o, dR, X, wgt, one, Jc);
    jac_point(o, c.R, wgt, sp, Jp);
    // A = Jp Vinv Jp' (2 x 2), then B = (I - A) Jc: the row's share of the
    // diagonal block, Jc'Jc - Jc'A Jc, formed per row so that the sum
    // carries no cancellation between two large sums.  A comes from the
    // FP64 inverse of the FP64-accumulated V of these same Jacobian rows, so
    // I - A is positive semidefinite up to FP64 rounding.
    double Jd[6], G[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) Jd[q] = (double)Jp[q];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        G[3 * r + j] = Jd[3 * r] * Vi[j] + Jd[3 * r + 1] * Vi[3 + j] + Jd[3 * r + 2] * Vi[6 + j];
    const T m00 = (T)(1.0 - (G[0] * Jd[0] + G[1] * Jd[1] + G[2] * Jd[2]));
    const T m01 = (T)(-(G[0] * Jd[3] + G[1] * Jd[4] + G[2] * Jd[5]));
    const T m11 = (T)(1.0 - (G[3] * Jd[3] + G[4] * Jd[4] + G[5] * Jd[5]));
    T B0[9], B1[9];
#pragma unroll
    for (int b = 0; b < 9; ++b) {
      B0[b] = m00 * Jc[b] + m01 * Jc[9 + b];
      B1[b] = m01 * Jc[b] + m11 * Jc[9 + b];
    }
    int k = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
      for (int b = a; b < 9; ++b) fma2(acc[k++], Jc[a], B0[b], Jc[9 + a], B1[b]);
  }
  const T sum = block_reduce<T, 45, B>(acc, lds);
  if (threadIdx.x < 45) partial[(size_t)kPartial * ch + threadIdx.x] = sum;
}

// Tsj_f = s_a s_b (sum of the F block's SCHUR_JACOBI partials), 45 per block,
// kept in FP64: the block is factored in FP64 and can be ill-conditioned
// enough that an FP32 rounding of it is no longer positive definite.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sj_finalize(int nf, const int32_t* __restrict__ f_lo,
                                                        const int32_t* __restrict__ f_hi,
                                                        const T* __restrict__ partial, const T* __restrict__ cscale,
                                                        Win win, double* __restrict__ Tsj) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int f = (int)(t / 45), k = (int)(t % 45);
  if (f >= nf || win_idle_f(win, f)) return;
  int a = 0;
  while (a < 8 && k >= ut9(a + 1, a + 1)) ++a;
  const int b = a + (k - ut9(a, a));
  const double v = sum_partials<double>(partial, f_lo[f], f_hi[f], k);
  Tsj[45 * (size_t)f + k] = v * (double)cscale[9 * (size_t)f + a] * (double)cscale[9 * (size_t)f + b];
}

// Re-damp the camera preconditioner after a rejected step.
template <typename T>
__global__ __launch_bounds__(64) void k_cam_damp(int nf, const T* __restrict__ U,
                                                 const double* __restrict__ diag_c, Win win,
                                                 T* __restrict__ Minv, const double* __restrict__ Tsj) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf || win_idle_f(win, f)) return;
  const int wi = win_of_f(win, f);
  const double inv_radius = win_inv_r(win, wi);
  double a[45];
#pragma unroll
  for (int q = 0; q < 45; ++q) a[q] = Tsj ? Tsj[45 * (size_t)f + q] : (double)U[45 * (size_t)f + q];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[ut9(i, i)] += diag_c[9 * (size_t)f + i] * inv_radius;
  double l[45];
  const bool ok = chol9(a, l);
  if (!ok) win.flag[wi] = 1;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    double e[9];
    if (ok) {
      chol9_inv_col(l, c, e);
    } else {
#pragma unroll
      for (int i = 0; i < 9; ++i) e[i] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Minv[81 * (size_t)f + 9 * i + c] = (T)e[i];
  }
}

// ---------------------------------------------------------------- matvecs
// Direction table for a camera vector x (F order, 9 per block, or NULL):
// xs = sign * cscale .* x;  M = sum_k xs_k dR_k, dt = xs_t, dintr.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cam_dir(int nc, const int32_t* __restrict__ cam_to_f,
                                                    const T* __restrict__ cscale,
                                                    const T* __restrict__ ctab,
                                                    const T* __restrict__ x, T sign,
                                                    T* __restrict__ dtab, const int32_t* __restrict__ all_done,
                                                    Win win) {
  if (all_done && *all_done) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  T* o = dtab + (size_t)kDirTab * c;
  const int f = cam_to_f[c];
  if (f >= 0 && win_idle_f(win, f)) return;
  if (f < 0 || x == nullptr) {
    for (int i = 0; i < kDirTab; ++i) o[i] = T(0);
    return;
  }
  T xs[9];
  for (int i = 0; i < 9; ++i) xs[i] = sign * cscale[9 * (size_t)f + i] * x[9 * (size_t)f + i];
  const T* dR = ctab + (size_t)kCamTab * c + kTabDR;
  for (int i = 0; i < 9; ++i) o[i] = xs[0] * dR[i] + xs[1] * dR[9 + i] + xs[2] * dR[18 + i];
  for (int i = 0; i < 6; ++i) o[9 + i] = xs[3 + i];
  o[15] = T(0);
}

// Per E block: q = sum_rows Jp'(beta r + Jc x) (x via dtab); z = Vinv q
// (apply_vinv) or q; pk's v part (if pk) becomes s_p .* z.
template <typename T, int L, typename Cams>
__device__ __forceinline__ void mv_point_block(
    SegPos qp, const Cams& cams, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const T* __restrict__ row_obs, const T* __restrict__ pkX, const T* __restrict__ pscale,
    const T* __restrict__ Vinv, T huber_a, T beta, int apply_vinv, T* __restrict__ zout, T* __restrict__ pk,
    const Win& win, const T* __restrict__ wrow, const T* __restrict__ rres, const double* __restrict__ Vinv64,
    int e_fallback) {
  if (qp.valid && win_idle_e(win, qp.e)) qp.valid = false;
  const int e = qp.valid ? qp.e : e_fallback;
  T X[3], unused[3];
  load_pk(pkX, e, X, unused);
  const T sp[3] = {pscale[3 * (size_t)e], pscale[3 * (size_t)e + 1], pscale[3 * (size_t)e + 2]};
  T q[3] = {0, 0, 0};
  // wrow (optional): the row's loss weight from the last pass; rres
  // (optional): its residual evaluated in FP64 (beta != 0 only).
  auto row = [&](int r, int cam, T ox, T oy) {
    CamProj<T> c;
    cams.proj(cam, c);
    Proj<T> o;
    project<T, L>(c, X, ox, oy, o);
    if (rres) {
      o.rx = rres[2 * (size_t)r];
      o.ry = rres[2 * (size_t)r + 1];
    }
    const T wgt = wrow ? wrow[r] : huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T Jp[6];
    jac_point(o, c.R, wgt, sp, Jp);
    T v[2];
    jac_cam_dir<T, L>(o, cams.dir(cam), X, wgt, v);
    v[0] += beta * wgt * o.rx;
    v[1] += beta * wgt * o.ry;
    fma2(q[0], Jp[0], v[0], Jp[3], v[1]);
    fma2(q[1], Jp[1], v[0], Jp[4], v[1]);
    fma2(q[2], Jp[2], v[0], Jp[5], v[1]);
  };
  if (qp.valid) {
    const int r1 = pt_start[e + 1], ln = qp.lanes;
    int r = pt_start[e] + qp.sub;
    // Two rows per iteration: both rows' loads are in flight together.
    for (; r + ln < r1; r += 2 * ln) {
      const int c0 = row_cam[r], c1 = row_cam[r + ln];
      const T o0x = row_obs[2 * (size_t)r], o0y = row_obs[2 * (size_t)r + 1];
      const T o1x = row_obs[2 * (size_t)(r + ln)], o1y = row_obs[2 * (size_t)(r + ln) + 1];
      row(r, c0, o0x, o0y);
      row(r + ln, c1, o1x, o1y);
    }
    if (r < r1) row(r, row_cam[r], row_obs[2 * (size_t)r], row_obs[2 * (size_t)r + 1]);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) q[i] = group_sum(q[i], qp.lanes);
  if (!qp.valid || qp.sub != 0) return;
  T zv[3] = {q[0], q[1], q[2]};
  if (apply_vinv && Vinv64) {
    // The FP64 inverse of the FP64-accumulated V: an FP32 copy of the
    // inverse of an ill-conditioned point block is too coarse to keep the
    // implicit Schur complement positive definite.
    const double* iv = Vinv64 + 6 * (size_t)e;
    const double q0 = q[0], q1 = q[1], q2 = q[2];
    zv[0] = (T)(iv[0] * q0 + iv[1] * q1 + iv[2] * q2);
    zv[1] = (T)(iv[1] * q0 + iv[3] * q1 + iv[4] * q2);
    zv[2] = (T)(iv[2] * q0 + iv[4] * q1 + iv[5] * q2);
  } else if (apply_vinv) {
    const T* iv = Vinv + 6 * (size_t)e;
    zv[0] = iv[0] * q[0] + iv[1] * q[1] + iv[2] * q[2];
    zv[1] = iv[1] * q[0] + iv[3] * q[1] + iv[4] * q[2];
    zv[2] = iv[2] * q[0] + iv[4] * q[1] + iv[5] * q[2];
  }
  for (int i = 0; i < 3; ++i) zout[3 * (size_t)e + i] = zv[i];
  if (pk) {
    const T sz[3] = {sp[0] * zv[0], sp[1] * zv[1], sp[2] * zv[2]};
    store_pk_v(pk, e, sz);
  }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_mv_point(
    Seg sg, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const T* __restrict__ row_obs, const T* __restrict__ pkX, const T* __restrict__ pscale,
    const T* __restrict__ ctab, const T* __restrict__ dtab, const T* __restrict__ Vinv, T huber_a,
    T beta, int apply_vinv, T* __restrict__ zout, T* __restrict__ pk, const int32_t* __restrict__ all_done,
    Win win, const T* __restrict__ wrow, const T* __restrict__ rres, const double* __restrict__ Vinv64) {
  if (all_done && *all_done) return;
  mv_point_block<T, L>(seg_locate(sg), GlobalCams<T, L>{ctab, dtab}, pt_start, row_cam, row_obs, pkX, pscale,
                       Vinv, huber_a, beta, apply_vinv, zout, pk, win, wrow, rres, Vinv64, sg.e0[0]);
}

// The same with every camera's record and direction staged in LDS (FP32,
// packed); persistent as k_point_pass_lds.
__global__ __launch_bounds__(kStageThreads) void k_mv_point_lds(
    Seg sg, int nc, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const float* __restrict__ row_obs, const float* __restrict__ pkX, const float* __restrict__ pscale,
    const float* __restrict__ ctab, const float* __restrict__ dtab, const float* __restrict__ Vinv, float huber_a,
    float beta, int apply_vinv, float* __restrict__ zout, float* __restrict__ pk, const int32_t* __restrict__ all_done,
    Win win, const float* __restrict__ wrow, const float* __restrict__ rres, const double* __restrict__ Vinv64) {
  extern __shared__ float4 stage4[];
  if (all_done && *all_done) return;
  stage_cams<8>(stage4, nc, ctab, dtab);
  const LdsCams<8> cams{reinterpret_cast<const float*>(stage4)};
  const int j = threadIdx.x / kBlock, tid = threadIdx.x % kBlock;
  const int nblk = sg.blk0[sg.n];
  for (int vb = blockIdx.x * 4 + j; vb < nblk; vb += 4 * gridDim.x)
    mv_point_block<float, 0>(seg_locate_at(sg, vb, tid, kBlock), cams, pt_start, row_cam, row_obs, pkX, pscale,
                             Vinv, huber_a, beta, apply_vinv, zout, pk, win, wrow, rres, Vinv64, sg.e0[0]);
}

// Per chunk: partial_c = sum_rows Jc'(Jp z_e) (9 values).
template <typename T, int B>
__global__ __launch_bounds__(B) void k_mv_cam(
    const int32_t* __restrict__ chunk_f, const int32_t* __restrict__ chunk_begin,
    const int32_t* __restrict__ chunk_end, const int32_t* __restrict__ f_to_cam,
    const int32_t* __restrict__ cm_e, const T* __restrict__ cm_obs, const T* __restrict__ pk,
    const T* __restrict__ cscale, const T* __restrict__ ctab, T huber_a, T* __restrict__ partial,
    const int32_t* __restrict__ all_done, Win win, const int32_t* __restrict__ slot_chunk,
    const T* __restrict__ wcm, const T* __restrict__ gtab) {
  // FP32: rows summed in the local rotation basis (jac_cam_local), taken to
  // the parameter basis and scaled once per chunk.
  constexpr bool kLocal = std::is_same<T, float>::value;
  __shared__ T lds[9 * (B / 64)];
  if (all_done && *all_done) return;
  const int ch = chunk_of_slot(slot_chunk);
  if (ch < 0) return;
  const int f = chunk_f[ch];
  if (win_idle_f(win, f)) return;
  const int cam = f_to_cam[f];
  const T* tab = ctab + (size_t)kCamTab * cam;
  CamProj<T> c;
  load_proj(tab, c);
  T dR[27];
  T cs[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) cs[i] = cscale[9 * (size_t)f + i];
  if constexpr (!kLocal) {
#pragma unroll
    for (int i = 0; i < 27; ++i) dR[i] = tab[kTabDR + i];
  }
  T acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = T(0);
  const int end = chunk_end[ch];
  // wcm (optional): the row's loss weight from the last pass.  Without it
  // and with a loss the weight needs the residual, hence the observation.
  const bool need_obs = !wcm && huber_a > T(0);
  auto row = [&](int i, const T X[3], const T sz[3], T ox, T oy) {
    Proj<T> o;
    if constexpr (kLocal) {
      T y[3];
      project_y(c, X, ox, oy, o, y);
      const T wgt = wcm ? wcm[i] : huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
      T u[2];
      jac_point_apply(o, c.R, sz, u);
      const T w2 = wgt * wgt;
      const T s0 = w2 * u[0], s1 = w2 * u[1];
      const T v0 = s0 * o.P[0] + s1 * o.P[3], v1 = s0 * o.P[1] + s1 * o.P[4], v2 = s0 * o.P[2] + s1 * o.P[5];
      acc[0] += y[1] * v2 - y[2] * v1;
      acc[1] += y[2] * v0 - y[0] * v2;
      acc[2] += y[0] * v1 - y[1] * v0;
      acc[3] += v0;
      acc[4] += v1;
      acc[5] += v2;
      fma2(acc[6], s0, o.jf[0], s1, o.jf[1]);
      fma2(acc[7], s0, o.jk1[0], s1, o.jk1[1]);
      fma2(acc[8], s0, o.jk2[0], s1, o.jk2[1]);
    } else {
      project(c, X, ox, oy, o);
      const T wgt = wcm ? wcm[i] : huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
      T Jc[18];
      jac_cam(o, dR, X, wgt, cs, Jc);
      T u[2];
      jac_point_apply(o, c.R, sz, u);
      u[0] *= wgt;
      u[1] *= wgt;
#pragma unroll
      for (int a = 0; a < 9; ++a) fma2(acc[a], Jc[a], u[0], Jc[9 + a], u[1]);
    }
  };
  int i = chunk_begin[ch] + threadIdx.x;
  // Two rows per iteration: both rows' point-record gathers in flight.
  for (; i + B < end; i += 2 * B) {
    const int e0 = cm_e[i], e1 = cm_e[i + B];
    T o0x = T(0), o0y = T(0), o1x = T(0), o1y = T(0);
    if (need_obs) {
      o0x = cm_obs[2 * (size_t)i];
      o0y = cm_obs[2 * (size_t)i + 1];
      o1x = cm_obs[2 * (size_t)(i + B)];
      o1y = cm_obs[2 * (size_t)(i + B) + 1];
    }
    T X0[3], s0[3], X1[3], s1[3];
    load_pk(pk, e0, X0, s0);
    load_pk(pk, e1, X1, s1);
    row(i, X0, s0, o0x, o0y);
    row(i + B, X1, s1, o1x, o1y);
  }
  if (i < end) {
    T X0[3], s0[3];
    load_pk(pk, cm_e[i], X0, s0);
    row(i, X0, s0, need_obs ? cm_obs[2 * (size_t)i] : T(0), need_obs ? cm_obs[2 * (size_t)i + 1] : T(0));
  }
  T s = block_reduce<T, 9, B>(acc, lds);
  if constexpr (kLocal) {
    if (threadIdx.x < 64) s = chunk_to_param_basis9(s, (int)threadIdx.x, gtab + 9 * (size_t)cam, cs);
  }
  if (threadIdx.x < 9) partial[(size_t)kPartial * ch + threadIdx.x] = s;
}

// y_f = [u_term] (U + D^2) x_f - sum partials (L = 0), or D^2 x_f + sum
// partials (L = 1, partials of F'(F x - E z)).
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_mv_finalize(
    int nf, const int32_t* __restrict__ f_lo, const int32_t* __restrict__ f_hi, const T* __restrict__ partial,
    const T* __restrict__ U, const double* __restrict__ diag_c, Win win,
    const T* __restrict__ x, int u_term, T* __restrict__ y, const int32_t* __restrict__ all_done) {
  if (all_done && *all_done) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf || win_idle_f(win, f)) return;
  const double inv_radius = win_inv_r(win, win_of_f(win, f));
  T s[9];
  for (int k = 0; k < 9; ++k) s[k] = T(0);
  for (int ch = f_lo[f]; ch < f_hi[f]; ++ch)
    for (int k = 0; k < 9; ++k) s[k] += partial[(size_t)kPartial * ch + k];
  const T* xf = x + 9 * (size_t)f;
  if (L == 1) {
    for (int i = 0; i < 9; ++i) y[9 * (size_t)f + i] = (T)(diag_c[9 * (size_t)f + i] * inv_radius) * xf[i] + s[i];
  } else if (u_term) {
    const T* u = U + 45 * (size_t)f;
    for (int i = 0; i < 9; ++i) {
      T a = (T)(diag_c[9 * (size_t)f + i] * inv_radius) * xf[i];
      for (int j = 0; j < 9; ++j) a += u[i <= j ? ut9(i, j) : ut9(j, i)] * xf[j];
      y[9 * (size_t)f + i] = a - s[i];
    }
  } else {
    for (int i = 0; i < 9; ++i) y[9 * (size_t)f + i] = -s[i];
  }
}

// rhs_f = g_f - sum partials (L = 0: partials of Jc'Jp w; L = 1: partials
// already carry the sign, -F'E w).
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_rhs_finalize(int nf, const int32_t* __restrict__ f_lo,
                                                         const int32_t* __restrict__ f_hi,
                                                         const T* __restrict__ partial,
                                                         const T* __restrict__ gc,
                                                         T* __restrict__ rhs, Win win) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf || win_idle_f(win, f)) return;
  T s[9];
  for (int k = 0; k < 9; ++k) s[k] = T(0);
  for (int ch = f_lo[f]; ch < f_hi[f]; ++ch)
    for (int k = 0; k < 9; ++k) s[k] += partial[(size_t)kPartial * ch + k];
  for (int k = 0; k < 9; ++k) rhs[9 * (size_t)f + k] = L == 0 ? gc[9 * (size_t)f + k] - s[k] : gc[9 * (size_t)f + k] + s[k];
}

// Model cost change terms: sum over rows of m.(r + m/2) with m = J step
// (camera part via dtab, point part from step_p); FP64 accumulation.
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_model_cost(
    Seg sg, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const T* __restrict__ row_obs, const T* __restrict__ pkX, int xstride, const T* __restrict__ pscale,
    const T* __restrict__ ctab, const T* __restrict__ dtab, const T* __restrict__ step_p,
    T huber_a, double* __restrict__ partial, Win win) {
  __shared__ double lds[kBlock / 64];
  SegPos qp = seg_locate(sg);
  if (qp.valid && win_idle_e(win, qp.e)) qp.valid = false;
  double acc[1] = {0.0};
  if (qp.valid) {
    const int e = qp.e;
    // Point position: a packed record (xstride kPk) or FP64 positions (3).
    const T X[3] = {pkX[(size_t)xstride * e], pkX[(size_t)xstride * e + 1], pkX[(size_t)xstride * e + 2]};
    const T sp[3] = {pscale[3 * (size_t)e], pscale[3 * (size_t)e + 1], pscale[3 * (size_t)e + 2]};
    const T dp[3] = {step_p[3 * (size_t)e], step_p[3 * (size_t)e + 1], step_p[3 * (size_t)e + 2]};
    for (int r = pt_start[e] + qp.sub; r < pt_start[e + 1]; r += qp.lanes) {
      const int cam = row_cam[r];
      CamProj<T> c;
      load_proj<T, L>(ctab + (size_t)kCamTab * cam, c);
      Proj<T> o;
      project<T, L>(c, X, row_obs[2 * (size_t)r], row_obs[2 * (size_t)r + 1], o);
      const T wgt = huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
      T Jp[6];
      jac_point(o, c.R, wgt, sp, Jp);
      T m[2];
      jac_cam_dir<T, L>(o, dtab + (size_t)dir_stride<L>() * cam, X, wgt, m);
      m[0] += Jp[0] * dp[0] + Jp[1] * dp[1] + Jp[2] * dp[2];
      m[1] += Jp[3] * dp[0] + Jp[4] * dp[1] + Jp[5] * dp[2];
      const double m0 = m[0], m1 = m[1];
      acc[0] += m0 * ((double)(wgt * o.rx) + 0.5 * m0) + m1 * ((double)(wgt * o.ry) + 0.5 * m1);
    }
  }
  if (win.acc) {
    win_atomic_add(win.acc, 1, 0, qp.valid ? win_of_e(win, qp.e) : -1, acc[0]);
    return;
  }
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

// x_out = x + step .* scale (FP64), with partial sums of |x - x_out|^2.
__global__ __launch_bounds__(kBlock) void k_plus(int64_t n, const double* __restrict__ x,
                                                 const double* __restrict__ step,
                                                 const double* __restrict__ scale,
                                                 double* __restrict__ out,
                                                 double* __restrict__ partial, Win win) {
  __shared__ double lds[kBlock / 64];
  double acc[1] = {0.0};
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const bool on = i < n && !win_idle_e(win, (int)(i / 3));
  if (on) {
    const double d = step[i] * scale[i];
    const double v = x[i] + d;
    out[i] = v;
    const double df = x[i] - v;
    acc[0] = df * df;
  }
  if (win.acc) {
    win_atomic_add(win.acc, 1, 0, on ? win_of_e(win, (int)(i / 3)) : -1, acc[0]);
    return;
  }
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

// Sum of squares of [points (3 per E block) | cameras (9 per F block)].
__global__ __launch_bounds__(kBlock) void k_sumsq(int64_t n, int64_t npts3, const double* __restrict__ x,
                                                  double* __restrict__ partial, Win win) {
  __shared__ double lds[kBlock / 64];
  double acc[1] = {0.0};
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i < n) acc[0] = x[i] * x[i];
  if (win.acc) {
    int wi = -1;
    if (i < n) {
      const bool pt = i < npts3;
      const int blk = pt ? (int)(i / 3) : (int)((i - npts3) / 9);
      if (!(pt ? win_idle_e(win, blk) : win_idle_f(win, blk))) wi = pt ? win_of_e(win, blk) : win_of_f(win, blk);
    }
    win_atomic_add(win.acc, 1, 0, wi, acc[0]);
    return;
  }
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

template <typename Ts, typename Td>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const Ts* __restrict__ src,
                                                    Td* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    dst[i] = (Td)src[i];
}

// Layout set-up: the E block of every point-major row.
__global__ __launch_bounds__(kBlock) void k_row_e(int ne, const int32_t* __restrict__ pt_start,
                                                  int32_t* __restrict__ row_e) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  for (int r = pt_start[e]; r < pt_start[e + 1]; ++r) row_e[r] = e;
}

// Layout set-up: camera-major rows gathered from the point-major rows.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_cm(int ncm, const int32_t* __restrict__ cm_src,
                                                      const int32_t* __restrict__ row_e,
                                                      const double* __restrict__ row_obs64,
                                                      int32_t* __restrict__ cm_e, T* __restrict__ cm_obs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncm) return;
  const int r = cm_src[k];
  cm_e[k] = row_e[r];
  cm_obs[2 * (size_t)k] = (T)row_obs64[2 * (size_t)r];
  cm_obs[2 * (size_t)k + 1] = (T)row_obs64[2 * (size_t)r + 1];
}

// Scatter FP64 values at `idx` positions (gather = 0) or gather.
// (xmask, if given: 0 for slots outside the parameter block -> 0).
__global__ __launch_bounds__(kBlock) void k_gather_cams(int nf, const int32_t* __restrict__ f_to_cam,
                                                        const double* __restrict__ cams,
                                                        double* __restrict__ out,
                                                        const uint8_t* __restrict__ xmask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * nf) return;
  out[i] = (xmask && !xmask[i]) ? 0.0 : cams[9 * (size_t)f_to_cam[i / 9] + i % 9];
}
__global__ __launch_bounds__(kBlock) void k_scatter_cams(int nf, const int32_t* __restrict__ f_to_cam,
                                                         const double* __restrict__ src,
                                                         double* __restrict__ cams) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * nf) return;
  cams[9 * (size_t)f_to_cam[i / 9] + i % 9] = src[i];
}

// -------------------------------------------------------------------- PCG
// Conjugate gradients of conjugate_gradients_solver.cc:66-245 with the state
// on the device.  F-space vectors, one thread per 9-entry F block; block
// partial dot products reduced by single-thread scalar kernels.

// PCG iteration, five launches: k_pcg_p_dir (p and its direction table),
// k_mv_point, k_mv_cam, k_mv_finalize_alpha (q = S p, p.q, alpha) and
// k_pcg_update_scalar (x, r, z = Minv r, the reductions, termination, rho,
// beta).  The scalar steps run in whichever workgroup finishes last (a
// counter the last one resets); single-problem partial sums are added in
// block order, so results do not depend on scheduling.

// Last-workgroup election once the block's partials are written: one
// device-scope fence per workgroup (thread 0), not one per wave.
__device__ __forceinline__ bool last_block(unsigned int* counter) {
  __shared__ int is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    is_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
    if (is_last) __threadfence();
  }
  __syncthreads();
  return is_last != 0;
}

// One window's alpha = rho / p.q with the reference's failure tests
// (conjugate_gradients_solver.cc:155-170).
__device__ __forceinline__ void pcg_alpha_state(PcgState& s, double pq) {
  s.pq = pq;
  if (pq <= 0.0 || isinf(pq)) {  // indefinite: NO_CONVERGENCE, keep x
    s.status = 1;
    s.done = 1;
    return;
  }
  s.alpha = s.rho / pq;
  if (isinf(s.alpha)) {
    s.status = -1;
    s.done = 1;
  }
}

// One window's termination tests and next rho / beta from r.z, x.(b + r)
// and r.r (conjugate_gradients_solver.cc:102-153, 201-240; `first`: x = 0,
// r = b).  Returns true when the window stops.
__device__ __forceinline__ bool pcg_step_state(PcgState& s, double rz, double xbr, double rr, bool first, double eta,
                                               int min_iters, int max_iters) {
  s.norm_r = sqrt(rr);
  bool stop = false;
  if (first) {
    s.norm_b = sqrt(rr);  // r = b at x = 0
    s.tol_r = s.tol_r * s.norm_b;
    if (s.norm_b == 0.0) {
      stop = true;
    } else if (min_iters == 0 && s.norm_r <= s.tol_r) {
      stop = true;
    } else {
      s.Q0 = -xbr;
      s.rho = 1.0;
    }
  } else {
    const double Q1 = -xbr;
    const double zeta = s.iter * (Q1 - s.Q0) / Q1;
    s.zeta = zeta;
    if (zeta < eta && s.iter >= min_iters) {
      stop = true;
    } else {
      s.Q0 = Q1;
      if (s.norm_r <= s.tol_r && s.iter >= min_iters) {
        stop = true;
      } else if (s.iter >= max_iters) {
        stop = true;
        s.status = 1;
      } else {
        s.iter += 1;
      }
    }
  }
  if (!stop) {
    // Next iteration's rho = r.z and beta.
    const double last = s.rho;
    s.rho_last = last;
    s.rho = rz;
    if (rz == 0.0 || isinf(rz)) {
      stop = true;
      s.status = -1;
    } else if (s.iter > 1) {
      s.beta = rz / last;
      if (s.beta == 0.0 || isinf(s.beta)) {
        stop = true;
        s.status = -1;
      }
    }
  }
  if (stop) s.done = 1;
  return stop;
}

// alpha = rho / p.q per window, with the reference's failure tests.  The
// per-window accumulators (windowed runs) are cleared as they are read.
__device__ __forceinline__ void pcg_alpha(const double* __restrict__ partial, int nb, int nw,
                                          PcgState* __restrict__ st, double* __restrict__ wacc) {
  if (!wacc && nw == 1) {
    // Single problem: the block partials summed by the first wave (strided
    // lanes, then the wave sum), not by one thread in turn.
    if (threadIdx.x < 64) {
      double v = 0.0;
      for (int i = threadIdx.x; i < nb; i += 64) v += partial[i];
      v = wave_sum(v);
      if (threadIdx.x == 0 && !st[0].done) pcg_alpha_state(st[0], v);
    }
    return;
  }
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    PcgState& s = st[w];
    if (s.done) continue;
    double pq = 0.0;
    if (wacc) {
      pq = wacc[w];
      wacc[w] = 0.0;  // ready for the next iteration's dot
    } else {
      for (int i = 0; i < nb; ++i) pq += partial[i];
    }
    pcg_alpha_state(s, pq);
  }
}

// Per window: termination tests and the next rho / beta; then the
// device-wide all-done flag.  `first`: the initial call (x = 0, r = b).
__device__ __forceinline__ void pcg_scalar(const double* __restrict__ partial, int nb, int nw,
                                           PcgState* __restrict__ st, double* __restrict__ wacc, double eta,
                                           int min_iters, int max_iters, int first, int32_t* __restrict__ all_done,
                                           int32_t* __restrict__ host_done) {
  __shared__ int any_active;
  if (threadIdx.x == 0) any_active = 0;
  __syncthreads();
  if (!wacc && nw == 1) {
    // Single problem: the three sums by the first wave, as in pcg_alpha.
    if (threadIdx.x < 64) {
      double v[3] = {0.0, 0.0, 0.0};
      for (int i = threadIdx.x; i < nb; i += 64)
        for (int k = 0; k < 3; ++k) v[k] += partial[3 * i + k];
      for (int k = 0; k < 3; ++k) v[k] = wave_sum(v[k]);
      if (threadIdx.x == 0 && !st[0].done &&
          !pcg_step_state(st[0], v[0], v[1], v[2], first != 0, eta, min_iters, max_iters))
        any_active = 1;
    }
    nw = 0;  // the loop below has nothing left to do
  }
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    PcgState& s = st[w];
    if (s.done) continue;
    double rz = 0.0, xbr = 0.0, rr = 0.0;
    if (wacc) {
      rz = wacc[3 * (size_t)w];
      xbr = wacc[3 * (size_t)w + 1];
      rr = wacc[3 * (size_t)w + 2];
      for (int k = 0; k < 3; ++k) wacc[3 * (size_t)w + k] = 0.0;
    } else {
      for (int i = 0; i < nb; ++i) {
        rz += partial[3 * i];
        xbr += partial[3 * i + 1];
        rr += partial[3 * i + 2];
      }
    }
    if (!pcg_step_state(s, rz, xbr, rr, first != 0, eta, min_iters, max_iters)) any_active = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *all_done = any_active ? 0 : 1;
    // Host-visible copy (pinned host memory) the host polls between launches.
    if (!any_active && host_done) *host_done = 1;
  }
}

// --------------------------------------------- persistent window PCG
// Batched windows (packed layout, at most kPcgWinMaxCams cameras each): one
// workgroup runs one window's whole PCG -- the iterations of
// pcg_async (direction, point-major and camera-major implicit S p, alpha,
// update, preconditioner, Ceres's termination tests, residual reset every
// 10 iterations) -- in one launch for all windows.  The window's camera
// projection tables, direction tables and camera vectors live in LDS; point
// records go through global memory (written and read by this workgroup
// only, ordered by barriers).  Sums are deterministic: the point-major sum
// of a point by one thread, a camera's camera-major sum by one wave.
// Workgroup size: 16 waves in FP32, 8 in FP64 (whose registers allow no
// more); a window's rows are latency-bound, so more waves per window run
// its matvecs faster (measured on MI355X: 256 -> 1024 threads cut the
// batched linear solve of 512 windows from 19.0 to 12.0 ms).
constexpr int kPcgWinMaxCams = 64;
template <typename T>
constexpr int pcg_win_threads() {
  return sizeof(T) == 4 ? 1024 : 512;
}

template <typename T>
struct PcgWinArgs {
  const int32_t* order;     // workgroup -> window (largest windows first)
  const int32_t* win_e_lo;  // W + 1: window w's device E blocks are
  const int32_t* win_e;     //   win_e[win_e_lo[w] .. win_e_lo[w + 1])
  const int32_t* f_off;     // W + 1: window F blocks
  const int32_t* cam_off;   // W + 1: window cameras (global ids)
  const int32_t* f_to_cam;
  const int32_t* cam_to_f;  // -1 for constant cameras
  const int32_t* fcm_lo;    // per F block: its camera-major rows
  const int32_t* fcm_hi;
  const int32_t* pt_start;
  const int32_t* row_cam;
  const T* row_obs;
  const int32_t* cm_e;
  const T* cm_obs;
  T* pk;
  const T* pscale;
  const T* cscale;
  const T* ctab;
  const double* Vinv64;
  const T* wrow;
  const T* wcm;
  const T* U;
  const T* Minv;
  const double* diag_c;
  const double* inv_r;
  const T* b;
  T* x;
  PcgState* st;
  const uint8_t* active;
  T huber_a;
  double eta, tol_r;
  int min_iters, max_iters;
};

template <typename T>
struct PcgWinShared {
  T tab[kPcgWinMaxCams][16];  // R 9 | t 3 | f k1 k2 | pad
  T dir[kPcgWinMaxCams][16];  // direction table of the multiplied vector
  T x[kPcgWinMaxCams * 9], r[kPcgWinMaxCams * 9], z[kPcgWinMaxCams * 9], p[kPcgWinMaxCams * 9];
  T q[kPcgWinMaxCams * 9], bv[kPcgWinMaxCams * 9], part[kPcgWinMaxCams * 9];
  double red[3 * (pcg_win_threads<T>() / 64)];
  PcgState st;
};

// Block-wide sums of three doubles (all threads call); the result is
// valid in every thread after the call.
template <typename T>
__device__ __forceinline__ void pcg_win_sum3(PcgWinShared<T>& sh, double v[3]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double t = wave_sum(v[j]);
    if (lane == 0) sh.red[3 * wave + j] = t;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double t = 0.0;
    for (int k = 0; k < pcg_win_threads<T>() / 64; ++k) t += sh.red[3 * k + j];
    v[j] = t;
  }
  __syncthreads();
}

// sh.dir = direction table of the window vector xs (9 per F block), in the
// layout of k_cam_dir: M = sum_k (s .* x)_k dR_k, then the t, f, k1, k2
// entries; constant cameras get zeros.
template <typename T>
__device__ __forceinline__ void pcg_win_dir(const PcgWinArgs<T>& a, PcgWinShared<T>& sh, const T* xs, int cam0,
                                            int ncam, int f0) {
  for (int i = threadIdx.x; i < ncam * 16; i += pcg_win_threads<T>()) {
    const int lc = i >> 4, k = i & 15;
    const int f = a.cam_to_f[cam0 + lc];
    T v = T(0);
    if (f >= 0 && k < 15) {
      const int lf = f - f0;
      const T* cs = a.cscale + 9 * (size_t)f;
      if (k < 9) {
        const T* dR = a.ctab + (size_t)kCamTab * (cam0 + lc) + kTabDR;
        v = cs[0] * xs[9 * lf] * dR[k] + cs[1] * xs[9 * lf + 1] * dR[9 + k] + cs[2] * xs[9 * lf + 2] * dR[18 + k];
      } else {
        v = cs[k - 6] * xs[9 * lf + k - 6];
      }
    }
    sh.dir[lc][k] = v;
  }
  __syncthreads();
}

// out = S xs for the window (implicit: (U + D) xs - F'E V^-1 E'F xs), with
// sh.dir the direction table of xs.
template <typename T>
__device__ void pcg_win_matvec(const PcgWinArgs<T>& a, PcgWinShared<T>& sh, const T* xs, T* out, int w, int cam0,
                               int f0, int nf, double inv_radius) {
  // Point-major: z_e = V^-1 E'(F xs) into the point records.
  const int e_lo = a.win_e_lo[w], e_hi = a.win_e_lo[w + 1];
  for (int i = e_lo + (int)threadIdx.x; i < e_hi; i += pcg_win_threads<T>()) {
    const int e = a.win_e[i];
    T X[3], unused[3];
    load_pk(a.pk, e, X, unused);
    const T sp[3] = {a.pscale[3 * (size_t)e], a.pscale[3 * (size_t)e + 1], a.pscale[3 * (size_t)e + 2]};
    T q[3] = {T(0), T(0), T(0)};
    for (int r = a.pt_start[e]; r < a.pt_start[e + 1]; ++r) {
      const int lc = a.row_cam[r] - cam0;
      if (a.cam_to_f[cam0 + lc] < 0) continue;  // constant camera: no direction
      CamProj<T> c;
      load_proj(&sh.tab[lc][0], c);
      Proj<T> o;
      const bool need_obs = !a.wrow && a.huber_a > T(0);
      project(c, X, need_obs ? a.row_obs[2 * (size_t)r] : T(0), need_obs ? a.row_obs[2 * (size_t)r + 1] : T(0), o);
      const T wgt = a.wrow ? a.wrow[r] : huber_weight(o.rx * o.rx + o.ry * o.ry, a.huber_a);
      T Jp[6];
      jac_point(o, c.R, wgt, sp, Jp);
      T v[2];
      jac_cam_dir<T, 0>(o, &sh.dir[lc][0], X, wgt, v);
      fma2(q[0], Jp[0], v[0], Jp[3], v[1]);
      fma2(q[1], Jp[1], v[0], Jp[4], v[1]);
      fma2(q[2], Jp[2], v[0], Jp[5], v[1]);
    }
    const double* iv = a.Vinv64 + 6 * (size_t)e;
    const double q0 = q[0], q1 = q[1], q2 = q[2];
    const T zv[3] = {(T)(iv[0] * q0 + iv[1] * q1 + iv[2] * q2), (T)(iv[1] * q0 + iv[3] * q1 + iv[4] * q2),
                     (T)(iv[2] * q0 + iv[4] * q1 + iv[5] * q2)};
    const T sz[3] = {sp[0] * zv[0], sp[1] * zv[1], sp[2] * zv[2]};
    store_pk_v(a.pk, e, sz);
  }
  __syncthreads();
  // Camera-major: sum over each F block's rows of Jc'(Jp z_e), one wave per
  // F block.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int lf = wave; lf < nf; lf += pcg_win_threads<T>() / 64) {
    const int f = f0 + lf, cam = a.f_to_cam[f];
    CamProj<T> c;
    load_proj(&sh.tab[cam - cam0][0], c);
    const T* tb = a.ctab + (size_t)kCamTab * cam + kTabDR;
    T dR[27], cs[9];
#pragma unroll
    for (int k = 0; k < 27; ++k) dR[k] = tb[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) cs[k] = a.cscale[9 * (size_t)f + k];
    T acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = T(0);
    for (int i = a.fcm_lo[f] + lane; i < a.fcm_hi[f]; i += 64) {
      T X[3], sz[3];
      load_pk(a.pk, a.cm_e[i], X, sz);
      Proj<T> o;
      const bool need_obs = !a.wcm && a.huber_a > T(0);
      project(c, X, need_obs ? a.cm_obs[2 * (size_t)i] : T(0), need_obs ? a.cm_obs[2 * (size_t)i + 1] : T(0), o);
      const T wgt = a.wcm ? a.wcm[i] : huber_weight(o.rx * o.rx + o.ry * o.ry, a.huber_a);
      T Jc[18];
      jac_cam(o, dR, X, wgt, cs, Jc);
      T u[2];
      jac_point_apply(o, c.R, sz, u);
      u[0] *= wgt;
      u[1] *= wgt;
#pragma unroll
      for (int k = 0; k < 9; ++k) fma2(acc[k], Jc[k], u[0], Jc[9 + k], u[1]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const T t = wave_sum(acc[k]);
      if (lane == 0) sh.part[9 * lf + k] = t;
    }
  }
  __syncthreads();
  // (U + D) xs - partial
  for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) {
    const int lf = i / 9, k = i % 9, f = f0 + lf;
    const T* u = a.U + 45 * (size_t)f;
    T v = (T)(a.diag_c[9 * (size_t)f + k] * inv_radius) * xs[9 * lf + k];
#pragma unroll
    for (int j = 0; j < 9; ++j) v += u[k <= j ? ut9(k, j) : ut9(j, k)] * xs[9 * lf + j];
    out[i] = v - sh.part[i];
  }
  __syncthreads();
}

// z = Minv r; dots r.z, x.(b + r), r.r; then the state's termination tests.
template <typename T>
__device__ __forceinline__ void pcg_win_precond(const PcgWinArgs<T>& a, PcgWinShared<T>& sh, int f0, int nf,
                                                bool first) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) {
    const int lf = i / 9, k = i % 9;
    const T* M = a.Minv + 81 * (size_t)(f0 + lf) + 9 * k;
    T v = T(0);
#pragma unroll
    for (int j = 0; j < 9; ++j) v += M[j] * sh.r[9 * lf + j];
    sh.z[i] = v;
    acc[0] += (double)sh.r[i] * (double)v;
    acc[1] += (double)sh.x[i] * ((double)sh.bv[i] + (double)sh.r[i]);
    acc[2] += (double)sh.r[i] * (double)sh.r[i];
  }
  pcg_win_sum3(sh, acc);
  if (threadIdx.x == 0) pcg_step_state(sh.st, acc[0], acc[1], acc[2], first, a.eta, a.min_iters, a.max_iters);
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(pcg_win_threads<T>()) void k_pcg_window(PcgWinArgs<T> a) {
  __shared__ PcgWinShared<T> sh;
  const int w = a.order[blockIdx.x];
  const int f0 = a.f_off[w], nf = a.f_off[w + 1] - f0;
  const int cam0 = a.cam_off[w], ncam = a.cam_off[w + 1] - cam0;
  if (threadIdx.x == 0) {
    PcgState z;
    z.rho = z.rho_last = z.Q0 = z.alpha = z.beta = z.pq = z.norm_b = z.zeta = z.norm_r = 0.0;
    z.tol_r = a.tol_r;
    z.iter = 1;
    z.done = (a.active && !a.active[w]) ? 1 : 0;
    z.status = 0;
    z.pad = 0;
    sh.st = z;
  }
  __syncthreads();
  if (sh.st.done || nf == 0) {
    if (threadIdx.x == 0) a.st[w] = sh.st;
    return;
  }
  const double inv_radius = a.inv_r[w];
  for (int i = threadIdx.x; i < ncam * 16; i += pcg_win_threads<T>()) {
    const int lc = i >> 4, k = i & 15;
    sh.tab[lc][k] = k < 15 ? a.ctab[(size_t)kCamTab * (cam0 + lc) + k] : T(0);
  }
  for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) {
    sh.bv[i] = a.b[9 * (size_t)f0 + i];
    sh.x[i] = T(0);
    sh.r[i] = sh.bv[i];
    sh.p[i] = T(0);
  }
  __syncthreads();
  pcg_win_precond(a, sh, f0, nf, true);
  while (!sh.st.done) {
    const bool first = sh.st.iter == 1;
    const T beta = (T)sh.st.beta;
    for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) sh.p[i] = first ? sh.z[i] : sh.z[i] + beta * sh.p[i];
    __syncthreads();
    pcg_win_dir(a, sh, sh.p, cam0, ncam, f0);
    pcg_win_matvec(a, sh, sh.p, sh.q, w, cam0, f0, nf, inv_radius);
    double pq[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) pq[0] += (double)sh.p[i] * (double)sh.q[i];
    pcg_win_sum3(sh, pq);
    if (threadIdx.x == 0) pcg_alpha_state(sh.st, pq[0]);
    __syncthreads();
    if (sh.st.done) break;
    const T alpha = (T)sh.st.alpha;
    const bool reset = sh.st.iter % 10 == 0;  // residual reset: r = b - S x
    for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) {
      sh.x[i] = sh.x[i] + alpha * sh.p[i];
      if (!reset) sh.r[i] = sh.r[i] - alpha * sh.q[i];
    }
    __syncthreads();
    if (reset) {
      pcg_win_dir(a, sh, sh.x, cam0, ncam, f0);
      pcg_win_matvec(a, sh, sh.x, sh.q, w, cam0, f0, nf, inv_radius);
      for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) sh.r[i] = sh.bv[i] - sh.q[i];
      __syncthreads();
    }
    pcg_win_precond(a, sh, f0, nf, false);
  }
  for (int i = threadIdx.x; i < 9 * nf; i += pcg_win_threads<T>()) a.x[9 * (size_t)f0 + i] = sh.x[i];
  if (threadIdx.x == 0) a.st[w] = sh.st;
}

// PCG start: every window's state cleared, windows not iterating
// (active[w] == 0) marked stopped; all_done cleared.
__global__ __launch_bounds__(kBlock) void k_pcg_init(int nw, PcgState* __restrict__ st,
                                                     const uint8_t* __restrict__ active, double tol_r,
                                                     int32_t* __restrict__ all_done,
                                                     unsigned int* __restrict__ progress) {
  if (threadIdx.x == 0 && progress) *progress = 0u;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    PcgState z;
    z.rho = z.rho_last = z.Q0 = z.alpha = z.beta = z.pq = z.norm_b = z.zeta = z.norm_r = 0.0;
    z.tol_r = tol_r;
    z.iter = 1;
    z.done = (active && !active[w]) ? 1 : 0;
    z.status = 0;
    z.pad = 0;
    st[w] = z;
  }
  if (threadIdx.x == 0) *all_done = 0;
}

// The small PCG kernels give each F block (camera) a group of 16 lanes,
// lane k < 9 owning component k; a group's values are exchanged with
// shuffles inside the wave.
constexpr int kGrp = 16;

// p = z + beta p (p = z on iteration 1) for the camera's F block and its
// direction-table row.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_p_dir(int nc, const int32_t* __restrict__ cam_to_f,
                                                      const PcgState* __restrict__ st, Win win,
                                                      const T* __restrict__ z, T* __restrict__ p,
                                                      const T* __restrict__ cscale, const T* __restrict__ ctab,
                                                      T* __restrict__ dtab, const int32_t* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int c = (int)(t / kGrp), k = (int)(t % kGrp);
  const int base = (threadIdx.x & 63) & ~(kGrp - 1);
  const int f = c < nc ? cam_to_f[c] : -1;
  bool on = f >= 0;
  T beta = T(0);
  int first = 0;
  if (on) {
    const PcgState& s = st[win_of_f(win, f)];
    on = !s.done;
    beta = (T)s.beta;
    first = s.iter == 1;
  }
  T xs = T(0);
  if (on && k < 9) {
    const size_t i = 9 * (size_t)f + k;
    const T pv = first ? z[i] : z[i] + beta * p[i];
    p[i] = pv;
    xs = cscale[i] * pv;
  }
  T x9[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) x9[j] = __shfl(xs, base + j, 64);
  if (c >= nc) return;
  T* o = dtab + (size_t)kDirTab * c;
  if (f < 0) {  // constant camera: no direction
    o[k] = T(0);
    return;
  }
  if (!on) return;
  T v = T(0);
  if (k < 9) {
    const T* dR = ctab + (size_t)kCamTab * c + kTabDR;
    v = x9[0] * dR[k] + x9[1] * dR[9 + k] + x9[2] * dR[18 + k];
  } else if (k < 15) {
    v = x9[k - 6];
  }
  o[k] = v;
}

// q_f = (U + D^2) p_f - sum partials (L = 0), or D^2 p_f + sum partials
// (L = 1: the partials hold F'(F p - E z) in full), p.q, then alpha (last
// workgroup).
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_mv_finalize_alpha(
    int nf, const int32_t* __restrict__ f_lo, const int32_t* __restrict__ f_hi, const T* __restrict__ partial,
    const T* __restrict__ U, const double* __restrict__ diag_c, Win win, const T* __restrict__ x,
    T* __restrict__ y, const int32_t* __restrict__ all_done, PcgState* __restrict__ st, int nw,
    double* __restrict__ bpart, unsigned int* __restrict__ counter) {
  __shared__ double lds[kBlock / 64];
  if (*all_done) return;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int f = (int)(t / kGrp), i = (int)(t % kGrp);
  const int base = (threadIdx.x & 63) & ~(kGrp - 1);
  const bool blk = f < nf && !win_idle_f(win, f);
  const T xi = (blk && i < 9) ? x[9 * (size_t)f + i] : T(0);
  T xv[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) xv[j] = __shfl(xi, base + j, 64);
  double acc[1] = {0.0};
  int wi = -1;
  if (blk && i < 9) {
    wi = win_of_f(win, f);
    const T sp = sum_partials<T>(partial, f_lo[f], f_hi[f], i);
    T a = (T)(diag_c[9 * (size_t)f + i] * win_inv_r(win, wi)) * xv[i];
    T yv;
    if constexpr (L == 0) {
      const T* u = U + 45 * (size_t)f;
#pragma unroll
      for (int j = 0; j < 9; ++j) a += u[i <= j ? ut9(i, j) : ut9(j, i)] * xv[j];
      yv = a - sp;
    } else {
      yv = a + sp;
    }
    y[9 * (size_t)f + i] = yv;
    acc[0] = (double)xi * (double)yv;
  }
  if (win.acc) {
    win_atomic_add(win.acc, 1, 0, wi, acc[0]);
  } else {
    block_sum<double, 1>(acc, lds);
    if (threadIdx.x == 0) {
      double v = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) v += lds[w];
      bpart[blockIdx.x] = v;
    }
  }
  if (!last_block(counter)) return;
  pcg_alpha(bpart, gridDim.x, nw, st, win.acc);
  if (threadIdx.x == 0) *counter = 0u;
}

// PCG vector update.  mode 0: x += alpha p, r -= alpha q;  1: x += alpha p
// only;  2: r = b - tmp;  3: r = b (x = 0).  Modes 0, 2, 3 then apply
// z = Minv r, reduce r.z, x.(b + r), r.r, and the last workgroup runs the
// scalar step (`first` with mode 3).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update_scalar(
    int nf, PcgState* __restrict__ st, Win win, int mode, const T* __restrict__ p, const T* __restrict__ q,
    const T* __restrict__ b, const T* __restrict__ tmp, const T* __restrict__ Minv, T* __restrict__ x,
    T* __restrict__ r, T* __restrict__ z, double* __restrict__ bpart, int32_t* __restrict__ all_done, int nw,
    double eta, int min_iters, int max_iters, unsigned int* __restrict__ counter, int32_t* __restrict__ host_done,
    unsigned int* __restrict__ progress) {
  __shared__ double lds[3 * (kBlock / 64)];
  if (*all_done) return;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int f = (int)(t / kGrp), k = (int)(t % kGrp);
  const int base = (threadIdx.x & 63) & ~(kGrp - 1);
  int wi = -1;
  T alpha = T(0);
  if (f < nf) {
    const int w = win_of_f(win, f);
    const PcgState& s = st[w];
    if (!s.done) {
      wi = w;
      alpha = (T)s.alpha;
    }
  }
  const bool on = wi >= 0 && k < 9;
  const size_t idx = 9 * (size_t)f + k;
  T xv = T(0), rv = T(0);
  if (on) {
    if (mode == 3) {
      xv = T(0);
      x[idx] = xv;
    } else if (mode <= 1) {
      xv = x[idx] + alpha * p[idx];
      x[idx] = xv;
    } else {
      xv = x[idx];
    }
    if (mode != 1) {
      rv = (mode == 0) ? r[idx] - alpha * q[idx] : (mode == 2 ? b[idx] - tmp[idx] : b[idx]);
      r[idx] = rv;
    }
  }
  if (mode == 1) return;
  T r9[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) r9[j] = __shfl(rv, base + j, 64);
  double acc[3] = {0.0, 0.0, 0.0};
  if (on) {
    const T* M = Minv + 81 * (size_t)f + 9 * k;
    T a = T(0);
#pragma unroll
    for (int j = 0; j < 9; ++j) a += M[j] * r9[j];
    z[idx] = a;
    acc[0] = (double)rv * (double)a;
    acc[1] = (double)xv * ((double)b[idx] + (double)rv);
    acc[2] = (double)rv * (double)rv;
  }
  const int wl = on ? wi : -1;
  if (win.acc) {
    for (int j = 0; j < 3; ++j) win_atomic_add(win.acc, 3, j, wl, acc[j]);
  } else {
    block_sum<double, 3>(acc, lds);
    if (threadIdx.x == 0)
      for (int j = 0; j < 3; ++j) {
        double v = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) v += lds[w * 3 + j];
        bpart[3 * (size_t)blockIdx.x + j] = v;
      }
  }
  if (!last_block(counter)) return;
  pcg_scalar(bpart, gridDim.x, nw, st, win.acc, eta, min_iters, max_iters, mode == 3, all_done, host_done);
  if (threadIdx.x == 0) {
    *counter = 0u;
    // Completed updates, for the host's look-ahead (host_done[1], pinned).
    if (progress && host_done) host_done[1] = (int32_t)(++*progress);
  }
}

// Inverse of the camera-major order: slot_of[src[k]] = k.
__global__ __launch_bounds__(kBlock) void k_slot_of_row(int n, const int32_t* __restrict__ src,
                                                        int32_t* __restrict__ slot_of) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) slot_of[src[k]] = k;
}

// A single problem's FP64 observations in row order, from the caller's
// order uploaded as is (src: each row's observation index).
__global__ __launch_bounds__(kBlock) void k_gather_obs(int64_t n, const int32_t* __restrict__ src,
                                                       const double2* __restrict__ in, double2* __restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r < n) out[r] = in[src[r]];
}

// Per-window flagged copy of trial parameters into the current ones.
__global__ __launch_bounds__(kBlock) void k_accept_points(int ne, Win win, const uint8_t* __restrict__ accept,
                                                          const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 3 * (int64_t)ne) return;
  if (accept[win_of_e(win, (int)(i / 3))]) dst[i] = src[i];
}
__global__ __launch_bounds__(kBlock) void k_accept_cams(int nf, Win win, const int32_t* __restrict__ f_to_cam,
                                                        const uint8_t* __restrict__ accept,
                                                        const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 9 * (int64_t)nf) return;
  const int f = (int)(i / 9);
  if (!accept[win_of_f(win, f)]) return;
  const size_t j = 9 * (size_t)f_to_cam[f] + i % 9;
  dst[j] = src[j];
}

// ----------------------------------------------------------- explicit S
// Dense S (column-major, n = 9 Nf, both triangles):
//   S(fa, fb) = [fa == fb] (U + D^2) - sum over observation rows a, b of a
//               common point with cameras fa >= fb of  W_a Vinv W_b',
// W_r = Jc_r' Jp_r (9 x 3, Jacobi-scaled columns).  In the grouped layout
// the rows are virtual rows (one per observation and F block it touches).

// Per point-major row: W_r and Y_r = W_r Vinv_e, kWY values per row
// (W at 0, Y at kWYY, row-major 9 x 3, padded for 16-byte loads).
constexpr int kWY = 56, kWYY = 28;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_s_rows(
    Seg sg, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const int32_t* __restrict__ cam_to_f, const T* __restrict__ row_obs, const T* __restrict__ pkX,
    const T* __restrict__ pscale, const T* __restrict__ cscale, const T* __restrict__ ctab,
    const T* __restrict__ Vinv, T huber_a, T* __restrict__ WY, Win win) {
  const SegPos q = seg_locate(sg);
  if (!q.valid || win_idle_e(win, q.e)) return;
  const int e = q.e;
  T X[3], unused[3];
  load_pk(pkX, e, X, unused);
  const T sp[3] = {pscale[3 * (size_t)e], pscale[3 * (size_t)e + 1], pscale[3 * (size_t)e + 2]};
  const T* iv = Vinv + 6 * (size_t)e;
  const T Vi[9] = {iv[0], iv[1], iv[2], iv[1], iv[3], iv[4], iv[2], iv[4], iv[5]};
  for (int r = pt_start[e] + q.sub; r < pt_start[e + 1]; r += q.lanes) {
    const int cam = row_cam[r];
    const int f = cam_to_f[cam];
    if (f < 0) continue;  // constant camera: no column in S
    const T* tab = ctab + (size_t)kCamTab * cam;
    CamProj<T> c;
    load_proj(tab, c);
    Proj<T> o;
    project(c, X, row_obs[2 * (size_t)r], row_obs[2 * (size_t)r + 1], o);
    const T wgt = huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T Jp[6], Jc[18], cs[9];
    for (int k = 0; k < 9; ++k) cs[k] = cscale[9 * (size_t)f + k];
    jac_point(o, c.R, wgt, sp, Jp);
    jac_cam(o, tab + kTabDR, X, wgt, cs, Jc);
    T* out = WY + (size_t)kWY * r;
    // Both halves (27 values and a zero pad) as 16-byte stores.
    T w[kWYY], y[kWYY];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) w[3 * i + j] = Jc[i] * Jp[j] + Jc[9 + i] * Jp[3 + j];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        y[3 * i + j] = w[3 * i] * Vi[j] + w[3 * i + 1] * Vi[3 + j] + w[3 * i + 2] * Vi[6 + j];
    w[27] = y[27] = T(0);
    using V16 = typename std::conditional<std::is_same<T, float>::value, float4, double2>::type;
    constexpr int nv = kWYY * (int)sizeof(T) / 16;
#pragma unroll
    for (int v = 0; v < nv; ++v) {
      V16 a, b;
      __builtin_memcpy(&a, &w[v * 16 / sizeof(T)], 16);
      __builtin_memcpy(&b, &y[v * 16 / sizeof(T)], 16);
      reinterpret_cast<V16*>(out)[v] = a;
      reinterpret_cast<V16*>(out + kWYY)[v] = b;
    }
  }
}

// Entry (i, j) of block (fa, fb) of S from the accumulated sum of its row
// pairs: written with its mirror above the diagonal, so S is stored whole.
// Diagonal blocks add U + D^2.  Grouped layout: a camera / group block pair
// also carries the F'F coupling of their common rows, cross[36 c + 6 a + b]
// (camera slot a, group slot b); blk_cross = 2 c (+1 if the group block is
// fa), or -1.
// Where the blocks of S go: one dense n x n column-major matrix (woff NULL),
// or one per window (batched explicit solve): window w's of order
// 9 (f_off[w+1] - f_off[w]) at S + woff[w], its F blocks numbered from
// f_off[w]; of_f gives an F block's window.
struct SDest {
  int64_t n;
  const int64_t* woff;
  const int32_t* f_off;
  const int32_t* of_f;
};

template <typename T>
__device__ __forceinline__ void s_put(int blk, int fa, int fb, int i, int j, T acc, const T* __restrict__ U,
                                      const double* __restrict__ diag_c, double inv_radius, const SDest& sd,
                                      T* __restrict__ S, const int32_t* __restrict__ blk_cross,
                                      const T* __restrict__ cross) {
  T base = T(0);
  if (fa == fb) {
    base = U[45 * (size_t)fa + (i <= j ? ut9(i, j) : ut9(j, i))];
    if (i == j) base += (T)(diag_c[9 * (size_t)fa + i] * inv_radius);
  } else if (blk_cross && i < 6 && j < 6) {
    const int xc = blk_cross[blk];
    if (xc >= 0) {
      const T* x = cross + 36 * (size_t)(xc >> 1);
      base = (xc & 1) ? x[6 * j + i] : x[6 * i + j];
    }
  }
  int64_t off = 0, ld = sd.n, la = fa, lb = fb;
  if (sd.woff) {
    const int w = sd.of_f[fa], f0 = sd.f_off[w];
    off = sd.woff[w];
    ld = 9 * (int64_t)(sd.f_off[w + 1] - f0);
    la = fa - f0;
    lb = fb - f0;
  }
  const int64_t row = 9 * la + i, col = 9 * lb + j;
  S[off + row + col * ld] = base - acc;
  if (fa != fb) S[off + col + row * ld] = base - acc;
}

// One 64-lane workgroup per work item: a block (fa, fb), fa >= fb, that
// some point couples, or a slice of at most kSplitPairs of its row pairs.
// The pairs (a, b) are staged 64 at a time in LDS and lane l sums entries l
// and l + 64 of the 9 x 9 product.  A whole block is written at once (no
// atomics, fixed summation order); a slice writes its 81 sums to scratch
// slot item_slot, and k_s_split_final adds a split block's slots in order.
constexpr int64_t kSplitPairs = 4096;
template <typename T>
__global__ __launch_bounds__(64) void k_s_blocks(
    const int32_t* __restrict__ item_blk, const int64_t* __restrict__ item_p0, const int64_t* __restrict__ item_p1,
    const int32_t* __restrict__ item_slot, const int32_t* __restrict__ blk_fa, const int32_t* __restrict__ blk_fb,
    const int32_t* __restrict__ pair_a, const int32_t* __restrict__ pair_b, const T* __restrict__ WY,
    const T* __restrict__ U, const double* __restrict__ diag_c, Win win, SDest sd, T* __restrict__ S,
    const int32_t* __restrict__ blk_cross, const T* __restrict__ cross, T* __restrict__ scratch) {
  __shared__ T Ys[64][kWYY];
  __shared__ T Ws[64][kWYY];
  const int item = blockIdx.x;
  const int lane = threadIdx.x;
  const int blk = item_blk[item];
  const int64_t p0 = item_p0[item], p1 = item_p1[item];
  const int fa = blk_fa[blk], fb = blk_fb[blk];
  if (win_idle_f(win, fa)) return;
  const double inv_radius = win_inv_r(win, win_of_f(win, fa));
  const int i0 = lane / 9, j0 = lane % 9;
  const int e1 = lane + 64, i1 = e1 / 9, j1 = e1 % 9;
  const bool two = e1 < 81;
  T acc0 = T(0), acc1 = T(0);
  if constexpr (std::is_same<T, float>::value) {
    // FP32: lane l takes pairs p0 + l, p0 + l + 64, ... and keeps all 81
    // sums in registers (its pair's two rows as 16-byte loads: both halves
    // of a row are 28 floats, 16-byte aligned); the wave then adds the 64
    // lanes' sums in lane order through LDS.  No LDS traffic per pair.
    float s[81];
#pragma unroll
    for (int e = 0; e < 81; ++e) s[e] = 0.0f;
    for (int64_t q = p0 + lane; q < p1; q += 64) {
      const float4* ya = reinterpret_cast<const float4*>(WY + (size_t)kWY * pair_a[q] + kWYY);
      const float4* wb = reinterpret_cast<const float4*>(WY + (size_t)kWY * pair_b[q]);
      float y[28], w[28];
#pragma unroll
      for (int v = 0; v < 7; ++v) {
        const float4 a4 = ya[v], b4 = wb[v];
        y[4 * v] = a4.x; y[4 * v + 1] = a4.y; y[4 * v + 2] = a4.z; y[4 * v + 3] = a4.w;
        w[4 * v] = b4.x; w[4 * v + 1] = b4.y; w[4 * v + 2] = b4.z; w[4 * v + 3] = b4.w;
      }
#pragma unroll
      for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j)
          s[9 * i + j] += y[3 * i] * w[3 * j] + y[3 * i + 1] * w[3 * j + 1] + y[3 * i + 2] * w[3 * j + 2];
    }
    // Two rounds through 41 x 65 floats of LDS: entries 0..40 (lane e sums
    // entry e), then 41..80 (lanes 41..63 entries 41..63, lanes 0..16
    // entries 64..80 as acc1).
    __shared__ float red[41 * 65];
#pragma unroll
    for (int e = 0; e < 41; ++e) red[e * 65 + lane] = s[e];
    __syncthreads();
    if (lane < 41)
      for (int m = 0; m < 64; ++m) acc0 += red[lane * 65 + m];
    __syncthreads();
#pragma unroll
    for (int e = 41; e < 81; ++e) red[(e - 41) * 65 + lane] = s[e];
    __syncthreads();
    if (lane >= 41)
      for (int m = 0; m < 64; ++m) acc0 += red[(lane - 41) * 65 + m];
    if (two)
      for (int m = 0; m < 64; ++m) acc1 += red[(e1 - 41) * 65 + m];
  } else
  for (int64_t base = p0; base < p1; base += 64) {
    const int np = (int)min((int64_t)64, p1 - base);
    if (lane < np) {
      const T* ya = WY + (size_t)kWY * pair_a[base + lane] + kWYY;
      const T* wb = WY + (size_t)kWY * pair_b[base + lane];
#pragma unroll
      for (int k = 0; k < 27; ++k) {
        Ys[lane][k] = ya[k];
        Ws[lane][k] = wb[k];
      }
    }
    __syncthreads();
    for (int p = 0; p < np; ++p) {
      acc0 += Ys[p][3 * i0] * Ws[p][3 * j0] + Ys[p][3 * i0 + 1] * Ws[p][3 * j0 + 1] +
              Ys[p][3 * i0 + 2] * Ws[p][3 * j0 + 2];
      if (two)
        acc1 += Ys[p][3 * i1] * Ws[p][3 * j1] + Ys[p][3 * i1 + 1] * Ws[p][3 * j1 + 1] +
                Ys[p][3 * i1 + 2] * Ws[p][3 * j1 + 2];
    }
    __syncthreads();
  }
  const int slot = item_slot[item];
  if (slot >= 0) {
    scratch[81 * (size_t)slot + lane] = acc0;
    if (two) scratch[81 * (size_t)slot + e1] = acc1;
    return;
  }
  s_put(blk, fa, fb, i0, j0, acc0, U, diag_c, inv_radius, sd, S, blk_cross, cross);
  if (two) s_put(blk, fa, fb, i1, j1, acc1, U, diag_c, inv_radius, sd, S, blk_cross, cross);
}

// Split blocks: sum slots [split_lo, split_hi) in order, then write.
template <typename T>
__global__ __launch_bounds__(64) void k_s_split_final(
    const int32_t* __restrict__ split_blk, const int32_t* __restrict__ split_lo,
    const int32_t* __restrict__ split_hi, const int32_t* __restrict__ blk_fa, const int32_t* __restrict__ blk_fb,
    const T* __restrict__ U, const double* __restrict__ diag_c, Win win, SDest sd, T* __restrict__ S,
    const int32_t* __restrict__ blk_cross, const T* __restrict__ cross, const T* __restrict__ scratch) {
  const int k = blockIdx.x, lane = threadIdx.x;
  const int blk = split_blk[k];
  const int fa = blk_fa[blk], fb = blk_fb[blk];
  if (win_idle_f(win, fa)) return;
  const double inv_radius = win_inv_r(win, win_of_f(win, fa));
  const int e1 = lane + 64;
  const bool two = e1 < 81;
  T acc0 = T(0), acc1 = T(0);
  for (int sl = split_lo[k]; sl < split_hi[k]; ++sl) {
    acc0 += scratch[81 * (size_t)sl + lane];
    if (two) acc1 += scratch[81 * (size_t)sl + e1];
  }
  s_put(blk, fa, fb, lane / 9, lane % 9, acc0, U, diag_c, inv_radius, sd, S, blk_cross, cross);
  if (two) s_put(blk, fa, fb, e1 / 9, e1 % 9, acc1, U, diag_c, inv_radius, sd, S, blk_cross, cross);
}

// Batched explicit solve: one workgroup per window factors the
// window's S (order n = 9 nf_w <= kWinMaxN, column-major at S + woff[w],
// lower triangle, in place) by right-looking column steps (column j scaled
// and cached in LDS, then the trailing lower triangle updated), then solves
// S x = rhs on the window's slice of the F-ordered vectors by forward and
// backward substitution.  A non-positive pivot flags the window (its step
// is then invalid) and is replaced by 1.
constexpr int kWinMaxF = 64, kWinMaxN = 9 * kWinMaxF, kWinThreads = 512;
template <typename T>
__global__ __launch_bounds__(kWinThreads) void k_win_chol_solve(const int64_t* __restrict__ woff,
                                                        const int32_t* __restrict__ f_off, T* __restrict__ S,
                                                        const T* __restrict__ rhs, T* __restrict__ x, Win win) {
  __shared__ T colj[kWinMaxN];
  __shared__ T v[kWinMaxN];
  __shared__ T piv;
  __shared__ int bad;
  const int w = blockIdx.x, t = threadIdx.x;
  if (win_idle(win, w)) return;
  const int f0 = f_off[w];
  const int n = 9 * (f_off[w + 1] - f0);
  T* A = S + woff[w];
  if (t == 0) bad = 0;
  for (int j = 0; j < n; ++j) {
    __syncthreads();
    if (t == 0) {
      T d = A[j + (int64_t)j * n];
      if (!(d > T(0))) {
        bad = 1;
        d = T(1);
      }
      piv = sqrt(d);
      A[j + (int64_t)j * n] = piv;
    }
    __syncthreads();
    const T sj = piv;
    for (int r = j + 1 + t; r < n; r += kWinThreads) {
      const T l = A[r + (int64_t)j * n] / sj;
      A[r + (int64_t)j * n] = l;
      colj[r] = l;
    }
    __syncthreads();
    for (int c = j + 1 + (t >> 5); c < n; c += kWinThreads / 32) {
      const T lc = colj[c];
      T* Ac = A + (int64_t)c * n;
      for (int r = c + (t & 31); r < n; r += 32) Ac[r] -= colj[r] * lc;
    }
  }
  __syncthreads();
  if (t == 0 && bad) win.flag[w] = 1;
  for (int r = t; r < n; r += kWinThreads) v[r] = rhs[9 * (size_t)f0 + r];
  // L y = rhs
  for (int j = 0; j < n; ++j) {
    __syncthreads();
    if (t == 0) v[j] = v[j] / A[j + (int64_t)j * n];
    __syncthreads();
    const T vj = v[j];
    for (int r = j + 1 + t; r < n; r += kWinThreads) v[r] -= A[r + (int64_t)j * n] * vj;
  }
  // L' x = y
  for (int j = n - 1; j >= 0; --j) {
    __syncthreads();
    if (t == 0) v[j] = v[j] / A[j + (int64_t)j * n];
    __syncthreads();
    const T vj = v[j];
    for (int r = t; r < j; r += kWinThreads) v[r] -= A[j + (int64_t)r * n] * vj;
  }
  __syncthreads();
  for (int r = t; r < n; r += kWinThreads) x[9 * (size_t)f0 + r] = v[r];
}

// ------------------------------------------------- dense Cholesky (lower)
// Right-looking blocked LLT of a column-major SPD matrix, NB = 64, in outer
// steps of two blocks:
//   k_potrf_diag   factors a 64 x 64 diagonal block and inverts it;
//   k_panel_col0   L21 = A21 L11^-T as a GEMM with the inverted block (for
//                  the outer step's first block also A22 -= L21 L21' on the
//                  second block's columns, after which workgroup 0 factors
//                  the second block);
//   k_syrk2_mfma   A22 -= L21a L21a' + L21b L21b' with both blocks' panels
//                  on the trailing tiles (K = 128).
// The triangular solves with the factor run one launch per direction.
constexpr int kNB = 64;


// Scale column J of the diagonal block (this lane's entry aj, row r) by
// its pivot and publish it in LDS (colp[J % 2], stored as [r % 4][r / 4]).
template <typename T, int J>
__device__ __forceinline__ void potrf_pivot(T& aj, T (*colp)[4][16], int r, int& bad) {
  T d = lane_bcast(aj, J);
  if (!(d > T(0))) {
    bad = bad ? bad : J + 1;
    d = T(1);
  }
  // FP32: the hardware reciprocal square root (~1 ulp) instead of an IEEE
  // square root and division on this chain (8.6 vs 9.5 ms per factorisation
  // at n = 9,000); the FP32 factor is refined in FP64.
  T v;
  if constexpr (std::is_same<T, float>::value) {
    const float is = __builtin_amdgcn_rsqf(d);
    v = r == J ? d * is : (r > J ? aj * is : T(0));
  } else {
    const T sj = sqrt(d);
    v = r == J ? sj : (r > J ? aj / sj : T(0));
  }
  aj = v;
  colp[J & 1][r & 3][r >> 2] = v;
}

// Factor step J (column J published before the barrier): the wave owning
// column J + 1 applies step J's update to that column first and publishes
// it, so the next pivot is ready while the other columns are updated.
// Finished columns (c <= J) are left alone.  A template on J, force-inlined
// into the kernel, keeps every register index a compile-time constant.
template <typename T, int J>
__device__ __forceinline__ void potrf_step(T (&a)[16], T (*colp)[4][16], int r, int w, int& bad) {
  constexpr int jn = J + 1, own = jn & 3, qn = (jn >> 2) & 15;
  __syncthreads();
  T(*cb)[16] = colp[J & 1];
  const T lr = cb[r & 3][r >> 2];
  if (jn < kNB && w == own) {
    a[qn] = fma(-lr, cb[w][qn], a[qn]);
    potrf_pivot<T, (jn < kNB ? jn : 0)>(a[qn], colp, r, bad);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    if (4 * q + 3 > J) {
      const T m = (w + 4 * q > J) ? cb[w][q] : T(0);
      if (jn < kNB && q == qn) {
        if (w != own) a[q] = fma(-lr, m, a[q]);
      } else {
        a[q] = fma(-lr, m, a[q]);
      }
    }
  }
}
template <typename T, int J>
struct PotrfSteps {
  static __device__ __forceinline__ void run(T (&a)[16], T (*colp)[4][16], int r, int w, int& bad) {
    potrf_step<T, J>(a, colp, r, w, bad);
    PotrfSteps<T, J + 1>::run(a, colp, r, w, bad);
  }
};
template <typename T>
struct PotrfSteps<T, kNB> {
  static __device__ __forceinline__ void run(T (&)[16], T (*)[4][16], int, int, int&) {}
};

// Four waves: factor the kb x kb diagonal block at (k, k), write it back,
// and write L11^-1 (64 x 64 column-major, zero-padded) and its transpose
// (kept per block for the triangular solves).  Rows and columns past kb are
// padded with the identity.  Thread (wave w, lane r) holds row r
// of columns c = w + 4 q (q < 16) in registers.  Factor step j: the wave
// owning column j takes the pivot from lane j, scales the column and
// publishes it in LDS (two buffers, so one barrier per step; stored as
// colp[c % 4][c / 4] so that a wave's 16 columns are contiguous); every
// wave then applies the rank-1 update to its unfinished columns (columns
// c <= j in a group of four with one c > j get a zero multiplier; entries
// above the diagonal are updated too and cleared when their column is
// scaled).  Steps are pipelined (potrf_step): column j + 1 is updated and
// published before the rest of step j's update.
// Inverse: the factor goes to LDS and each wave forms its 16 columns of
// L11^-1 by column-oriented forward substitution, lane r holding row r and
// X[j][c] broadcast from lane j times 1 / L[j][j]; step j touches only the
// columns c <= j + 3 that can be nonzero there.
template <typename T>
// slab (optional): a 64 x 64 column-major copy of the block's first
// sub-diagonal slab (rows k + kb .. k + kb + 63 of its columns, zero past
// row n), which k_panel_col0's workgroups read as their operand while its
// first workgroup overwrites the slab in A with L21.
// The factor of one diagonal block by a 256-thread workgroup, with its LDS
// arrays passed in (k_panel_col0 lends its tiles).
__device__ __forceinline__ void potrf_block(T* __restrict__ A, int64_t n, int k, int kb, T* __restrict__ Linv,
                                            T* __restrict__ LinvT, int32_t* __restrict__ info,
                                            T* __restrict__ slab, T (*colp)[4][16], T (*Ls)[kNB + 1],
                                            T* rinv, int* bad_w) {
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  // The slab's and the block's loads in one batch (one memory round trip).
  T a[16], sv[16];
  const int64_t srow = (int64_t)k + kb + r;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = w + 4 * q;
    if (slab) sv[q] = (srow < n && c < kb) ? A[srow + (int64_t)(k + c) * n] : T(0);
    a[q] = (r < kb && c < kb) ? A[(k + r) + (int64_t)(k + c) * n] : (r == c ? T(1) : T(0));
  }
  if (slab) {
#pragma unroll
    for (int q = 0; q < 16; ++q) slab[r + kNB * (w + 4 * q)] = sv[q];
  }
  int bad = 0;  // first non-positive pivot among this wave's columns, + 1
  if (w == 0) potrf_pivot<T, 0>(a[0], colp, r, bad);
  PotrfSteps<T, 0>::run(a, colp, r, w, bad);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = w + 4 * q;
    Ls[r][c] = c <= r ? a[q] : T(0);
    if (r < kb && c < kb && c <= r) A[(k + r) + (int64_t)(k + c) * n] = a[q];
  }
  if (r == 0) bad_w[w] = bad;
  __syncthreads();
  if (w == 0) rinv[r] = T(1) / Ls[r][r];
  if (threadIdx.x == 0) {
    int first = 0;
    for (int v = 0; v < 4; ++v)
      if (bad_w[v] && (!first || bad_w[v] < first)) first = bad_w[v];
    if (first && first <= kb && *info == 0) *info = k + first;
  }
  __syncthreads();
  T x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = (r == w + 4 * q) ? T(1) : T(0);
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    const T rj = rinv[j];
    const T lrj = Ls[r][j];
    // Columns c = w + 4 q > j + 3 hold zeros in rows <= j: nothing to do.
#pragma unroll
    for (int q = 0; q <= (j >> 2); ++q) {
      const T xj = lane_bcast(x[q], j) * rj;
      x[q] = r == j ? xj : (r > j ? fma(-lrj, xj, x[q]) : x[q]);
    }
  }
  // LinvT goes through LDS so that both stores run along columns.
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = w + 4 * q;
    const T v = (r < kb && c < kb) ? x[q] : T(0);
    Linv[r + kNB * c] = v;
    Ls[r][c] = v;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = w + 4 * q;
    LinvT[r + kNB * c] = Ls[c][r];
  }
}
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(T* __restrict__ A, int64_t n, int k, int kb,
                                                    T* __restrict__ Linv, T* __restrict__ LinvT,
                                                    int32_t* __restrict__ info, T* __restrict__ slab) {
  __shared__ T colp[2][4][16];
  __shared__ T Ls[kNB][kNB + 1];
  __shared__ T rinv[kNB];
  __shared__ int bad_w[4];
  potrf_block<T>(A, n, k, kb, Linv, LinvT, info, slab, colp, Ls, rinv, bad_w);
}

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// 64 x 64 tile product D = As Bs' (As, Bs: 64 x kpad in LDS, row-major) on
// MFMA, four waves, one 32 x 32 quadrant each; `put(row, col, v)` receives
// every element of the wave's quadrant.
//   FP32: v_mfma_f32_32x32x2_f32 -- A operand lane l = A[l % 32][k + l / 32],
//         B operand lane l = B[k + l / 32][l % 32], result register r of lane
//         l = D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
//   FP64: v_mfma_f64_16x16x4_f64 -- A lane l = A[l % 16][k + l / 16], B lane
//         l = B[k + l / 16][l % 16], result r = D[l / 16 + 4 r][l % 16]
//         (both result layouts checked on gfx950 with a lane probe).
// One wave's 32 x 32 quadrant of a 64 x 64 tile product, in MFMA registers.
template <typename T>
struct TileAcc;
template <>
struct TileAcc<float> {
  floatx16 v;
};
template <>
struct TileAcc<double> {
  doublex4 v[2][2];
};

__device__ __forceinline__ void tile_mma(const float (*As)[kNB + 1], const float (*Bs)[kNB + 1], int kb,
                                         TileAcc<float>& acc, bool accumulate = false) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = w >> 1, qj = w & 1;
  if (!accumulate)
    for (int r = 0; r < 16; ++r) acc.v[r] = 0.0f;
  const int kpad = (kb + 1) & ~1;
  for (int kk = 0; kk < kpad; kk += 2) {
    const float a = As[qi * 32 + (lane & 31)][kk + (lane >> 5)];
    const float b = Bs[qj * 32 + (lane & 31)][kk + (lane >> 5)];
    acc.v = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc.v, 0, 0, 0);
  }
}
template <typename F>
__device__ __forceinline__ void tile_put(const TileAcc<float>& acc, F&& put) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = w >> 1, qj = w & 1;
#pragma unroll
  for (int r = 0; r < 16; ++r)
    put(qi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), qj * 32 + (lane & 31), acc.v[r]);
}
__device__ __forceinline__ void tile_mma(const double (*As)[kNB + 1], const double (*Bs)[kNB + 1], int kb,
                                         TileAcc<double>& acc, bool accumulate = false) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = w >> 1, qj = w & 1;
  if (!accumulate)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int r = 0; r < 4; ++r) acc.v[a][b][r] = 0.0;
  const int kpad = (kb + 3) & ~3;
  for (int kk = 0; kk < kpad; kk += 4) {
    double av[2], bv[2];
    for (int q = 0; q < 2; ++q) {
      av[q] = As[qi * 32 + q * 16 + (lane & 15)][kk + (lane >> 4)];
      bv[q] = Bs[qj * 32 + q * 16 + (lane & 15)][kk + (lane >> 4)];
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        acc.v[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc.v[a][b], 0, 0, 0);
  }
}
template <typename F>
__device__ __forceinline__ void tile_put(const TileAcc<double>& acc, F&& put) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = w >> 1, qj = w & 1;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        put(qi * 32 + a * 16 + (lane >> 4) + 4 * r, qj * 32 + b * 16 + (lane & 15), acc.v[a][b][r]);
}
template <typename T, typename F>
__device__ __forceinline__ void tile_nt(const T (*As)[kNB + 1], const T (*Bs)[kNB + 1], int kb, F&& put) {
  TileAcc<T> acc;
  tile_mma(As, Bs, kb, acc);
  tile_put(acc, put);
}

// A[r0 + j][c0 + i] -= element (i, j) of a transposed tile product, lower
// triangle only (row >= col).  All 16 reads of a lane are issued before the
// first write: a read-modify-write per element would wait one memory round
// trip per element.
template <typename T>
__device__ __forceinline__ void tile_sub_lower(const TileAcc<T>& acc, T* __restrict__ A, int64_t n, int64_t r0,
                                               int64_t c0) {
  T val[16];
  int q = 0;
  if (r0 + kNB <= n && r0 >= c0 + kNB) {
    // Whole tile inside and strictly below the diagonal (uniform): no
    // per-element guards, so the stores issue back to back.
    tile_put(acc, [&](int i, int j, T) { val[q++] = A[(r0 + j) + (c0 + i) * n]; });
    q = 0;
    tile_put(acc, [&](int, int, T v) {
      val[q] -= v;
      ++q;
    });
    q = 0;
    tile_put(acc, [&](int i, int j, T) { A[(r0 + j) + (c0 + i) * n] = val[q++]; });
    return;
  }
  tile_put(acc, [&](int i, int j, T) {
    const int64_t row = r0 + j, col = c0 + i;
    val[q++] = (row < n && col < n && row >= col) ? A[row + col * n] : T(0);
  });
  q = 0;
  tile_put(acc, [&](int i, int j, T v) {
    const int64_t row = r0 + j, col = c0 + i;
    if (row < n && col < n && row >= col) A[row + col * n] = val[q] - v;
    ++q;
  });
}

// Rows r0.. of columns k..k+kb of A into a 64 x 64 LDS tile (zero-padded).
template <typename T>
__device__ __forceinline__ void load_cols(T (*Ts)[kNB + 1], const T* __restrict__ A, int64_t n, int k, int kb,
                                          int64_t r0) {
  // 256 threads: all 16 loads of a thread in flight before the LDS stores.
  T v[kNB * kNB / 256];
  if (kb == kNB && r0 + kNB <= n) {  // whole tile in range (uniform): no per-load guards
#pragma unroll
    for (int q = 0; q < kNB * kNB / 256; ++q) {
      const int idx = threadIdx.x + 256 * q, i = idx % kNB, l = idx / kNB;
      v[q] = A[(r0 + i) + (int64_t)(k + l) * n];
    }
  } else {
#pragma unroll
    for (int q = 0; q < kNB * kNB / 256; ++q) {
      const int idx = threadIdx.x + 256 * q, i = idx % kNB, l = idx / kNB;
      v[q] = (l < kb && r0 + i < n) ? A[(r0 + i) + (int64_t)(k + l) * n] : T(0);
    }
  }
#pragma unroll
  for (int q = 0; q < kNB * kNB / 256; ++q) {
    const int idx = threadIdx.x + 256 * q;
    Ts[idx % kNB][idx / kNB] = v[q];
  }
}

// Panel step: workgroup ti writes its 64-row slab of L21 = A21 L11^-T and,
// with col0, then applies the trailing update to tile (ti, 0), the first
// block column after the panel (the next block's diagonal and panel
// columns).  Every workgroup forms the first slab's L21 rows itself (same
// operands, same MFMA sequence: the values the panel writes), so no
// workgroup waits for another; they read that slab's A21 rows from the
// copy k_potrf_diag made (slab), since workgroup 0 overwrites them in A and
// a workgroup may start after it has.  col0 = 0: the panel alone (its
// trailing update is left to k_syrk2_mfma).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_col0(T* __restrict__ A, int64_t n, int k, int kb,
                                                    const T* __restrict__ Linv, int col0,
                                                    const T* __restrict__ slab, T* __restrict__ Linv2,
                                                    T* __restrict__ LinvT2, int32_t* __restrict__ info) {
  __shared__ T As[kNB][kNB + 1];
  __shared__ T Bs[kNB][kNB + 1];
  __shared__ T Cs[kNB][kNB + 1];
  const int64_t base = (int64_t)k + kb;
  const int64_t r0 = base + (int64_t)blockIdx.x * kNB;
  load_cols(As, A, n, k, kb, r0);
  load_cols(Bs, Linv, kNB, 0, kNB, 0);
  if (col0) load_cols(Cs, slab, kNB, 0, kb, 0);
  __syncthreads();
  TileAcc<T> xi, x0;
  tile_mma(Bs, As, kb, xi);  // element (i, j) = L21[r0 + j][k + i]
  if (col0) tile_mma(Bs, Cs, kb, x0);  // element (i, j) = L21[base + j][k + i]
  tile_put(xi, [&](int i, int j, T v) {
    if (r0 + j < n && i < kb) A[(r0 + j) + (int64_t)(k + i) * n] = v;
  });
  if (!col0) return;
  __syncthreads();
  tile_put(xi, [&](int i, int j, T v) { As[j][i] = v; });
  tile_put(x0, [&](int i, int j, T v) { Cs[j][i] = v; });
  __syncthreads();
  TileAcc<T> up;
  tile_mma(Cs, As, kb, up);  // element (i, j) = sum_l L21[base + i][l] L21[r0 + j][l]
  tile_sub_lower(up, A, n, r0, base);
  if (Linv2 == nullptr || blockIdx.x != 0) return;
  // Workgroup 0 has just finished the next block's diagonal tile (no other
  // workgroup writes it): factor that block here instead of in a launch of
  // its own.  Its stores are complete and visible to the workgroup after
  // the barrier; the LDS tiles are free.
  __syncthreads();
  potrf_block<T>(A, n, (int)base, (int)(n - base < kNB ? n - base : kNB), Linv2, LinvT2, info, nullptr,
                 reinterpret_cast<T(*)[4][16]>(&Bs[0][0]), As, &Cs[0][0], reinterpret_cast<int*>(&Cs[2][0]));
}

__device__ __forceinline__ void syrk_tile_coords(int t, int& ti, int& tj) {
  ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  tj = t - ti * (ti + 1) / 2;
}

// Trailing update after two 64-column steps k and k + 64 (both full):
// A22 -= L21a L21a' + L21b L21b' for the tiles after base = k + 128, the two
// products summed in the MFMA accumulators and the tile read and written
// once (one read-modify-write of the trailing matrix per 128 columns).
// cols = 1: tiles (ti, tj) of the first two block columns (tj < 2, ti >= tj;
// nt tile rows after base), t < nt -> (t, 0), else (t - nt + 1, 1);
// cols = 0: the lower tiles with ti >= tj >= 2.
template <typename T>
__global__ __launch_bounds__(256) void k_syrk2_mfma(T* __restrict__ A, int64_t n, int k, int nt, int cols) {
  __shared__ T As[kNB][kNB + 1];
  __shared__ T Bs[kNB][kNB + 1];
  int ti, tj;
  const int t = (int)blockIdx.x;
  if (cols) {
    tj = t < nt ? 0 : 1;
    ti = t < nt ? t : t - nt + 1;
  } else {
    syrk_tile_coords(t, ti, tj);
    ti += 2;
    tj += 2;
  }
  const int64_t base = (int64_t)k + 2 * kNB;
  const int64_t r0 = base + (int64_t)ti * kNB, c0 = base + (int64_t)tj * kNB;
  TileAcc<T> up;
  load_cols(As, A, n, k, kNB, r0);
  load_cols(Bs, A, n, k, kNB, c0);
  __syncthreads();
  tile_mma(Bs, As, kNB, up);
  __syncthreads();
  load_cols(As, A, n, k + kNB, kNB, r0);
  load_cols(Bs, A, n, k + kNB, kNB, c0);
  __syncthreads();
  tile_mma(Bs, As, kNB, up, true);
  tile_sub_lower(up, A, n, r0, c0);
}

// Triangular solves with the factor in one launch per direction: one
// workgroup of kFlowThreads threads per 64-row block.  A block's solution
// is handed to the workgroups that need it as 64-bit words (32 bits of the
// value, the solve's epoch), each stored and polled with agent-scope
// relaxed atomics: a word that carries the epoch is this solve's value, so
// there is no flag, fence or barrier between the blocks.  A workgroup waits
// only on blocks that lower-numbered workgroups own, and workgroups are
// dispatched in order, so every block it waits on is owned by a workgroup
// already resident or finished.  Each wait is bounded (spin_max polls,
// kFlowSpinMax by default):
// a wait that runs out stores the epoch into err, the other waits then give
// up too, and the blocks written after that are NaN, so a failed solve
// reaches the LM as a non-finite step instead of a hang.  Per-thread
// partial sums over all earlier blocks are reduced once, at the end.
constexpr int kFlowThreads = 4 * 64;
constexpr unsigned kFlowSpinMax = 1u << 20;

template <typename T>
__host__ __device__ constexpr int flow_words() { return (int)(sizeof(T) / 4); }

template <typename T>
__device__ __forceinline__ void flow_put(uint64_t* dst, T v, unsigned epoch) {
  unsigned u[flow_words<T>()];
  __builtin_memcpy(u, &v, sizeof(T));
#pragma unroll
  for (int c = 0; c < flow_words<T>(); ++c)
    __hip_atomic_store(dst + c, ((uint64_t)epoch << 32) | u[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Polls the N values at src (flow_words<T>() words each) until every word
// carries epoch; false when the wait ran out (spin_max polls; 0: it gives up
// without polling) or another one had (err == epoch).
template <typename T, int N>
__device__ __forceinline__ bool flow_get(const uint64_t* src, unsigned epoch, uint64_t* err, unsigned spin_max,
                                         T out[N]) {
  constexpr int C = flow_words<T>();
  uint64_t w[N * C];
#pragma unroll
  for (int q = 0; q < N * C; ++q) w[q] = 0;
  bool good = spin_max != 0;
  for (unsigned spins = 0; good; ++spins) {
    bool all = true;
#pragma unroll
    for (int q = 0; q < N * C; ++q) {
      w[q] = __hip_atomic_load(src + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      all = all && (unsigned)(w[q] >> 32) == epoch;
    }
    if (all) break;
    if (spins >= spin_max ||
        ((spins & 63u) == 63u && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch))
      good = false;
    else
      __builtin_amdgcn_s_sleep(1);
  }
  if (!good) __hip_atomic_store(err, (uint64_t)epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    unsigned u[C];
#pragma unroll
    for (int c = 0; c < C; ++c) u[c] = (unsigned)w[j * C + c];
    __builtin_memcpy(&out[j], u, sizeof(T));
  }
  return good;
}

// The thread's 16 entries of the block's stored 64 x 64 inverse M
// (column-major): row lane, columns 16w..16w+15.  Loaded at the start, off
// the chain of blocks.
template <typename T>
__device__ __forceinline__ void flow_block_load(const T* __restrict__ M, T m[16]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 16; ++j) m[j] = M[lane + kNB * (16 * w + j)];
}

// d = M s: wave w takes columns 16w..16w+15 (m from flow_block_load), the
// four partial column sums meet in LDS.  s in LDS (zero past the block's
// rows); every thread calls it.
template <typename T>
__device__ __forceinline__ T flow_block_gemv(const T m[16], const T* sv, T (*part)[kNB]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T a = T(0);
#pragma unroll
  for (int j = 0; j < 16; ++j) a = fma(m[j], sv[16 * w + j], a);
  part[w][lane] = a;
  __syncthreads();
  return (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// Wave 0 writes the block's values (rows < nrows) to out and, as epoch
// words, to hand.
template <typename T>
__device__ __forceinline__ void flow_publish(T* out, uint64_t* hand, int nrows, T v, unsigned epoch) {
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < 64 && lane < nrows) {
    out[lane] = v;
    flow_put(hand + (size_t)flow_words<T>() * lane, v, epoch);
  }
}

// Forward L y = b, two blocks per workgroup: workgroup g solves blocks
// a = 2g and b = 2g + 1 (if b < nb) with y_i = L_ii^-1 (b_i - sum_{k<i}
// L_ik y_k).  Thread (wave w, lane) accumulates row lane of columns
// 16w..16w+15 of every L_ak and L_bk, k < a (coalesced column reads), loading
// the tiles before polling for y_k; block a is then finished and published,
// and block b adds L_ba y_a with y_a passed through LDS (one hand-off between
// workgroups per two blocks).  The sums run over k in the same order as with
// one block per workgroup.  hand: n values as epoch words.
template <typename T>
__global__ __launch_bounds__(kFlowThreads) void k_trsv_lower_flow(const T* __restrict__ L, const T* __restrict__ Linv,
                                                                  int64_t n, int nb, const T* __restrict__ b,
                                                                  T* __restrict__ y, uint64_t* hand, unsigned epoch,
                                                                  uint64_t* err, unsigned spin_max) {
  constexpr int C = flow_words<T>();
  __shared__ T part[4][kNB];
  __shared__ T sv[kNB];
  const int ia = 2 * (int)blockIdx.x, ibb = ia + 1;
  const bool two = ibb < nb;  // then block a is full
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t a0 = (int64_t)ia * kNB, b0 = a0 + kNB;
  const int na = (int)min((int64_t)kNB, n - a0);
  const int nbr = two ? (int)min((int64_t)kNB, n - b0) : 0;
  const int64_t rowa = a0 + (lane < na ? lane : 0);
  const int64_t rowb = two ? b0 + (lane < nbr ? lane : 0) : rowa;
  T ma[16], mb[16];
  flow_block_load(Linv + (size_t)ia * kNB * kNB, ma);
  if (two) flow_block_load(Linv + (size_t)ibb * kNB * kNB, mb);
  const T ba = w == 0 && lane < na ? b[a0 + lane] : T(0);
  const T bb = w == 0 && lane < nbr ? b[b0 + lane] : T(0);
  T acca = T(0), accb = T(0);
  bool good = true;
  for (int k = 0; k < ia && good; ++k) {
    // Block k < a is full.
    const int64_t c0 = (int64_t)k * kNB + 16 * w;
    const T* cola = L + rowa + c0 * n;
    const T* colb = L + rowb + c0 * n;
    T la[16], lb[16], yk[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      la[j] = cola[j * n];
      lb[j] = two ? colb[j * n] : T(0);
    }
    good = flow_get<T, 16>(hand + C * c0, epoch, err, spin_max, yk);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      acca = fma(la[j], yk[j], acca);
      accb = fma(lb[j], yk[j], accb);
    }
  }
  // L_ba: row lane of block b, columns 16w..16w+15 of block a.
  T lba[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) lba[j] = two ? L[rowb + (a0 + 16 * w + j) * n] : T(0);
  part[w][lane] = acca;
  good = __syncthreads_and(good) != 0;
  if (w == 0) sv[lane] = lane < na ? ba - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) : T(0);
  __syncthreads();
  T ya = flow_block_gemv(ma, sv, part);
  if (!good) ya = T(NAN);
  flow_publish(y + a0, hand + C * a0, na, ya, epoch);
  if (!two) return;
  __syncthreads();  // the gemv's reads of sv and part are done
  if (w == 0) sv[lane] = ya;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) accb = fma(lba[j], sv[16 * w + j], accb);
  part[w][lane] = accb;
  __syncthreads();
  if (w == 0) sv[lane] = lane < nbr ? bb - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) : T(0);
  __syncthreads();
  T yb = flow_block_gemv(mb, sv, part);
  if (!good) yb = T(NAN);
  flow_publish(y + b0, hand + C * b0, nbr, yb, epoch);
}

// Backward L' x = y, two blocks per workgroup: workgroup g solves blocks
// h = nb - 1 - 2g and l = h - 1 (if l >= 0) with x_i = L_ii^-T (y_i -
// sum_{k>i} L_ki' x_k).  Thread (wave w, lane) accumulates lane's share of
// rows 16w..16w+15 of every L_kh' and L_kl', k > h (columns of L read along
// their rows k0 + lane: coalesced), summed over the wave at the end; block h
// is finished and published first, and block l adds L_hl' x_h, each thread
// holding x_h's row lane from the block product.
template <typename T>
__device__ __forceinline__ T flow_upper_finish(T acc[16], T yv, int ib, const T m[16], T* sv, T (*part)[kNB]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const T v = wave_sum(acc[r]);
    if (lane == r) acc[0] = v;  // lane r keeps row 16w + r
  }
  if (lane < 16) sv[16 * w + lane] = 16 * w + lane < ib ? yv - acc[0] : T(0);
  __syncthreads();
  return flow_block_gemv(m, sv, part);
}
template <typename T>
__global__ __launch_bounds__(kFlowThreads) void k_trsv_upper_flow(const T* __restrict__ L, const T* __restrict__ LinvT,
                                                                  int64_t n, int nb, const T* __restrict__ y,
                                                                  T* __restrict__ x, uint64_t* hand, unsigned epoch,
                                                                  uint64_t* err, unsigned spin_max) {
  constexpr int C = flow_words<T>();
  __shared__ T part[4][kNB];
  __shared__ T sv[kNB];
  const int ih = nb - 1 - 2 * (int)blockIdx.x, il = ih - 1;
  const bool two = il >= 0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t h0 = (int64_t)ih * kNB, l0 = h0 - kNB;
  const int nh = (int)min((int64_t)kNB, n - h0);  // block l is full
  T mh[16], ml[16];
  flow_block_load(LinvT + (size_t)ih * kNB * kNB, mh);
  if (two) flow_block_load(LinvT + (size_t)il * kNB * kNB, ml);
  const T yh = lane < 16 && 16 * w + lane < nh ? y[h0 + 16 * w + lane] : T(0);
  const T yl = two && lane < 16 ? y[l0 + 16 * w + lane] : T(0);
  T acch[16], accl[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acch[r] = accl[r] = T(0);
  bool good = true;
  for (int k = nb - 1; k > ih && good; --k) {
    // Blocks h, l < k are full; block k is the last (possibly partial) one.
    const int64_t k0 = (int64_t)k * kNB;
    const bool on = lane < (int)min((int64_t)kNB, n - k0);
    const T* ph = L + (k0 + (on ? lane : 0)) + (h0 + 16 * w) * n;
    const T* pl = L + (k0 + (on ? lane : 0)) + (l0 + 16 * w) * n;
    T lh[16], ll[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      lh[r] = on ? ph[r * n] : T(0);
      ll[r] = on && two ? pl[r * n] : T(0);
    }
    T xk[1] = {T(0)};
    if (on) good = flow_get<T, 1>(hand + C * (k0 + lane), epoch, err, spin_max, xk);
    good = __all(good);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      acch[r] = fma(lh[r], xk[0], acch[r]);
      accl[r] = fma(ll[r], xk[0], accl[r]);
    }
  }
  // L_hl: rows h0 + lane (block h, possibly partial), columns l0 + 16w + r.
  const bool onh = lane < nh;
  T lhl[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) lhl[r] = two && onh ? L[(h0 + lane) + (l0 + 16 * w + r) * n] : T(0);
  good = __syncthreads_and(good) != 0;
  T xh = flow_upper_finish(acch, yh, nh, mh, sv, part);
  if (!good) xh = T(NAN);
  flow_publish(x + h0, hand + C * h0, nh, xh, epoch);
  if (!two) return;
  const T xo = onh ? xh : T(0);
#pragma unroll
  for (int r = 0; r < 16; ++r) accl[r] = fma(lhl[r], xo, accl[r]);
  __syncthreads();  // the block product's reads of sv and part are done
  T xl = flow_upper_finish(accl, yl, kNB, ml, sv, part);
  if (!good) xl = T(NAN);
  flow_publish(x + l0, hand + C * l0, kNB, xl, epoch);
}

// FP64 residual r = b - S x (S column-major, both triangles stored) in two
// passes: partial row sums over kGemvSplit column ranges (coalesced down
// the columns), then r = b - the partials summed in a fixed order.
constexpr int kGemvSplit = 32;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_partial(const T* __restrict__ S, int64_t n,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ part) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t per = (n + kGemvSplit - 1) / kGemvSplit;
  const int64_t j0 = blockIdx.y * per, j1 = min(n, j0 + per);
  if (i >= n) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t j = j0;
  for (; j + 4 <= j1; j += 4) {
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += (double)S[i + (j + q) * n] * x[j + q];
  }
  for (; j < j1; ++j) acc[0] += (double)S[i + j * n] * x[j];
  part[blockIdx.y * n + i] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}
__global__ __launch_bounds__(kBlock) void k_gemv_final(int64_t n, const double* __restrict__ part,
                                                       const double* __restrict__ b, double* __restrict__ r) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int q = 0; q < kGemvSplit; ++q) s += part[q * n + i];
  r[i] = b[i] - s;
}

// ------------------------------------------------------- small helpers
__global__ __launch_bounds__(kBlock) void k_axpy_inplace(int64_t n, const double* __restrict__ dx,
                                                         double* __restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] += dx[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_negate(int64_t n, const T* __restrict__ src, T* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = -src[i];
}

__global__ __launch_bounds__(kBlock) void k_copy64(int64_t n, const double* __restrict__ src,
                                                   double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_fill64(int64_t n, double v, double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// Trial cameras: out[cam] = cams[cam] + step .* scale for active cameras
// (F order step), with partial sums of |delta|^2.
__global__ __launch_bounds__(kBlock) void k_plus_cams(int nf, const int32_t* __restrict__ f_to_cam,
                                                      const double* __restrict__ cams,
                                                      const double* __restrict__ step,
                                                      const double* __restrict__ scale,
                                                      double* __restrict__ out,
                                                      double* __restrict__ partial, Win win) {
  __shared__ double lds[kBlock / 64];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  const bool on = i < 9 * (int64_t)nf && !win_idle_f(win, (int)(i / 9));
  if (on) {
    const size_t j = 9 * (size_t)f_to_cam[i / 9] + i % 9;
    const double v = cams[j] + step[i] * scale[i];
    out[j] = v;
    const double df = cams[j] - v;
    acc[0] = df * df;
  }
  if (win.acc) {
    win_atomic_add(win.acc, 1, 0, on ? win_of_f(win, (int)(i / 9)) : -1, acc[0]);
    return;
  }
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

// Jacobi scaling 1 / (1 + |column|) from an unscaled pass
// (TrustRegionMinimizer::EstimateScale, trust_region_minimizer.cc:103-109).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_make_scale(int ne, int nf, const double* __restrict__ V,
                                                       const T* __restrict__ U,
                                                       double* __restrict__ pscale,
                                                       double* __restrict__ cscale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) {
    const int di[3] = {0, 3, 5};
    for (int j = 0; j < 3; ++j) pscale[3 * (size_t)i + j] = 1.0 / (1.0 + sqrt(V[6 * (size_t)i + di[j]]));
  }
  if (i < nf)
    for (int j = 0; j < 9; ++j) cscale[9 * (size_t)i + j] = 1.0 / (1.0 + sqrt((double)U[45 * (size_t)i + ut9(j, j)]));
}

// CostFunction::Evaluate per observation in input order (no loss, no
// scaling): residual, row-major 2x9 camera and 2x3 point Jacobians, in a
// grid-stride loop.  Observations and points are read as `In` and rounded
// once to the kernel precision T.  With cost_partial, each block writes the
// FP64 sum of 1/2 rho(|r|^2) (HuberLoss with threshold huber_a; <= 0: none)
// of its observations to cost_partial[blockIdx.x].  With nc > 0, an
// observation whose camera or point index is outside [0, nc) / [0, np) is
// never read through: its outputs are NaN.
template <typename T, typename In = T>
__global__ __launch_bounds__(kBlock) void k_eval_obs(int64_t n, int nc, int np,
                                                     const int32_t* __restrict__ cam_idx,
                                                     const int32_t* __restrict__ pt_idx,
                                                     const In* __restrict__ obs, const T* __restrict__ ctab,
                                                     const In* __restrict__ pts, double* __restrict__ res,
                                                     double* __restrict__ jc, double* __restrict__ jp,
                                                     double huber_a, double* __restrict__ cost_partial) {
  __shared__ double lds[kBlock / 64];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const int cam = cam_idx[i], q = pt_idx[i];
    if (nc > 0 && (cam < 0 || cam >= nc || q < 0 || q >= np)) {
      const double nan = __builtin_nan("");
      if (res) res[2 * i] = res[2 * i + 1] = nan;
      if (jc)
        for (int k = 0; k < 18; ++k) jc[18 * i + k] = nan;
      if (jp)
        for (int k = 0; k < 6; ++k) jp[6 * i + k] = nan;
      acc[0] += nan;
      continue;
    }
    const T* tab = ctab + (size_t)kCamTab * cam;
    CamProj<T> c;
    load_proj(tab, c);
    const T X[3] = {(T)pts[3 * (size_t)q], (T)pts[3 * (size_t)q + 1], (T)pts[3 * (size_t)q + 2]};
    Proj<T> o;
    project(c, X, (T)obs[2 * i], (T)obs[2 * i + 1], o);
    if (res) {
      res[2 * i] = (double)o.rx;
      res[2 * i + 1] = (double)o.ry;
    }
    if (cost_partial) {
      const double rx = (double)o.rx, ry = (double)o.ry;
      acc[0] += huber_cost(rx * rx + ry * ry, huber_a);
    }
    const T one[9] = {T(1), T(1), T(1), T(1), T(1), T(1), T(1), T(1), T(1)};
    if (jc) {
      T J[18];
      jac_cam(o, tab + kTabDR, X, T(1), one, J);
      for (int k = 0; k < 18; ++k) jc[18 * i + k] = (double)J[k];
    }
    if (jp) {
      T J[6];
      jac_point(o, c.R, T(1), one, J);
      for (int k = 0; k < 6; ++k) jp[6 * i + k] = (double)J[k];
    }
  }
  if (!cost_partial) return;
  block_sum<double, 1>(acc, lds);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    cost_partial[blockIdx.x] = s;
  }
}

}  // namespace mba

#endif  // MBA_DEVICE_H_
