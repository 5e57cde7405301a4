// Kernels of the grouped layout (L = 1): residual blocks (intrinsics,
// camera, point) of libs/sfm's incremental engine
// (sfm_incremental_engine.cpp:1413-1446, ErrorFunc_Refine_Camera_3DPoints in
// libs/sfm/include/mvg/sfm/pinhole_brown_rt_ceres_functor.h:42-110).
//
// A row touches two F blocks: its camera (slots [aa, t]) and its camera's
// intrinsic group (slots [f, ppx, ppy, k1, k2, k3]).  Camera-major chunks
// are ordered by group, so the chunks of one group's cameras are contiguous;
// each chunk writes two partial rows, its camera's at `ch` and its group's at
// `nchunks + ch`, and an F block sums the partial rows [f_lo, f_hi).
//
// FP64 parameters ("slot blocks", 9 doubles each): camera c at c, group g at
// nc + g; unused slots hold 0.
#ifndef MBA_DEVICE_GROUPED_H_
#define MBA_DEVICE_GROUPED_H_

#include "mba_device.h"

namespace mba {

// Camera table of camera c from its slot block and its group's
// (cam_grp[c] < 0: a camera without observations, intrinsics 0).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cam_table_g(int nc, const double* __restrict__ slots,
                                                        const int32_t* __restrict__ cam_grp,
                                                        T* __restrict__ tab) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double* p = slots + 9 * (size_t)c;
  double R[9], dR[27];
  const double w[3] = {p[0], p[1], p[2]};
  rodrigues_with_jacobian(w, R, dR);
  const int g = cam_grp[c];
  double in[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (g >= 0)
    for (int i = 0; i < 6; ++i) in[i] = slots[9 * ((size_t)nc + g) + i];
  T* o = tab + (size_t)kCamTab * c;
  for (int i = 0; i < 9; ++i) o[i] = (T)R[i];
  for (int i = 0; i < 3; ++i) o[kTabT + i] = (T)p[3 + i];
  o[kTabF] = (T)in[0];
  o[kTabF + 1] = (T)in[3];
  o[kTabF + 2] = (T)in[4];
  o[kTabK3] = (T)in[5];
  for (int i = 0; i < 27; ++i) o[kTabDR + i] = (T)dR[i];
  o[kTabPP] = (T)in[1];
  o[kTabPP + 1] = (T)in[2];
  for (int i = kTabPP + 2; i < kCamTab; ++i) o[i] = (T)0;
}

// Extrinsic Jacobian (2 x 6 row-major: aa, t) and intrinsic Jacobian (2 x 6:
// f, ppx, ppy, k1, k2, k3), weight and per-column scales applied.
template <typename T>
__device__ __forceinline__ void jac_ext(const Proj<T>& o, const T* dR, const T X[3], T wgt, const T cs[6],
                                        T J[12]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T* D = dR + 9 * k;
    const T v0 = D[0] * X[0] + D[1] * X[1] + D[2] * X[2];
    const T v1 = D[3] * X[0] + D[4] * X[1] + D[5] * X[2];
    const T v2 = D[6] * X[0] + D[7] * X[1] + D[8] * X[2];
    J[k] = wgt * cs[k] * (o.P[0] * v0 + o.P[1] * v1 + o.P[2] * v2);
    J[6 + k] = wgt * cs[k] * (o.P[3] * v0 + o.P[4] * v1 + o.P[5] * v2);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    J[3 + k] = wgt * cs[3 + k] * o.P[k];
    J[9 + k] = wgt * cs[3 + k] * o.P[3 + k];
  }
}

template <typename T>
__device__ __forceinline__ void jac_intr(const Proj<T>& o, T wgt, const T cs[6], T J[12]) {
  J[0] = wgt * cs[0] * o.jf[0];
  J[1] = wgt * cs[1];
  J[2] = T(0);
  J[3] = wgt * cs[3] * o.jk1[0];
  J[4] = wgt * cs[4] * o.jk2[0];
  J[5] = wgt * cs[5] * o.jk3[0];
  J[6] = wgt * cs[0] * o.jf[1];
  J[7] = T(0);
  J[8] = wgt * cs[2];
  J[9] = wgt * cs[3] * o.jk1[1];
  J[10] = wgt * cs[4] * o.jk2[1];
  J[11] = wgt * cs[5] * o.jk3[1];
}

__device__ __forceinline__ int ut6(int i, int j) { return i * 6 - i * (i - 1) / 2 + (j - i); }

// Accumulator layout of the grouped camera pass: U_ee 21 | U_ii 21 | g_e 6 |
// h_e 6 | g_i 6 | h_i 6 (camera columns unscaled).
constexpr int kAccG = 66;

// Value `q` (0..62) of a 9-slot partial row (U 45 | g 9 | h 9) of part 0
// (camera) or 1 (group), from the reduced accumulators in LDS.
template <typename T>
__device__ __forceinline__ T slot_value_g(const T* lds, int part, int q) {
  int idx = -1;
  if (q < 45) {
    int a = 0;
    while (a < 8 && q >= ut9(a + 1, a + 1)) ++a;
    const int b = a + (q - ut9(a, a));
    if (a < 6 && b < 6) idx = 21 * part + ut6(a, b);
  } else if (q < 54) {
    const int a = q - 45;
    if (a < 6) idx = 42 + 12 * part + a;
  } else {
    const int a = q - 54;
    if (a < 6) idx = 48 + 12 * part + a;
  }
  if (idx < 0) return T(0);
  T s = T(0);
  for (int w = 0; w < kBlock / 64; ++w) s += lds[w * kAccG + idx];
  return s;
}

// One workgroup per camera-major chunk: partial U, g and
// h = sum J'(r - Jp w_e) of the chunk's camera block and of its group block.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cam_pass_g(
    int nchunks, const int32_t* __restrict__ chunk_cam, const int32_t* __restrict__ chunk_begin,
    const int32_t* __restrict__ chunk_end, const int32_t* __restrict__ cm_e, const T* __restrict__ cm_obs,
    const T* __restrict__ pk, const T* __restrict__ ctab, T huber_a, T* __restrict__ partial) {
  __shared__ T lds[kAccG * (kBlock / 64)];
  const int ch = blockIdx.x;
  const T* tab = ctab + (size_t)kCamTab * chunk_cam[ch];
  CamProj<T> c;
  load_proj<T, 1>(tab, c);
  T dR[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) dR[i] = tab[kTabDR + i];
  const T one[6] = {T(1), T(1), T(1), T(1), T(1), T(1)};
  T acc[kAccG];
#pragma unroll
  for (int i = 0; i < kAccG; ++i) acc[i] = T(0);
  auto row = [&](const T X[3], const T sw[3], T ox, T oy) {
    Proj<T> o;
    project<T, 1>(c, X, ox, oy, o);
    const T wgt = huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T Je[12], Ji[12];
    jac_ext(o, dR, X, wgt, one, Je);
    jac_intr(o, wgt, one, Ji);
    const T rx = wgt * o.rx, ry = wgt * o.ry;
    T u[2];
    jac_point_apply(o, c.R, sw, u);
    const T tx = rx - wgt * u[0], ty = ry - wgt * u[1];
    int k = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a; b < 6; ++b) fma2(acc[k++], Je[a], Je[b], Je[6 + a], Je[6 + b]);
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a; b < 6; ++b) fma2(acc[k++], Ji[a], Ji[b], Ji[6 + a], Ji[6 + b]);
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      fma2(acc[42 + a], Je[a], rx, Je[6 + a], ry);
      fma2(acc[48 + a], Je[a], tx, Je[6 + a], ty);
      fma2(acc[54 + a], Ji[a], rx, Ji[6 + a], ry);
      fma2(acc[60 + a], Ji[a], tx, Ji[6 + a], ty);
    }
  };
  const int end = chunk_end[ch];
  int i = chunk_begin[ch] + threadIdx.x;
  for (; i + kBlock < end; i += 2 * kBlock) {
    const int e0 = cm_e[i], e1 = cm_e[i + kBlock];
    const T o0x = cm_obs[2 * (size_t)i], o0y = cm_obs[2 * (size_t)i + 1];
    const T o1x = cm_obs[2 * (size_t)(i + kBlock)], o1y = cm_obs[2 * (size_t)(i + kBlock) + 1];
    T X0[3], s0[3], X1[3], s1[3];
    load_pk(pk, e0, X0, s0);
    load_pk(pk, e1, X1, s1);
    row(X0, s0, o0x, o0y);
    row(X1, s1, o1x, o1y);
  }
  if (i < end) {
    T X0[3], s0[3];
    load_pk(pk, cm_e[i], X0, s0);
    row(X0, s0, cm_obs[2 * (size_t)i], cm_obs[2 * (size_t)i + 1]);
  }
  block_sum<T, kAccG>(acc, lds);
  for (int k = threadIdx.x; k < 126; k += kBlock) {
    const int part = k / 63, q = k % 63;
    partial[(size_t)kPartial * (part ? (size_t)nchunks + ch : (size_t)ch) + q] = slot_value_g(lds, part, q);
  }
}

// A group block's partial rows span all chunks of its cameras (hundreds to
// thousands of rows): one 1024-thread workgroup per group block sums rows
// [g_lo, g_hi) (first `width` values) in a fixed order into row out_row[k],
// so the finalize kernels read one row per group.
constexpr int kGroupReduceThreads = 1024;
template <typename T>
__global__ __launch_bounds__(kGroupReduceThreads) void k_group_rows(const int32_t* __restrict__ g_lo,
                                                                   const int32_t* __restrict__ g_hi,
                                                                   const int32_t* __restrict__ out_row, int width,
                                                                   T* __restrict__ partial,
                                                                   const int32_t* __restrict__ all_done) {
  __shared__ T lds[kGroupReduceThreads];
  if (all_done && *all_done) return;
  const int k = blockIdx.x;
  const int v = threadIdx.x & 63, stripe = threadIdx.x >> 6, nstripe = kGroupReduceThreads / 64;
  T s = T(0);
  if (v < width)
    for (int r = g_lo[k] + stripe; r < g_hi[k]; r += nstripe) s += partial[(size_t)kPartial * r + v];
  lds[threadIdx.x] = s;
  __syncthreads();
  if (stripe == 0 && v < width) {
    T t = T(0);
    for (int w = 0; w < nstripe; ++w) t += lds[64 * w + v];
    partial[(size_t)kPartial * out_row[k] + v] = t;
  }
}

// Direction table of a grouped camera for an F-space vector x (or zeros):
// M = sum_k xs_k dR_k, dt = xs_t from its camera block, and its group
// block's [df, dppx, dppy, dk1, dk2, dk3]; xs = sign * cscale .* x.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cam_dir_g(int nc, const int32_t* __restrict__ cam_to_f,
                                                      const int32_t* __restrict__ cam_gf,
                                                      const T* __restrict__ cscale, const T* __restrict__ ctab,
                                                      const T* __restrict__ x, T sign, T* __restrict__ dtab,
                                                      const int32_t* __restrict__ all_done) {
  if (all_done && *all_done) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  T* o = dtab + (size_t)kDirTabG * c;
  const int f = cam_to_f[c], g = cam_gf[c];
  T xs[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, xi[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (x && f >= 0)
    for (int i = 0; i < 6; ++i) xs[i] = sign * cscale[9 * (size_t)f + i] * x[9 * (size_t)f + i];
  if (x && g >= 0)
    for (int i = 0; i < 6; ++i) xi[i] = sign * cscale[9 * (size_t)g + i] * x[9 * (size_t)g + i];
  const T* dR = ctab + (size_t)kCamTab * c + kTabDR;
  for (int i = 0; i < 9; ++i) o[i] = xs[0] * dR[i] + xs[1] * dR[9 + i] + xs[2] * dR[18 + i];
  for (int i = 0; i < 3; ++i) o[9 + i] = xs[3 + i];
  for (int i = 0; i < 6; ++i) o[12 + i] = xi[i];
}

// p = z + beta p (p = z on iteration 1), one thread per F slot.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_p(int nf, const PcgState* __restrict__ st, const T* __restrict__ z,
                                                  T* __restrict__ p, const int32_t* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 9 * (int64_t)nf || st[0].done) return;
  const T beta = (T)st[0].beta;
  p[i] = st[0].iter == 1 ? z[i] : z[i] + beta * p[i];
}

// Per chunk: partials of F'(F x - E z) for the chunk's camera block (row ch)
// and its group block (row nchunks + ch), rows of the chunk's camera; F x
// through the camera's direction-table row (dtab NULL: x = 0), E z through
// the packed point records (v = s_p .* z).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mv_cam_g(
    int nchunks, const int32_t* __restrict__ chunk_cam, const int32_t* __restrict__ chunk_begin,
    const int32_t* __restrict__ chunk_end, const int32_t* __restrict__ cam_to_f,
    const int32_t* __restrict__ cam_gf, const int32_t* __restrict__ cm_e, const T* __restrict__ cm_obs,
    const T* __restrict__ pk, const T* __restrict__ cscale, const T* __restrict__ ctab,
    const T* __restrict__ dtab, T huber_a, T* __restrict__ partial, const int32_t* __restrict__ all_done) {
  __shared__ T lds[12 * (kBlock / 64)];
  if (all_done && *all_done) return;
  const int ch = blockIdx.x;
  const int cam = chunk_cam[ch];
  const int f = cam_to_f[cam], g = cam_gf[cam];
  const T* tab = ctab + (size_t)kCamTab * cam;
  CamProj<T> c;
  load_proj<T, 1>(tab, c);
  T dR[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) dR[i] = tab[kTabDR + i];
  T cs[6], ci[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    cs[i] = f >= 0 ? cscale[9 * (size_t)f + i] : T(0);
    ci[i] = g >= 0 ? cscale[9 * (size_t)g + i] : T(0);
  }
  T dv[18];
  const bool has_x = dtab != nullptr;
#pragma unroll
  for (int i = 0; i < 18; ++i) dv[i] = has_x ? dtab[(size_t)kDirTabG * cam + i] : T(0);
  T acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = T(0);
  auto row = [&](const T X[3], const T sz[3], T ox, T oy) {
    Proj<T> o;
    project<T, 1>(c, X, ox, oy, o);
    const T wgt = huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T u[2];
    jac_point_apply(o, c.R, sz, u);
    u[0] *= -wgt;
    u[1] *= -wgt;
    if (has_x) {
      T v[2];
      jac_cam_dir<T, 1>(o, dv, X, wgt, v);
      u[0] += v[0];
      u[1] += v[1];
    }
    T Je[12], Ji[12];
    jac_ext(o, dR, X, wgt, cs, Je);
    jac_intr(o, wgt, ci, Ji);
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      fma2(acc[a], Je[a], u[0], Je[6 + a], u[1]);
      fma2(acc[6 + a], Ji[a], u[0], Ji[6 + a], u[1]);
    }
  };
  const int end = chunk_end[ch];
  int i = chunk_begin[ch] + threadIdx.x;
  for (; i + kBlock < end; i += 2 * kBlock) {
    const int e0 = cm_e[i], e1 = cm_e[i + kBlock];
    const T o0x = cm_obs[2 * (size_t)i], o0y = cm_obs[2 * (size_t)i + 1];
    const T o1x = cm_obs[2 * (size_t)(i + kBlock)], o1y = cm_obs[2 * (size_t)(i + kBlock) + 1];
    T X0[3], s0[3], X1[3], s1[3];
    load_pk(pk, e0, X0, s0);
    load_pk(pk, e1, X1, s1);
    row(X0, s0, o0x, o0y);
    row(X1, s1, o1x, o1y);
  }
  if (i < end) {
    T X0[3], s0[3];
    load_pk(pk, cm_e[i], X0, s0);
    row(X0, s0, cm_obs[2 * (size_t)i], cm_obs[2 * (size_t)i + 1]);
  }
  block_sum<T, 12>(acc, lds);
  const int k = threadIdx.x;
  if (k < 18) {
    const int part = k / 9, a = k % 9;
    T s = T(0);
    if (a < 6)
      for (int w = 0; w < kBlock / 64; ++w) s += lds[w * 12 + 6 * part + a];
    partial[(size_t)kPartial * (part ? (size_t)nchunks + ch : (size_t)ch) + a] = s;
  }
}

// ------------------------------------------------------------ explicit S
// Each observation row r has two virtual rows of S: its camera's (v = r)
// and its group's (v = nrows + r), with W_v = J_v' J_p (9-slot rows, kWY
// layout of the packed kernels) and Y_v = W_v Vinv; the packed block
// assembly then runs over virtual-row pairs.  The camera-group block of F'F
// (same observation, two F blocks) comes from k_cross_g / k_cross_final.
template <typename T>
__device__ __forceinline__ void write_wy(const T J[12], const T Jp[6], const T Vi[9], T* __restrict__ out) {
  T w[18];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) w[3 * i + j] = J[i] * Jp[j] + J[6 + i] * Jp[3 + j];
#pragma unroll
  for (int k = 0; k < 18; ++k) out[k] = w[k];
#pragma unroll
  for (int k = 18; k < 27; ++k) out[k] = T(0);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      out[kWYY + 3 * i + j] = w[3 * i] * Vi[j] + w[3 * i + 1] * Vi[3 + j] + w[3 * i + 2] * Vi[6 + j];
#pragma unroll
  for (int k = 18; k < 27; ++k) out[kWYY + k] = T(0);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_s_rows_g(
    Seg sg, const int32_t* __restrict__ pt_start, const int32_t* __restrict__ row_cam,
    const int32_t* __restrict__ cam_to_f, const int32_t* __restrict__ cam_gf, const T* __restrict__ row_obs,
    const T* __restrict__ pkX, const T* __restrict__ pscale, const T* __restrict__ cscale,
    const T* __restrict__ ctab, const T* __restrict__ Vinv, T huber_a, int nrows, T* __restrict__ WY) {
  const SegPos q = seg_locate(sg);
  if (!q.valid) return;
  const int e = q.e;
  T X[3], unused[3];
  load_pk(pkX, e, X, unused);
  const T sp[3] = {pscale[3 * (size_t)e], pscale[3 * (size_t)e + 1], pscale[3 * (size_t)e + 2]};
  const T* iv = Vinv + 6 * (size_t)e;
  const T Vi[9] = {iv[0], iv[1], iv[2], iv[1], iv[3], iv[4], iv[2], iv[4], iv[5]};
  for (int r = pt_start[e] + q.sub; r < pt_start[e + 1]; r += q.lanes) {
    const int cam = row_cam[r];
    const int f = cam_to_f[cam], g = cam_gf[cam];
    if (f < 0 && g < 0) continue;
    const T* tab = ctab + (size_t)kCamTab * cam;
    CamProj<T> c;
    load_proj<T, 1>(tab, c);
    Proj<T> o;
    project<T, 1>(c, X, row_obs[2 * (size_t)r], row_obs[2 * (size_t)r + 1], o);
    const T wgt = huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T Jp[6], J[12], cs[6];
    jac_point(o, c.R, wgt, sp, Jp);
    if (f >= 0) {
      for (int k = 0; k < 6; ++k) cs[k] = cscale[9 * (size_t)f + k];
      jac_ext(o, tab + kTabDR, X, wgt, cs, J);
      write_wy(J, Jp, Vi, WY + (size_t)kWY * r);
    }
    if (g >= 0) {
      for (int k = 0; k < 6; ++k) cs[k] = cscale[9 * (size_t)g + k];
      jac_intr(o, wgt, cs, J);
      write_wy(J, Jp, Vi, WY + (size_t)kWY * ((size_t)nrows + r));
    }
  }
}

// Merged group rows: row 2 nrows + m of WY = the sum of the group parts
// (rows nrows + r) listed in m_row[m_start[m] .. m_start[m + 1]), in order.
// One wave per merged row, lane k sums entry k.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_merge_rows(int64_t nm, const int32_t* __restrict__ m_start,
                                                      const int32_t* __restrict__ m_row, int nrows,
                                                      T* __restrict__ WY) {
  const int64_t m = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int k = threadIdx.x & 63;
  if (m >= nm || k >= kWY) return;
  T v = T(0);
  for (int q = m_start[m]; q < m_start[m + 1]; ++q) v += WY[(size_t)kWY * ((size_t)nrows + m_row[q]) + k];
  WY[(size_t)kWY * (2 * (size_t)nrows + m) + k] = v;
}

// Per camera-major chunk: sum over its rows of J_e' J_i (6 x 6, unscaled).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cross_g(
    const int32_t* __restrict__ chunk_cam, const int32_t* __restrict__ chunk_begin,
    const int32_t* __restrict__ chunk_end, const int32_t* __restrict__ cm_e, const T* __restrict__ cm_obs,
    const T* __restrict__ pk, const T* __restrict__ ctab, T huber_a, T* __restrict__ partial) {
  __shared__ T lds[36 * (kBlock / 64)];
  const int ch = blockIdx.x;
  const T* tab = ctab + (size_t)kCamTab * chunk_cam[ch];
  CamProj<T> c;
  load_proj<T, 1>(tab, c);
  T dR[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) dR[i] = tab[kTabDR + i];
  const T one[6] = {T(1), T(1), T(1), T(1), T(1), T(1)};
  T acc[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) acc[i] = T(0);
  for (int i = chunk_begin[ch] + threadIdx.x; i < chunk_end[ch]; i += kBlock) {
    T X[3], unused[3];
    load_pk(pk, cm_e[i], X, unused);
    Proj<T> o;
    project<T, 1>(c, X, cm_obs[2 * (size_t)i], cm_obs[2 * (size_t)i + 1], o);
    const T wgt = huber_weight(o.rx * o.rx + o.ry * o.ry, huber_a);
    T Je[12], Ji[12];
    jac_ext(o, dR, X, wgt, one, Je);
    jac_intr(o, wgt, one, Ji);
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b) fma2(acc[6 * a + b], Je[a], Ji[b], Je[6 + a], Ji[6 + b]);
  }
  block_sum<T, 36>(acc, lds);
  if (threadIdx.x < 36) {
    T s = T(0);
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w * 36 + threadIdx.x];
    partial[(size_t)kPartial * ch + threadIdx.x] = s;
  }
}

// cross[36 c + 6 a + b] = s_e[a] s_i[b] (sum of camera c's chunk partials)
// for cameras whose block and group block are both free.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cross_final(int nc, const int32_t* __restrict__ cam_to_f,
                                                        const int32_t* __restrict__ cam_gf,
                                                        const int32_t* __restrict__ f_lo,
                                                        const int32_t* __restrict__ f_hi,
                                                        const T* __restrict__ partial, const T* __restrict__ cscale,
                                                        T* __restrict__ cross) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int c = (int)(t / 36), k = (int)(t % 36);
  if (c >= nc) return;
  const int f = cam_to_f[c], g = cam_gf[c];
  if (f < 0 || g < 0) return;
  const T s = sum_partials<T>(partial, f_lo[f], f_hi[f], k);
  cross[36 * (size_t)c + k] = s * cscale[9 * (size_t)f + k / 6] * cscale[9 * (size_t)g + k % 6];
}

// Zero the kernel-precision scale of slots outside a block's local
// parameterisation (unused slots and SubsetParameterization constants), so
// their Jacobian columns vanish; mask[i] is 1 for a free slot.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_mask(int64_t n, const uint8_t* __restrict__ mask,
                                                       T* __restrict__ scale) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && !mask[i]) scale[i] = T(0);
}

// CostFunction::Evaluate of ErrorFunc_Refine_Camera_3DPoints per
// observation in input order (no loss, no scaling): residual, row-major 2x6
// intrinsic, 2x6 extrinsic and 2x3 point Jacobians.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_eval_obs_g(int64_t n, const int32_t* __restrict__ cam_idx,
                                                       const int32_t* __restrict__ pt_idx,
                                                       const T* __restrict__ obs, const T* __restrict__ ctab,
                                                       const T* __restrict__ pts, double* __restrict__ res,
                                                       double* __restrict__ ji, double* __restrict__ jc,
                                                       double* __restrict__ jp) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* tab = ctab + (size_t)kCamTab * cam_idx[i];
  CamProj<T> c;
  load_proj<T, 1>(tab, c);
  const int q = pt_idx[i];
  const T X[3] = {pts[3 * (size_t)q], pts[3 * (size_t)q + 1], pts[3 * (size_t)q + 2]};
  Proj<T> o;
  project<T, 1>(c, X, obs[2 * i], obs[2 * i + 1], o);
  if (res) {
    res[2 * i] = (double)o.rx;
    res[2 * i + 1] = (double)o.ry;
  }
  const T one[6] = {T(1), T(1), T(1), T(1), T(1), T(1)};
  T J[12];
  if (ji) {
    jac_intr(o, T(1), one, J);
    for (int k = 0; k < 12; ++k) ji[12 * i + k] = (double)J[k];
  }
  if (jc) {
    jac_ext(o, tab + kTabDR, X, T(1), one, J);
    for (int k = 0; k < 12; ++k) jc[12 * i + k] = (double)J[k];
  }
  if (jp) {
    T Jp[6];
    jac_point(o, c.R, T(1), one, Jp);
    for (int k = 0; k < 6; ++k) jp[6 * i + k] = (double)Jp[k];
  }
}

}  // namespace mba

#endif  // MBA_DEVICE_GROUPED_H_
