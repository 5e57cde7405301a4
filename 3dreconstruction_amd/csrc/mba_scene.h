/*
 * Seeded synthetic bundle-adjustment scenes.
 *
 * Header-only C (usable from C, C++ and HIP host code).  Every random draw
 * comes from a counter-based splitmix64 stream keyed by (seed, entity), so a
 * scene is bit-identical whatever order or thread count fills it.
 *
 * Camera model: the libs/sfm angle-axis pinhole with polynomial radial
 * distortion (libs/sfm/include/mvg/sfm/pinhole_brown_rt_ceres_functor.h:60-107,
 * libs/camera/include/mvg/camera/brown_pinhole_camera.h:14-41) packed as nine
 * parameters [aa(3), t(3), f, k1, k2] with principal point 0 and k3 = 0.
 *
 * Observation order is the order a libs/sfm engine hands residual blocks to
 * Problem::AddResidualBlock: point by point in point-id order, and within a
 * point by ascending camera id (std::map track order,
 * libs/sfm/src/sfm_incremental_engine.cpp:1357-1388).
 *
 * Distributions:
 *   aa ~ N(0, 0.1^2) per component, t ~ N(0, 1), f ~ U(450, 550),
 *   k1 ~ N(0, 0.01^2), k2 ~ N(0, 0.001^2);
 *   points uniform in [-2,2] x [-2,2] x [4,10];
 *   a (camera, point) pair whose depth is <= 0.5, or that falls outside a
 *   90-degree field of view (|x/z| > 1 or |y/z| > 1: an image of +-f
 *   pixels), is redrawn -- a camera records only points inside its image;
 *   pixel noise N(0, noise^2) per coordinate; a fraction of observations is
 *   shifted by outlier_px in a uniformly random direction;
 *   the initial guess multiplies every parameter by (1 + N(0, perturb^2)).
 *
 * Grouped-intrinsics variant (the libs/sfm incremental engine's layout,
 * libs/sfm/src/sfm_incremental_engine.cpp:1273-1393): cameras are [aa(3),
 * t(3)] (the same draws as the packed scene of the same seed), camera c uses
 * intrinsic group c mod G, a group is [f, ppx, ppy, k1, k2, k3] with
 * f ~ U(450, 550), ppx, ppy ~ N(0, 10^2), k1 ~ N(0, 0.01^2), k2 ~ N(0,
 * 0.001^2), k3 ~ N(0, 0.0001^2); the full Brown model projects the points.
 */
#ifndef MBA_SCENE_H_
#define MBA_SCENE_H_

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MBA_TRACK_UNIFORM = 0, MBA_TRACK_POWER_LAW = 1 };

typedef struct mba_scene_spec {
  int32_t num_cameras;
  int32_t num_points;
  int32_t track_kind;   /* MBA_TRACK_UNIFORM or MBA_TRACK_POWER_LAW */
  int32_t track_min;    /* shortest track (>= 1) */
  int32_t track_max;    /* longest track; clamped to num_cameras */
  int32_t pad_;
  double alpha;         /* power-law exponent, p(t) ~ t^-alpha */
  double noise_px;      /* pixel noise sigma */
  double outlier_frac;  /* fraction of observations shifted */
  double outlier_px;    /* shift length of an outlier, pixels */
  double perturb;       /* relative sigma of the initial-guess perturbation */
  uint64_t seed;
} mba_scene_spec;

/* ---------------------------------------------------------------- RNG --- */
typedef struct { uint64_t s; } mba_rng;

static inline uint64_t mba_rng_next(mba_rng* r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Stream for (seed, domain, index): hashes the key so neighbouring indices
 * give unrelated streams. */
static inline mba_rng mba_rng_stream(uint64_t seed, uint64_t domain,
                                     uint64_t index) {
  mba_rng r;
  r.s = seed * 0xD1B54A32D192ED03ULL ^ (domain << 56) ^ index;
  (void)mba_rng_next(&r);
  r.s ^= mba_rng_next(&r);
  return r;
}

/* Uniform in [0, 1) with 53 random bits. */
static inline double mba_uniform(mba_rng* r) {
  return (double)(mba_rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

/* Standard normal by Box-Muller (one value per call, the second discarded so
 * the stream position does not depend on call history). */
static inline double mba_normal(mba_rng* r) {
  double u1 = mba_uniform(r);
  double u2 = mba_uniform(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

/* ------------------------------------------------------- camera model --- */
/* Rodrigues rotation with the first-order branch for theta^2 <= DBL_EPSILON,
 * as ceres::AngleAxisRotatePoint (3rdparty/ceres/include/ceres/rotation.h:565-625). */
static inline void mba_scene_rotate(const double aa[3], const double p[3],
                                    double out[3]) {
  const double th2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  if (th2 > 2.220446049250313080847e-16) {
    const double th = sqrt(th2), c = cos(th), s = sin(th), it = 1.0 / th;
    const double w[3] = {aa[0] * it, aa[1] * it, aa[2] * it};
    const double wx[3] = {w[1] * p[2] - w[2] * p[1], w[2] * p[0] - w[0] * p[2],
                          w[0] * p[1] - w[1] * p[0]};
    const double tmp = (w[0] * p[0] + w[1] * p[1] + w[2] * p[2]) * (1.0 - c);
    for (int k = 0; k < 3; ++k) out[k] = p[k] * c + wx[k] * s + w[k] * tmp;
  } else {
    out[0] = p[0] + aa[1] * p[2] - aa[2] * p[1];
    out[1] = p[1] + aa[2] * p[0] - aa[0] * p[2];
    out[2] = p[2] + aa[0] * p[1] - aa[1] * p[0];
  }
}

/* Projection of a point by a packed 9-parameter camera; returns the depth
 * (or -1 when the point lies outside the 90-degree field of view). */
static inline double mba_scene_project(const double cam[9], const double X[3],
                                       double uv[2]) {
  double x[3];
  mba_scene_rotate(cam, X, x);
  x[0] += cam[3];
  x[1] += cam[4];
  x[2] += cam[5];
  const double xe = x[0] / x[2], ye = x[1] / x[2];
  const double r2 = xe * xe + ye * ye;
  const double d = 1.0 + cam[7] * r2 + cam[8] * r2 * r2;
  uv[0] = cam[6] * xe * d;
  uv[1] = cam[6] * ye * d;
  if (fabs(xe) > 1.0 || fabs(ye) > 1.0) return -1.0;
  return x[2];
}

/* ------------------------------------------------------ track lengths --- */
enum { MBA_DOM_CAM = 1, MBA_DOM_PT = 2, MBA_DOM_TRACK = 3, MBA_DOM_GROUP = 4 };

/* Projection with the full libs/sfm Brown model: camera [aa(3), t(3)],
 * intrinsics [f, ppx, ppy, k1, k2, k3]; returns the depth (or -1 outside the
 * 90-degree field of view). */
static inline double mba_scene_project_brown(const double cam[6], const double intr[6], const double X[3],
                                             double uv[2]) {
  double x[3];
  mba_scene_rotate(cam, X, x);
  x[0] += cam[3];
  x[1] += cam[4];
  x[2] += cam[5];
  const double xe = x[0] / x[2], ye = x[1] / x[2];
  const double r2 = xe * xe + ye * ye;
  const double d = 1.0 + intr[3] * r2 + intr[4] * r2 * r2 + intr[5] * r2 * r2 * r2;
  uv[0] = intr[1] + intr[0] * xe * d;
  uv[1] = intr[2] + intr[0] * ye * d;
  if (fabs(xe) > 1.0 || fabs(ye) > 1.0) return -1.0;
  return x[2];
}

/* Camera models the fill core projects with. */
typedef struct {
  const double* cams;   /* 9 per camera (packed) or 6 (grouped) */
  const double* intr;   /* 6 per group, NULL for packed */
  int32_t groups;
} mba_scene_model;

static inline double mba_scene_project_model(const mba_scene_model* m, int32_t c, const double X[3],
                                             double uv[2]) {
  if (!m->intr) return mba_scene_project(m->cams + 9 * (size_t)c, X, uv);
  return mba_scene_project_brown(m->cams + 6 * (size_t)c, m->intr + 6 * (size_t)(c % m->groups), X, uv);
}

/* Track length of every point (first pass; O(num_points)).  A power law is
 * drawn by inverse CDF over the truncated discrete distribution. */
static inline int64_t mba_scene_count(const mba_scene_spec* sp,
                                      int32_t* track_len) {
  int lo = sp->track_min, hi = sp->track_max;
  if (hi > sp->num_cameras) hi = sp->num_cameras;
  if (lo > hi) lo = hi;
  double cdf[1024];
  int nb = hi - lo + 1;
  if (nb > 1024) nb = 1024;
  double z = 0.0;
  for (int i = 0; i < nb; ++i) {
    z += (sp->track_kind == MBA_TRACK_UNIFORM) ? 1.0 : pow((double)(lo + i), -sp->alpha);
    cdf[i] = z;
  }
  int64_t n = 0;
  for (int32_t p = 0; p < sp->num_points; ++p) {
    mba_rng r = mba_rng_stream(sp->seed, MBA_DOM_TRACK, (uint64_t)p);
    const double u = mba_uniform(&r) * z;
    int i = 0;
    while (i < nb - 1 && u >= cdf[i]) ++i;
    track_len[p] = lo + i;
    n += track_len[p];
  }
  return n;
}

/* Points, tracks and observations of a scene whose cameras are `m`.  The
 * draws depend only on (seed, point), never on the camera model's layout. */
static inline int64_t mba_scene_fill_points(const mba_scene_spec* sp, const int32_t* track_len,
                                            const mba_scene_model* m, double* pt_true, double* pt_init,
                                            double* obs, int32_t* cam_idx, int32_t* pt_idx,
                                            int64_t* obs_offset) {
  const int32_t nc = sp->num_cameras, np = sp->num_points;
  obs_offset[0] = 0;
  for (int32_t p = 0; p < np; ++p) obs_offset[p + 1] = obs_offset[p] + track_len[p];

  for (int32_t p = 0; p < np; ++p) {
    mba_rng r = mba_rng_stream(sp->seed, MBA_DOM_PT, (uint64_t)p);
    double X[3];
    X[0] = -2.0 + 4.0 * mba_uniform(&r);
    X[1] = -2.0 + 4.0 * mba_uniform(&r);
    X[2] = 4.0 + 6.0 * mba_uniform(&r);
    if (pt_true) memcpy(pt_true + 3 * (size_t)p, X, sizeof(X));
    for (int k = 0; k < 3; ++k)
      pt_init[3 * (size_t)p + k] = X[k] * (1.0 + sp->perturb * mba_normal(&r));
    const int t = track_len[p];
    const int64_t base = obs_offset[p];
    /* Distinct cameras with depth > 0.5, drawn by rejection. */
    int n = 0, guard = 0;
    while (n < t) {
      int32_t c = (int32_t)(mba_uniform(&r) * (double)nc);
      if (c >= nc) c = nc - 1;
      int dup = 0;
      for (int j = 0; j < n; ++j) dup |= (cam_idx[base + j] == c);
      double uv[2];
      const double depth = mba_scene_project_model(m, c, X, uv);
      if (!dup && (depth > 0.5 || ++guard > 1000)) cam_idx[base + n++] = c;
    }
    /* Ascending camera id within a track (insertion sort, t is small). */
    for (int i = 1; i < t; ++i) {
      int32_t v = cam_idx[base + i];
      int j = i - 1;
      while (j >= 0 && cam_idx[base + j] > v) {
        cam_idx[base + j + 1] = cam_idx[base + j];
        --j;
      }
      cam_idx[base + j + 1] = v;
    }
    for (int i = 0; i < t; ++i) {
      const int64_t o = base + i;
      double uv[2];
      mba_scene_project_model(m, cam_idx[o], X, uv);
      uv[0] += sp->noise_px * mba_normal(&r);
      uv[1] += sp->noise_px * mba_normal(&r);
      if (mba_uniform(&r) < sp->outlier_frac) {
        const double a = 6.283185307179586476925 * mba_uniform(&r);
        uv[0] += sp->outlier_px * cos(a);
        uv[1] += sp->outlier_px * sin(a);
      }
      obs[2 * o] = uv[0];
      obs[2 * o + 1] = uv[1];
      pt_idx[o] = p;
    }
  }
  return obs_offset[np];
}

/* Packed 9-parameter cameras (true and perturbed). */
static inline void mba_scene_cameras(const mba_scene_spec* sp, double* ct, double* cam_init) {
  for (int32_t c = 0; c < sp->num_cameras; ++c) {
    mba_rng r = mba_rng_stream(sp->seed, MBA_DOM_CAM, (uint64_t)c);
    double* cam = ct + 9 * (size_t)c;
    for (int k = 0; k < 3; ++k) cam[k] = 0.1 * mba_normal(&r);
    for (int k = 3; k < 6; ++k) cam[k] = mba_normal(&r);
    cam[6] = 450.0 + 100.0 * mba_uniform(&r);
    cam[7] = 0.01 * mba_normal(&r);
    cam[8] = 0.001 * mba_normal(&r);
    for (int k = 0; k < 9; ++k)
      cam_init[9 * (size_t)c + k] = cam[k] * (1.0 + sp->perturb * mba_normal(&r));
  }
}

/* Fill a scene.  Arrays (caller-owned):
 *   cam_true, cam_init : 9 * num_cameras
 *   pt_true,  pt_init  : 3 * num_points
 *   obs                : 2 * num_obs
 *   cam_idx, pt_idx    : num_obs
 *   obs_offset         : num_points + 1 (filled: CSR start of each point)
 *   track_len          : num_points (from mba_scene_count)
 * Any of cam_true / pt_true may be NULL.  Returns the observation count. */
static inline int64_t mba_scene_fill(const mba_scene_spec* sp,
                                     const int32_t* track_len,
                                     double* cam_true, double* cam_init,
                                     double* pt_true, double* pt_init,
                                     double* obs, int32_t* cam_idx,
                                     int32_t* pt_idx, int64_t* obs_offset) {
  const int32_t nc = sp->num_cameras;
  double* ct = cam_true ? cam_true : (double*)malloc(sizeof(double) * 9 * (size_t)nc);
  mba_scene_cameras(sp, ct, cam_init);
  const mba_scene_model m = {ct, NULL, 0};
  const int64_t n = mba_scene_fill_points(sp, track_len, &m, pt_true, pt_init, obs, cam_idx, pt_idx, obs_offset);
  if (!cam_true) free(ct);
  return n;
}

/* Grouped-intrinsics scene.  Arrays (caller-owned):
 *   cam_true, cam_init   : 6 * num_cameras ([aa, t])
 *   intr_true, intr_init : 6 * groups ([f, ppx, ppy, k1, k2, k3])
 *   intr_idx             : num_obs (group of the observing camera)
 *   the rest as in mba_scene_fill.  cam_true / intr_true / pt_true may be
 *   NULL.  Returns the observation count. */
static inline int64_t mba_scene_fill_groups(const mba_scene_spec* sp, int32_t groups,
                                            const int32_t* track_len, double* cam_true,
                                            double* cam_init, double* intr_true, double* intr_init,
                                            double* pt_true, double* pt_init, double* obs,
                                            int32_t* cam_idx, int32_t* intr_idx, int32_t* pt_idx,
                                            int64_t* obs_offset) {
  const int32_t nc = sp->num_cameras;
  double* c9 = (double*)malloc(sizeof(double) * 9 * (size_t)nc);
  double* c9i = (double*)malloc(sizeof(double) * 9 * (size_t)nc);
  mba_scene_cameras(sp, c9, c9i);
  double* ct = cam_true ? cam_true : (double*)malloc(sizeof(double) * 6 * (size_t)nc);
  for (int32_t c = 0; c < nc; ++c)
    for (int k = 0; k < 6; ++k) {
      ct[6 * (size_t)c + k] = c9[9 * (size_t)c + k];
      cam_init[6 * (size_t)c + k] = c9i[9 * (size_t)c + k];
    }
  double* it = intr_true ? intr_true : (double*)malloc(sizeof(double) * 6 * (size_t)groups);
  for (int32_t g = 0; g < groups; ++g) {
    mba_rng r = mba_rng_stream(sp->seed, MBA_DOM_GROUP, (uint64_t)g);
    double* q = it + 6 * (size_t)g;
    q[0] = 450.0 + 100.0 * mba_uniform(&r);
    q[1] = 10.0 * mba_normal(&r);
    q[2] = 10.0 * mba_normal(&r);
    q[3] = 0.01 * mba_normal(&r);
    q[4] = 0.001 * mba_normal(&r);
    q[5] = 0.0001 * mba_normal(&r);
    for (int k = 0; k < 6; ++k) intr_init[6 * (size_t)g + k] = q[k] * (1.0 + sp->perturb * mba_normal(&r));
  }
  const mba_scene_model m = {ct, it, groups};
  const int64_t n = mba_scene_fill_points(sp, track_len, &m, pt_true, pt_init, obs, cam_idx, pt_idx, obs_offset);
  for (int64_t o = 0; o < n; ++o) intr_idx[o] = cam_idx[o] % groups;
  free(c9);
  free(c9i);
  if (!cam_true) free(ct);
  if (!intr_true) free(it);
  return n;
}

/* The benchmark scenes (index 1..6; 4 is one window of the batched set,
 * selected by `window`; 6 is a grouped-intrinsics scene with
 * MBA_SCENE_GROUPS groups, built by mba_scene_fill_groups). */
#define MBA_SCENE_GROUPS 2
static inline mba_scene_spec mba_scene_preset(int config, int window) {
  mba_scene_spec s;
  memset(&s, 0, sizeof(s));
  s.noise_px = 0.5;
  s.perturb = 0.01;
  s.seed = (uint64_t)config;
  switch (config) {
    case 1:
      s.num_cameras = 16; s.num_points = 2000;
      s.track_kind = MBA_TRACK_UNIFORM; s.track_min = 2; s.track_max = 6;
      break;
    case 2:
      s.num_cameras = 200; s.num_points = 100000;
      s.track_kind = MBA_TRACK_POWER_LAW; s.track_min = 2; s.track_max = 30;
      s.alpha = 1.4727;
      break;
    case 3:
      s.num_cameras = 1000; s.num_points = 500000;
      s.track_kind = MBA_TRACK_POWER_LAW; s.track_min = 2; s.track_max = 50;
      s.alpha = 1.6036; s.outlier_frac = 0.05; s.outlier_px = 20.0;
      break;
    case 4: {
      /* Window sizes: cameras U{5..40}, points U{200..5000}, obs/pt U{2..8}. */
      mba_rng r = mba_rng_stream(4, 7, (uint64_t)window);
      s.num_cameras = 5 + (int32_t)(mba_uniform(&r) * 36.0);
      s.num_points = 200 + (int32_t)(mba_uniform(&r) * 4801.0);
      s.track_kind = MBA_TRACK_UNIFORM; s.track_min = 2; s.track_max = 8;
      s.seed = 4000003ULL + (uint64_t)window;
      break;
    }
    case 5:
      s.num_cameras = 10000; s.num_points = 4000000;
      s.track_kind = MBA_TRACK_POWER_LAW; s.track_min = 2; s.track_max = 100;
      s.alpha = 2.0482;
      break;
    case 6:
      /* Incremental-engine scale, grouped intrinsics (MBA_SCENE_GROUPS
       * groups): a few hundred images of one or two camera models. */
      s.num_cameras = 300; s.num_points = 100000;
      s.track_kind = MBA_TRACK_UNIFORM; s.track_min = 2; s.track_max = 10;
      s.outlier_frac = 0.05; s.outlier_px = 20.0;
      break;
    default:
      break;
  }
  return s;
}

#ifdef __cplusplus
}
#endif
#endif /* MBA_SCENE_H_ */
