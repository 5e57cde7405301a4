// Host-side data formats and boundary conversions of the bundle-adjustment
// library (no device code):
//   * rotation matrix <-> angle-axis, the conversions the incremental engine
//     applies on entry and exit (sfm_incremental_engine.cpp:1306,1505;
//     include/ceres/rotation.h:357-415);
//   * the libmv bundle-adjustment problem files
//     (3rdparty/ceres/examples/libmv_bundle_adjuster.cc:255-437, real
//     camera-refinement problems under 3rdparty/ceres/data/libmv-ba-problems);
//   * BAL text problems (3rdparty/ceres/examples/bal_problem.cc:72-175).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mba.h"

namespace mba {
// Shared with the engine's translation unit (thread-local error text).
void set_last_error(const std::string& msg);
}  // namespace mba

namespace {

mba_status fail(mba_status s, const std::string& msg) {
  mba::set_last_error(msg);
  return s;
}

// Rotation matrix (column-major) to a unit quaternion [w, x, y, z]: the
// trace form when the trace is non-negative, otherwise the form pivoted on
// the largest diagonal entry, so no division by a small number
// (Shepperd's method; the branch structure of rotation.h:321-351).
void matrix_to_quaternion(const double* Rc, double q[4]) {
  auto R = [&](int r, int c) { return Rc[r + 3 * c]; };
  const double tr = R(0, 0) + R(1, 1) + R(2, 2);
  if (tr >= 0.0) {
    const double s = std::sqrt(tr + 1.0);
    q[0] = 0.5 * s;
    const double h = 0.5 / s;
    q[1] = (R(2, 1) - R(1, 2)) * h;
    q[2] = (R(0, 2) - R(2, 0)) * h;
    q[3] = (R(1, 0) - R(0, 1)) * h;
    return;
  }
  int i = R(1, 1) > R(0, 0) ? 1 : 0;
  if (R(2, 2) > R(i, i)) i = 2;
  const int j = (i + 1) % 3, k = (j + 1) % 3;
  const double s = std::sqrt(R(i, i) - R(j, j) - R(k, k) + 1.0);
  q[i + 1] = 0.5 * s;
  const double h = 0.5 / s;
  q[0] = (R(k, j) - R(j, k)) * h;
  q[j + 1] = (R(j, i) + R(i, j)) * h;
  q[k + 1] = (R(k, i) + R(i, k)) * h;
}

// Unit quaternion to angle-axis with the angle wrapped into [-pi, pi]
// (rotation.h:263-303): angle 2 atan2(|v|, w), taking the shorter way round
// when w < 0; first order (2 v) for a zero vector part.
void quaternion_to_angle_axis(const double q[4], double aa[3]) {
  const double s2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  if (s2 > 0.0) {
    const double s = std::sqrt(s2), c = q[0];
    const double two_theta = 2.0 * (c < 0.0 ? std::atan2(-s, -c) : std::atan2(s, c));
    const double k = two_theta / s;
    for (int i = 0; i < 3; ++i) aa[i] = q[i + 1] * k;
  } else {
    for (int i = 0; i < 3; ++i) aa[i] = 2.0 * q[i + 1];
  }
}

// Rodrigues' formula (rotation.h:368-404): R = c I + s [w]x + (1 - c) w w',
// first order I + [aa]x when theta^2 <= DBL_EPSILON.  Column-major out.
void angle_axis_to_matrix(const double aa[3], double* Rc) {
  auto R = [&](int r, int c) -> double& { return Rc[r + 3 * c]; };
  const double th2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  if (th2 > 2.220446049250313e-16) {
    const double th = std::sqrt(th2), wx = aa[0] / th, wy = aa[1] / th, wz = aa[2] / th;
    const double c = std::cos(th), s = std::sin(th), oc = 1.0 - c;
    R(0, 0) = c + wx * wx * oc;
    R(1, 0) = wz * s + wx * wy * oc;
    R(2, 0) = -wy * s + wx * wz * oc;
    R(0, 1) = wx * wy * oc - wz * s;
    R(1, 1) = c + wy * wy * oc;
    R(2, 1) = wx * s + wy * wz * oc;
    R(0, 2) = wy * s + wx * wz * oc;
    R(1, 2) = -wx * s + wy * wz * oc;
    R(2, 2) = c + wz * wz * oc;
  } else {
    R(0, 0) = 1.0;     R(0, 1) = -aa[2]; R(0, 2) = aa[1];
    R(1, 0) = aa[2];   R(1, 1) = 1.0;    R(1, 2) = -aa[0];
    R(2, 0) = -aa[1];  R(2, 1) = aa[0];  R(2, 2) = 1.0;
  }
}

// Reads the libmv problem format: a byte 'v' (little endian) or 'V' (big),
// a byte 'P' (markers in pixels) or 'N' (normalised), 8 float intrinsics,
// then counted lists of cameras (int image, 9 float column-major R, 3 float
// t), points (int track, 3 float X) and markers (int image, int track,
// float x, float y).
struct LibmvFile {
  bool image_space = false;
  float intr[8];
  struct Cam { int32_t image; float R[9], t[3]; };
  struct Pt { int32_t track; float X[3]; };
  struct Marker { int32_t image, track; float x, y; };
  std::vector<Cam> cams;
  std::vector<Pt> pts;
  std::vector<Marker> markers;
};

class Reader {
 public:
  explicit Reader(FILE* f) : f_(f) {}
  bool ok() const { return ok_; }
  void set_swap(bool s) { swap_ = s; }
  uint8_t byte() {
    uint8_t b = 0;
    ok_ = ok_ && std::fread(&b, 1, 1, f_) == 1;
    return b;
  }
  uint32_t word() {
    uint8_t b[4] = {0, 0, 0, 0};
    ok_ = ok_ && std::fread(b, 1, 4, f_) == 4;
    if (swap_) std::swap(b[0], b[3]), std::swap(b[1], b[2]);
    uint32_t v;
    std::memcpy(&v, b, 4);
    return v;
  }
  int32_t i32() {
    const uint32_t w = word();
    int32_t v;
    std::memcpy(&v, &w, 4);
    return v;
  }
  float f32() {
    const uint32_t w = word();
    float v;
    std::memcpy(&v, &w, 4);
    return v;
  }

 private:
  FILE* f_;
  bool ok_ = true, swap_ = false;
};

mba_status read_libmv_file(const char* path, LibmvFile* out) {
  if (!path) return fail(MBA_ERR_INVALID_ARGUMENT, "path is NULL");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(MBA_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  Reader r(f);
  const uint8_t endian = r.byte();
  const uint32_t probe = 1;
  const bool host_little = *reinterpret_cast<const uint8_t*>(&probe) == 1;
  mba_status st = MBA_OK;
  std::string err;
  if (endian != 'v' && endian != 'V') {
    st = MBA_ERR_INVALID_ARGUMENT;
    err = "unknown endian flag";
  } else {
    r.set_swap((endian == 'v') != host_little);
    const uint8_t space = r.byte();
    if (space != 'P' && space != 'N') {
      st = MBA_ERR_INVALID_ARGUMENT;
      err = "unknown marker space flag";
    } else {
      out->image_space = space == 'P';
      for (float& v : out->intr) v = r.f32();
      const int32_t nc = r.i32();
      if (!r.ok() || nc < 0 || nc > (1 << 26)) {
        st = MBA_ERR_INVALID_ARGUMENT;
        err = "bad camera count";
      }
      for (int32_t i = 0; st == MBA_OK && i < nc && r.ok(); ++i) {
        LibmvFile::Cam c;
        c.image = r.i32();
        for (float& v : c.R) v = r.f32();
        for (float& v : c.t) v = r.f32();
        out->cams.push_back(c);
      }
      const int32_t np = st == MBA_OK ? r.i32() : 0;
      if (st == MBA_OK && (!r.ok() || np < 0 || np > (1 << 28))) {
        st = MBA_ERR_INVALID_ARGUMENT;
        err = "bad point count";
      }
      for (int32_t i = 0; st == MBA_OK && i < np && r.ok(); ++i) {
        LibmvFile::Pt p;
        p.track = r.i32();
        for (float& v : p.X) v = r.f32();
        out->pts.push_back(p);
      }
      const int32_t nm = st == MBA_OK ? r.i32() : 0;
      if (st == MBA_OK && (!r.ok() || nm < 0)) {
        st = MBA_ERR_INVALID_ARGUMENT;
        err = "bad marker count";
      }
      for (int32_t i = 0; st == MBA_OK && i < nm && r.ok(); ++i) {
        LibmvFile::Marker m;
        m.image = r.i32();
        m.track = r.i32();
        m.x = r.f32();
        m.y = r.f32();
        out->markers.push_back(m);
      }
      if (st == MBA_OK && !r.ok()) {
        st = MBA_ERR_INVALID_ARGUMENT;
        err = "file ends early";
      }
    }
  }
  std::fclose(f);
  return st == MBA_OK ? MBA_OK : fail(st, std::string(path) + ": " + err);
}

}  // namespace

extern "C" {

mba_status mba_rotation_matrix_to_angle_axis(const double* R, double* angle_axis, int64_t n) {
  if (n < 0 || (n > 0 && (!R || !angle_axis))) return fail(MBA_ERR_INVALID_ARGUMENT, "bad arguments");
  for (int64_t i = 0; i < n; ++i) {
    double q[4];
    matrix_to_quaternion(R + 9 * i, q);
    quaternion_to_angle_axis(q, angle_axis + 3 * i);
  }
  return MBA_OK;
}

mba_status mba_angle_axis_to_rotation_matrix(const double* angle_axis, double* R, int64_t n) {
  if (n < 0 || (n > 0 && (!R || !angle_axis))) return fail(MBA_ERR_INVALID_ARGUMENT, "bad arguments");
  for (int64_t i = 0; i < n; ++i) angle_axis_to_matrix(angle_axis + 3 * i, R + 9 * i);
  return MBA_OK;
}

mba_status mba_read_libmv(const char* path, int32_t* num_cameras, int32_t* num_points,
                          int64_t* num_observations, int32_t* image_space, double* intrinsics,
                          double* cameras, double* points, double* observations, int32_t* camera_index,
                          int32_t* point_index, int32_t* image_of_camera, int32_t* track_of_point) {
  LibmvFile lf;
  const mba_status st = read_libmv_file(path, &lf);
  if (st != MBA_OK) return st;
  if (lf.intr[6] != 0.0f || lf.intr[7] != 0.0f)
    return fail(MBA_ERR_UNSUPPORTED, "tangential distortion (p1, p2) is not part of the libs/sfm camera model");
  // Cameras by image and points by track (later entries of a repeated id
  // replace earlier ones, as the reference's resize-and-assign does); dense
  // indices in ascending id order.
  std::map<int32_t, size_t> cam_of, pt_of;
  for (size_t i = 0; i < lf.cams.size(); ++i)
    if (lf.cams[i].image >= 0) cam_of[lf.cams[i].image] = i;
  for (size_t i = 0; i < lf.pts.size(); ++i)
    if (lf.pts[i].track >= 0) pt_of[lf.pts[i].track] = i;
  std::map<int32_t, int32_t> cam_idx, pt_idx;
  for (const auto& kv : cam_of) cam_idx.emplace(kv.first, (int32_t)cam_idx.size());
  for (const auto& kv : pt_of) pt_idx.emplace(kv.first, (int32_t)pt_idx.size());
  int64_t nobs = 0;
  for (const auto& m : lf.markers)
    if (cam_idx.count(m.image) && pt_idx.count(m.track)) ++nobs;
  if (num_cameras) *num_cameras = (int32_t)cam_idx.size();
  if (num_points) *num_points = (int32_t)pt_idx.size();
  if (num_observations) *num_observations = nobs;
  if (image_space) *image_space = lf.image_space ? 1 : 0;
  if (intrinsics) {
    // [f, ppx, ppy, k1, k2, k3] of the grouped layout (the file's fx, cx,
    // cy, k1, k2, k3; p1 = p2 = 0 checked above).
    for (int k = 0; k < 6; ++k) intrinsics[k] = (double)lf.intr[k];
  }
  if (cameras || image_of_camera) {
    for (const auto& kv : cam_idx) {
      const LibmvFile::Cam& c = lf.cams[cam_of[kv.first]];
      if (image_of_camera) image_of_camera[kv.second] = kv.first;
      if (cameras) {
        double R[9];
        for (int k = 0; k < 9; ++k) R[k] = (double)c.R[k];
        double q[4];
        matrix_to_quaternion(R, q);
        quaternion_to_angle_axis(q, cameras + 6 * (size_t)kv.second);
        for (int k = 0; k < 3; ++k) cameras[6 * (size_t)kv.second + 3 + k] = (double)c.t[k];
      }
    }
  }
  if (points || track_of_point)
    for (const auto& kv : pt_idx) {
      if (track_of_point) track_of_point[kv.second] = kv.first;
      if (points)
        for (int k = 0; k < 3; ++k) points[3 * (size_t)kv.second + k] = (double)lf.pts[pt_of[kv.first]].X[k];
    }
  if (observations || camera_index || point_index) {
    int64_t o = 0;
    for (const auto& m : lf.markers) {
      const auto ci = cam_idx.find(m.image);
      const auto pi = pt_idx.find(m.track);
      if (ci == cam_idx.end() || pi == pt_idx.end()) continue;
      if (observations) {
        observations[2 * o] = (double)m.x;
        observations[2 * o + 1] = (double)m.y;
      }
      if (camera_index) camera_index[o] = ci->second;
      if (point_index) point_index[o] = pi->second;
      ++o;
    }
  }
  return MBA_OK;
}

// BAL text (bal_problem.cc:72-175): "Nc Np Nobs", Nobs lines "camera
// point x y", then 9 values per camera [aa, t, f, k1, k2] and 3 per point,
// whitespace separated.  BAL cameras look down -z (Snavely's model,
// examples/snavely_reprojection_error.h:70-90: p = -P / P.z); the libs/sfm
// model projects P / P.z, so the same camera vector predicts the negated
// image point.  Observations are negated on reading and on writing: the
// problem keeps its cost and normal equations (residuals change sign).
mba_status mba_read_bal(const char* path, int32_t* num_cameras, int32_t* num_points, int64_t* num_observations,
                        double* cameras, double* points, double* observations, int32_t* camera_index,
                        int32_t* point_index) {
  if (!path) return fail(MBA_ERR_INVALID_ARGUMENT, "path is NULL");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(MBA_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  int nc = 0, np = 0;
  long long no = 0;
  std::string err;
  if (std::fscanf(f, "%d %d %lld", &nc, &np, &no) != 3 || nc < 0 || np < 0 || no < 0) err = "bad header";
  if (err.empty()) {
    if (num_cameras) *num_cameras = nc;
    if (num_points) *num_points = np;
    if (num_observations) *num_observations = no;
    const bool want = cameras && points && observations && camera_index && point_index;
    for (long long i = 0; want && err.empty() && i < no; ++i) {
      int c, q;
      double x, y;
      if (std::fscanf(f, "%d %d %lf %lf", &c, &q, &x, &y) != 4) err = "observation " + std::to_string(i);
      else if (c < 0 || c >= nc || q < 0 || q >= np) err = "observation " + std::to_string(i) + " index out of range";
      camera_index[i] = c;
      point_index[i] = q;
      observations[2 * i] = -x;
      observations[2 * i + 1] = -y;
    }
    for (long long i = 0; want && err.empty() && i < 9LL * nc; ++i)
      if (std::fscanf(f, "%lf", cameras + i) != 1) err = "camera values";
    for (long long i = 0; want && err.empty() && i < 3LL * np; ++i)
      if (std::fscanf(f, "%lf", points + i) != 1) err = "point values";
  }
  std::fclose(f);
  return err.empty() ? MBA_OK : fail(MBA_ERR_INVALID_ARGUMENT, std::string(path) + ": " + err);
}

// Same text as BALProblem::WriteToFile (bal_problem.cc:136-175): header,
// observations with %g, parameters one per line with %.16g.
mba_status mba_write_bal(const char* path, const mba_problem* p) {
  if (!path || !p) return fail(MBA_ERR_INVALID_ARGUMENT, "NULL argument");
  if (p->layout != MBA_LAYOUT_PACKED9) return fail(MBA_ERR_UNSUPPORTED, "BAL files hold packed cameras");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(MBA_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
  std::fprintf(f, "%d %d %lld\n", p->num_cameras, p->num_points, (long long)p->num_observations);
  for (int64_t i = 0; i < p->num_observations; ++i)
    std::fprintf(f, "%d %d %g %g\n", p->camera_index[i], p->point_index[i], -p->observations[2 * i],
                 -p->observations[2 * i + 1]);
  for (int64_t i = 0; i < 9LL * p->num_cameras; ++i) std::fprintf(f, "%.16g\n", p->cameras[i]);
  for (int64_t i = 0; i < 3LL * p->num_points; ++i) std::fprintf(f, "%.16g\n", p->points[i]);
  const bool ok = std::fclose(f) == 0;
  return ok ? MBA_OK : fail(MBA_ERR_INVALID_ARGUMENT, std::string("error writing ") + path);
}

}  // extern "C"
