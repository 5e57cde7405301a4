// Launch shapes of the point-major and camera-major kernels by problem size,
// as chosen on MI355X by 3dreconstruction_amd/autotune.py (the sweep and
// its timings are recorded in profiles/autotune_mi355x.json).  Used when
// mba_options leaves point_rows_per_lane / camera_chunk_rows at 0.
#ifndef MBA_TUNING_H_
#define MBA_TUNING_H_

#include <cstdint>

namespace mba {

struct TunedShape {
  int64_t max_rows;             // applies to problems with at most this many observations
  int32_t point_rows_per_lane;  // rows per lane of the point-major kernels
  int32_t camera_chunk_rows;    // rows per camera-major chunk (0: ~2048 chunks of 64..2048 rows)
};

constexpr TunedShape kTunedShapes[] = {
    {int64_t(1) << 18, 4, 0},
    {int64_t(2097152), 4, 1024},  // adjoint_check_697k
    {int64_t(8388608), 8, 1536},  // mixed_precision_4.0M
    {INT64_MAX, 4, 1536},  // large_scene_25M
};

inline TunedShape tuned_shape(int64_t rows) {
  for (const TunedShape& t : kTunedShapes)
    if (rows <= t.max_rows) return t;
  return kTunedShapes[sizeof(kTunedShapes) / sizeof(kTunedShapes[0]) - 1];
}

}  // namespace mba

#endif  // MBA_TUNING_H_
