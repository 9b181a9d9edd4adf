// sw2d_launch.hpp -- host-side launchers of the stencil kernels, one set per
// state precision (sw2d_f64.hip, sw2d_f32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sw2d_device.hpp"

namespace sw2d {

// default rows per strip segment: the segment length that gives about
// TARGET_ITEMS work items (8 workgroups per CU), at least 8 rows
constexpr int TARGET_ITEMS = 2048;
inline int default_seg_rows(int rows, int strips) {
  long long s = ((long long)rows * strips + TARGET_ITEMS - 1) / TARGET_ITEMS;
  s = (s + 3) / 4 * 4;
  return (int)(s < 8 ? 8 : (s > rows ? ((rows + 3) / 4 * 4) : s));
}

struct IterLaunch {
  const void *grids;   // device Grid<T> array
  const Tile *work;    // tile work-list, or nullptr: all THMAX-row tiles of grid 0
  int n_work;          // work items (tiles)
  int blocks;          // workgroups launched (<= n_work; persistent loop)
  Ctrl *const *ctrl;   // device array of control-block pointers
  int n_ctrl;
  const hipEvent_t *ev;  // optional: recorded after control, momentum, continuity
};

hipError_t launch_control(Ctrl *const *ctrl, int n, int commit_only, hipStream_t s);
hipError_t launch_status(Ctrl *const *ctrl, int n, int *done, long long *iters, hipStream_t s);
hipError_t launch_iteration_f64(const IterLaunch &a, hipStream_t s);
hipError_t launch_iteration_f32(const IterLaunch &a, hipStream_t s);
hipError_t launch_cfl_state(int precision, const void *grids, int gid, int blocks, hipStream_t s);
hipError_t launch_cell_velocity(int precision, const void *grids, int gid, double *uu, double *vv,
                                int blocks, hipStream_t s);
hipError_t launch_volume(int precision, const void *grids, int gid, double *partial, int blocks,
                         hipStream_t s);

}  // namespace sw2d
