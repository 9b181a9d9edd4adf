// sw2d_launch.hpp -- host-side launchers of the stencil kernels, one set per
// state precision (sw2d_f64.hip, sw2d_f32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sw2d_device.hpp"

namespace sw2d {

// tile height (rows) of each precision's kernels; tiles are 64 columns wide
constexpr int TILE_H_F64 = 8;
constexpr int TILE_H_F32 = 16;

struct IterLaunch {
  const void *grids;   // device Grid<T> array
  const int2 *work;    // (grid, tile) work-list, or nullptr for grid 0
  int n_work;          // work items (tiles)
  int blocks;          // workgroups launched (<= n_work; persistent loop)
  Ctrl *ctrl;          // device control blocks
  int n_ctrl;
  const hipEvent_t *ev;  // optional: recorded after control, momentum, continuity
};

hipError_t launch_control(Ctrl *ctrl, int n, int commit_only, hipStream_t s);
hipError_t launch_iteration_f64(const IterLaunch &a, hipStream_t s);
hipError_t launch_iteration_f32(const IterLaunch &a, hipStream_t s);
hipError_t launch_cfl_state(int precision, const void *grids, int gid, int blocks, hipStream_t s);
hipError_t launch_cell_velocity(int precision, const void *grids, int gid, double *uu, double *vv,
                                int blocks, hipStream_t s);
hipError_t launch_volume(int precision, const void *grids, int gid, double *partial, int blocks,
                         hipStream_t s);

}  // namespace sw2d
