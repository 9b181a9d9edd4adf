// sw2d_launch.hpp -- host-side launchers of the stencil kernels, one set per
// state precision (sw2d_f64.hip, sw2d_f32.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "sw2d_device.hpp"

namespace sw2d {

// default rows per strip segment: 32 (measured on 8192^2 grids: 32 and 64
// rows are within 2 %, 128 and 256 lose 11 % and 26 % to the last round of
// workgroups), halved to 16 while the grid would give fewer than MIN_ITEMS
// items, so small grids still fill the chip; a segment pays 7 lead-in steps,
// so it never gets shorter than 16 rows.  SW2D_SEG_ROWS in the environment
// overrides it (a multiple of 4 in [4, 4096]).
constexpr int MIN_ITEMS = 8192;
inline int default_seg_rows(int rows, int strips) {
  if (const char *e = getenv("SW2D_SEG_ROWS")) {
    const int v = atoi(e);
    if (v >= 4 && v <= 4096 && v % 4 == 0) return v;
  }
  int s = 32;
  while (s > 16 && (long long)strips * ((rows + s - 1) / s) < MIN_ITEMS) s /= 2;
  return s;
}

struct IterLaunch {
  const void *grids;   // device Grid<T> array
  const Tile *work;    // tile work-list, or nullptr: all THMAX-row tiles of grid 0
  int n_work;          // work items (tiles)
  int blocks;          // workgroups launched (<= n_work; persistent loop)
  Ctrl *const *ctrl;   // device array of control-block pointers
  int n_ctrl;
  const hipEvent_t *ev;  // optional: recorded after control, after the step kernel (twice)
};

hipError_t launch_control(Ctrl *const *ctrl, int n, int commit_only, hipStream_t s);
hipError_t launch_status(Ctrl *const *ctrl, int n, int *done, long long *iters, hipStream_t s);
hipError_t launch_iteration_f64(const IterLaunch &a, hipStream_t s);
hipError_t launch_iteration_f32(const IterLaunch &a, hipStream_t s);
// the stencil step alone (no control kernel): with an iteration decided and
// not committed, repeated launches recompute the same scratch state
hipError_t launch_step_only(int precision, const IterLaunch &a, hipStream_t s);
hipError_t launch_step_only_f32(const IterLaunch &a, hipStream_t s);
hipError_t launch_cfl_state(int precision, const void *grids, int gid, int blocks, hipStream_t s);
hipError_t launch_cell_velocity(int precision, const void *grids, int gid, double *uu, double *vv,
                                int blocks, hipStream_t s);
hipError_t launch_volume(int precision, const void *grids, int gid, double *partial, int blocks,
                         hipStream_t s);

}  // namespace sw2d
