// sw2d_f32.hip -- FP32-state instantiation of the stencil kernels.
#include "sw2d_kernels.inc"
#include "sw2d_launch.hpp"

namespace sw2d {

hipError_t launch_control(Ctrl *const *ctrl, int n, int commit_only, hipStream_t s);

hipError_t launch_iteration_f32(const IterLaunch &a, hipStream_t s) {
  hipError_t e = launch_control(a.ctrl, a.n_ctrl, 0, s);
  if (e != hipSuccess) return e;
  const Grid<float> *g = static_cast<const Grid<float> *>(a.grids);
  if (a.ev) (void)hipEventRecord(a.ev[0], s);
  hipLaunchKernelGGL(k_step<float>, dim3(a.blocks), dim3(WAVE), 0, s, g, a.work, a.n_work);
  if (a.ev) {
    (void)hipEventRecord(a.ev[1], s);
    (void)hipEventRecord(a.ev[2], s);
  }
  return hipGetLastError();
}

hipError_t launch_step_only_f32(const IterLaunch &a, hipStream_t s) {
  hipLaunchKernelGGL(k_step<float>, dim3(a.blocks), dim3(WAVE), 0, s,
                     static_cast<const Grid<float> *>(a.grids), a.work, a.n_work);
  return hipGetLastError();
}

hipError_t launch_cfl_state_f64(const void *grids, int gid, int blocks, hipStream_t s);
hipError_t launch_cell_velocity_f64(const void *grids, int gid, double *uu, double *vv, double *hout,
                                    int blocks, hipStream_t s);
hipError_t launch_volume_f64(const void *grids, int gid, double *partial, int blocks, hipStream_t s);

hipError_t launch_cfl_state(int precision, const void *grids, int gid, int blocks, hipStream_t s) {
  if (precision == 0) return launch_cfl_state_f64(grids, gid, blocks, s);
  hipLaunchKernelGGL(k_cfl_state<float>, dim3(blocks), dim3(BLOCK), 0, s,
                     static_cast<const Grid<float> *>(grids), gid);
  return hipGetLastError();
}
hipError_t launch_cell_velocity(int precision, const void *grids, int gid, double *uu, double *vv,
                                double *hout, int blocks, hipStream_t s) {
  if (precision == 0) return launch_cell_velocity_f64(grids, gid, uu, vv, hout, blocks, s);
  hipLaunchKernelGGL(k_cell_velocity<float>, dim3(blocks), dim3(BLOCK), 0, s,
                     static_cast<const Grid<float> *>(grids), gid, uu, vv, hout);
  return hipGetLastError();
}
hipError_t launch_volume(int precision, const void *grids, int gid, double *partial, int blocks,
                         hipStream_t s) {
  if (precision == 0) return launch_volume_f64(grids, gid, partial, blocks, s);
  hipLaunchKernelGGL(k_volume<float>, dim3(blocks), dim3(BLOCK), 0, s,
                     static_cast<const Grid<float> *>(grids), gid, partial);
  return hipGetLastError();
}

}  // namespace sw2d
