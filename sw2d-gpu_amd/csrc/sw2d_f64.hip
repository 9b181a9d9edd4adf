// sw2d_f64.hip -- FP64 instantiation of the stencil kernels (parity path).
// Built with -ffp-contract=off: every product and sum is rounded as written.
#define SW2D_WITH_CONTROL 1
#include "sw2d_kernels.inc"
#include "sw2d_launch.hpp"

namespace sw2d {

hipError_t launch_control(Ctrl *const *ctrl, int n, int commit_only, hipStream_t s) {
  hipLaunchKernelGGL(k_control, dim3(n), dim3(WAVE), 0, s, ctrl, n, commit_only);
  return hipGetLastError();
}

hipError_t launch_status(Ctrl *const *ctrl, int n, int *done, long long *iters, hipStream_t s) {
  hipLaunchKernelGGL(k_status, dim3((n + 63) / 64), dim3(64), 0, s, ctrl, n, done, iters);
  return hipGetLastError();
}

hipError_t launch_iteration_f64(const IterLaunch &a, hipStream_t s) {
  hipLaunchKernelGGL(k_control, dim3(a.n_ctrl), dim3(WAVE), 0, s, a.ctrl, a.n_ctrl, 0);
  const Grid<double> *g = static_cast<const Grid<double> *>(a.grids);
  if (a.ev) (void)hipEventRecord(a.ev[0], s);
  hipLaunchKernelGGL(k_step<double>, dim3(a.blocks), dim3(WAVE), 0, s, g, a.work, a.n_work);
  if (a.ev) {
    (void)hipEventRecord(a.ev[1], s);
    (void)hipEventRecord(a.ev[2], s);
  }
  return hipGetLastError();
}

hipError_t launch_step_only(int precision, const IterLaunch &a, hipStream_t s) {
  if (precision != 0) return launch_step_only_f32(a, s);
  hipLaunchKernelGGL(k_step<double>, dim3(a.blocks), dim3(WAVE), 0, s,
                     static_cast<const Grid<double> *>(a.grids), a.work, a.n_work);
  return hipGetLastError();
}

hipError_t launch_cfl_state_f64(const void *grids, int gid, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_cfl_state<double>, dim3(blocks), dim3(BLOCK), 0, s,
                     static_cast<const Grid<double> *>(grids), gid);
  return hipGetLastError();
}
hipError_t launch_cell_velocity_f64(const void *grids, int gid, double *uu, double *vv, double *hout,
                                    int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_velocity<double>, dim3(blocks), dim3(BLOCK), 0, s,
                     static_cast<const Grid<double> *>(grids), gid, uu, vv, hout);
  return hipGetLastError();
}
hipError_t launch_volume_f64(const void *grids, int gid, double *partial, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_volume<double>, dim3(blocks), dim3(BLOCK), 0, s,
                     static_cast<const Grid<double> *>(grids), gid, partial);
  return hipGetLastError();
}

}  // namespace sw2d
