// sw2d_internal.hpp -- host-side context of one scenario, shared by the
// single-grid runtime (sw2d_api.hip) and the batch/serving runtime
// (sw2d_batch.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "sw2d.h"
#include "sw2d_launch.hpp"

using sw2d::Ctrl;
using sw2d::Grid;

// message slot behind sw2d_last_error (defined in sw2d_api.hip)
void sw2d_set_error_text(const std::string &m);

namespace {

inline int fail(int code, const std::string &msg) {
  sw2d_set_error_text(msg);
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(SW2D_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Comp {
  double s = 0, c = 0;
  void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

// A captured run of GRAPH_ITERS iterations, replayed with one launch.
constexpr int GRAPH_ITERS = 32;
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
};

inline hipError_t launch_one(int prec, const sw2d::IterLaunch &a, hipStream_t s) {
  return prec == SW2D_FP64 ? sw2d::launch_iteration_f64(a, s) : sw2d::launch_iteration_f32(a, s);
}

// Enqueue n iterations; with a cache, whole chunks replay a captured graph
// (the launch parameters are fixed: everything that changes per iteration
// lives in the device control blocks).
inline int enqueue_iterations(int prec, hipStream_t s, const sw2d::IterLaunch &a, int64_t n,
                              GraphCache *gc) {
  int64_t k = 0;
  if (gc && n >= GRAPH_ITERS && !std::getenv("SW2D_NO_GRAPH")) {
    if (!gc->exec) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t le = hipSuccess;
      for (int i = 0; i < GRAPH_ITERS && le == hipSuccess; i++) le = launch_one(prec, a, s);
      hipError_t ce = hipStreamEndCapture(s, &g);
      if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(SW2D_E_DEVICE, std::string("graph capture: ") +
                                       hipGetErrorString(le != hipSuccess ? le : ce));
      }
      hipError_t ie = hipGraphInstantiate(&gc->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        gc->exec = nullptr;
        return fail(SW2D_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
      }
    }
    for (; k + GRAPH_ITERS <= n; k += GRAPH_ITERS) HIPCHK(hipGraphLaunch(gc->exec, s));
  }
  for (; k < n; k++) {
    hipError_t e = launch_one(prec, a, s);
    if (e != hipSuccess) return fail(SW2D_E_DEVICE, std::string("launch: ") + hipGetErrorString(e));
  }
  return SW2D_OK;
}

}  // namespace

struct sw2d_ctx {
  int prec = 0;
  size_t esz = 8;
  int R = 0, C = 0, P = 0, TH = 0, tiles_x = 0, tiles_y = 0, n_tiles = 0;
  int RS = 0, RSS = 0;  // row strides of the state and static blocks (elements)
  int n_acc = 0;  // accumulator slots: enough for the shortest segment (4 rows)
  sw2d_params prm{};
  double nodata = 0;
  double dx = 0, dy = 0;
  hipStream_t stream = nullptr;
  // A batch scenario's work runs on the batch stream: every call on the
  // context waits for it as well (fetch_ctrl), and the context cannot be
  // stepped on its own.
  hipStream_t owner = nullptr;
  std::vector<void *> allocs;
  void *slab = nullptr;  // the state block (row 0 of its first piece)
  // row 0 of each state array inside the state block; the bed (its block's
  // allocation) and the raster inside the static block
  void *h[2]{}, *U[2]{}, *V[2]{}, *z = nullptr, *ncell = nullptr;
  float *datum = nullptr;
  void *rain_map = nullptr;  // rain multiplier map, in the state precision
  double *acc = nullptr, *partial = nullptr;
  double *d_uu = nullptr, *d_vv = nullptr;
  Ctrl *d_ctrl = nullptr;
  Ctrl **d_ctrl_list = nullptr;  // one-entry pointer array for k_control
  void *d_grid = nullptr;
  double *d_rain = nullptr, *d_evapo = nullptr, *d_q = nullptr, *d_dsched = nullptr,
         *d_dthist = nullptr, *d_ql = nullptr;
  int *d_src_col = nullptr;
  long long *d_src_cell = nullptr;
  int vol_blocks = 0;
  int dthist_cap = 0;
  std::vector<unsigned char> active;  // rows*cols
  unsigned char *d_active = nullptr;  // its device copy
  std::vector<float> datum_host;
  int datum_cols = 0;
  // kernel timing: 4 events per iteration (before control, after control,
  // after momentum, after continuity)
  std::vector<hipEvent_t> prof_ev;
  int prof_cap = 0, prof_used = 0;

  GraphCache graph;

  // output snapshots (sw2d_output_begin / sw2d_output_wait): the device
  // snapshot of (h, vv, uu), its two pinned host copies and the side stream
  // that copies it while the context stream goes on stepping
  double *d_snap = nullptr;
  double *h_snap[2] = {};
  hipEvent_t snap_done[2] = {};
  hipEvent_t snap_ready = nullptr;
  hipStream_t side = nullptr;
  int snap_next = 0;
  bool snap_used[2] = {};

  ~sw2d_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    for (int k = 0; k < 2; k++) {
      if (snap_done[k]) (void)hipEventDestroy(snap_done[k]);
      if (h_snap[k]) (void)hipHostFree(h_snap[k]);
    }
    if (snap_ready) (void)hipEventDestroy(snap_ready);
    if (side) (void)hipStreamDestroy(side);
    graph.reset();
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    for (void *p : allocs) (void)hipFree(base_of(p));
    if (stream) (void)hipStreamDestroy(stream);
  }

  // Pointers handed out `shift` bytes into their allocation (see alloc);
  // allocs keeps the handed-out pointers.
  std::vector<std::pair<void *, size_t>> shifted;
  void *base_of(void *p) const {
    for (const auto &s : shifted)
      if (s.first == p) return (char *)p - s.second;
    return p;
  }
  // `shift`: the pointer handed out lies that many bytes into a block as much
  // larger (the pitched statics use it to move their row reads onto sector
  // boundaries)
  int alloc(void **p, size_t bytes, size_t shift = 0) {
    void *raw = nullptr;
    hipError_t e = hipMalloc(&raw, (bytes ? bytes : 8) + shift);
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(SW2D_E_MEMORY, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    *p = (char *)raw + shift;
    if (shift) shifted.emplace_back(*p, shift);
    allocs.push_back(*p);
    return SW2D_OK;
  }
  void release(void *p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it == allocs.end()) return;  // not this context's allocation
    allocs.erase(it);
    (void)hipFree(base_of(p));
    shifted.erase(std::remove_if(shifted.begin(), shifted.end(),
                                 [p](const std::pair<void *, size_t> &s) { return s.first == p; }),
                  shifted.end());
  }

  int fetch_ctrl(Ctrl &c) {
    if (owner) HIPCHK(hipStreamSynchronize(owner));
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(&c, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    return SW2D_OK;
  }
  int store_ctrl(const Ctrl &c) {
    HIPCHK(hipMemcpy(d_ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice));
    return SW2D_OK;
  }
  int blocks_for(long long n) const {
    long long b = (n + sw2d::BLOCK - 1) / sw2d::BLOCK;
    return (int)std::max(1LL, std::min(b, 4096LL));
  }
  int recompute_cfl() {
    Ctrl c;
    int rc = fetch_ctrl(c);
    if (rc) return rc;
    c.smax_bits = 0;
    std::memset(c.smax_slot, 0, sizeof(c.smax_slot));
    rc = store_ctrl(c);
    if (rc) return rc;
    HIPCHK(sw2d::launch_cfl_state(prec, d_grid, 0, blocks_for((long long)R * C), stream));
    HIPCHK(sw2d::launch_control(d_ctrl_list, 1, 1, stream));  // fold the slots
    HIPCHK(hipStreamSynchronize(stream));
    return SW2D_OK;
  }
  // Copy a device array (row stride `pitch` elements) into a dense host
  // double array: the conversion runs on the device (clamp0: negative values
  // read as 0), a chunk of rows at a time through a stream-ordered staging
  // buffer of at most sw2d::STAGING_MAX bytes.
  // With `stage` (a device buffer of stage_n doubles) the rows pass through
  // it instead, and nothing is allocated.
  int read_pitched(const void *src, int rows, int cols, double *out, int pitch, bool clamp0 = false,
                   double *stage = nullptr, size_t stage_n = 0) {
    const size_t n = (size_t)rows * cols;
    if (n == 0) return SW2D_OK;
    int cr = sw2d::staging_rows(rows, cols);
    double *d_out = stage;
    if (stage) {
      if (stage_n < (size_t)cols) return fail(SW2D_E_INVALID, "a row is wider than the staging buffer");
      cr = (int)std::min<size_t>((size_t)rows, stage_n / (size_t)cols);
    } else {
      HIPCHK(hipMallocAsync((void **)&d_out, (size_t)cr * cols * sizeof(double), stream));
    }
    hipError_t e = hipSuccess;
    for (int a = 0; a < rows && e == hipSuccess; a += cr) {
      const int nr = std::min(cr, rows - a);
      e = sw2d::launch_unpack(prec, d_out, (const char *)src + (size_t)a * pitch * esz, nr, cols, (size_t)pitch,
                              clamp0 ? 1 : 0, stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(out + (size_t)a * cols, d_out, (size_t)nr * cols * sizeof(double),
                           hipMemcpyDeviceToHost, stream);
    }
    const hipError_t f = stage ? hipSuccess : hipFreeAsync(d_out, stream);
    const hipError_t y = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = f != hipSuccess ? f : y;
    if (e != hipSuccess) return fail(SW2D_E_DEVICE, std::string("read: ") + hipGetErrorString(e));
    return SW2D_OK;
  }
  // Write a dense host array (rows x cols) into a device array with row
  // stride `pitch` elements; only the grid's values are written (the padding
  // is set once at creation).  The array goes up a chunk of rows at a time,
  // through a stream-ordered staging buffer of at most sw2d::STAGING_MAX
  // bytes, and the conversion (mode: sw2d::PackMode) runs on the device.
  int pack(void *dst, int rows, int cols, const double *in, int mode, int pitch, double g = 0.0) {
    const int cr = sw2d::staging_rows(rows, cols);
    double *d_in = nullptr;
    HIPCHK(hipMallocAsync((void **)&d_in, std::max<size_t>((size_t)cr * cols, 1) * sizeof(double), stream));
    hipError_t e = hipSuccess;
    for (int a = 0; a < rows && e == hipSuccess; a += cr) {
      const int nr = std::min(cr, rows - a);
      e = hipMemcpyAsync(d_in, in + (size_t)a * cols, (size_t)nr * cols * sizeof(double), hipMemcpyHostToDevice,
                         stream);
      // the masked modes convert cell arrays (cols == C); the mask is not
      // read otherwise
      const bool masked = mode == sw2d::PACK_CELLS || mode == sw2d::PACK_BED;
      if (e == hipSuccess)
        e = sw2d::launch_pack(prec, (char *)dst + (size_t)a * pitch * esz, d_in,
                              masked ? d_active + (size_t)a * cols : nullptr, nr, cols, (size_t)pitch, mode, datum,
                              datum_cols, g, a, stream);
    }
    const hipError_t f = hipFreeAsync(d_in, stream);
    const hipError_t y = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = f != hipSuccess ? f : y;
    if (e != hipSuccess) return fail(SW2D_E_DEVICE, std::string("write: ") + hipGetErrorString(e));
    return SW2D_OK;
  }
  // The state arrays: the grid's values of h (inactive cells as -1), U and V.
  int write_h(void *dst, const double *in) { return pack(dst, R, C, in, sw2d::PACK_CELLS, RS); }
  int write_U(void *dst, const double *in) { return pack(dst, R + 1, C, in, sw2d::PACK_PLAIN, RS); }
  int write_V(void *dst, const double *in) { return pack(dst, R, C + 1, in, sw2d::PACK_PLAIN, RS); }
};

