// sw2d_launch.hpp -- host-side launchers of the stencil kernels, one set per
// state precision (sw2d_f64.hip, sw2d_f32.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "sw2d_device.hpp"

namespace sw2d {

// Default rows per strip segment.  A segment pays 7 lead-in steps, so long
// segments save work, but the launch runs in rounds of `slots` concurrent
// wavefronts (CUs x 4 SIMDs x waves per SIMD) and a partly filled last round
// idles the chip.  The default takes the longest segments that still give
// about `rounds` full rounds, of at least 16 rows on grids large enough:
//  * FP32 (4,096 slots), three rounds: whole 8192^2 runs, interleaved on one
//    MI355X, took 100 rows (2.94 rounds) to 115.0 Gcell/s, 76 rows (3.9
//    rounds) to 114.2 and 152 rows (1.94 rounds) to 112.0; 296 rows (one
//    round) and 32 rows lost 3-6 % and 4 % on an earlier form of the step;
//  * FP64 (3,072 slots), six rounds: 64 rows gave 42.6 Gcell/s, 32 rows
//    41.8, 148 rows 38.3 and 296 rows 34.7 on an earlier form of the step;
//    on the current one, from a wet state, 64 rows 2.024 ms per launch, 32
//    rows 2.087, 44 rows 2.065, 88 rows 2.098, 128 rows 2.032 and 200 rows
//    2.139; the FP64 rows are twice as wide and long segments lose their
//    halo reuse in L2.
// On larger grids the number of rounds grows until the segments are no
// longer than `max_rows` (160 FP32, 64 FP64), the longest measured to pay.
// A grid too small for `rounds` rounds of 16-row segments leaves most of the
// chip idle: there an iteration lasts about one wavefront's walk, and the
// default is the shortest segment (at least 4 rows) whose items fit in one
// round.  Per iteration, 16-row segments against this rule: the 200 x 100
// FP64 dam break 59 against 27 us (4 rows), a 1024^2 FP32 catchment 36.7
// against 31.0 us (8 rows; 4 rows, 1.2 rounds, 38.9), the same in FP64 90
// against 75 us (8 rows), a 256^2 FP32 catchment 28.7 against 18.2 us (4
// rows).  SW2D_SEG_ROWS in the environment overrides the default (a
// multiple of 4 in [4, 4096]).
inline int default_seg_rows(int rows, int strips, long long slots, int rounds, int max_rows) {
  if (const char *e = getenv("SW2D_SEG_ROWS")) {
    const int v = atoi(e);
    if (v >= 4 && v <= 4096 && v % 4 == 0) return v;
  }
  const long long st = strips > 0 ? strips : 1;
  long long s = 16;
  for (long long k = rounds; k <= 4096; k++) {
    const long long segs = k * slots / st;
    s = segs > 0 ? (rows + segs - 1) / segs : rows;
    s = (s + 3) / 4 * 4;
    if (s <= max_rows) break;
  }
  if (s < 16) {  // a small grid: one round of the shortest segments
    const long long segs = slots / st;
    s = segs > 0 ? (rows + segs - 1) / segs : rows;
    s = (s + 3) / 4 * 4;
    if (s < 4) s = 4;
  }
  return (int)(s > 4096 ? 4096 : s);
}

// Longest segment (a multiple of 4, at most 4096 rows) whose reach in the
// state block -- its rows, the 4 above and the 5 below -- stays below 2^31
// bytes for a block row of `row_bytes`: the step kernel addresses a segment's
// rows with 32-bit offsets and drops stores at offset 2^31.  0 when not even
// 4 rows fit.
inline int reach_rows_cap(long long row_bytes) {
  const long long n = ((1LL << 31) / (row_bytes > 0 ? row_bytes : 1) - 10) / 4 * 4;
  return (int)(n < 4 ? 0 : n > 4096 ? 4096 : n);
}

// Segment rows of the scenarios of a batch whose live set holds
// `strip_rows` strip rows in all (sum over scenarios of strips x rows): the
// height that gives the whole set about twelve rounds of wave slots, at
// least 16 and at most `max_rows`; a live set too small for that takes the
// shortest segments (at least 4 rows) that fit in one round, as a small
// single grid does (default_seg_rows).  A batch mixes segments of different
// cost (wet and dry grids, short last segments), so it needs more rounds
// than one uniform grid to keep the end of a launch busy.  On the 256
// scenarios of 64 grids (sides 259-1955, 3.36 M strip rows, FP32), whole
// 2 h runs took 17.8 s at this height (68 rows), 17.9 s at 52 and 64 rows,
// 18.4 s at 128, 18.6 s at 32 and 20.4 s at 256 rows; the single-grid
// heights of the scenarios (16-36 rows) took 19.3 s.  0 when
// SW2D_SEG_ROWS is set (the override then applies to every scenario).
inline int batch_seg_rows(long long strip_rows, long long slots, int max_rows) {
  if (getenv("SW2D_SEG_ROWS")) return 0;
  const long long sl = slots > 0 ? slots : 1;
  long long s = (strip_rows + 12 * sl - 1) / (12 * sl);
  s = (s + 3) / 4 * 4;
  if (s < 16) {
    s = (strip_rows + sl - 1) / sl;
    s = (s + 3) / 4 * 4;
    if (s < 4) s = 4;
  }
  return (int)(s > max_rows ? max_rows : s);
}

// concurrent single-wavefront workgroups of the step kernel on the current
// device: CUs x 4 SIMDs x waves per SIMD (4 for the FP32 kernel, 3 for FP64,
// set by their register use)
inline long long step_slots(bool fp64) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  return (long long)cus * 4 * (fp64 ? 3 : 4);
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
// cell velocities uu, vv (uu_vv L:472-479) and, with hout, the depth
hipError_t launch_cell_velocity(int precision, const void *grids, int gid, double *uu, double *vv,
                                double *hout, int blocks, hipStream_t s);
hipError_t launch_volume(int precision, const void *grids, int gid, double *partial, int blocks,
                         hipStream_t s);

// Dense double arrays <-> pitched state arrays, converted on the device.
// Uploads and read-backs pass whole rows through a device staging buffer of
// dense doubles, at most STAGING_MAX bytes (and at least one row).
constexpr size_t STAGING_MAX = size_t(32) << 20;
inline int staging_rows(int rows, int cols) {
  const size_t per = (size_t)(cols > 0 ? cols : 1) * sizeof(double);
  size_t r = STAGING_MAX / per;
  r = r < 1 ? 1 : r;
  return (int)(r < (size_t)rows ? r : (size_t)rows);
}
inline size_t staging_bytes(int rows, int cols) {
  return (size_t)staging_rows(rows, cols) * (size_t)cols * sizeof(double);
}
// pack: element (i, j) of a rows x cols dense array goes to row i, column j
// of a pitched array (row stride `pitch` elements); nothing else is written.
// Modes: PACK_PLAIN; PACK_CELLS (cell state: inactive cells of `mask` are
// written as -1); PACK_BED (0 on inactive cells; FP32: minus the cell's block
// datum); PACK_GN2 (g n^2).  The host restatement of each mode rounds alike
// (the same double operations, then one conversion to the state precision).
// dst, src and mask point at row row0 of their arrays (a chunk of rows);
// row0 places the rows in the bed datum blocks.
enum PackMode { PACK_PLAIN = 0, PACK_CELLS = 1, PACK_BED = 2, PACK_GN2 = 3 };
hipError_t launch_pack(int precision, void *dst, const double *src, const unsigned char *mask, int rows, int cols,
                       size_t pitch, int mode, const float *datum, int datum_cols, double g, int row0,
                       hipStream_t s);
// the padding pattern of the state block (see STATE_ARRAYS) over `rows` rows
// of six pieces of P elements, from dst on
hipError_t launch_slab_init(int precision, void *dst, size_t rows, int P, hipStream_t s);
// unpack: the rows x cols part of a pitched array into a dense double array;
// clamp0 turns negative values (the -1 of inactive cells) into 0.
hipError_t launch_unpack(int precision, double *dst, const void *src, int rows, int cols, size_t pitch, int clamp0,
                         hipStream_t s);

}  // namespace sw2d
