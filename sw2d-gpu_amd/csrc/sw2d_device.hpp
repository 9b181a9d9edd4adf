// sw2d_device.hpp -- device-side data structures shared by the step kernels
// and the host runtime of the MI355X SW2D core.
//
// Memory layout in HBM (one scenario):
//   all arrays row-major, row pieces of P >= cols+9 elements (a multiple of
//   64), interleaved by row in two blocks (see STATE_ARRAYS), with padding
//   rows above and below the grid
//   h[2]      rows x P     depth, ping-pong; inactive cells hold -1
//   U[2]      rows+1 x P   discharge on U faces (row i = face north of cell i)
//   V[2]      rows x P     discharge on V faces (column j = face west of j,
//                          column cols = east boundary face)
//   z         rows   x P   bed (FP32 form: minus a per-64x64-block datum)
//   ncell     rows   x P   optional Manning raster (FP32 form: g*n^2)
// The padding rows and the columns beyond the grid (from cols, or cols+1 for
// V) hold the values outside the grid: h = -1 (no cell), U = V = 0 (closed),
// bed and raster 0.  The step kernel reads neighbours outside the grid there
// instead of masking them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sw2d {

// Output columns of a stencil work item.  FP32: 56 (a row piece is 224 B,
// whole 32-byte sectors; measured 3 % faster than 57 over the whole 8192^2
// run).  FP64: 57 (the FP64 step is bound by its arithmetic, and 56 columns
// mean 2 % more strips: 4.5 % slower).
constexpr int STRIP_W32 = 56, STRIP_W64 = 57;
__host__ __device__ constexpr int strip_width(bool fp64) { return fp64 ? STRIP_W64 : STRIP_W32; }
constexpr int WAVE = 64;            // threads of a stencil workgroup (one wavefront)
constexpr int BLOCK = 256;          // threads of the whole-grid helper kernels
constexpr int MAX_TILE_SRC = 256;   // point sources inside one segment and its halo
constexpr int DATUM_BLOCK = 64;     // FP32 bed datum block (cells)
constexpr int ACC_N = 5;            // rain, source, evap, outflow, clamp
constexpr int SMAX_SLOTS = 64;      // CFL maximum: workgroups spread their atomics
constexpr int EDGE_N = 1, EDGE_S = 2, EDGE_W = 4, EDGE_E = 8;

// Width P of one array's row piece: the V east face column cols, then at
// least eight padding columns, rounded to 64 elements.
__host__ __device__ inline size_t state_pitch(int cols) { return ((size_t)cols + 9 + 63) / 64 * 64; }

// The six state arrays live in one block, interleaved by row: row r holds
// the row pieces of h0 h1 U0 U1 V0 V1 (STATE_ARRAYS pieces of P elements, row
// stride RS = 6 P); bed and raster likewise share a second block (row stride
// P, or 2 P with a raster).  Each block holds PAD_ROWS_ABOVE rows above row 0
// and rows_after(rows) rows from row 0 on: the grid rows, then padding rows.
// The padding rows and the columns beyond the grid hold the values outside
// the grid (h = -1, U = V = 0, bed and raster 0), except the last four
// columns of each piece, which hold those of the NEXT piece (cyclically, V1
// before the next row's h0): the step kernel reads a row from column j0 - 4
// of its strip, so on the first strip columns -4 .. -1 of a piece are the
// last four columns of the piece before it.  A lane whose column lies beyond
// cols + 4 reads the following piece's first columns; such a lane is more
// than the stencil's three columns away from every output column.
constexpr int STATE_ARRAYS = 6;
constexpr int PAD_ROWS_ABOVE = 5, PAD_ROWS_BELOW = 6;
__host__ __device__ inline size_t rows_after(int rows) { return (size_t)rows + PAD_ROWS_BELOW; }
__host__ __device__ inline size_t slab_rows(int rows) { return PAD_ROWS_ABOVE + rows_after(rows); }
// piece of field f (0 = h, 1 = U, 2 = V) in ping-pong buffer b
__host__ __device__ inline int state_piece(int f, int b) { return 2 * f + b; }
constexpr int ERR_NONFINITE = 1, ERR_TOO_MANY_SRC = 2;

// Per-scenario control block, resident in device memory.  The control
// kernel fills the per-iteration fields; the stencil kernels only read them.
struct Ctrl {
  // clock
  double t;            // time at the start of the pending iteration
  double t_next;       // time after the pending iteration
  double t_end;        // loop runs while t + dth <= t_end (L:1256)
  double dt_setup;     // setup.dat dt
  double cfl;          // 0: fixed step
  // values of the pending iteration
  double D, dth;
  double rr_a, rr_b;   // rain rate [m/s] at t and t + dth
  double ev_a, ev_b;   // evaporation rate [m/s]
  long long iter;      // committed iterations
  int parity;          // h/U/V buffer holding the committed state
  int run;             // 1 while an iteration is pending
  int done;            // end condition reached
  int error;           // ERR_* bits
  unsigned long long smax_bits;  // CFL speed max of the newest state (folded)
  unsigned long long smax_slot[SMAX_SLOTS];  // per-workgroup atomics land here
  // forcing tables (device pointers)
  const double *rain; int n_rain; double dtrain;
  const double *evapo; int n_evapo; int evaporation_on;
  const double *q; int n_q; int q_cols; double dtoq;
  int nsrc;                      // unique source cells
  const int *src_col;            // Q-table column of each source
  const long long *src_cell;     // linear cell index i*cols + j
  double *ql_a, *ql_b;           // source rates [m/s] at t and t + dth
  const double *dsched; int n_dsched;
  double *dt_hist; int dt_hist_cap;
  int forcing_mode;
  double dxdy;
};

// One work item of the stencil kernels: rows [i0, i0+th) of the strip_width-column
// strip starting at column j0 of grid `gid`; its volume accumulators live in
// slot `slot` of that grid.
struct Tile {
  int gid, i0, j0, th, slot, pad0, pad1, pad2;
};

template <typename T>
struct Grid {
  int R, C, P;
  int RS, RSS;                   // row strides of the state and static blocks (elements)
  int tiles_x, tiles_y;          // strip_width-column strips, row segments
  int seg_rows;                  // rows per segment when no work-list is given
  int edge_open;
  int has_ncell;
  int rain_block, rain_cols;     // rain multiplier map geometry (0 = none)
  int rain_shift;                // log2(rain_block)
  int datum_cols;                // FP32: datum blocks per row
  double dx, dy, g, th, manning;
  double INT, INF, LWL, EVmin;
  T *h[2];
  T *U[2];
  T *V[2];
  const T *z;
  const T *ncell;
  const float *datum;            // FP32: datum per 64x64 block
  const T *rain_map;              // rain multiplier per block, in the state precision
  Ctrl *ctrl;
  double *acc;                   // per tile: ACC_N x (sum, compensation)
  T unit;                        // 1: the rain multiplier read where there is no map
};

// Interpolation shared by rain, point sources and evaporation
// (stream_flow L:533-538, L:547-552; evaporation_load L:592-597).
// mode 1: piecewise constant.  Indices past the table hold the last entry.
__host__ __device__ inline double forcing_value(double time, const double *table,
                                                int n, double dT, int mode) {
  if (n <= 0) return 0.0;
  if (mode == 1) {
    int it = (int)(time / dT);
    it = it < 0 ? 0 : (it > n - 1 ? n - 1 : it);
    return table[it];
  }
  if (time <= 1.00) return table[0] * time;
  int it = (int)(time / dT);
  int i0 = it < 0 ? 0 : (it > n - 1 ? n - 1 : it);
  int i1 = it + 1 < 0 ? 0 : (it + 1 > n - 1 ? n - 1 : it + 1);
  return table[i0] + (table[i1] - table[i0]) / (dT * (time - dT * (it)));
}

// fma(a, b, k) for a constant k: one correctly rounded fused multiply-add on
// either side.  On the device k is passed in a scalar register pair (two
// scalar moves): left to itself the compiler builds each such constant in a
// vector register pair for v_fmac_f64, two vector moves per Horner step.
static __device__ inline double fma_k(double a, double b, double k) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(k));
  return r;
}
static __host__ inline double fma_k(double a, double b, double k) { return fma(a, b, k); }

// x^1.3333330 (the friction exponent of L:195, L:312) for 0 < x < 1e300,
// from IEEE operations only: +, -, *, / and fma (each correctly rounded),
// frexp, ldexp and rint (exact).  The same sequence therefore gives the same
// bits on the device and on the host, wherever contraction is off (the FP64
// kernels, the C ABI helper sw2d_friction_pow); the library pow functions of
// the two sides differ in the last bit for some arguments.  Error below
// 0.65 ulp: log x = e ln2 + 2 atanh(s), s = (m - 1)/(m + 1), m in [1/sqrt2,
// sqrt2), and exp(t) = 2^k exp(r), |r| <= ln2/2, by series, each carried in
// double-double where its rounding would matter (Dekker/Knuth two-sum and
// fma two-product).
// Out of line on the device: inlined at its four call sites in the FP64 step,
// its temporaries took the kernel past 168 VGPRs (three waves per SIMD).
static __host__ __device__ __attribute__((noinline)) double pow_fric_f64(double x) {
  const double Y = 1.3333330;
  const double LN2_HI = 0x1.62e42fee00000p-1;   // ln 2, 32 significant bits
  const double LN2_LO = 0x1.a39ef35793c76p-33;  // ln 2 - LN2_HI
  const double INV_LN2 = 0x1.71547652b82fep+0;
  int e;
  double m = frexp(x, &e);
  if (m < 0x1.6a09e667f3bcdp-1) {  // 1/sqrt(2)
    m = m * 2.0;
    e = e - 1;
  }
  // s_hi + s_lo = (m - 1)/(m + 1); m - 1 is exact, b + bl = m + 1 exactly.
  // One division: s_hi = a (1/b) is within an ulp of a/b, the remainder
  // a - s_hi b is exact, and s_lo carries it (its own relative error, 2^-52,
  // is far below what the result keeps).
  const double a = m - 1.0;
  const double b = m + 1.0;
  const double bv = b - m;
  const double bl = (m - (b - bv)) + (1.0 - bv);
  const double ib = 1.0 / b;
  const double s_hi = a * ib;
  const double p = s_hi * b;
  const double pe = fma(s_hi, b, -p);
  const double s_lo = (((a - p) - pe) - s_hi * bl) * ib;
  // log m = 2 s + 2 s^3 (1/3 + z/5 + ... + z^11/25), z = s^2 <= 0.0295
  // (Horner steps fused, here and in the exponential below)
  const double z = s_hi * s_hi;
  double P = 0x1.47ae147ae147bp-5;  // 1/25
  P = fma_k(P, z, 0x1.642c8590b2164p-5);  // 1/23
  P = fma_k(P, z, 0x1.8618618618618p-5);  // 1/21
  P = fma_k(P, z, 0x1.af286bca1af28p-5);  // 1/19
  P = fma_k(P, z, 0x1.e1e1e1e1e1e1ep-5);  // 1/17
  P = fma_k(P, z, 0x1.1111111111111p-4);  // 1/15
  P = fma_k(P, z, 0x1.3b13b13b13b14p-4);  // 1/13
  P = fma_k(P, z, 0x1.745d1745d1746p-4);  // 1/11
  P = fma_k(P, z, 0x1.c71c71c71c71cp-4);  // 1/9
  P = fma_k(P, z, 0x1.2492492492492p-3);  // 1/7
  P = fma_k(P, z, 0x1.999999999999ap-3);  // 1/5
  P = fma_k(P, z, 0x1.5555555555555p-2);  // 1/3
  const double L_hi = 2.0 * s_hi;
  const double L_lo = 2.0 * s_lo + 2.0 * (s_hi * z) * P;
  // log x = e ln2 + log m (e LN2_HI is exact)
  const double E_hi = (double)e * LN2_HI;
  const double E_lo = (double)e * LN2_LO;
  const double S = E_hi + L_hi;
  const double Sv = S - E_hi;
  const double Se = (E_hi - (S - Sv)) + (L_hi - Sv);
  const double lo = Se + (E_lo + L_lo);
  const double lx_hi = S + lo;
  const double lx_lo = lo - (lx_hi - S);
  // t = Y log x
  const double t_hi = Y * lx_hi;
  const double t_lo = fma(Y, lx_hi, -t_hi) + Y * lx_lo;
  // exp t = 2^k exp r; t_hi - k LN2_HI is exact
  const double k = rint(t_hi * INV_LN2);
  const double r_hi = t_hi - k * LN2_HI;
  const double r_lo = t_lo - k * LN2_LO;
  const double r = r_hi + r_lo;
  const double r_e = r_lo - (r - r_hi);
  double q = 0x1.ae7f3e733b81fp-41;  // 1/15!
  q = fma_k(q, r, 0x1.93974a8c07c9dp-37);  // 1/14!
  q = fma_k(q, r, 0x1.6124613a86d09p-33);  // 1/13!
  q = fma_k(q, r, 0x1.1eed8eff8d898p-29);  // 1/12!
  q = fma_k(q, r, 0x1.ae64567f544e4p-26);  // 1/11!
  q = fma_k(q, r, 0x1.27e4fb7789f5cp-22);  // 1/10!
  q = fma_k(q, r, 0x1.71de3a556c734p-19);  // 1/9!
  q = fma_k(q, r, 0x1.a01a01a01a01ap-16);  // 1/8!
  q = fma_k(q, r, 0x1.a01a01a01a01ap-13);  // 1/7!
  q = fma_k(q, r, 0x1.6c16c16c16c17p-10);  // 1/6!
  q = fma_k(q, r, 0x1.1111111111111p-7);   // 1/5!
  q = fma_k(q, r, 0x1.5555555555555p-5);   // 1/4!
  q = fma_k(q, r, 0x1.5555555555555p-3);   // 1/3!
  q = fma_k(q, r, 0x1.0000000000000p-1);   // 1/2!
  q = (r * r) * q;
  const double one_r = 1.0 + r;
  const double one_e = (1.0 - one_r) + r;  // |r| < 1
  const double res = one_r + (one_e + (r_e + q));
  return ldexp(res, (int)k);
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of tiles so stencil halos shared by neighbouring tiles
// stay in that XCD's L2.
__device__ inline int xcd_tile(int b, int n) {
  int q = n >> 3;
  if (b >= (q << 3)) return b;
  return (b & 7) * q + (b >> 3);
}

}  // namespace sw2d
