// sw2d_api.hip -- host runtime behind include/sw2d.h: device memory,
// uploads, the per-iteration launch sequence, reads, stats and snapshots.
//
// Replaces the reference's main()-resident set-up and loop
// (SW2D-GPU_linux/src/SW2D-GPU.cu L:1105-1336): 16 FP64 + 4 int arrays and
// 9 launches + 9 device syncs + 4 H2D copies per iteration become ping-pong
// state buffers, device-resident forcing tables and three stream-ordered
// launches per iteration with no host synchronisation.
#include <memory>
#include <mutex>

#include "sw2d_internal.hpp"

namespace {
thread_local std::string g_err;

// Dense <-> pitched conversions on the device (see sw2d::launch_pack).
// rows x cols values of a dense double array into a pitched array (row
// stride `pitch`); dst, src and mask start at row `row0` of their arrays,
// which places the rows in the bed datum blocks
template <typename T>
__global__ void k_pack(T *__restrict__ dst, const double *__restrict__ src, const unsigned char *__restrict__ mask,
                       int rows, int cols, size_t pitch, int mode, const float *__restrict__ datum, int datum_cols,
                       double g, int row0) {
  const size_t n = (size_t)rows * cols;
  for (size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(s / (size_t)cols), j = (int)(s % (size_t)cols);
    double v = src[s];
    if (mode == sw2d::PACK_CELLS) {
      if (!mask[s]) v = -1.0;
    } else if (mode == sw2d::PACK_BED) {
      if (!mask[s]) v = 0.0;
      else if (datum)
        v -= (double)datum[(size_t)((row0 + i) / sw2d::DATUM_BLOCK) * datum_cols + j / sw2d::DATUM_BLOCK];
    } else if (mode == sw2d::PACK_GN2) {
      v = g * v * v;
    }
    dst[(size_t)i * pitch + j] = (T)v;
  }
}

// the padding pattern of the state block over `rows` block rows of six
// pieces of P elements (see sw2d::STATE_ARRAYS): h pieces -1, U and V pieces
// 0, and the last four columns of each piece the value of the next piece
template <typename T>
__global__ void k_slab_init(T *__restrict__ dst, size_t rows, int P) {
  const size_t RS = (size_t)sw2d::STATE_ARRAYS * P, n = rows * RS;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const int e = (int)(k % RS), piece = e / P, j = e % P;
    const int p = j >= P - 4 ? (piece + 1) % sw2d::STATE_ARRAYS : piece;
    dst[k] = p / 2 == 0 ? (T)-1 : (T)0;
  }
}

template <typename T>
__global__ void k_unpack(double *__restrict__ dst, const T *__restrict__ src, int rows, int cols, size_t pitch,
                         int clamp0) {
  const size_t n = (size_t)rows * cols;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / (size_t)cols), j = (int)(k % (size_t)cols);
    double v = (double)src[(size_t)i * pitch + j];
    if (clamp0 && v < 0.0) v = 0.0;
    dst[k] = v;
  }
}

int conv_blocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

}  // namespace

namespace sw2d {
hipError_t launch_pack(int precision, void *dst, const double *src, const unsigned char *mask, int rows, int cols,
                       size_t pitch, int mode, const float *datum, int datum_cols, double g, int row0,
                       hipStream_t s) {
  const int nb = conv_blocks((size_t)rows * cols);
  if (precision == SW2D_FP64)
    hipLaunchKernelGGL(k_pack<double>, dim3(nb), dim3(256), 0, s, (double *)dst, src, mask, rows, cols, pitch, mode,
                       datum, datum_cols, g, row0);
  else
    hipLaunchKernelGGL(k_pack<float>, dim3(nb), dim3(256), 0, s, (float *)dst, src, mask, rows, cols, pitch, mode,
                       datum, datum_cols, g, row0);
  return hipGetLastError();
}
hipError_t launch_slab_init(int precision, void *dst, size_t rows, int P, hipStream_t s) {
  const int nb = conv_blocks(rows * (size_t)STATE_ARRAYS * P);
  if (precision == SW2D_FP64)
    hipLaunchKernelGGL(k_slab_init<double>, dim3(nb), dim3(256), 0, s, (double *)dst, rows, P);
  else
    hipLaunchKernelGGL(k_slab_init<float>, dim3(nb), dim3(256), 0, s, (float *)dst, rows, P);
  return hipGetLastError();
}
hipError_t launch_unpack(int precision, double *dst, const void *src, int rows, int cols, size_t pitch, int clamp0,
                         hipStream_t s) {
  const int nb = conv_blocks((size_t)rows * cols);
  if (precision == SW2D_FP64)
    hipLaunchKernelGGL(k_unpack<double>, dim3(nb), dim3(256), 0, s, dst, (const double *)src, rows, cols, pitch,
                       clamp0);
  else
    hipLaunchKernelGGL(k_unpack<float>, dim3(nb), dim3(256), 0, s, dst, (const float *)src, rows, cols, pitch, clamp0);
  return hipGetLastError();
}
}  // namespace sw2d

void sw2d_set_error_text(const std::string &m) { g_err = m; }

extern "C" {

const char *sw2d_last_error(void) { return g_err.c_str(); }

void sw2d_params_default(sw2d_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->dt = 0.01;
  p->g = 9.80;        // L:809
  p->th = 1.0e-4;     // L:1130
  p->manning = 0.012;
  p->precision = SW2D_FP64;
  p->forcing_mode = SW2D_FORCING_REFERENCE;
}

size_t sw2d_footprint_bytes(int rows, int cols, int precision, int has_n) {
  const size_t esz = precision == SW2D_FP64 ? 8 : 4;
  const size_t P = sw2d::state_pitch(cols);
  const size_t sw = (size_t)sw2d::strip_width(precision == SW2D_FP64);
  const size_t tiles = ((size_t)cols + sw - 1) / sw * (((size_t)rows + 3) / 4);
  // the state block (h, U, V ping-pong pairs) and the static block (bed and
  // raster), with their padding rows
  size_t b = esz * sw2d::slab_rows(rows) * P * (sw2d::STATE_ARRAYS + (has_n ? 2 : 1)) + 4 * esz;
  b += tiles * 2 * sw2d::ACC_N * sizeof(double);
  b += 4 * (((size_t)rows + 63) / 64) * (((size_t)cols + 63) / 64);
  // the active-cell mask (one byte per cell), and the transient staging
  // buffer that uploads and read-backs convert through
  b += (size_t)rows * cols + sw2d::staging_bytes(rows, cols);
  return b + 65536;
}

double sw2d_friction_pow(double x) { return sw2d::pow_fric_f64(x); }

int sw2d_device_memory(size_t *free_bytes, size_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(SW2D_E_INVALID, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(SW2D_E_DEVICE, "no HIP device available");
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
  return SW2D_OK;
}

double sw2d_forcing_value(double time, const double *table, int n, double dT, int mode) {
  return sw2d::forcing_value(time, table, n, dT, mode);
}

int sw2d_evaporation_table(double albedo, const double *T, const double *Rg, int n, double dtrain,
                           double *out) {
  if (n < 0 || (n > 0 && (!T || !Rg || !out))) return fail(SW2D_E_INVALID, "evaporation table args");
  for (int i = 0; i < n; i++) {
    // gpu_evaporation_calc L:561-582 (net radiation, ITS-90 water density,
    // Henderson-Sellers latent heat, energy-balance evaporation)
    const double t = T[i];
    const double Rs = (1 - albedo) * Rg[i];
    const double pw = 999.85308 + 6.32693 * (0.01) * t - 8.523829 * (0.001) * t * t +
                      6.943248 * (0.00001) * t * t * t - 3.821216 * (0.0000001) * t * t * t * t;
    const double lv = (((t + 273.15) / ((t + 273.15) - 33.91)) * ((t + 273.15) / ((t + 273.15) - 33.91))) *
                      1.91846 * 1000000;
    double e = (Rs / (lv * pw)) * dtrain * 1000.0;
    if (e < 0) e = 0;
    out[i] = e;
  }
  return SW2D_OK;
}

int sw2d_create(const sw2d_grid *grid, const sw2d_params *params, sw2d_ctx **out) {
  if (!grid || !params || !out) return fail(SW2D_E_INVALID, "null argument");
  *out = nullptr;
  const int R = grid->rows, C = grid->cols;
  if (R <= 0 || C <= 0 || !grid->z || !grid->h0) return fail(SW2D_E_INVALID, "grid dims or arrays missing");
  if ((long long)R * C > (1LL << 33)) return fail(SW2D_E_INVALID, "grid too large");
  if (!(grid->dx > 0) || !(grid->dy > 0)) return fail(SW2D_E_INVALID, "cell size must be > 0");
  if (params->precision != SW2D_FP64 && params->precision != SW2D_FP32)
    return fail(SW2D_E_INVALID, "precision must be SW2D_FP64 or SW2D_FP32");
  if (!(params->dt > 0)) return fail(SW2D_E_INVALID, "dt must be > 0");
  if (params->cfl < 0) return fail(SW2D_E_INVALID, "cfl must be >= 0");
  // the FP32 form's friction power takes face depths above th: th must be a
  // normal float, so that such depths are too
  if (params->precision == SW2D_FP32 && !(params->th >= 1e-30))
    return fail(SW2D_E_INVALID, "th must be >= 1e-30 in the FP32 form");
  // the FP64 step takes square roots of g h for depths h > th without the
  // range handling of arguments below 2^-767
  if (params->precision == SW2D_FP64 && !(params->g * params->th >= 0x1p-760))
    return fail(SW2D_E_INVALID, "g * th must be >= 2^-760 in the FP64 form");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SW2D_E_DEVICE, "no HIP device available");

  sw2d_ctx *x = new sw2d_ctx();
  x->prec = params->precision;
  x->esz = x->prec == SW2D_FP64 ? 8 : 4;
  x->R = R;
  x->C = C;
  x->P = (int)sw2d::state_pitch(C);
  const int sw = sw2d::strip_width(x->prec == SW2D_FP64);
  x->tiles_x = (C + sw - 1) / sw;
  x->TH = sw2d::default_seg_rows(R, x->tiles_x, sw2d::step_slots(x->prec == SW2D_FP64),
                                 x->prec == SW2D_FP64 ? 6 : 3, x->prec == SW2D_FP64 ? 64 : 160);
  {
    const int cap = sw2d::reach_rows_cap((long long)sw2d::STATE_ARRAYS * x->P * x->esz);
    if (cap == 0) {
      delete x;
      return fail(SW2D_E_INVALID, "grid too wide");
    }
    x->TH = std::min(x->TH, cap);
  }
  x->tiles_y = (R + x->TH - 1) / x->TH;
  x->n_tiles = x->tiles_x * x->tiles_y;
  x->n_acc = x->tiles_x * ((R + 3) / 4);
  x->prm = *params;
  x->nodata = grid->nodata;
  x->dx = grid->dx;
  x->dy = grid->dy;
  auto bad = [&](int rc) {
    delete x;
    return rc;
  };
  if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess)
    return bad(fail(SW2D_E_DEVICE, "hipStreamCreate failed"));

  // Two blocks (see sw2d::STATE_ARRAYS): the state arrays interleaved by row
  // (row stride RS = 6 P) and the bed and raster (RSS = P or 2 P), each with
  // PAD_ROWS_ABOVE rows above row 0.  The step reads every row piece from
  // column j0 - 4 of its strip (j0 a multiple of the strip width) and writes
  // from column j0.  State rows start on a 256-byte boundary, so the written
  // row pieces are whole 32-byte sectors in FP32; the static block is only
  // read, and starts STATIC_SHIFT_COLS elements further into its allocation,
  // so its row reads are.
  x->RS = sw2d::STATE_ARRAYS * x->P;
  x->RSS = (grid->n_cell ? 2 : 1) * x->P;
  const size_t srows = sw2d::slab_rows(R);
  constexpr size_t STATIC_SHIFT_COLS = 4;
  int rc = x->alloc(&x->slab, srows * x->RS * x->esz, (size_t)sw2d::PAD_ROWS_ABOVE * x->RS * x->esz);
  rc = rc ? rc : x->alloc(&x->z, srows * x->RSS * x->esz,
                          ((size_t)sw2d::PAD_ROWS_ABOVE * x->RSS + STATIC_SHIFT_COLS) * x->esz);
  if (!rc) {
    for (int b = 0; b < 2; b++) {
      x->h[b] = (char *)x->slab + (size_t)sw2d::state_piece(0, b) * x->P * x->esz;
      x->U[b] = (char *)x->slab + (size_t)sw2d::state_piece(1, b) * x->P * x->esz;
      x->V[b] = (char *)x->slab + (size_t)sw2d::state_piece(2, b) * x->P * x->esz;
    }
    if (grid->n_cell) x->ncell = (char *)x->z + (size_t)x->P * x->esz;
  }
  rc = rc ? rc : x->alloc((void **)&x->acc, (size_t)x->n_acc * 2 * sw2d::ACC_N * sizeof(double));
  x->vol_blocks = x->blocks_for((long long)R * C);
  rc = rc ? rc : x->alloc((void **)&x->partial, (size_t)x->vol_blocks * 2 * sizeof(double));
  rc = rc ? rc : x->alloc((void **)&x->d_ctrl, sizeof(Ctrl));
  rc = rc ? rc : x->alloc((void **)&x->d_ctrl_list, sizeof(Ctrl *));
  rc = rc ? rc : x->alloc(&x->d_grid, x->prec == SW2D_FP64 ? sizeof(Grid<double>) : sizeof(Grid<float>));
  if (rc) return bad(rc);

  // active mask (init_inf L:39-44) and statics
  x->active.resize((size_t)R * C);
  for (size_t k = 0; k < (size_t)R * C; k++) x->active[k] = grid->z[k] != grid->nodata;
  rc = x->alloc((void **)&x->d_active, (size_t)R * C);
  if (rc) return bad(rc);
  if (hipMemcpy(x->d_active, x->active.data(), (size_t)R * C, hipMemcpyHostToDevice) != hipSuccess)
    return bad(fail(SW2D_E_DEVICE, "mask upload"));
  x->datum_cols = (C + sw2d::DATUM_BLOCK - 1) / sw2d::DATUM_BLOCK;
  const int datum_rows = (R + sw2d::DATUM_BLOCK - 1) / sw2d::DATUM_BLOCK;
  if (x->prec == SW2D_FP32) {
    // bed datum per 64x64 block: a multiple of 1/16 m at or below the lowest
    // active bed, so datum differences between blocks are exact in FP32
    x->datum_host.assign((size_t)datum_rows * x->datum_cols, 0.0f);
    for (int bi = 0; bi < datum_rows; bi++)
      for (int bj = 0; bj < x->datum_cols; bj++) {
        double m = std::numeric_limits<double>::infinity();
        for (int i = bi * 64; i < std::min(R, bi * 64 + 64); i++)
          for (int j = bj * 64; j < std::min(C, bj * 64 + 64); j++)
            if (x->active[(size_t)i * C + j]) m = std::min(m, grid->z[(size_t)i * C + j]);
        x->datum_host[(size_t)bi * x->datum_cols + bj] = std::isfinite(m) ? (float)(std::floor(m * 16.0) / 16.0) : 0.0f;
      }
    rc = x->alloc((void **)&x->datum, x->datum_host.size() * sizeof(float));
    if (rc) return bad(rc);
    if (hipMemcpy(x->datum, x->datum_host.data(), x->datum_host.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return bad(fail(SW2D_E_DEVICE, "datum upload"));
  }
  // the static block: 0 in the padding; the bed 0 on inactive cells and,
  // in FP32, relative to its block's datum; the raster (FP32: g n^2)
  if (hipMemsetAsync((char *)x->z - ((size_t)sw2d::PAD_ROWS_ABOVE * x->RSS + STATIC_SHIFT_COLS) * x->esz, 0,
                     srows * x->RSS * x->esz, x->stream) != hipSuccess)
    return bad(fail(SW2D_E_DEVICE, "bed clear"));
  rc = x->pack(x->z, R, C, grid->z, sw2d::PACK_BED, x->RSS);
  if (rc) return bad(rc);
  if (grid->n_cell) {
    rc = x->pack(x->ncell, R, C, grid->n_cell, x->prec == SW2D_FP64 ? sw2d::PACK_PLAIN : sw2d::PACK_GN2, x->RSS,
                 params->g);
    if (rc) return bad(rc);
  }
  // the state block: padding values everywhere (written once: the kernels
  // never write the padding), U = V = 0, then the initial depth into both
  // ping-pong buffers
  if (sw2d::launch_slab_init(x->prec, (char *)x->slab - (size_t)sw2d::PAD_ROWS_ABOVE * x->RS * x->esz,
                             srows, x->P, x->stream) != hipSuccess)
    return bad(fail(SW2D_E_DEVICE, "state clear"));
  rc = x->write_h(x->h[0], grid->h0);
  rc = rc ? rc : x->write_h(x->h[1], grid->h0);
  if (rc) return bad(rc);
  if (hipMemsetAsync(x->acc, 0, (size_t)x->n_acc * 2 * sw2d::ACC_N * sizeof(double), x->stream) != hipSuccess ||
      hipStreamSynchronize(x->stream) != hipSuccess)
    return bad(fail(SW2D_E_DEVICE, "state init"));

  if (hipMemcpy(x->d_ctrl_list, &x->d_ctrl, sizeof(Ctrl *), hipMemcpyHostToDevice) != hipSuccess)
    return bad(fail(SW2D_E_DEVICE, "control list upload"));
  Ctrl c;
  std::memset(&c, 0, sizeof(c));
  c.t = params->t0;
  c.t_end = params->t_end > 0 ? params->t_end : std::numeric_limits<double>::infinity();
  c.dt_setup = params->dt;
  c.cfl = params->cfl;
  c.forcing_mode = params->forcing_mode;
  c.evaporation_on = params->evaporation_on;
  c.dxdy = grid->dx * grid->dy;
  c.dtrain = 1.0;
  c.dtoq = 1.0;
  rc = x->store_ctrl(c);
  if (rc) return bad(rc);

  auto fill = [&](auto &G) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(G.h[0])>>;
    G.R = R;
    G.C = C;
    G.P = x->P;
    G.RS = x->RS;
    G.RSS = x->RSS;
    G.tiles_x = x->tiles_x;
    G.tiles_y = x->tiles_y;
    G.seg_rows = x->TH;
    G.edge_open = params->edge_open;
    G.has_ncell = grid->n_cell != nullptr;
    G.rain_block = 0;
    G.rain_cols = 0;
    G.datum_cols = x->datum_cols;
    G.dx = grid->dx;
    G.dy = grid->dy;
    G.g = params->g;
    G.th = params->th;
    G.manning = params->manning;
    G.INT = params->INT;
    G.INF = params->INF;
    G.LWL = params->LWL;
    G.EVmin = params->EV_WL_min;
    for (int b = 0; b < 2; b++) {
      G.h[b] = (T *)x->h[b];
      G.U[b] = (T *)x->U[b];
      G.V[b] = (T *)x->V[b];
    }
    G.z = (const T *)x->z;
    G.ncell = (const T *)x->ncell;
    G.datum = x->datum;
    G.rain_map = nullptr;
    G.unit = (T)1;
    G.ctrl = x->d_ctrl;
    G.acc = x->acc;
    return hipMemcpy(x->d_grid, &G, sizeof(G), hipMemcpyHostToDevice);
  };
  hipError_t ge;
  if (x->prec == SW2D_FP64) {
    Grid<double> G{};
    ge = fill(G);
  } else {
    Grid<float> G{};
    ge = fill(G);
  }
  if (ge != hipSuccess) return bad(fail(SW2D_E_DEVICE, "grid descriptor upload"));
  if (params->cfl > 0) {
    rc = x->recompute_cfl();
    if (rc) return bad(rc);
  }
  if (hipStreamSynchronize(x->stream) != hipSuccess) return bad(fail(SW2D_E_DEVICE, "init sync"));
  *out = x;
  return SW2D_OK;
}

void sw2d_destroy(sw2d_ctx *ctx) { delete ctx; }

static int update_grid_rain(sw2d_ctx *x, void *map, int block, int cols) {
  int shift = 0;
  while (block > 0 && (1 << shift) < block) shift++;
  // patch the rain map fields of the device grid descriptor
  if (x->prec == SW2D_FP64) {
    Grid<double> G;
    HIPCHK(hipMemcpy(&G, x->d_grid, sizeof(G), hipMemcpyDeviceToHost));
    G.rain_map = (const double *)map;
    G.rain_block = block;
    G.rain_cols = cols;
    G.rain_shift = shift;
    HIPCHK(hipMemcpy(x->d_grid, &G, sizeof(G), hipMemcpyHostToDevice));
  } else {
    Grid<float> G;
    HIPCHK(hipMemcpy(&G, x->d_grid, sizeof(G), hipMemcpyDeviceToHost));
    G.rain_map = (const float *)map;
    G.rain_block = block;
    G.rain_cols = cols;
    G.rain_shift = shift;
    HIPCHK(hipMemcpy(x->d_grid, &G, sizeof(G), hipMemcpyHostToDevice));
  }
  return SW2D_OK;
}

int sw2d_set_rain(sw2d_ctx *x, const double *table, int n, double dtrain, const double *block_map, int block) {
  if (!x || n < 0 || (n > 0 && !table)) return fail(SW2D_E_INVALID, "rain table");
  if (n > 0 && !(dtrain > 0)) return fail(SW2D_E_INVALID, "dtrain must be > 0");
  if (block_map && (block <= 0 || 64 % block != 0)) return fail(SW2D_E_INVALID, "rain block must divide 64");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  x->release(x->d_rain);
  x->d_rain = nullptr;
  if (n > 0) {
    rc = x->alloc((void **)&x->d_rain, (size_t)n * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipMemcpy(x->d_rain, table, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  }
  c.rain = x->d_rain;
  c.n_rain = n;
  if (n > 0) c.dtrain = dtrain;
  rc = x->store_ctrl(c);
  if (rc) return rc;
  x->release(x->rain_map);
  x->rain_map = nullptr;
  int cols = 0;
  if (block_map) {
    const int br = (x->R + block - 1) / block;
    cols = (x->C + block - 1) / block;
    const size_t nm = (size_t)br * cols;
    rc = x->alloc(&x->rain_map, nm * x->esz);
    if (rc) return rc;
    if (x->prec == SW2D_FP64) {
      HIPCHK(hipMemcpy(x->rain_map, block_map, nm * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> mf(block_map, block_map + nm);
      HIPCHK(hipMemcpy(x->rain_map, mf.data(), nm * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  return update_grid_rain(x, x->rain_map, block_map ? block : 0, cols);
}

int sw2d_set_sources(sw2d_ctx *x, const double *q, int n_rows, int nst, double dtoq, const int64_t *cell_index) {
  if (!x || n_rows < 0 || nst < 0 || (n_rows * nst > 0 && !q) || (nst > 0 && !cell_index))
    return fail(SW2D_E_INVALID, "source table");
  if (nst > 0 && !(dtoq > 0)) return fail(SW2D_E_INVALID, "dtoq must be > 0");
  // unique cells, the last source on a cell wins (assignment order L:545)
  std::vector<long long> cells;
  std::vector<int> cols;
  const long long ncell = (long long)x->R * x->C;
  for (int k = 0; k < nst; k++) {
    const long long cidx = cell_index[k];
    if (cidx < 0 || cidx >= ncell) continue;
    auto it = std::find(cells.begin(), cells.end(), cidx);
    if (it != cells.end()) cols[it - cells.begin()] = k;
    else {
      cells.push_back(cidx);
      cols.push_back(k);
    }
  }
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  for (void *p : {(void *)x->d_q, (void *)x->d_src_col, (void *)x->d_src_cell, (void *)x->d_ql}) x->release(p);
  x->d_q = x->d_ql = nullptr;
  x->d_src_col = nullptr;
  x->d_src_cell = nullptr;
  const int ns = (int)cells.size();
  if (ns > 0) {
    rc = x->alloc((void **)&x->d_q, (size_t)n_rows * nst * sizeof(double));
    rc = rc ? rc : x->alloc((void **)&x->d_src_col, ns * sizeof(int));
    rc = rc ? rc : x->alloc((void **)&x->d_src_cell, ns * sizeof(long long));
    rc = rc ? rc : x->alloc((void **)&x->d_ql, 2 * (size_t)ns * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipMemcpy(x->d_q, q, (size_t)n_rows * nst * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->d_src_col, cols.data(), ns * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->d_src_cell, cells.data(), ns * sizeof(long long), hipMemcpyHostToDevice));
  }
  c.q = x->d_q;
  c.n_q = n_rows;
  c.q_cols = nst;
  c.dtoq = nst > 0 ? dtoq : 1.0;
  c.nsrc = ns;
  c.src_col = x->d_src_col;
  c.src_cell = x->d_src_cell;
  c.ql_a = x->d_ql;
  c.ql_b = x->d_ql ? x->d_ql + ns : nullptr;
  return x->store_ctrl(c);
}

int sw2d_set_evaporation(sw2d_ctx *x, const double *evapo, int n) {
  if (!x || n < 0 || (n > 0 && !evapo)) return fail(SW2D_E_INVALID, "evaporation table");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  x->release(x->d_evapo);
  x->d_evapo = nullptr;
  if (n > 0) {
    rc = x->alloc((void **)&x->d_evapo, (size_t)n * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipMemcpy(x->d_evapo, evapo, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  }
  c.evapo = x->d_evapo;
  c.n_evapo = n;
  return x->store_ctrl(c);
}

int sw2d_set_dt_schedule(sw2d_ctx *x, const double *D, int n) {
  if (!x || n < 0 || (n > 0 && !D)) return fail(SW2D_E_INVALID, "dt schedule");
  for (int i = 0; i < n; i++)
    if (!(D[i] > 0)) return fail(SW2D_E_INVALID, "dt schedule entries must be > 0");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  x->release(x->d_dsched);
  x->d_dsched = nullptr;
  if (n > 0) {
    rc = x->alloc((void **)&x->d_dsched, (size_t)n * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipMemcpy(x->d_dsched, D, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  }
  c.dsched = x->d_dsched;
  c.n_dsched = n;
  return x->store_ctrl(c);
}

int sw2d_record_dt(sw2d_ctx *x, int capacity) {
  if (!x || capacity < 0) return fail(SW2D_E_INVALID, "dt history capacity");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  x->release(x->d_dthist);
  x->d_dthist = nullptr;
  x->dthist_cap = 0;
  if (capacity > 0) {
    rc = x->alloc((void **)&x->d_dthist, (size_t)capacity * sizeof(double));
    if (rc) return rc;
    x->dthist_cap = capacity;
  }
  c.dt_hist = x->d_dthist;
  c.dt_hist_cap = capacity;
  return x->store_ctrl(c);
}

int sw2d_get_dt_history(sw2d_ctx *x, double *out, int capacity, int *n_out) {
  if (!x || !out || !n_out) return fail(SW2D_E_INVALID, "dt history args");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  const long long n = std::min<long long>(c.iter, x->dthist_cap);
  if (capacity < n) return fail(SW2D_E_INVALID, "dt history buffer too small");
  if (n > 0) HIPCHK(hipMemcpy(out, x->d_dthist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  *n_out = (int)n;
  return SW2D_OK;
}

int sw2d_step(sw2d_ctx *x, int64_t n_iter) {
  if (!x || n_iter < 0) return fail(SW2D_E_INVALID, "step count");
  if (x->owner) return fail(SW2D_E_STATE, "a batch scenario is stepped by sw2d_batch_step");
  if (n_iter == 0) return SW2D_OK;
  sw2d::IterLaunch a;
  a.grids = x->d_grid;
  a.work = nullptr;
  a.n_work = x->n_tiles;
  a.blocks = x->n_tiles;
  a.ctrl = x->d_ctrl_list;
  a.n_ctrl = 1;
  a.ev = nullptr;
  if (x->prof_used >= x->prof_cap) {
    int rc = enqueue_iterations(x->prec, x->stream, a, n_iter, &x->graph);
    if (rc) return rc;
  } else {
    for (int64_t k = 0; k < n_iter; k++) {
      if (x->prof_used < x->prof_cap) {
        hipEvent_t *ev = &x->prof_ev[(size_t)4 * x->prof_used++];
        HIPCHK(hipEventRecord(ev[0], x->stream));
        a.ev = ev + 1;
      } else {
        a.ev = nullptr;
      }
      hipError_t e = launch_one(x->prec, a, x->stream);
      if (e != hipSuccess) return fail(SW2D_E_DEVICE, std::string("launch: ") + hipGetErrorString(e));
    }
  }
  HIPCHK(sw2d::launch_control(x->d_ctrl_list, 1, 1, x->stream));
  return SW2D_OK;
}

int sw2d_profile_begin(sw2d_ctx *x, int max_iters) {
  if (!x || max_iters < 0) return fail(SW2D_E_INVALID, "profile capacity");
  HIPCHK(hipStreamSynchronize(x->stream));
  while ((int)x->prof_ev.size() < 4 * max_iters) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    x->prof_ev.push_back(e);
  }
  x->prof_cap = max_iters;
  x->prof_used = 0;
  return SW2D_OK;
}

int sw2d_profile_end(sw2d_ctx *x, double ms[3], int64_t *n_iter) {
  if (!x || !ms || !n_iter) return fail(SW2D_E_INVALID, "null argument");
  HIPCHK(hipStreamSynchronize(x->stream));
  ms[0] = ms[1] = ms[2] = 0.0;
  for (int k = 0; k < x->prof_used; k++) {
    hipEvent_t *ev = &x->prof_ev[(size_t)4 * k];
    for (int q = 0; q < 3; q++) {
      float t = 0.0f;
      HIPCHK(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
      ms[q] += t;
    }
  }
  *n_iter = x->prof_used;
  x->prof_cap = x->prof_used = 0;
  return SW2D_OK;
}

int sw2d_sync(sw2d_ctx *x) {
  if (!x) return fail(SW2D_E_INVALID, "null context");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  if (c.error & sw2d::ERR_NONFINITE) return fail(SW2D_E_NONFINITE, "non-finite discharge in the state");
  if (c.error & sw2d::ERR_TOO_MANY_SRC) return fail(SW2D_E_STATE, "too many point sources in one tile");
  return SW2D_OK;
}

}  // extern "C"

namespace {
// One field into host memory.  stage: a device buffer of stage_n doubles
// that the conversion passes through (nothing is allocated), or null.
int read_field(sw2d_ctx *x, int field, double *out, double *stage, size_t stage_n) {
  if (!x || !out) return fail(SW2D_E_INVALID, "null argument");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  const int p = c.parity, R = x->R, C = x->C;
  switch (field) {
    case SW2D_FIELD_H:  // inactive cells (-1) read as 0
      return x->read_pitched(x->h[p], R, C, out, x->RS, true, stage, stage_n);
    case SW2D_FIELD_U:
      return x->read_pitched(x->U[p], R + 1, C, out, x->RS, false, stage, stage_n);
    case SW2D_FIELD_V:
      return x->read_pitched(x->V[p], R, C + 1, out, x->RS, false, stage, stage_n);
    case SW2D_FIELD_UU:
    case SW2D_FIELD_VV: {
      if (stage) return fail(SW2D_E_INVALID, "cell velocities are read with sw2d_read");
      if (!x->d_uu) {
        rc = x->alloc((void **)&x->d_uu, (size_t)R * C * sizeof(double));
        rc = rc ? rc : x->alloc((void **)&x->d_vv, (size_t)R * C * sizeof(double));
        if (rc) return rc;
      }
      HIPCHK(sw2d::launch_cell_velocity(x->prec, x->d_grid, 0, x->d_uu, x->d_vv, nullptr, x->vol_blocks,
                                        x->stream));
      HIPCHK(hipStreamSynchronize(x->stream));
      HIPCHK(hipMemcpy(out, field == SW2D_FIELD_UU ? x->d_uu : x->d_vv, (size_t)R * C * sizeof(double),
                       hipMemcpyDeviceToHost));
      return SW2D_OK;
    }
    case SW2D_FIELD_Z:
      rc = x->read_pitched(x->z, R, C, out, x->RSS, false, stage, stage_n);
      if (!rc && x->prec == SW2D_FP32)
        for (int i = 0; i < R; i++)
          for (int j = 0; j < C; j++)
            if (x->active[(size_t)i * C + j])
              out[(size_t)i * C + j] += x->datum_host[(size_t)(i / 64) * x->datum_cols + j / 64];
      if (!rc)
        for (size_t k = 0; k < (size_t)R * C; k++)
          if (!x->active[k]) out[k] = x->nodata;
      return rc;
    default:
      return fail(SW2D_E_INVALID, "unknown field");
  }
}

constexpr size_t NOALLOC_STAGE_DOUBLES = size_t(2) << 20;  // 16 MiB
}  // namespace

// The staging buffer of sw2d_read_noalloc: a fixed array in device memory,
// one copy per GPU for the life of the process.  Host code never uses the
// array's name as an address (on the host it only names the symbol): the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[NOALLOC_STAGE_DOUBLES];

namespace {
struct ScratchTable {
  std::mutex m;
  std::vector<double *> ptr;   // per device, resolved on first use
  std::vector<std::unique_ptr<std::mutex>> use;  // serialises calls per device
};
ScratchTable &scratch_table() {
  static ScratchTable t;
  return t;
}
// device address of the current device's scratch array and its lock
int scratch_for_current_device(double **p, std::mutex **lock) {
  int dev = 0, ndev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipGetDeviceCount(&ndev));
  ScratchTable &t = scratch_table();
  std::lock_guard<std::mutex> g(t.m);
  if ((int)t.ptr.size() < ndev) {
    t.ptr.resize(ndev, nullptr);
    while ((int)t.use.size() < ndev) t.use.emplace_back(new std::mutex());
  }
  if (!t.ptr[dev]) {
    void *a = nullptr;
    const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a)
      return fail(SW2D_E_DEVICE, std::string("staging symbol address: ") + hipGetErrorString(e));
    t.ptr[dev] = (double *)a;
  }
  *p = t.ptr[dev];
  *lock = t.use[dev].get();
  return SW2D_OK;
}
}  // namespace

extern "C" {

int sw2d_read(sw2d_ctx *x, int field, double *out) { return read_field(x, field, out, nullptr, 0); }

int sw2d_read_noalloc(sw2d_ctx *x, int field, double *out) {
  if (!x || !out) return fail(SW2D_E_INVALID, "null argument");
  if (field == SW2D_FIELD_UU || field == SW2D_FIELD_VV)
    return fail(SW2D_E_INVALID, "sw2d_read_noalloc reads h, U, V and z (cell velocities: sw2d_read)");
  const int cols = field == SW2D_FIELD_V ? x->C + 1 : x->C;
  if ((size_t)cols > NOALLOC_STAGE_DOUBLES)
    return fail(SW2D_E_INVALID, "sw2d_read_noalloc: a row of more than 2,097,152 values");
  double *p = nullptr;
  std::mutex *lock = nullptr;
  int rc = scratch_for_current_device(&p, &lock);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(*lock);
  return read_field(x, field, out, p, NOALLOC_STAGE_DOUBLES);
}

int sw2d_output_begin(sw2d_ctx *x, int *slot) {
  if (!x || !slot) return fail(SW2D_E_INVALID, "null argument");
  if (x->owner) return fail(SW2D_E_STATE, "a batch scenario is read with sw2d_read");
  const size_t n = (size_t)x->R * x->C;
  if (!x->d_snap) {
    int rc = x->alloc((void **)&x->d_snap, 3 * n * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipStreamCreateWithFlags(&x->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&x->snap_ready, hipEventDisableTiming));
    for (int k = 0; k < 2; k++) {
      HIPCHK(hipEventCreateWithFlags(&x->snap_done[k], hipEventDisableTiming));
      HIPCHK(hipHostMalloc((void **)&x->h_snap[k], 3 * n * sizeof(double), hipHostMallocDefault));
    }
  }
  const int k = x->snap_next;
  // the device snapshot is rewritten only after the other slot's copy read it
  if (x->snap_used[k ^ 1]) HIPCHK(hipStreamWaitEvent(x->stream, x->snap_done[k ^ 1], 0));
  HIPCHK(sw2d::launch_cell_velocity(x->prec, x->d_grid, 0, x->d_snap + 2 * n, x->d_snap + n, x->d_snap,
                                    x->vol_blocks, x->stream));
  HIPCHK(hipEventRecord(x->snap_ready, x->stream));
  HIPCHK(hipStreamWaitEvent(x->side, x->snap_ready, 0));
  HIPCHK(hipMemcpyAsync(x->h_snap[k], x->d_snap, 3 * n * sizeof(double), hipMemcpyDeviceToHost, x->side));
  HIPCHK(hipEventRecord(x->snap_done[k], x->side));
  x->snap_used[k] = true;
  x->snap_next = k ^ 1;
  *slot = k;
  return SW2D_OK;
}

int sw2d_output_wait(sw2d_ctx *x, int slot, const double **h, const double **vv, const double **uu) {
  if (!x || slot < 0 || slot > 1 || !x->snap_used[slot]) return fail(SW2D_E_INVALID, "output slot");
  HIPCHK(hipEventSynchronize(x->snap_done[slot]));
  const size_t n = (size_t)x->R * x->C;
  if (h) *h = x->h_snap[slot];
  if (vv) *vv = x->h_snap[slot] + n;
  if (uu) *uu = x->h_snap[slot] + 2 * n;
  return SW2D_OK;
}

int sw2d_write_state(sw2d_ctx *x, const double *h, const double *U, const double *V) {
  if (!x || !h || !U || !V) return fail(SW2D_E_INVALID, "null argument");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  const int p = c.parity;
  rc = x->write_h(x->h[p], h);
  rc = rc ? rc : x->write_U(x->U[p], U);
  rc = rc ? rc : x->write_V(x->V[p], V);
  if (rc) return rc;
  if (x->prm.cfl > 0) return x->recompute_cfl();
  return SW2D_OK;
}

int sw2d_get_stats(sw2d_ctx *x, sw2d_stats *st) {
  if (!x || !st) return fail(SW2D_E_INVALID, "null argument");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  HIPCHK(sw2d::launch_volume(x->prec, x->d_grid, 0, x->partial, x->vol_blocks, x->stream));
  std::vector<double> part((size_t)x->vol_blocks * 2);
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(part.data(), x->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  Comp vol;
  for (int b = 0; b < x->vol_blocks; b++) {
    vol.add(part[2 * b]);
    vol.add(part[2 * b + 1]);
  }
  std::vector<double> acc((size_t)x->n_acc * 2 * sw2d::ACC_N);
  HIPCHK(hipMemcpy(acc.data(), x->acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
  Comp q[sw2d::ACC_N];
  for (int t = 0; t < x->n_acc; t++)
    for (int k = 0; k < sw2d::ACC_N; k++) {
      q[k].add(acc[(size_t)t * 2 * sw2d::ACC_N + 2 * k]);
      q[k].add(acc[(size_t)t * 2 * sw2d::ACC_N + 2 * k + 1]);
    }
  const double a = x->dx * x->dy;  // depth-equivalent sums -> m3
  st->t = c.t;
  st->iterations = c.iter;
  st->last_D = c.D;
  st->done = c.done;
  st->volume = vol.value() * a;
  st->rain_in = q[0].value() * a;
  st->source_in = q[1].value() * a;
  st->evap_out = q[2].value() * a;
  st->outflow = q[3].value() * a;
  st->clamp_gain = q[4].value() * a;
  return SW2D_OK;
}

namespace {
struct SnapHeader {
  uint32_t magic, version;
  int32_t prec, R, C, done;
  double t;
  int64_t iter;
  uint64_t smax_bits;
  double last_D;
};
constexpr uint32_t SNAP_MAGIC = 0x44325753u;  // "SW2D"
}  // namespace

int sw2d_snapshot(sw2d_ctx *x, void *buf, size_t *size) {
  if (!x || !size) return fail(SW2D_E_INVALID, "null argument");
  const size_t nh = (size_t)x->R * x->C, nu = (size_t)(x->R + 1) * x->C, nv = (size_t)x->R * (x->C + 1);
  const size_t nacc = (size_t)x->n_acc * 2 * sw2d::ACC_N;
  const size_t need = sizeof(SnapHeader) + (nh + nu + nv) * x->esz + nacc * sizeof(double);
  if (!buf) {
    *size = need;
    return SW2D_OK;
  }
  if (*size < need) return fail(SW2D_E_INVALID, "snapshot buffer too small");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  SnapHeader hd{SNAP_MAGIC, 1, x->prec, x->R, x->C, c.done, c.t, c.iter, c.smax_bits, c.D};
  unsigned char *o = (unsigned char *)buf;
  std::memcpy(o, &hd, sizeof(hd));
  o += sizeof(hd);
  auto grab = [&](void *src, int rows, int cols) -> int {
    HIPCHK(hipMemcpy2D(o, (size_t)cols * x->esz, src, (size_t)x->RS * x->esz, (size_t)cols * x->esz, rows,
                       hipMemcpyDeviceToHost));
    o += (size_t)rows * cols * x->esz;
    return SW2D_OK;
  };
  rc = grab(x->h[c.parity], x->R, x->C);
  rc = rc ? rc : grab(x->U[c.parity], x->R + 1, x->C);
  rc = rc ? rc : grab(x->V[c.parity], x->R, x->C + 1);
  if (rc) return rc;
  HIPCHK(hipMemcpy(o, x->acc, nacc * sizeof(double), hipMemcpyDeviceToHost));
  *size = need;
  return SW2D_OK;
}

int sw2d_restore(sw2d_ctx *x, const void *buf, size_t size) {
  if (!x || !buf) return fail(SW2D_E_INVALID, "null argument");
  SnapHeader hd;
  if (size < sizeof(hd)) return fail(SW2D_E_INVALID, "snapshot too small");
  std::memcpy(&hd, buf, sizeof(hd));
  if (hd.magic != SNAP_MAGIC || hd.prec != x->prec || hd.R != x->R || hd.C != x->C)
    return fail(SW2D_E_INVALID, "snapshot does not match this context");
  const size_t nh = (size_t)x->R * x->C, nu = (size_t)(x->R + 1) * x->C, nv = (size_t)x->R * (x->C + 1);
  const size_t nacc = (size_t)x->n_acc * 2 * sw2d::ACC_N;
  if (size < sizeof(hd) + (nh + nu + nv) * x->esz + nacc * sizeof(double))
    return fail(SW2D_E_INVALID, "snapshot truncated");
  Ctrl c;
  int rc = x->fetch_ctrl(c);
  if (rc) return rc;
  const unsigned char *in = (const unsigned char *)buf + sizeof(hd);
  // only the grid's rows and columns: the padding holds the values outside
  // the grid from creation on
  auto put = [&](void *dst, int rows, int cols) -> int {
    HIPCHK(hipMemcpy2D(dst, (size_t)x->RS * x->esz, in, (size_t)cols * x->esz, (size_t)cols * x->esz, rows,
                       hipMemcpyHostToDevice));
    in += (size_t)rows * cols * x->esz;
    return SW2D_OK;
  };
  rc = put(x->h[0], x->R, x->C);
  rc = rc ? rc : put(x->U[0], x->R + 1, x->C);
  rc = rc ? rc : put(x->V[0], x->R, x->C + 1);
  if (rc) return rc;
  HIPCHK(hipMemcpy(x->acc, in, nacc * sizeof(double), hipMemcpyHostToDevice));
  c.parity = 0;
  c.run = 0;
  c.t = hd.t;
  c.iter = hd.iter;
  c.done = hd.done;
  c.smax_bits = hd.smax_bits;
  std::memset(c.smax_slot, 0, sizeof(c.smax_slot));
  c.D = hd.last_D;
  c.error = 0;
  return x->store_ctrl(c);
}

}  // extern "C"
