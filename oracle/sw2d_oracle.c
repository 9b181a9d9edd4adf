/*
 * sw2d_oracle.c -- FP64 host restatement of the SW2D-GPU time step.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle + OpenMP CPU baseline).
 * Only the test suite, the smoke check and bench.py's CPU-baseline leg load
 * this code.  The product library never links it.
 *
 * It restates, stage by stage and with the reference's operand order, the
 * kernels and the per-iteration loop of SW2D-GPU_linux/src/SW2D-GPU.cu:
 *   derive_faces   <- hm_hn L:363-399, uu1_vv1 L:405-453, uua_vva L:490-523
 *   flux_stage     <- flux L:98-354
 *   continuity     <- continuity L:603-665
 *   treat_error    <- treat_error L:695-725 (owner-computes form)
 *   cell_velocity  <- uu_vv L:459-484
 *   iteration      <- loop body L:1256-1336
 *   forcing        <- stream_flow L:526-555, evaporation_load L:589-601,
 *                     gpu_evaporation_calc L:557-587
 * Decisions where the reference reads memory outside its arrays (ghost half
 * row L:829, south faces past the allocation, east faces aliasing the next
 * row): faces outside the grid are closed walls with U = V = 0 and derived
 * face velocities 0, which is what the reference computes whenever that
 * memory reads as zero.
 *
 * Extensions beyond the reference (each reduces to it when switched off):
 *   - adaptive CFL step D <= 2*dt (cfl > 0)
 *   - open (free-outflow) edges: an outside ghost cell with bed -inf and
 *     depth 0, so the reference dry-bed rule 1a/1b (L:130-145) sets the flux
 *   - Manning raster: n_face = sqrt(0.5*(na*na + nb*nb)) (equals n when
 *     both cells carry the same n)
 *   - rain block multiplier map, piecewise-constant forcing mode
 *   - FP64 compensated volume accounting (diagnostic only)
 * Build with -ffp-contract=off so every product is rounded as written.
 */
#include "sw2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FRICTION_EXP 1.3333330 /* L:195, L:312 (not 4/3) */

/* The friction power hf^1.3333330 of L:195 and L:312.  The reference calls
 * the CUDA math library's pow (not in the reference; CUDA 10.1, its ulp
 * results unpinned).  The library pow functions of glibc and of the GPU
 * device library differ from each other in the last bit for some arguments,
 * and over a long run with wetting and drying such a bit grows through the
 * switching dry-bed rules.  This oracle and the device FP64 step therefore
 * both evaluate the power with the same sequence of IEEE operations (each
 * correctly rounded or exact), written out here on its own: log x = e ln2 +
 * 2 atanh(s) with s = (m - 1)/(m + 1), m in [1/sqrt2, sqrt2), and exp(t) =
 * 2^k exp(r) with |r| <= ln2/2, both by their Taylor series, carried in
 * double-double (Knuth two-sum, fma two-product) where a rounding would
 * reach the result.  Error below 0.65 ulp (tests/test_oracle_pins.py checks
 * it against a 60-digit decimal evaluation and against glibc pow). */
double orc_pow_fric(double x) {
  const double Y = FRICTION_EXP;
  const double LN2_HI = 0x1.62e42fee00000p-1;  /* ln 2 rounded to 32 bits */
  const double LN2_LO = 0x1.a39ef35793c76p-33; /* ln 2 - LN2_HI */
  const double INV_LN2 = 0x1.71547652b82fep+0;
  /* 1/(2j+1) for j = 12 .. 1 and 1/n! for n = 15 .. 2 */
  static const double ATANH_C[12] = {
      0x1.47ae147ae147bp-5, 0x1.642c8590b2164p-5, 0x1.8618618618618p-5, 0x1.af286bca1af28p-5,
      0x1.e1e1e1e1e1e1ep-5, 0x1.1111111111111p-4, 0x1.3b13b13b13b14p-4, 0x1.745d1745d1746p-4,
      0x1.c71c71c71c71cp-4, 0x1.2492492492492p-3, 0x1.999999999999ap-3, 0x1.5555555555555p-2};
  static const double EXP_C[14] = {
      0x1.ae7f3e733b81fp-41, 0x1.93974a8c07c9dp-37, 0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29,
      0x1.ae64567f544e4p-26, 0x1.27e4fb7789f5cp-22, 0x1.71de3a556c734p-19, 0x1.a01a01a01a01ap-16,
      0x1.a01a01a01a01ap-13, 0x1.6c16c16c16c17p-10, 0x1.1111111111111p-7,  0x1.5555555555555p-5,
      0x1.5555555555555p-3,  0x1.0000000000000p-1};
  int e;
  double m = frexp(x, &e); /* x = m 2^e, m in [1/2, 1) */
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = m * 2.0;
    e = e - 1;
  }
  double a = m - 1.0; /* exact */
  double b = m + 1.0; /* b + bl = m + 1 exactly */
  double bv = b - m;
  double bl = (m - (b - bv)) + (1.0 - bv);
  double ib = 1.0 / b; /* one division: s_hi within an ulp of a/b, s_lo corrects it */
  double s_hi = a * ib;
  double p = s_hi * b;
  double pe = fma(s_hi, b, -p);
  double s_lo = (((a - p) - pe) - s_hi * bl) * ib;
  double z = s_hi * s_hi;
  double P = ATANH_C[0];
  for (int j = 1; j < 12; j++) P = fma(P, z, ATANH_C[j]); /* Horner steps fused */
  double L_hi = 2.0 * s_hi;
  double L_lo = 2.0 * s_lo + 2.0 * (s_hi * z) * P;
  double E_hi = (double)e * LN2_HI; /* exact */
  double E_lo = (double)e * LN2_LO;
  double S = E_hi + L_hi;
  double Sv = S - E_hi;
  double Se = (E_hi - (S - Sv)) + (L_hi - Sv);
  double lo = Se + (E_lo + L_lo);
  double lx_hi = S + lo;
  double lx_lo = lo - (lx_hi - S);
  double t_hi = Y * lx_hi;
  double t_lo = fma(Y, lx_hi, -t_hi) + Y * lx_lo;
  double k = rint(t_hi * INV_LN2);
  double r_hi = t_hi - k * LN2_HI; /* exact */
  double r_lo = t_lo - k * LN2_LO;
  double r = r_hi + r_lo;
  double r_e = r_lo - (r - r_hi);
  double q = EXP_C[0];
  for (int j = 1; j < 14; j++) q = fma(q, r, EXP_C[j]);
  q = (r * r) * q;
  double one_r = 1.0 + r;
  double one_e = (1.0 - one_r) + r;
  double res = one_r + (one_e + (r_e + q));
  return ldexp(res, (int)k);
}

struct orc_sim {
  orc_params p;
  int R, C;
  /* statics */
  double *z;
  unsigned char *act;
  double *ncell; /* may be NULL */
  /* cell arrays */
  double *h, *ho, *h1, *uu, *vv;
  /* U faces (R+1) x C */
  double *U, *Uo, *hm, *uu1, *vva;
  /* V faces R x (C+1) */
  double *V, *Vo, *hn, *vv1, *uua;
  /* forcing */
  double *rain;
  int n_rain;
  double dtrain;
  double *rain_map;
  int rain_block;
  double *q;
  int n_q, nst;
  double dtoq;
  long long *src_cell;
  double *evapo;
  int n_evapo;
  double *dsched;
  int n_dsched;
  /* clock and accounting */
  double t;
  long long iters;
  double last_D;
  double acc[5][2]; /* rain, source, evap, outflow, clamp: Neumaier sum+c */
};

#define CI(i, j) ((size_t)(i) * (size_t)s->C + (size_t)(j))
#define UI(i, j) ((size_t)(i) * (size_t)s->C + (size_t)(j))
#define VI(i, j) ((size_t)(i) * (size_t)(s->C + 1) + (size_t)(j))

static void neumaier_add(double *sum_c, double x) {
  double sum = sum_c[0];
  double t = sum + x;
  if (fabs(sum) >= fabs(x))
    sum_c[1] += (sum - t) + x;
  else
    sum_c[1] += (x - t) + sum;
  sum_c[0] = t;
}

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

orc_sim *orc_create(const orc_params *p, const double *z,
                    const unsigned char *active, const double *h0,
                    const double *n_cell) {
  if (!p || p->rows <= 0 || p->cols <= 0 || !z || !active || !h0) return NULL;
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->p = *p;
  s->R = p->rows;
  s->C = p->cols;
  size_t nc = (size_t)s->R * s->C, nu = (size_t)(s->R + 1) * s->C,
         nv = (size_t)s->R * (s->C + 1);
  s->z = dalloc(nc);
  s->act = (unsigned char *)calloc(nc, 1);
  memcpy(s->z, z, nc * sizeof(double));
  memcpy(s->act, active, nc);
  if (n_cell) {
    s->ncell = dalloc(nc);
    memcpy(s->ncell, n_cell, nc * sizeof(double));
  }
  s->h = dalloc(nc); s->ho = dalloc(nc); s->h1 = dalloc(nc);
  s->uu = dalloc(nc); s->vv = dalloc(nc);
  s->U = dalloc(nu); s->Uo = dalloc(nu); s->hm = dalloc(nu);
  s->uu1 = dalloc(nu); s->vva = dalloc(nu);
  s->V = dalloc(nv); s->Vo = dalloc(nv); s->hn = dalloc(nv);
  s->vv1 = dalloc(nv); s->uua = dalloc(nv);
  memcpy(s->h, h0, nc * sizeof(double));
  /* init_inf L:37: ho = h.  Host zero-init of U/V and derived L:1137-1155.
   * initiald L:62-87: hm/hn from the initial depth. */
  memcpy(s->ho, h0, nc * sizeof(double));
  s->t = p->t0;
  s->rain_block = 0;
  /* derived faces of the initial state: hm, hn as initiald; uu1 = vv1 =
   * uua = vva = 0 (host zero-init), which is what derive gives with U=V=0 */
  for (int i = 0; i <= s->R; i++)
    for (int j = 0; j < s->C; j++) {
      double v;
      if (i == 0) v = s->h[CI(0, j)];
      else if (i == s->R) v = s->h[CI(s->R - 1, j)];
      else v = 0.50 * (s->h[CI(i, j)] + s->h[CI(i - 1, j)]);
      s->hm[UI(i, j)] = v;
    }
  for (int i = 0; i < s->R; i++)
    for (int j = 0; j <= s->C; j++) {
      double v;
      if (j == 0) v = s->h[CI(i, 0)];
      else if (j == s->C) v = s->h[CI(i, s->C - 1)];
      else v = 0.50 * (s->h[CI(i, j)] + s->h[CI(i, j - 1)]);
      s->hn[VI(i, j)] = v;
    }
  return s;
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  double *ptrs[] = {s->z, s->ncell, s->h, s->ho, s->h1, s->uu, s->vv, s->U,
                    s->Uo, s->hm, s->uu1, s->vva, s->V, s->Vo, s->hn, s->vv1,
                    s->uua, s->rain, s->rain_map, s->q, s->evapo, s->dsched};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
  free(s->act);
  free(s->src_cell);
  free(s);
}

int orc_set_rain(orc_sim *s, const double *table, int n, double dtrain,
                 const double *block_map, int block) {
  free(s->rain);
  s->rain = dalloc((size_t)n);
  if (n > 0) memcpy(s->rain, table, (size_t)n * sizeof(double));
  s->n_rain = n;
  s->dtrain = dtrain;
  free(s->rain_map);
  s->rain_map = NULL;
  s->rain_block = 0;
  if (block_map && block > 0) {
    int br = (s->R + block - 1) / block, bc = (s->C + block - 1) / block;
    s->rain_map = dalloc((size_t)br * bc);
    memcpy(s->rain_map, block_map, (size_t)br * bc * sizeof(double));
    s->rain_block = block;
  }
  return 0;
}

int orc_set_sources(orc_sim *s, const double *q, int n_rows, int nst,
                    double dtoq, const long long *cell_index) {
  free(s->q);
  free(s->src_cell);
  s->q = dalloc((size_t)n_rows * nst);
  if (n_rows * nst > 0) memcpy(s->q, q, (size_t)n_rows * nst * sizeof(double));
  s->n_q = n_rows;
  s->nst = nst;
  s->dtoq = dtoq;
  s->src_cell = (long long *)calloc(nst > 0 ? nst : 1, sizeof(long long));
  for (int k = 0; k < nst; k++) s->src_cell[k] = cell_index[k];
  return 0;
}

int orc_set_evaporation(orc_sim *s, const double *evapo, int n) {
  free(s->evapo);
  s->evapo = dalloc((size_t)n);
  if (n > 0) memcpy(s->evapo, evapo, (size_t)n * sizeof(double));
  s->n_evapo = n;
  return 0;
}

int orc_set_dt_schedule(orc_sim *s, const double *D, int n) {
  free(s->dsched);
  s->dsched = NULL;
  s->n_dsched = 0;
  if (n > 0) {
    s->dsched = dalloc((size_t)n);
    memcpy(s->dsched, D, (size_t)n * sizeof(double));
    s->n_dsched = n;
  }
  return 0;
}

/* ---------------------------------------------------------------------
 * Forcing: stream_flow L:526-555 and evaporation_load L:589-601 share one
 * interpolation form, a[it] + (a[it+1]-a[it]) / (dT*(t - dT*it)), with the
 * t <= 1 ramp a[0]*t.  Mode 1 is the piecewise-constant a[int(t/dT)].
 * Indices past the table hold the last entry (the reference reads past
 * its allocation there).
 * ------------------------------------------------------------------- */
double orc_stream_rate(double time, const double *table, int n, double dT,
                       int mode) {
  if (n <= 0) return 0.0;
  if (mode == 1) {
    int it = (int)(time / dT);
    if (it < 0) it = 0;
    if (it > n - 1) it = n - 1;
    return table[it];
  }
  if (time <= 1.00) return table[0] * time; /* L:533-534, L:547-548 */
  int it = (int)(time / dT);                /* L:537, L:551 */
  int i0 = it < 0 ? 0 : (it > n - 1 ? n - 1 : it);
  int i1 = it + 1 < 0 ? 0 : (it + 1 > n - 1 ? n - 1 : it + 1);
  return table[i0] + (table[i1] - table[i0]) / (dT * (time - dT * (it)));
}

/* gpu_evaporation_calc L:561-582, evaluated once per meteorological record */
double orc_evap_table_entry(double albedo, double T, double Rg, double dtrain) {
  double Rs = (1 - albedo) * Rg;
  double pw = 999.85308 + 6.32693 * (0.01) * T - 8.523829 * (0.001) * T * T +
              6.943248 * (0.00001) * T * T * T -
              3.821216 * (0.0000001) * T * T * T * T;
  double lv = (((T + 273.15) / ((T + 273.15) - 33.91)) *
               ((T + 273.15) / ((T + 273.15) - 33.91))) *
              1.91846 * 1000000;
  double e = (Rs / (lv * pw)) * dtrain * 1000.0;
  if (e < 0) e = 0;
  return e;
}

typedef struct {
  double rr;  /* rain [m/s] (L:554) */
  double ev;  /* evaporation [m/s] (L:599) */
} forcing_t;

static forcing_t forcing_at(orc_sim *s, double time, double *ql_cell /*R*C*/) {
  forcing_t f = {0.0, 0.0};
  int mode = s->p.forcing_mode;
  /* point sources L:532-546: assign in source order, last one wins */
  for (int k = 0; k < s->nst; k++) {
    double col[4096];
    double *tab = s->n_q <= 4096 ? col : dalloc((size_t)s->n_q);
    for (int r = 0; r < s->n_q; r++) tab[r] = s->q[(size_t)r * s->nst + k];
    double ql = orc_stream_rate(time, tab, s->n_q, s->dtoq, mode);
    if (tab != col) free(tab);
    ql = ql / (s->p.dx * s->p.dy); /* L:540 */
    long long c = s->src_cell[k];
    if (c >= 0 && c < (long long)s->R * s->C) ql_cell[c] = ql;
  }
  if (s->n_rain > 0) {
    double r = orc_stream_rate(time, s->rain, s->n_rain, s->dtrain, mode);
    f.rr = r / (s->dtrain * 1000.0); /* L:554 */
  }
  if (s->p.evaporation_on == 1 && s->n_evapo > 0) {
    double e = orc_stream_rate(time, s->evapo, s->n_evapo, s->dtrain, mode);
    f.ev = e / (s->dtrain * 1000.0); /* L:599 */
  }
  return f;
}

/* --------------------------------------------------------------------- */
/* hm_hn L:374-394, uu1_vv1 L:418-447, uua_vva L:502-519                  */
static void derive_faces(orc_sim *s) {
  const int R = s->R, C = s->C;
  const double th = s->p.th;
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= R; i++)
    for (int j = 0; j < C; j++) {
      double v;
      if (i == 0) v = s->h[CI(0, j)];
      else if (i == R) v = s->h[CI(R - 1, j)];
      else v = 0.50 * (s->h[CI(i, j)] + s->h[CI(i - 1, j)]);
      s->hm[UI(i, j)] = v;
      s->uu1[UI(i, j)] = (v >= th) ? s->U[UI(i, j)] / v : 0.00;
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; i++)
    for (int j = 0; j <= C; j++) {
      double v;
      if (j == 0) v = s->h[CI(i, 0)];
      else if (j == C) v = s->h[CI(i, C - 1)];
      else v = 0.50 * (s->h[CI(i, j)] + s->h[CI(i, j - 1)]);
      s->hn[VI(i, j)] = v;
      s->vv1[VI(i, j)] = (v >= th) ? s->V[VI(i, j)] / v : 0.0;
    }
  /* vva on U faces (L:505-506, L:514-518) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= R; i++)
    for (int j = 0; j < C; j++) {
      double a;
      if (i == 0)
        a = 0.50 * (s->vv1[VI(0, j)] + s->vv1[VI(0, j + 1)]);
      else if (i == R)
        a = 0.50 * (s->vv1[VI(R - 1, j)] + s->vv1[VI(R - 1, j + 1)]);
      else
        a = 0.250 * (s->vv1[VI(i, j)] + s->vv1[VI(i, j + 1)] +
                     s->vv1[VI(i - 1, j)] + s->vv1[VI(i - 1, j + 1)]);
      s->vva[UI(i, j)] = a;
    }
  /* uua on V faces (L:502-503, L:508-512) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; i++)
    for (int j = 0; j <= C; j++) {
      double a;
      if (j == 0)
        a = 0.50 * (s->uu1[UI(i, 0)] + s->uu1[UI(i + 1, 0)]);
      else if (j == C)
        a = 0.50 * (s->uu1[UI(i, C - 1)] + s->uu1[UI(i + 1, C - 1)]);
      else
        a = 0.250 * (s->uu1[UI(i, j)] + s->uu1[UI(i + 1, j)] +
                     s->uu1[UI(i, j - 1)] + s->uu1[UI(i + 1, j - 1)]);
      s->uua[VI(i, j)] = a;
    }
}

/* uu_vv L:472-479 (output only) */
static void cell_velocity(orc_sim *s) {
  const int R = s->R, C = s->C;
  const double th = s->p.th;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; i++)
    for (int j = 0; j < C; j++) {
      if (s->h[CI(i, j)] >= th) {
        s->uu[CI(i, j)] = (s->uu1[UI(i + 1, j)] + s->uu1[UI(i, j)]) / 2.00;
        s->vv[CI(i, j)] = (s->vv1[VI(i, j + 1)] + s->vv1[VI(i, j)]) / 2.00;
      } else {
        s->uu[CI(i, j)] = 0.00;
        s->vv[CI(i, j)] = 0.00;
      }
    }
}

static inline double face_n(const orc_sim *s, size_t a, size_t b) {
  if (!s->ncell) return s->p.manning;
  double na = s->ncell[a], nb = s->ncell[b];
  return sqrt(0.5 * (na * na + nb * nb));
}

/* Dry-bed rules 1a/1b/2 and the momentum update for one face, written for
 * the U direction (L:116-225); the V direction (L:232-342) is the same
 * sequence with the roles of the two cells and face arrays swapped, and the
 * caller supplies the transverse term. */
static inline int dry_bed(double hA /*cell i-1 or j-1*/, double zA,
                          double hB /*cell i or j*/, double zB, double g,
                          double th, double *out) {
  double hhe = hB + zB; /* L:121 */
  double hhw = hA + zA; /* L:122 */
  double hhep = hB - th;
  double hhwp = hA - th;
  if (hhe < zA) { /* rule 1a L:130-137 */
    *out = (hA > th) ? 0.5440 * hA * sqrt(g * hA) : 0;
    return 1;
  } else if (hhw < zB) { /* rule 1b L:138-145 */
    *out = (hB > th) ? -0.544 * hB * sqrt(g * hB) : 0;
    return 1;
  } else if (hhep * hhwp < 0) { /* rule 2 L:149-159 */
    if ((hB > 0) || (hA > 0)) {
      double hhan = hhep - hhwp;
      double sgnm = hhan / fabs(hhan);
      double hh3 = fmax((hB + zB), (hA + zA)) - fmax(zB, zA);
      *out = -sgnm * 0.350 * hh3 * sqrt(2.00 * g * hh3);
    } else {
      *out = 0;
    }
    return 1;
  }
  return 0;
}

/* open-edge flux: outside ghost cell with bed -inf and depth 0 makes rule
 * 1a (outside downstream of the cell) or 1b apply; sign points outward. */
static inline double open_flux(double h_in, double g, double th, double sign) {
  if (h_in > th) return sign * 0.5440 * h_in * sqrt(g * h_in);
  return 0;
}

static void flux_stage(orc_sim *s, double D) {
  const int R = s->R, C = s->C;
  const double g = s->p.g, th = s->p.th, dx = s->p.dx, dy = s->p.dy;
  const int eo = s->p.edge_open;
  /* U faces, L:116-225 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= R; i++)
    for (int j = 0; j < C; j++) {
      double um;
      if (i == 0 || i == R) {
        um = 0;
        if (i == 0 && (eo & ORC_EDGE_NORTH) && s->act[CI(0, j)])
          um = -open_flux(s->h[CI(0, j)], g, th, 1.0);
        if (i == R && (eo & ORC_EDGE_SOUTH) && s->act[CI(R - 1, j)])
          um = open_flux(s->h[CI(R - 1, j)], g, th, 1.0);
        s->U[UI(i, j)] = um;
        continue;
      }
      size_t cB = CI(i, j), cA = CI(i - 1, j);
      double hB = s->h[cB], hA = s->h[cA];
      if (!(s->act[cB] && s->act[cA] && ((hA > th) || (hB > th)))) {
        s->U[UI(i, j)] = 0; /* L:207-225 */
        continue;
      }
      if (dry_bed(hA, s->z[cA], hB, s->z[cB], g, th, &um)) {
        s->U[UI(i, j)] = um;
        continue;
      }
      const double *uu1 = s->uu1, *Uo = s->Uo, *vv1 = s->vv1;
      double hmv = s->hm[UI(i, j)];
      /* gravity L:166 */
      double u13 = g * hmv * (D / dx) * (hB + s->z[cB] - hA - s->z[cA]);
      /* convection L:172-176 */
      double uo_c = Uo[UI(i, j)];
      double uo_s = Uo[UI(i + 1, j)]; /* U(R,j) is the wall / open face */
      double uo_n = Uo[UI(i - 1, j)];
      double u11uur = 0.50 * (uu1[UI(i + 1, j)] + uu1[UI(i, j)]);
      double u11uul = 0.50 * (uu1[UI(i, j)] + uu1[UI(i - 1, j)]);
      double umr = u11uur * (uo_s + uo_c) * 0.50 + fabs(u11uur) * (uo_c - uo_s) * 0.50;
      double uml = u11uul * (uo_c + uo_n) * 0.50 + fabs(u11uul) * (uo_n - uo_c) * 0.50;
      double u11 = (D / dx) * (umr - uml);
      /* transverse L:178-190 */
      double u12;
      if (j == 0) {
        u12 = 0;
      } else {
        double uo_e = (j + 1 < C) ? Uo[UI(i, j + 1)] : 0.0;
        double uo_w = Uo[UI(i, j - 1)];
        double u12vvu = 0.50 * (vv1[VI(i, j + 1)] + vv1[VI(i - 1, j + 1)]);
        double u12vvd = 0.50 * (vv1[VI(i, j)] + vv1[VI(i - 1, j)]);
        double umu = u12vvu * (uo_e + uo_c) * 0.50 + fabs(u12vvu) * (uo_c - uo_e) * 0.50;
        double umd = u12vvd * (uo_w + uo_c) * 0.50 + fabs(u12vvd) * (uo_w - uo_c) * 0.50;
        u12 = (D / dy) * (umu - umd);
      }
      /* friction L:194-199 */
      double a = uu1[UI(i, j)], b = s->vva[UI(i, j)];
      double sqx = sqrt(a * a + b * b);
      double expx = orc_pow_fric(hmv); /* pow(hmv, 1.3333330) */
      double rn = face_n(s, cB, cA);
      double ram = g * rn * rn * sqx / expx;
      if (hmv <= th) ram = 0.00;
      /* L:203 */
      s->U[UI(i, j)] = ((1.00 - D * ram * 0.50) * uo_c + (-u11 - u12 - u13)) /
                       (1.00 + D * ram * 0.50);
    }
  /* V faces, L:232-342 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; i++)
    for (int j = 0; j <= C; j++) {
      double vn;
      if (j == 0 || j == C) {
        vn = 0;
        if (j == 0 && (eo & ORC_EDGE_WEST) && s->act[CI(i, 0)])
          vn = -open_flux(s->h[CI(i, 0)], g, th, 1.0);
        if (j == C && (eo & ORC_EDGE_EAST) && s->act[CI(i, C - 1)])
          vn = open_flux(s->h[CI(i, C - 1)], g, th, 1.0);
        s->V[VI(i, j)] = vn;
        continue;
      }
      size_t cB = CI(i, j), cA = CI(i, j - 1);
      double hB = s->h[cB], hA = s->h[cA];
      if (!(s->act[cB] && s->act[cA] && ((hA > th) || (hB > th)))) {
        s->V[VI(i, j)] = 0;
        continue;
      }
      if (dry_bed(hA, s->z[cA], hB, s->z[cB], g, th, &vn)) {
        s->V[VI(i, j)] = vn;
        continue;
      }
      const double *uu1 = s->uu1, *Vo = s->Vo, *vv1 = s->vv1;
      double hnv = s->hn[VI(i, j)];
      double v13 = g * hnv * (D / dy) * (hB + s->z[cB] - hA - s->z[cA]); /* L:283 */
      double vo_c = Vo[VI(i, j)];
      double v11;
      if (i == 0) {
        v11 = 0; /* L:289-291 */
      } else {
        double vo_s = (i + 1 < R) ? Vo[VI(i + 1, j)] : 0.0;
        double vo_n = Vo[VI(i - 1, j)];
        double v11uur = 0.50 * (uu1[UI(i + 1, j)] + uu1[UI(i + 1, j - 1)]);
        double v11uul = 0.50 * (uu1[UI(i, j)] + uu1[UI(i, j - 1)]);
        double vnr = v11uur * (vo_s + vo_c) * 0.50 + fabs(v11uur) * (vo_c - vo_s) * 0.50;
        double vnl = v11uul * (vo_c + vo_n) * 0.50 + fabs(v11uul) * (vo_n - vo_c) * 0.50;
        v11 = (D / dx) * (vnr - vnl);
      }
      /* L:303-307 */
      double vo_e = Vo[VI(i, j + 1)]; /* V(i,C): wall / open face */
      double vo_w = Vo[VI(i, j - 1)];
      double v12vvu = 0.50 * (vv1[VI(i, j + 1)] + vv1[VI(i, j)]);
      double v12vvd = 0.50 * (vv1[VI(i, j)] + vv1[VI(i, j - 1)]);
      double vnu = v12vvu * (vo_e + vo_c) * 0.50 + fabs(v12vvu) * (vo_c - vo_e) * 0.50;
      double vnd = v12vvd * (vo_w + vo_c) * 0.50 + fabs(v12vvd) * (vo_w - vo_c) * 0.50;
      double v12 = (D / dy) * (vnu - vnd);
      /* friction L:311-316 */
      double a = s->uua[VI(i, j)], b = vv1[VI(i, j)];
      double sqy = sqrt(a * a + b * b);
      double expy = orc_pow_fric(hnv); /* pow(hnv, 1.3333330) */
      double rn = face_n(s, cB, cA);
      double ram = g * rn * rn * sqy / expy;
      if (hnv <= th) ram = 0.00;
      s->V[VI(i, j)] = ((1.00 - D * ram * 0.50) * vo_c + (-v11 - v12 - v13)) /
                       (1.00 + D * ram * 0.50); /* L:320 */
    }
}

/* continuity L:627-660 over all cells; out = ho - D*(div - ql - rr*p + ev),
 * clamped at 0, zero on inactive cells.  With account != 0 the volume terms
 * of this update are accumulated (rows summed in order, Neumaier). */
static void continuity(orc_sim *s, double D, forcing_t f, const double *ql,
                       double *out, int account) {
  const int R = s->R, C = s->C;
  const double dx = s->p.dx, dy = s->p.dy;
  const double INT = s->p.INT, INF = s->p.INF, LWL = s->p.LWL,
               EVmin = s->p.EV_WL_min;
  double *rows_acc = account ? dalloc((size_t)R * 5) : NULL;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; i++) {
    double racc[5][2] = {{0}};
    for (int j = 0; j < C; j++) {
      size_t c = CI(i, j);
      double pp, evapo;
      if (s->ho[c] > EVmin) {
        pp = 1.00 - LWL;
        evapo = f.ev;
      } else {
        pp = 1.00 - (INT + INF);
        evapo = 0.000;
      }
      double rr = f.rr;
      if (s->rain_map) {
        int b = s->rain_block, bc = (C + b - 1) / b;
        rr = rr * s->rain_map[(size_t)(i / b) * bc + (j / b)];
      }
      double qlc = ql[c];
      double hv = s->ho[c] - D * ((s->U[UI(i + 1, j)] - s->U[UI(i, j)]) / dx +
                                  (s->V[VI(i, j + 1)] - s->V[VI(i, j)]) / dy -
                                  qlc - rr * pp + evapo);
      double hc = fmax(hv, 0.00);
      if (!s->act[c]) hc = 0.00;
      out[c] = hc;
      if (account && s->act[c]) {
        neumaier_add(racc[0], D * (rr * pp));
        neumaier_add(racc[1], D * qlc);
        neumaier_add(racc[2], D * evapo);
        neumaier_add(racc[4], hc - hv);
      }
    }
    if (account)
      for (int k = 0; k < 5; k++) rows_acc[(size_t)i * 5 + k] = racc[k][0] + racc[k][1];
  }
  if (account) {
    for (int i = 0; i < R; i++)
      for (int k = 0; k < 5; k++)
        if (k != 3) neumaier_add(s->acc[k], rows_acc[(size_t)i * 5 + k]);
    free(rows_acc);
    /* outflow through open edges, as depth-equivalent volume */
    double out_acc[2] = {0, 0};
    if (s->p.edge_open & ORC_EDGE_NORTH)
      for (int j = 0; j < C; j++) neumaier_add(out_acc, D * (-s->U[UI(0, j)]) / dx);
    if (s->p.edge_open & ORC_EDGE_SOUTH)
      for (int j = 0; j < C; j++) neumaier_add(out_acc, D * s->U[UI(R, j)] / dx);
    if (s->p.edge_open & ORC_EDGE_WEST)
      for (int i = 0; i < R; i++) neumaier_add(out_acc, D * (-s->V[VI(i, 0)]) / dy);
    if (s->p.edge_open & ORC_EDGE_EAST)
      for (int i = 0; i < R; i++) neumaier_add(out_acc, D * s->V[VI(i, C)] / dy);
    neumaier_add(s->acc[3], out_acc[0] + out_acc[1]);
  }
}

/* treat_error L:704-719, owner-computes: a face is zeroed when the cell it
 * drains is active and dry (h1 < th) and the flow leaves that cell. */
static void treat_error(orc_sim *s) {
  const int R = s->R, C = s->C;
  const double th = s->p.th;
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= R; i++)
    for (int j = 0; j < C; j++) {
      double u = s->U[UI(i, j)];
      if (u > 0 && i >= 1) { /* cell (i-1,j) exports southward, L:707-709 */
        size_t c = CI(i - 1, j);
        if (s->act[c] && s->h1[c] < th) u = 0;
      }
      if (u < 0.0 && i < R) { /* cell (i,j) exports northward, L:711 */
        size_t c = CI(i, j);
        if (s->act[c] && s->h1[c] < th) u = 0.00;
      }
      s->U[UI(i, j)] = u;
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; i++)
    for (int j = 0; j <= C; j++) {
      double v = s->V[VI(i, j)];
      if (v > 0 && j >= 1) { /* L:713-715 */
        size_t c = CI(i, j - 1);
        if (s->act[c] && s->h1[c] < th) v = 0;
      }
      if (v < 0 && j < C) { /* L:717 */
        size_t c = CI(i, j);
        if (s->act[c] && s->h1[c] < th) v = 0;
      }
      s->V[VI(i, j)] = v;
    }
}

/* CFL step: D = min(2*dt, cfl / max over wet faces of (|u|+sqrt(g*hface))/d).
 * A wet face has face depth > th and two active cells (open edge faces:
 * the inside cell).  Max is order-free, so the result is exact. */
static double cfl_step(orc_sim *s) {
  const int R = s->R, C = s->C;
  const double g = s->p.g, th = s->p.th, dx = s->p.dx, dy = s->p.dy;
  const int eo = s->p.edge_open;
  double smax = 0.0;
#pragma omp parallel for reduction(max : smax) schedule(static)
  for (int i = 0; i <= R; i++)
    for (int j = 0; j < C; j++) {
      int ok;
      if (i == 0) ok = (eo & ORC_EDGE_NORTH) && s->act[CI(0, j)];
      else if (i == R) ok = (eo & ORC_EDGE_SOUTH) && s->act[CI(R - 1, j)];
      else ok = s->act[CI(i, j)] && s->act[CI(i - 1, j)];
      double hf = s->hm[UI(i, j)];
      if (ok && hf > th) {
        double sp = (fabs(s->uu1[UI(i, j)]) + sqrt(g * hf)) / dx;
        if (sp > smax) smax = sp;
      }
    }
#pragma omp parallel for reduction(max : smax) schedule(static)
  for (int i = 0; i < R; i++)
    for (int j = 0; j <= C; j++) {
      int ok;
      if (j == 0) ok = (eo & ORC_EDGE_WEST) && s->act[CI(i, 0)];
      else if (j == C) ok = (eo & ORC_EDGE_EAST) && s->act[CI(i, C - 1)];
      else ok = s->act[CI(i, j)] && s->act[CI(i, j - 1)];
      double hf = s->hn[VI(i, j)];
      if (ok && hf > th) {
        double sp = (fabs(s->vv1[VI(i, j)]) + sqrt(g * hf)) / dy;
        if (sp > smax) smax = sp;
      }
    }
  double Dmax = 2.00 * s->p.dt;
  if (smax > 0.0) {
    double Dc = s->p.cfl / smax;
    if (Dc < Dmax) return Dc;
  }
  return Dmax;
}

/* one loop iteration, L:1256-1336 */
static void iteration(orc_sim *s, double *ql) {
  double D;
  if (s->n_dsched > 0)
    D = s->dsched[s->iters < s->n_dsched ? s->iters : s->n_dsched - 1];
  else if (s->p.cfl > 0)
    D = cfl_step(s);
  else
    D = 2.00 * s->p.dt; /* L:808 */
  double dth = (s->n_dsched > 0 || s->p.cfl > 0) ? 0.5 * D : s->p.dt;
  s->last_D = D;

  flux_stage(s, D);                               /* L:1268 */
  forcing_t f = forcing_at(s, s->t, ql);          /* L:1278-1285 */
  continuity(s, D, f, ql, s->h1, 0);              /* L:1287 */
  treat_error(s);                                 /* L:1294 */
  s->t = s->t + dth;                              /* L:1297 */
  f = forcing_at(s, s->t, ql);                    /* L:1302-1310 */
  continuity(s, D, f, ql, s->h, 1);               /* L:1313 */
  derive_faces(s);                                /* L:1319-1326 */
  cell_velocity(s);
  /* forward L:679-681 */
  memcpy(s->Uo, s->U, (size_t)(s->R + 1) * s->C * sizeof(double));
  memcpy(s->Vo, s->V, (size_t)s->R * (s->C + 1) * sizeof(double));
  memcpy(s->ho, s->h, (size_t)s->R * s->C * sizeof(double));
  s->t = s->t + dth; /* L:1335 */
  s->iters++;
}

int orc_set_state(orc_sim *s, const double *h, const double *U, const double *V) {
  size_t nc = (size_t)s->R * s->C, nu = (size_t)(s->R + 1) * s->C,
         nv = (size_t)s->R * (s->C + 1);
  for (size_t c = 0; c < nc; c++) s->h[c] = s->act[c] ? h[c] : 0.0;
  memcpy(s->U, U, nu * sizeof(double));
  memcpy(s->V, V, nv * sizeof(double));
  derive_faces(s);
  cell_velocity(s);
  memcpy(s->Uo, s->U, nu * sizeof(double));
  memcpy(s->Vo, s->V, nv * sizeof(double));
  memcpy(s->ho, s->h, nc * sizeof(double));
  return 0;
}

int orc_step(orc_sim *s, int n_iter) {
  if (!s || n_iter < 0) return -1;
  double *ql = dalloc((size_t)s->R * s->C); /* h_ql, zero-initialised L:1149 */
  for (int k = 0; k < n_iter; k++) iteration(s, ql);
  free(ql);
  return 0;
}

int orc_read(const orc_sim *s, int field, double *out) {
  size_t nc = (size_t)s->R * s->C, nu = (size_t)(s->R + 1) * s->C,
         nv = (size_t)s->R * (s->C + 1);
  const double *src;
  size_t n;
  switch (field) {
    case ORC_H: src = s->h; n = nc; break;
    case ORC_U: src = s->U; n = nu; break;
    case ORC_V: src = s->V; n = nv; break;
    case ORC_UU: src = s->uu; n = nc; break;
    case ORC_VV: src = s->vv; n = nc; break;
    case ORC_HM: src = s->hm; n = nu; break;
    case ORC_HN: src = s->hn; n = nv; break;
    case ORC_UU1: src = s->uu1; n = nu; break;
    case ORC_VV1: src = s->vv1; n = nv; break;
    case ORC_UUA: src = s->uua; n = nv; break;
    case ORC_VVA: src = s->vva; n = nu; break;
    default: return -1;
  }
  memcpy(out, src, n * sizeof(double));
  return 0;
}

int orc_get_stats(const orc_sim *s, orc_stats *st) {
  double vol[2] = {0, 0};
  for (size_t c = 0; c < (size_t)s->R * s->C; c++)
    if (s->act[c]) neumaier_add(vol, s->h[c]);
  double area = s->p.dx * s->p.dy;
  st->t = s->t;
  st->iterations = s->iters;
  st->last_D = s->last_D;
  st->volume = (vol[0] + vol[1]) * area;
  st->rain_in = (s->acc[0][0] + s->acc[0][1]) * area;
  st->source_in = (s->acc[1][0] + s->acc[1][1]) * area;
  st->evap_out = (s->acc[2][0] + s->acc[2][1]) * area;
  st->outflow = (s->acc[3][0] + s->acc[3][1]) * area;
  st->clamp_gain = (s->acc[4][0] + s->acc[4][1]) * area;
  return 0;
}
