/*
 * sw2d_oracle.h -- FP64 host restatement of the SW2D-GPU time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * OpenMP CPU baseline.  Only the test suite, the smoke check and the CPU
 * baseline leg of bench.py may load it; the product library never links it.
 *
 * Reference: SW2D-GPU_linux/src/SW2D-GPU.cu (kernels L:29-725, loop
 * L:1247-1336).  Every stage in sw2d_oracle.c cites the lines it restates.
 *
 * Layout (row-major, i = DEM row north->south, j = DEM column west->east):
 *   cell arrays   R x C       index i*C + j
 *   U faces       (R+1) x C   U(i,j) between cells (i-1,j) and (i,j)
 *   V faces       R x (C+1)   V(i,j) between cells (i,j-1) and (i,j)
 * Faces outside the grid are walls (value 0) unless the edge is open.
 */
#ifndef SW2D_ORACLE_H
#define SW2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* edge bits for open (free-outflow) boundaries */
#define ORC_EDGE_NORTH 1 /* U faces of row 0      */
#define ORC_EDGE_SOUTH 2 /* U faces of row R      */
#define ORC_EDGE_WEST 4  /* V faces of column 0   */
#define ORC_EDGE_EAST 8  /* V faces of column C   */

typedef struct {
  int rows, cols;
  double dx, dy;
  double dt;      /* setup.dat half step; one iteration advances 2*dt     */
  double g;       /* 9.80  (L:809)                                        */
  double th;      /* 1e-4  (L:1130)                                       */
  double manning; /* scalar Manning n, used when no raster is given        */
  double INT, INF, LWL, EV_WL_min;
  int evaporation_on;
  double cfl;     /* 0: fixed D = 2*dt; >0: Courant number, D <= 2*dt     */
  int edge_open;  /* ORC_EDGE_* bitmask                                   */
  int forcing_mode; /* 0: reference formula (L:538), 1: piecewise constant */
  double t0;
} orc_params;

typedef struct orc_sim orc_sim;

/* z: R*C bed; active: R*C (1 = active); h0: R*C initial depth;
 * n_cell: optional R*C Manning raster (NULL = scalar). Arrays are copied. */
orc_sim *orc_create(const orc_params *p, const double *z,
                    const unsigned char *active, const double *h0,
                    const double *n_cell);
void orc_destroy(orc_sim *s);

/* rain table: values in mm per dtrain (rain.dat); spatial multiplier map
 * (optional, per block of block x block cells, row-major over blocks). */
int orc_set_rain(orc_sim *s, const double *table, int n, double dtrain,
                 const double *block_map, int block);
/* point sources: q[n_rows*nst] in m3/s (row-major by time), dtoq, and the
 * linear cell index of each source computed as in L:542-545 (-1 = none). */
int orc_set_sources(orc_sim *s, const double *q, int n_rows, int nst,
                    double dtoq, const long long *cell_index);
/* evaporation table in mm per dtrain (output of the L:557-587 formula) */
int orc_set_evaporation(orc_sim *s, const double *evapo, int n);
/* replay a fixed schedule of D values (overrides CFL) */
int orc_set_dt_schedule(orc_sim *s, const double *D, int n);

/* overwrite the state (h: R*C, U: (R+1)*C, V: R*(C+1)); the derived face
 * fields are rebuilt from it, inactive cells keep h = 0 */
int orc_set_state(orc_sim *s, const double *h, const double *U, const double *V);

/* run n iterations; each iteration = one update of step D */
int orc_step(orc_sim *s, int n_iter);

/* field ids */
#define ORC_H 0   /* R*C */
#define ORC_U 1   /* (R+1)*C */
#define ORC_V 2   /* R*(C+1) */
#define ORC_UU 3  /* R*C cell-centre velocity along i (L:459-484) */
#define ORC_VV 4  /* R*C cell-centre velocity along j */
#define ORC_HM 5
#define ORC_HN 6
#define ORC_UU1 7
#define ORC_VV1 8
#define ORC_UUA 9
#define ORC_VVA 10
int orc_read(const orc_sim *s, int field, double *out);

typedef struct {
  double t;
  long long iterations;
  double last_D;
  double volume;      /* sum h * dx * dy over active cells [m3]           */
  double rain_in;     /* accumulated volumes [m3]                          */
  double source_in;
  double evap_out;
  double outflow;
  double clamp_gain;
} orc_stats;
int orc_get_stats(const orc_sim *s, orc_stats *st);

/* forcing functions restated from L:526-555 and L:589-601 */
/* hf^1.3333330, the friction power of L:195 / L:312 (see the source) */
double orc_pow_fric(double x);
double orc_stream_rate(double time, const double *table, int n, double dT,
                       int mode);
double orc_evap_table_entry(double albedo, double T, double Rg, double dtrain);

#ifdef __cplusplus
}
#endif
#endif
