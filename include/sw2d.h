/*
 * sw2d.h -- C ABI of the MI355X-native SW2D time-stepping core.
 *
 * The reference (LabHidro/SW2D-GPU, one CUDA translation unit) has no
 * library interface: its path is reached through the `SW2D-GPU` executable,
 * the case files it reads, and its kernel signatures.  Each entry point
 * below names the reference interface it replaces
 * (L:n = SW2D-GPU_linux/src/SW2D-GPU.cu line n).
 *
 * Conventions: every call returns 0 on success or a negative SW2D_E_* code;
 * sw2d_last_error() gives the message of the last failure on the calling
 * thread.  The library owns device memory; callers own host buffers.  Each
 * context has its own HIP stream; calls on one context are not thread-safe,
 * distinct contexts are independent.
 *
 * Grid layout seen through this ABI (row-major, double precision):
 *   cells   rows x cols          i = DEM row (north->south), j = column
 *   U faces (rows+1) x cols      U(i,j) between cells (i-1,j) and (i,j)
 *   V faces rows x (cols+1)      V(i,j) between cells (i,j-1) and (i,j)
 */
#ifndef SW2D_H
#define SW2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW2D_OK 0
#define SW2D_E_INVALID -1   /* bad argument or inconsistent sizes       */
#define SW2D_E_DEVICE -2    /* HIP runtime failure / no usable GPU      */
#define SW2D_E_IO -3        /* missing or malformed case file           */
#define SW2D_E_MEMORY -4    /* allocation failed or over the budget     */
#define SW2D_E_STATE -5     /* call not valid in the current state      */
#define SW2D_E_NONFINITE -6 /* a non-finite value reached the state     */

/* open (free-outflow) edges; closed walls otherwise */
#define SW2D_EDGE_NORTH 1 /* U faces of row 0    */
#define SW2D_EDGE_SOUTH 2 /* U faces of row rows */
#define SW2D_EDGE_WEST 4  /* V faces of column 0 */
#define SW2D_EDGE_EAST 8  /* V faces of column cols */

#define SW2D_FP64 0 /* all state and arithmetic in double (parity path)     */
#define SW2D_FP32 1 /* float state, bed relative to a per-tile datum, FP64
                       threshold tests and volume accumulators              */

#define SW2D_FORCING_REFERENCE 0 /* interpolation formula of L:538/L:552   */
#define SW2D_FORCING_STEP 1      /* piecewise constant table[int(t/dT)]    */

typedef struct sw2d_ctx sw2d_ctx;

/* Grid and statics.  Replaces the host arrays h_baseo / h_initial_condition
 * and the launch set-up of L:815-834, L:1022-1033, L:1163-1241. */
typedef struct {
  int rows, cols;
  double dx, dy;           /* cellsize (L:826, L:1236-1237)                */
  double nodata;           /* NODATA_value: z == nodata marks an inactive
                              cell (init_inf L:39-44)                      */
  const double *z;         /* rows*cols bed elevation                      */
  const double *h0;        /* rows*cols initial depth                      */
  const double *n_cell;    /* optional rows*cols Manning raster, or NULL   */
} sw2d_grid;

/* Scalars of setup.dat (L:788-809, L:1130) plus the extensions. */
typedef struct {
  double dt;               /* half step; one iteration advances 2*dt       */
  double g;                /* 9.80 in the reference (L:809)                */
  double th;               /* dry threshold, 1e-4 (L:1130); FP32: >= 1e-30,
                              FP64: g * th >= 2^-760                      */
  double manning;          /* scalar n (L:799), used without a raster      */
  double INT, INF, LWL, EV_WL_min; /* runoff fractions (L:627-634)         */
  int evaporation_on;
  double cfl;              /* 0: fixed step 2*dt; >0 Courant number with
                              the step capped at 2*dt                      */
  int edge_open;           /* SW2D_EDGE_* bitmask                          */
  int forcing_mode;        /* SW2D_FORCING_*                               */
  int precision;           /* SW2D_FP64 or SW2D_FP32                       */
  double t0;               /* start time (L:1082)                          */
  double t_end;            /* stop when t + dt > t_end (L:1256); <= 0: none */
} sw2d_params;

/* Fills p with the reference defaults (g = 9.80, th = 1e-4, FP64, ...). */
void sw2d_params_default(sw2d_params *p);

/* Allocates the device state and uploads the statics; replaces the
 * cudaMalloc/cudaMemcpy block L:1163-1215 and init kernels init_inf L:29,
 * initiald L:52. */
int sw2d_create(const sw2d_grid *grid, const sw2d_params *params,
                sw2d_ctx **out);
void sw2d_destroy(sw2d_ctx *ctx);

/* Forcing tables, kept on the device and evaluated there each half step.
 * Replaces stream_flow L:526-555 and the per-step H2D copies of d_ql/d_rr
 * (L:1278-1280, L:1302-1305).  table_mm: rain.dat values [mm per dtrain].
 * block_map (optional): rain multiplier per block x block cells, row-major
 * over ceil(rows/block) x ceil(cols/block) blocks; block must divide 64. */
int sw2d_set_rain(sw2d_ctx *ctx, const double *table_mm, int n, double dtrain,
                  const double *block_map, int block);
/* q: n_rows x nst discharges [m3/s] (Q_source_sink.dat), dtoq, and each
 * source's linear cell index computed as in L:542-545 (indices outside
 * 0..rows*cols-1 are ignored; a later source on the same cell wins). */
int sw2d_set_sources(sw2d_ctx *ctx, const double *q, int n_rows, int nst,
                     double dtoq, const int64_t *cell_index);
/* evapo_mm: evaporation per dtrain [mm] (gpu_evaporation_calc L:557-587
 * output, see sw2d_evaporation_table); replaces evaporation_load L:589. */
int sw2d_set_evaporation(sw2d_ctx *ctx, const double *evapo_mm, int n);
/* Replay a fixed sequence of steps D (overrides cfl); n = 0 clears it. */
int sw2d_set_dt_schedule(sw2d_ctx *ctx, const double *D, int n);
/* Record the step D of every iteration into a ring of `capacity` entries
 * (0 disables); read back with sw2d_get_dt_history. */
int sw2d_record_dt(sw2d_ctx *ctx, int capacity);
int sw2d_get_dt_history(sw2d_ctx *ctx, double *out, int capacity, int *n_out);

/* Enqueue n iterations of the loop body L:1256-1336 (flux, continuity,
 * treat_error, continuity, derived fields, forward) on the context stream.
 * Iterations past t_end are skipped on the device.  Asynchronous. */
int sw2d_step(sw2d_ctx *ctx, int64_t n_iter);
/* Kernel timing (no reference counterpart; the reference times the whole
 * loop with std::clock, L:1251-1252, L:1477).  After sw2d_profile_begin,
 * each sw2d_step records HIP events on the context stream around the two
 * launches of every iteration (up to max_iters iterations);
 * sw2d_profile_end returns the summed milliseconds of [control kernel,
 * stencil step kernel, 0 (reserved)] and the number of iterations timed. */
int sw2d_profile_begin(sw2d_ctx *ctx, int max_iters);
int sw2d_profile_end(sw2d_ctx *ctx, double ms_out[3], int64_t *n_iter);

/* Wait for queued work; returns SW2D_E_NONFINITE if the state went bad. */
int sw2d_sync(sw2d_ctx *ctx);

#define SW2D_FIELD_H 0  /* rows*cols depth (inactive cells read 0)         */
#define SW2D_FIELD_U 1  /* (rows+1)*cols unit discharge on U faces         */
#define SW2D_FIELD_V 2  /* rows*(cols+1) unit discharge on V faces         */
#define SW2D_FIELD_UU 3 /* rows*cols cell velocity along i (uu_vv L:459)   */
#define SW2D_FIELD_VV 4 /* rows*cols cell velocity along j                 */
#define SW2D_FIELD_Z 5  /* rows*cols bed as used on the device             */
/* Copies one field to host memory in double precision (synchronises). */
int sw2d_read(sw2d_ctx *ctx, int field, double *host_out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc).  Same as sw2d_read for SW2D_FIELD_H, _U, _V and _Z,
 * but the conversion passes through a fixed 16 MiB staging array in device
 * memory (one per GPU, held until the process ends) instead of a buffer
 * allocated per call: it calls no hipMalloc, hipFree or allocating library
 * routine.  Limit: rows of at most 2,097,152 values; the cell velocities
 * (SW2D_FIELD_UU, _VV) are refused (use sw2d_read).  Calls on one GPU share
 * the array: within a process they are serialised; calls that overlap on one
 * GPU from several processes race.  Synchronises. */
int sw2d_read_noalloc(sw2d_ctx *ctx, int field, double *host_out);
/* Output snapshots that overlap the copy to host memory with the next
 * iterations (the reference synchronises, copies h, uu, vv and writes its
 * files while the GPU idles, L:1339-1472).  sw2d_output_begin enqueues, on
 * the context stream after the work queued so far, the depth (inactive
 * cells 0) and the cell velocities uu_vv (L:459-484) of the state reached
 * into a device snapshot, whose copy into one of two pinned host slots runs
 * on a side stream; it returns at once with the slot.  sw2d_output_wait
 * waits for that copy and returns rows*cols arrays in the slot (valid until
 * sw2d_output_begin hands out the slot again, two calls later).  Not for
 * batch scenarios. */
int sw2d_output_begin(sw2d_ctx *ctx, int *slot);
int sw2d_output_wait(sw2d_ctx *ctx, int slot, const double **h, const double **vv, const double **uu);
/* Overwrites the state (h, U, V) from host arrays (synchronises). */
int sw2d_write_state(sw2d_ctx *ctx, const double *h, const double *U,
                     const double *V);

typedef struct {
  double t;                /* model time                                   */
  int64_t iterations;      /* completed loop iterations                    */
  double last_D;           /* step of the last iteration                   */
  int done;                /* t_end reached                                */
  double volume;           /* sum of h*dx*dy over active cells [m3]        */
  double rain_in, source_in, evap_out, outflow, clamp_gain; /* [m3]        */
} sw2d_stats;
int sw2d_get_stats(sw2d_ctx *ctx, sw2d_stats *out);

/* Snapshot of (h, U, V, t, iteration, accumulators) in a host buffer;
 * with buf == NULL only *size is set.  Restoring gives a bitwise
 * identical continuation. */
int sw2d_snapshot(sw2d_ctx *ctx, void *buf, size_t *size);
int sw2d_restore(sw2d_ctx *ctx, const void *buf, size_t size);

/* Free and total memory of the current device (hipMemGetInfo). */
int sw2d_device_memory(size_t *free_bytes, size_t *total_bytes);

/* Device bytes a context with these dimensions and precision needs: its
 * arrays, the active-cell mask and the peak of the transient staging buffer
 * (8 B per cell) that uploads and read-backs pass through. */
size_t sw2d_footprint_bytes(int rows, int cols, int precision, int has_n);

const char *sw2d_last_error(void);

/* Helpers restating host-side formulas of the reference. */
/* gpu_evaporation_calc L:561-582 over n meteorological records. */
int sw2d_evaporation_table(double albedo, const double *T, const double *Rg,
                           int n, double dtrain, double *evapo_mm_out);
/* x^1.3333330 for 0 < x < 1e300, the friction power of flux (L:195,
 * L:312), as the FP64 step computes it on the device: a sequence of IEEE
 * operations that gives the same bits on host and device (< 0.65 ulp). */
double sw2d_friction_pow(double x);
/* stream_flow / evaporation_load interpolation (L:533-538, L:547-552). */
double sw2d_forcing_value(double time, const double *table, int n, double dT,
                          int mode);

/* ---- batched ensembles (no reference counterpart: the reference runs the
 * cases of dir.txt one after another, L:775-1526) ---------------------- */
/* Many scenarios, each a full sw2d_ctx, advanced together: each iteration's
 * step launch covers every live scenario (a work-list of 56-column strip
 * segments of `rows` rows, chosen per scenario), and chunks of iterations
 * replay a captured HIP graph.  Every scenario keeps its own clock, step D
 * and end time. */
typedef struct sw2d_batch sw2d_batch;
int sw2d_batch_create(int precision, sw2d_batch **out);
void sw2d_batch_destroy(sw2d_batch *b);
/* Adds a scenario (same arguments as sw2d_create, params->precision must
 * match).  share_with >= 0 names an earlier scenario with the same bed and
 * Manning raster: its static arrays are shared, and tiles of the two are
 * scheduled next to each other so the statics are read once from HBM. */
int sw2d_batch_add(sw2d_batch *b, const sw2d_grid *grid, const sw2d_params *params,
                   int share_with, int *id_out);
/* The scenario's context (for set_rain/read/stats/snapshot).  Waits for
 * the batch stream.  Changing forcing tables needs sw2d_batch_touch. */
int sw2d_batch_scenario(sw2d_batch *b, int id, sw2d_ctx **ctx);
/* Call after changing any scenario through its context (forcing tables,
 * sw2d_restore, sw2d_write_state): rebuilds the launch descriptors and
 * returns finished scenarios to the work-list. */
int sw2d_batch_touch(sw2d_batch *b);
/* Removes a scenario from the work-list and frees its device memory. */
int sw2d_batch_remove(sw2d_batch *b, int id);
/* Takes a scenario out of the work-list without destroying it: *ctx_out is
 * its context, owned by the caller from then on (sw2d_destroy frees it),
 * with its state, clock and accumulators as they are; it can be read, or
 * stepped on its own.  Waits for the batch stream only: no device memory
 * is freed, so other streams are not stalled (freeing device memory waits
 * for the whole device).  Refused (SW2D_E_STATE) for a scenario whose bed
 * and raster belong to another scenario, or are shared by one. */
int sw2d_batch_detach(sw2d_batch *b, int id, sw2d_ctx **ctx_out);
/* Adds an existing context (one not in a batch, of the batch's precision)
 * as a scenario, with its state, clock, forcing and accumulators as they
 * are; the batch owns it from then on.  It joins the work-list at the next
 * sw2d_batch_step: nothing waits for the batch stream, so a context can be
 * created and given its forcing while the batch runs.  *id_out is the
 * lowest id that no scenario holds: ids freed by sw2d_batch_detach or
 * sw2d_batch_remove are given out again. */
int sw2d_batch_attach(sw2d_batch *b, sw2d_ctx *ctx, int *id_out);
/* Segment length (rows) of one scenario's work items: a multiple of 4 in
 * [4, 4096]; 0 = the autotuned choice for its size bucket, or before any
 * tuning the batch default: the scenario's single-grid length raised to the
 * length that gives the whole live set about twelve rounds of the device's
 * concurrent wavefronts (recomputed when the live set changes). */
int sw2d_batch_set_tile_rows(sw2d_batch *b, int id, int rows);
/* Chooses the segment length of each size bucket (floor(log2(cells)/2)):
 * with one iteration of every live scenario decided but not committed, the
 * whole batch's step kernel is timed over `iters` launches for the
 * bucket's current length (the incumbent) and each candidate (16, 32, 64,
 * 128 rows), one bucket at a time, largest bucket first.  A candidate
 * replaces the incumbent only when it is more than 3 % faster.  Control
 * blocks and volume accumulators are restored afterwards, so the batch
 * state is bitwise unchanged. */
int sw2d_batch_autotune(sw2d_batch *b, int iters);
int sw2d_batch_tile_rows(sw2d_batch *b, int id, int *rows);
int sw2d_batch_step(sw2d_batch *b, int64_t n_iter);
int sw2d_batch_sync(sw2d_batch *b);
/* *idle = 1 when the batch stream has no work left (does not wait). */
int sw2d_batch_idle(sw2d_batch *b, int *idle);
/* done[k] = 1 when scenario k (ids 0..n-1, removed ones report 1) has
 * reached its t_end; iters[k] = its committed iterations (either may be
 * NULL).  Synchronises.  Scenarios seen finished leave the work-list at
 * the next step (sw2d_batch_touch returns them). */
int sw2d_batch_status(sw2d_batch *b, int *done, int64_t *iters, int cap, int *n_out);

/* ---- serving loop (no reference counterpart) -----------------------------
 * Continuous batching of scenario requests on one device.  Requests queue
 * on submit (the server keeps host copies of their inputs) and are admitted
 * in arrival order while their device footprint (sw2d_footprint_bytes) fits
 * in the memory budget and a slot is free: admission creates the request's
 * context (its device state).  Admitted requests become active at once,
 * or, with a target set by sw2d_server_set_active_cells, least estimated
 * work (cells x simulated seconds, divided by 1 + seconds waited / 60)
 * first while the active ones hold fewer cells than the target (or none is
 * active).  Active scenarios run in one batch per size class (cells < 2^18,
 * < 2^21, larger), each batch on its own stream, so the classes advance
 * concurrently and each at its own rate: a small scenario is not held to
 * the launch rate of the large ones.  Each sw2d_server_advance waits until
 * a class has finished its last chunk of iterations, a request can become
 * active, or the head of the queue fits; then it takes the scenarios that
 * reached their t_end out of the idle classes, activates waiting requests,
 * enqueues each idle class's next chunk of iterations without waiting
 * (n_iter times 128 for the smallest class, 32 for the middle one, 1 for
 * the largest),
 * admits queued requests, and reads the finished scenarios' final h, U, V
 * and statistics into host memory.  A scenario's result equals the same
 * scenario run alone (sw2d_create + sw2d_step), bitwise in h, U and V. */
typedef struct sw2d_server sw2d_server;
#define SW2D_REQ_QUEUED 0
#define SW2D_REQ_RUNNING 1
#define SW2D_REQ_DONE 2
typedef struct {
  int state;               /* SW2D_REQ_*                                   */
  int rows, cols;
  double submit_s;         /* server clock (seconds since create)          */
  double admit_s;          /* -1 until admitted                            */
  double finish_s;         /* -1 until retired                             */
  int64_t iterations;      /* committed iterations (when retired)          */
  double start_s;          /* -1 until active (its first chunk of steps)   */
} sw2d_request_info;
/* mem_budget: device bytes the admitted scenarios may hold; max_slots:
 * concurrent scenarios (0 = no limit). */
int sw2d_server_create(int precision, size_t mem_budget, int max_slots, sw2d_server **out);
void sw2d_server_destroy(sw2d_server *s);
/* Queues a request: grid and parameters as sw2d_create (params->t_end > 0
 * required), rain table [mm per dtrain] as sw2d_set_rain (n_rain may be 0).
 * Fails with SW2D_E_MEMORY when the request alone exceeds the budget. */
int sw2d_server_submit(sw2d_server *s, const sw2d_grid *grid, const sw2d_params *params,
                       const double *rain_mm, int n_rain, double dtrain, int *req_id);
/* Activation target in cells: admitted requests wait until the active ones
 * hold fewer cells, then start least estimated work first (shorter median
 * time-to-result, at some cost in throughput); 0 (the default): every
 * admitted request is active. */
int sw2d_server_set_active_cells(sw2d_server *s, int64_t cells);
/* Retire, activate, step and admit (see above).  Optional outputs after
 * the call: requests admitted and not finished (active or waiting for the
 * device), and requests queued for admission. */
int sw2d_server_advance(sw2d_server *s, int64_t n_iter, int *n_running, int *n_queued);
int sw2d_server_info(sw2d_server *s, int req_id, sw2d_request_info *info);
/* Final field (SW2D_FIELD_H, _U or _V) or statistics of a retired request. */
int sw2d_server_read(sw2d_server *s, int req_id, int field, double *host_out);
int sw2d_server_stats(sw2d_server *s, int req_id, sw2d_stats *out);
/* Drops a retired request's host result. */
int sw2d_server_release(sw2d_server *s, int req_id);

/* ---- reference case directories (db/<case>/input, output) ------------ */
typedef struct sw2d_case sw2d_case;

typedef struct {
  int rows, cols;
  double xll, yll, cellsize, nodata;
  double tstart[4], tend[4];     /* day hour min sec                       */
  double dt, dpout, dkout, dtoq;
  int evaporation_on;
  double EV_WL_min, INT, INF, LWL, manning_coef;
  int out_velocity_x, out_velocity_y, out_elevation, out_depth,
      out_outlet_on;
  int nst, n_q_rows, n_rain, n_out;
  double dtrain, albedo;
  double time0, timmax;          /* L:1082-1087                            */
} sw2d_case_info;

/* Parses db/<case>/input/ exactly as L:779-1077 reads it. */
int sw2d_case_load(const char *case_dir, sw2d_case **out);
void sw2d_case_free(sw2d_case *c);
int sw2d_case_get_info(const sw2d_case *c, sw2d_case_info *info);

#define SW2D_CASE_DEM 0       /* rows*cols                                 */
#define SW2D_CASE_IC 1        /* rows*cols                                 */
#define SW2D_CASE_RAIN 2      /* n_rain                                    */
#define SW2D_CASE_Q 3         /* n_q_rows*nst                              */
#define SW2D_CASE_SRC_XY 4    /* 2*nst, divided by cellsize (L:853-854)    */
#define SW2D_CASE_OUT_XY 5    /* 2*n_out, divided by cellsize (L:1065)     */
#define SW2D_CASE_TEMP 6      /* n_rain                                    */
#define SW2D_CASE_RG 7        /* n_rain                                    */
#define SW2D_CASE_EVAPO 8     /* n_rain (derived, L:557-587)               */
/* Copies one parsed array; *n_out receives its length. */
int sw2d_case_get_array(const sw2d_case *c, int which, double *out,
                        size_t cap, size_t *n_out);
/* Linear cell index of each source as L:542-545 computes it. */
int sw2d_case_source_cells(const sw2d_case *c, int64_t *out, int cap);
/* Creates a context for the case with its setup.dat scalars and forcing
 * tables; params may override precision/cfl/edges (NULL = defaults). */
int sw2d_case_create(const sw2d_case *c, const sw2d_params *overrides,
                     sw2d_ctx **out);

/* Writes Results_<k>.vtk in the byte format of L:1364-1472. depth, vv, uu,
 * z are rows*cols arrays in DEM order. */
int sw2d_write_vtk(const char *path, const sw2d_case *c, const double *depth,
                   const double *vv, const double *uu, const double *z);

/* The drop-in driver: runs every case listed in <db_root>/dir.txt as the
 * reference main() (L:728-1530) does, writing the same output files.
 * Progress lines go to stdout.  Returns 0 or the first error. */
int sw2d_run_db(const char *db_root, const sw2d_params *overrides);

#ifdef __cplusplus
}
#endif
#endif
