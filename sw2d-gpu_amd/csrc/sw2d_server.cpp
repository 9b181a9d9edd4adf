// sw2d_server.cpp -- serving loop: continuous batching of flood-scenario
// requests over one batch per size class.  Built only on the public C ABI
// (sw2d.h); no reference counterpart (the reference runs one case directory
// at a time, L:772-1486).
//
// Requests wait in a FIFO queue with host copies of their inputs.  Admission
// takes the head of the queue while the admitted scenarios' device
// footprint stays within the budget and a slot is free (strict arrival
// order: a request that does not fit blocks those behind it); it creates the
// request's context on its own stream.  Admitted requests become active (join
// their size class's batch with sw2d_batch_attach) at once, or, under an
// activation target, least estimated work first while the active ones hold
// fewer cells than the target.  Each size class is a sw2d_batch with its own
// stream: a chunk of iterations of one class is enqueued without waiting,
// so the classes' kernels run concurrently and a class of small grids
// advances at its own launch rate instead of the large grids'.  When a
// class's stream is idle its finished scenarios leave the batch
// (sw2d_batch_detach), waiting requests are activated and the class's next
// chunk is enqueued at once; then, while the chunks run, queued requests are
// admitted and the finished scenarios' final fields and statistics are read
// into host memory.  A retired scenario's context is not destroyed at once:
// freeing device memory waits for the whole device, which would stall the
// other classes' work, so retired contexts are kept (their memory still
// charged to the budget) and destroyed together when an admission needs the
// memory, or with the server.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <limits>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "sw2d.h"

void sw2d_set_error_text(const std::string &m);  // shared message slot (sw2d_api.hip)

namespace {

int srv_fail(int code, const std::string &m) {
  sw2d_set_error_text("server: " + m);
  return code;
}

struct Request {
  int state = SW2D_REQ_QUEUED;
  int rows = 0, cols = 0;
  double dx = 0, dy = 0, nodata = 0;
  std::vector<double> z, h0, n;  // inputs, held until admission
  bool has_n = false;
  sw2d_params params{};
  std::vector<double> rain;
  double dtrain = 0;
  size_t footprint = 0;
  double submit_s = 0, admit_s = -1, start_s = -1, finish_s = -1;
  int64_t iterations = 0;
  int cls = 0;                   // size class
  long long cells = 0;
  double work = 0;               // estimated work: cells x simulated seconds
  sw2d_ctx *ctx = nullptr;       // admitted, waiting for the device: its context
  int scen = -1;                 // scenario id in its class's batch while active
  std::vector<double> H, U, V;   // final fields once retired
  sw2d_stats stats{};
};

// Size classes by cells; a class's chunk is n_iter iterations times its
// multiplier.  The largest class's short chunks set the period of the host
// loop, in which admissions and read-backs happen; a class idles from the
// end of its chunk until the loop comes back to it, so the classes of small
// grids, whose iterations are short, take long chunks (on the 200-request
// serving stream, multipliers 128 and 32 against 4 and 2 halved the median
// time-to-result at the same throughput and p99).
constexpr int N_CLASS = 3;
constexpr long long CLASS_CELLS[N_CLASS - 1] = {1LL << 18, 1LL << 21};
constexpr int CLASS_MULT[N_CLASS] = {128, 32, 1};
int class_of(int rows, int cols) {
  const long long c = (long long)rows * cols;
  int k = 0;
  while (k < N_CLASS - 1 && c >= CLASS_CELLS[k]) k++;
  return k;
}

// Activation: admitted requests join their class's batch, least estimated
// work first, while the active scenarios hold fewer cells than the server's
// target (sw2d_server_set_active_cells; by default every admitted request is
// active); a request's estimated work is divided by 1 + (seconds waited) /
// ACTIVE_AGING_S, so that a long request waiting behind a stream of short
// ones still starts.
constexpr double ACTIVE_AGING_S = 60.0;

struct Class {
  sw2d_batch *batch = nullptr;
  std::vector<int> running;  // active request ids
  int n_added = 0;           // scenario ids the batch has handed out (0 .. n_added-1)
};

}  // namespace

struct sw2d_server {
  int prec = SW2D_FP32;
  size_t budget = 0;
  int max_slots = 0;
  Class cls[N_CLASS];
  std::vector<Request> req;
  std::deque<int> queue;
  std::vector<int> waiting;      // admitted, not yet active (contexts created)
  long long active_target = std::numeric_limits<long long>::max();
  long long active_cells = 0;
  size_t used = 0;
  // retired scenarios not yet destroyed, and their footprints
  std::vector<std::pair<sw2d_ctx *, size_t>> retired;

  // Destroys the retired contexts (one device-wide wait for all of them)
  // and returns their memory to the budget.
  void free_retired() {
    for (auto &p : retired) {
      sw2d_destroy(p.first);
      used -= p.second;
    }
    retired.clear();
  }
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  int n_active() const {
    int n = 0;
    for (const Class &c : cls) n += (int)c.running.size();
    return n;
  }
  // admitted requests: active, or waiting for the device
  int n_running() const { return n_active() + (int)waiting.size(); }
  bool can_activate() const {
    return !waiting.empty() && (active_cells < active_target || n_active() == 0);
  }

  // Admit from the head of the queue (strict arrival order) while the
  // footprint fits the budget and a slot is free: the request's context is
  // created on its own stream and given its rain (the uploads run while the
  // classes' chunks do); it waits there until activate() puts it in its
  // class's batch.
  int admit() {
    while (!queue.empty()) {
      Request &r = req[queue.front()];
      if (max_slots > 0 && n_running() >= max_slots) break;
      if (used + r.footprint > budget && !retired.empty()) free_retired();
      if (used + r.footprint > budget) break;
      sw2d_grid g;
      g.rows = r.rows;
      g.cols = r.cols;
      g.dx = r.dx;
      g.dy = r.dy;
      g.nodata = r.nodata;
      g.z = r.z.data();
      g.h0 = r.h0.data();
      g.n_cell = r.has_n ? r.n.data() : nullptr;
      sw2d_ctx *ctx = nullptr;
      int rc = sw2d_create(&g, &r.params, &ctx);
      if (rc) return rc;
      if (!r.rain.empty())
        rc = sw2d_set_rain(ctx, r.rain.data(), (int)r.rain.size(), r.dtrain, nullptr, 0);
      if (rc) {
        // not admitted: the request stays at the head of the queue
        sw2d_destroy(ctx);
        return rc;
      }
      r.ctx = ctx;
      r.state = SW2D_REQ_RUNNING;
      r.admit_s = now();
      std::vector<double>().swap(r.z);
      std::vector<double>().swap(r.h0);
      std::vector<double>().swap(r.n);
      used += r.footprint;
      waiting.push_back(queue.front());
      queue.pop_front();
    }
    return SW2D_OK;
  }

  // Puts waiting requests in their classes' batches, least (aged) estimated
  // work first, while the active cells are below the target or nothing is
  // active.  sw2d_batch_attach waits for no batch stream: a request joins
  // the work-list at its class's next chunk.
  int activate() {
    const double t = now();
    while (can_activate()) {
      size_t best = 0;
      double best_key = 0;
      for (size_t j = 0; j < waiting.size(); j++) {
        const Request &r = req[waiting[j]];
        const double key = r.work / (1.0 + (t - r.submit_s) / ACTIVE_AGING_S);
        if (j == 0 || key < best_key) {
          best = j;
          best_key = key;
        }
      }
      const int q = waiting[best];
      Request &r = req[q];
      Class &c = cls[r.cls];
      int id = -1;
      const int rc = sw2d_batch_attach(c.batch, r.ctx, &id);
      if (rc) return rc;  // the request keeps waiting with its context
      r.ctx = nullptr;
      r.scen = id;
      r.start_s = t;
      c.n_added = std::max(c.n_added, id + 1);
      c.running.push_back(q);
      active_cells += r.cells;
      waiting.erase(waiting.begin() + (long)best);
    }
    return SW2D_OK;
  }

  // Takes the finished scenarios of an idle class out of its batch; their
  // results are read by finish() once the next chunks are enqueued.
  int detach_finished(int k, std::vector<std::pair<int, sw2d_ctx *>> &out) {
    Class &c = cls[k];
    if (c.running.empty()) return SW2D_OK;
    int n = 0;
    std::vector<int> done(std::max(c.n_added, 1));
    std::vector<int64_t> iters(std::max(c.n_added, 1));
    int rc = sw2d_batch_status(c.batch, done.data(), iters.data(), c.n_added, &n);
    if (rc) return rc;
    // after a failed detach the rest stay in the batch and in `running`, so
    // the class's lists still agree
    std::vector<int> still;
    for (int q : c.running) {
      Request &r = req[q];
      sw2d_ctx *ctx = nullptr;
      if (rc || !done[r.scen] || (rc = sw2d_batch_detach(c.batch, r.scen, &ctx))) {
        still.push_back(q);
        continue;
      }
      r.iterations = iters[r.scen];
      r.scen = -1;
      active_cells -= r.cells;
      out.emplace_back(q, ctx);
    }
    c.running.swap(still);
    return rc;
  }

  // Reads the final fields and statistics of detached scenarios into host
  // memory: the requests are done from then on.  Their contexts are kept
  // (memory still charged) until free_retired().
  int finish(std::vector<std::pair<int, sw2d_ctx *>> &list) {
    int rc = SW2D_OK;
    for (auto &qc : list) {
      Request &r = req[qc.first];
      sw2d_ctx *ctx = qc.second;
      retired.emplace_back(ctx, r.footprint);
      if (rc) continue;  // after a failure the rest are left unread, not done
      r.H.resize((size_t)r.rows * r.cols);
      r.U.resize((size_t)(r.rows + 1) * r.cols);
      r.V.resize((size_t)r.rows * (r.cols + 1));
      if ((rc = sw2d_read(ctx, SW2D_FIELD_H, r.H.data())) || (rc = sw2d_read(ctx, SW2D_FIELD_U, r.U.data())) ||
          (rc = sw2d_read(ctx, SW2D_FIELD_V, r.V.data())) || (rc = sw2d_get_stats(ctx, &r.stats)))
        continue;
      r.state = SW2D_REQ_DONE;
      r.finish_s = now();
    }
    list.clear();
    return rc;
  }

  int advance(int64_t n_iter) {
    bool idle[N_CLASS];
    // wait until a class with active scenarios has finished its last
    // chunk, a waiting request can become active, or the request at the
    // head of the queue fits (or nothing runs)
    for (;;) {
      bool ready = false, busy = false;
      for (int k = 0; k < N_CLASS; k++) {
        int v = 1;
        int rc = sw2d_batch_idle(cls[k].batch, &v);
        if (rc) return rc;
        idle[k] = v != 0;
        if (!idle[k]) busy = true;
        else if (!cls[k].running.empty()) ready = true;
      }
      if (!queue.empty() && (max_slots == 0 || n_running() < max_slots) &&
          (used + req[queue.front()].footprint <= budget || !retired.empty()))
        ready = true;
      if (can_activate()) ready = true;
      if (ready || !busy) break;
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    // 1. the idle classes' finished scenarios leave their batches
    std::vector<std::pair<int, sw2d_ctx *>> finished;
    for (int k = 0; k < N_CLASS; k++)
      if (idle[k]) {
        int rc = detach_finished(k, finished);
        if (rc) {
          (void)finish(finished);
          return rc;
        }
      }
    // 2. waiting requests take the freed room, and the idle classes' next
    // chunks go to the device
    int rc = activate();
    for (int k = 0; k < N_CLASS && !rc; k++) {
      Class &c = cls[k];
      if (idle[k] && !c.running.empty() && n_iter > 0) rc = sw2d_batch_step(c.batch, n_iter * CLASS_MULT[k]);
    }
    // 3. while they run: admissions (a request admitted now may become
    // active at once, for its class's next chunk) and the finished
    // scenarios' read-back
    if (!rc) rc = admit();
    if (!rc) rc = activate();
    const int rf = finish(finished);
    return rc ? rc : rf;
  }
};

int sw2d_server_create(int precision, size_t mem_budget, int max_slots, sw2d_server **out) {
  if (!out || (precision != SW2D_FP32 && precision != SW2D_FP64) || mem_budget == 0 || max_slots < 0)
    return srv_fail(SW2D_E_INVALID, "create arguments");
  sw2d_server *s = new sw2d_server();
  s->prec = precision;
  s->budget = mem_budget;
  s->max_slots = max_slots;
  for (Class &c : s->cls) {
    int rc = sw2d_batch_create(precision, &c.batch);
    if (rc) {
      sw2d_server_destroy(s);
      return rc;
    }
  }
  *out = s;
  return SW2D_OK;
}

void sw2d_server_destroy(sw2d_server *s) {
  if (!s) return;
  s->free_retired();
  for (int q : s->waiting) sw2d_destroy(s->req[q].ctx);
  for (Class &c : s->cls)
    if (c.batch) sw2d_batch_destroy(c.batch);
  delete s;
}

int sw2d_server_submit(sw2d_server *s, const sw2d_grid *grid, const sw2d_params *params,
                       const double *rain_mm, int n_rain, double dtrain, int *req_id) {
  if (!s || !grid || !params || !req_id || !grid->z || !grid->h0) return srv_fail(SW2D_E_INVALID, "null argument");
  if (grid->rows <= 0 || grid->cols <= 0) return srv_fail(SW2D_E_INVALID, "grid size");
  if (params->precision != s->prec) return srv_fail(SW2D_E_INVALID, "request precision differs from the server");
  if (!(params->t_end > 0)) return srv_fail(SW2D_E_INVALID, "a request needs t_end > 0");
  if (n_rain < 0 || (n_rain > 0 && (!rain_mm || !(dtrain > 0)))) return srv_fail(SW2D_E_INVALID, "rain table");
  Request r;
  r.rows = grid->rows;
  r.cols = grid->cols;
  r.dx = grid->dx;
  r.dy = grid->dy;
  r.nodata = grid->nodata;
  const size_t cells = (size_t)r.rows * r.cols;
  r.z.assign(grid->z, grid->z + cells);
  r.h0.assign(grid->h0, grid->h0 + cells);
  r.has_n = grid->n_cell != nullptr;
  if (r.has_n) r.n.assign(grid->n_cell, grid->n_cell + cells);
  r.params = *params;
  if (n_rain > 0) r.rain.assign(rain_mm, rain_mm + n_rain);
  r.dtrain = dtrain;
  r.footprint = sw2d_footprint_bytes(r.rows, r.cols, s->prec, r.has_n);
  r.cls = class_of(r.rows, r.cols);
  r.cells = (long long)r.rows * r.cols;
  r.work = (double)r.cells * std::max(0.0, params->t_end - params->t0);
  if (r.footprint > s->budget) return srv_fail(SW2D_E_MEMORY, "request footprint exceeds the memory budget");
  r.submit_s = s->now();
  s->req.push_back(std::move(r));
  *req_id = (int)s->req.size() - 1;
  s->queue.push_back(*req_id);
  return SW2D_OK;
}

int sw2d_server_set_active_cells(sw2d_server *s, int64_t cells) {
  if (!s || cells < 0) return srv_fail(SW2D_E_INVALID, "active cells");
  s->active_target = cells > 0 ? (long long)cells : std::numeric_limits<long long>::max();
  return SW2D_OK;
}

int sw2d_server_advance(sw2d_server *s, int64_t n_iter, int *n_running, int *n_queued) {
  if (!s || n_iter < 0) return srv_fail(SW2D_E_INVALID, "advance arguments");
  const int rc = s->advance(n_iter);
  if (n_running) *n_running = s->n_running();
  if (n_queued) *n_queued = (int)s->queue.size();
  return rc;
}

static Request *find_req(sw2d_server *s, int id) {
  if (!s || id < 0 || id >= (int)s->req.size()) return nullptr;
  return &s->req[id];
}

int sw2d_server_info(sw2d_server *s, int req_id, sw2d_request_info *info) {
  Request *r = find_req(s, req_id);
  if (!r || !info) return srv_fail(SW2D_E_INVALID, "request id");
  info->state = r->state;
  info->rows = r->rows;
  info->cols = r->cols;
  info->submit_s = r->submit_s;
  info->admit_s = r->admit_s;
  info->start_s = r->start_s;
  info->finish_s = r->finish_s;
  info->iterations = r->iterations;
  return SW2D_OK;
}

int sw2d_server_read(sw2d_server *s, int req_id, int field, double *host_out) {
  Request *r = find_req(s, req_id);
  if (!r || !host_out) return srv_fail(SW2D_E_INVALID, "request id");
  if (r->state != SW2D_REQ_DONE) return srv_fail(SW2D_E_STATE, "request not finished");
  const std::vector<double> *v = field == SW2D_FIELD_H ? &r->H : field == SW2D_FIELD_U ? &r->U
                                 : field == SW2D_FIELD_V ? &r->V : nullptr;
  if (!v) return srv_fail(SW2D_E_INVALID, "field (H, U or V)");
  if (v->empty()) return srv_fail(SW2D_E_STATE, "result released");
  std::memcpy(host_out, v->data(), v->size() * sizeof(double));
  return SW2D_OK;
}

int sw2d_server_stats(sw2d_server *s, int req_id, sw2d_stats *out) {
  Request *r = find_req(s, req_id);
  if (!r || !out) return srv_fail(SW2D_E_INVALID, "request id");
  if (r->state != SW2D_REQ_DONE) return srv_fail(SW2D_E_STATE, "request not finished");
  *out = r->stats;
  return SW2D_OK;
}

int sw2d_server_release(sw2d_server *s, int req_id) {
  Request *r = find_req(s, req_id);
  if (!r) return srv_fail(SW2D_E_INVALID, "request id");
  if (r->state != SW2D_REQ_DONE) return srv_fail(SW2D_E_STATE, "request not finished");
  std::vector<double>().swap(r->H);
  std::vector<double>().swap(r->U);
  std::vector<double>().swap(r->V);
  return SW2D_OK;
}
