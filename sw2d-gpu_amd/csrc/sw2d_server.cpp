// sw2d_server.cpp -- serving loop: continuous batching of flood-scenario
// requests over one sw2d_batch.  Built only on the public C ABI (sw2d.h);
// no reference counterpart (the reference runs one case directory at a
// time, L:772-1486).
//
// Requests wait in a FIFO queue with host copies of their inputs.  Each
// advance admits from the head of the queue while the admitted scenarios'
// device footprint stays within the budget and a slot is free (strict
// arrival order: a request that does not fit blocks those behind it), steps
// the batch, and retires finished scenarios: their final fields and
// statistics move to host memory and their device memory returns to the
// budget.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <string>
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
  double submit_s = 0, admit_s = -1, finish_s = -1;
  int64_t iterations = 0;
  int scen = -1;                 // batch scenario id while running
  std::vector<double> H, U, V;   // final fields once retired
  sw2d_stats stats{};
};

}  // namespace

struct sw2d_server {
  int prec = SW2D_FP32;
  size_t budget = 0;
  int max_slots = 0;
  sw2d_batch *batch = nullptr;
  std::vector<Request> req;
  std::deque<int> queue;
  std::vector<int> running;
  size_t used = 0;
  int n_added = 0;  // scenario ids the batch has handed out (ids are 0 .. n_added-1)
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  int admit() {
    bool any = false;
    while (!queue.empty()) {
      Request &r = req[queue.front()];
      if (max_slots > 0 && (int)running.size() >= max_slots) break;
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
      int id = -1;
      int rc = sw2d_batch_add(batch, &g, &r.params, -1, &id);
      if (rc) return rc;
      n_added = std::max(n_added, id + 1);
      sw2d_ctx *ctx = nullptr;
      rc = sw2d_batch_scenario(batch, id, &ctx);
      if (!rc && !r.rain.empty())
        rc = sw2d_set_rain(ctx, r.rain.data(), (int)r.rain.size(), r.dtrain, nullptr, 0);
      if (rc) {
        // the scenario is in the batch but not admitted: take it out again
        // (the request stays at the head of the queue)
        (void)sw2d_batch_remove(batch, id);
        return rc;
      }
      r.scen = id;
      r.state = SW2D_REQ_RUNNING;
      r.admit_s = now();
      std::vector<double>().swap(r.z);
      std::vector<double>().swap(r.h0);
      std::vector<double>().swap(r.n);
      used += r.footprint;
      running.push_back(queue.front());
      queue.pop_front();
      any = true;
    }
    return any ? sw2d_batch_touch(batch) : SW2D_OK;
  }

  int retire() {
    if (running.empty()) return SW2D_OK;
    int n = 0;
    std::vector<int> done(std::max(n_added, 1));
    std::vector<int64_t> iters(std::max(n_added, 1));
    int rc = sw2d_batch_status(batch, done.data(), iters.data(), n_added, &n);
    if (rc) return rc;
    const double t = now();
    std::vector<int> still;
    for (int k : running) {
      Request &r = req[k];
      if (!done[r.scen]) {
        still.push_back(k);
        continue;
      }
      sw2d_ctx *ctx = nullptr;
      rc = sw2d_batch_scenario(batch, r.scen, &ctx);
      if (rc) return rc;
      r.H.resize((size_t)r.rows * r.cols);
      r.U.resize((size_t)(r.rows + 1) * r.cols);
      r.V.resize((size_t)r.rows * (r.cols + 1));
      if ((rc = sw2d_read(ctx, SW2D_FIELD_H, r.H.data())) || (rc = sw2d_read(ctx, SW2D_FIELD_U, r.U.data())) ||
          (rc = sw2d_read(ctx, SW2D_FIELD_V, r.V.data())) || (rc = sw2d_get_stats(ctx, &r.stats)))
        return rc;
      if ((rc = sw2d_batch_remove(batch, r.scen))) return rc;
      r.iterations = iters[r.scen];
      r.scen = -1;
      r.state = SW2D_REQ_DONE;
      r.finish_s = t;
      used -= r.footprint;
    }
    running.swap(still);
    return SW2D_OK;
  }
};

int sw2d_server_create(int precision, size_t mem_budget, int max_slots, sw2d_server **out) {
  if (!out || (precision != SW2D_FP32 && precision != SW2D_FP64) || mem_budget == 0 || max_slots < 0)
    return srv_fail(SW2D_E_INVALID, "create arguments");
  sw2d_server *s = new sw2d_server();
  s->prec = precision;
  s->budget = mem_budget;
  s->max_slots = max_slots;
  int rc = sw2d_batch_create(precision, &s->batch);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return SW2D_OK;
}

void sw2d_server_destroy(sw2d_server *s) {
  if (!s) return;
  if (s->batch) sw2d_batch_destroy(s->batch);
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
  if (r.footprint > s->budget) return srv_fail(SW2D_E_MEMORY, "request footprint exceeds the memory budget");
  r.submit_s = s->now();
  s->req.push_back(std::move(r));
  *req_id = (int)s->req.size() - 1;
  s->queue.push_back(*req_id);
  return SW2D_OK;
}

int sw2d_server_advance(sw2d_server *s, int64_t n_iter, int *n_running, int *n_queued) {
  if (!s || n_iter < 0) return srv_fail(SW2D_E_INVALID, "advance arguments");
  int rc = s->admit();
  if (!rc && !s->running.empty() && n_iter > 0) rc = sw2d_batch_step(s->batch, n_iter);
  if (!rc) rc = s->retire();
  if (n_running) *n_running = (int)s->running.size();
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
