// sw2d_batch.hip -- batched ensembles: many scenarios of uneven size
// advanced by one launch per stage over a shared tile work-list, with a
// per-size-bucket tile-height autotuner and HIP-graph replay.
//
// Each scenario is a full sw2d_ctx (own arrays, control block, clock, step
// and end time).  The batch owns a device array of the scenarios' grid
// descriptors, an array of their control-block pointers and the work-list.
// Per-cell arithmetic does not depend on the tile a cell falls in or on the
// other scenarios, so a scenario's h, U, V are bitwise the same alone and in
// a batch; only the order of the volume-accounting sums changes.
#include <algorithm>
#include <map>

#include "sw2d_internal.hpp"

using sw2d::Ctrl;
using sw2d::Grid;
using sw2d::Tile;

struct sw2d_batch {
  int prec = 0;
  hipStream_t stream = nullptr;
  std::vector<sw2d_ctx *> scen;   // index = scenario id; nullptr once removed
  std::vector<int> share_of;      // id whose statics are shared, or -1
  std::vector<int> rows_override; // manual segment rows, 0 = bucket choice
  std::vector<char> parked;       // seen at its end time by status: left out of the work-list
  std::map<int, int> bucket_rows; // autotuner result per size bucket
  // device
  void *d_grids = nullptr;
  Ctrl **d_ctrls = nullptr;
  Tile *d_work = nullptr;
  int *d_done = nullptr;
  long long *d_iters = nullptr;
  size_t cap_grids = 0, cap_work = 0, cap_ctrls = 0, cap_done = 0, cap_iters = 0;
  int n_live = 0, n_work = 0, blocks = 0;
  std::vector<int> live_ids;      // grid index -> scenario id
  bool dirty = true;
  GraphCache graph;

  ~sw2d_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    graph.reset();
    for (auto *c : scen) delete c;
    for (void *p : {(void *)d_grids, (void *)d_ctrls, (void *)d_work, (void *)d_done, (void *)d_iters})
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int batch_rows = 0;              // batch-wide segment height of the live set (rebuild)
  // default segment rows of a scenario: its single-grid choice, raised to
  // the batch-wide height of the live set
  int default_rows(const sw2d_ctx *c) const { return std::max(c->TH, batch_rows); }
  static int bucket_of(const sw2d_ctx *c) {
    const double cells = (double)c->R * c->C;
    return (int)std::floor(std::log2(std::max(cells, 1.0)) / 2.0);
  }
  int rows_for(int id) const {
    if (rows_override[id]) return rows_override[id];
    auto it = bucket_rows.find(bucket_of(scen[id]));
    return it != bucket_rows.end() ? it->second : default_rows(scen[id]);
  }

  int grow(void **p, size_t &cap, size_t need) {
    if (need <= cap) return SW2D_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    size_t n = std::max(need, cap * 2);
    HIPCHK(hipMalloc(p, n));
    cap = n;
    return SW2D_OK;
  }

  // Tiles of the given scenarios (grid indices), siblings that share
  // statics interleaved tile by tile, then spread over the 8 XCDs so each
  // XCD walks a contiguous run of tiles.
  void build_tiles(const std::vector<int> &gidx, const std::vector<int> &rows, std::vector<Tile> &out) const {
    std::vector<Tile> list;
    std::vector<char> used(gidx.size(), 0);
    for (size_t a = 0; a < gidx.size(); a++) {
      if (used[a]) continue;
      std::vector<size_t> group{a};
      const int ida = live_ids[gidx[a]];
      for (size_t b2 = a + 1; b2 < gidx.size(); b2++) {
        const int idb = live_ids[gidx[b2]];
        if (!used[b2] && share_of[idb] >= 0 &&
            (share_of[idb] == ida || share_of[idb] == share_of[ida]) && rows[b2] == rows[a])
          group.push_back(b2);
      }
      for (size_t m : group) used[m] = 1;
      const sw2d_ctx *c = scen[ida];
      const int th = std::min(rows[a], sw2d::reach_rows_cap((long long)c->RS * c->esz));
      int slot = 0;
      for (int i0 = 0; i0 < c->R; i0 += th)
        for (int j0 = 0; j0 < c->C; j0 += sw2d::strip_width(prec == SW2D_FP64), slot++)
          for (size_t m : group) list.push_back(Tile{gidx[m], i0, j0, th, slot, 0, 0, 0});
    }
    const size_t n = list.size(), per = (n + 7) / 8;
    out.assign(per * 8, Tile{-1, 0, 0, 0, 0, 0, 0, 0});
    for (size_t k = 0; k < n; k++) {
      const size_t x = k / per, r = k % per;
      out[r * 8 + x] = list[k];
    }
  }

  int rebuild() {
    // iterations still in flight read d_grids, d_ctrls and d_work and may be
    // replaying the graph about to be destroyed: the null-stream copies
    // below do not order against the batch stream
    HIPCHK(hipStreamSynchronize(stream));
    live_ids.clear();
    for (size_t id = 0; id < scen.size(); id++)
      if (scen[id]) live_ids.push_back((int)id);
    n_live = (int)live_ids.size();
    const size_t gsz = prec == SW2D_FP64 ? sizeof(Grid<double>) : sizeof(Grid<float>);
    const size_t nl = std::max<size_t>(1, n_live);
    int rc = grow(&d_grids, cap_grids, nl * gsz);
    if (!rc) rc = grow((void **)&d_ctrls, cap_ctrls, nl * sizeof(Ctrl *));
    if (!rc) rc = grow((void **)&d_done, cap_done, nl * sizeof(int));
    if (!rc) rc = grow((void **)&d_iters, cap_iters, nl * sizeof(long long));
    if (rc) return rc;
    std::vector<unsigned char> grids(std::max<size_t>(1, n_live) * gsz);
    std::vector<Ctrl *> ctrls(std::max(1, n_live));
    for (int g = 0; g < n_live; g++) {
      sw2d_ctx *c = scen[live_ids[g]];
      HIPCHK(hipMemcpy(grids.data() + g * gsz, c->d_grid, gsz, hipMemcpyDeviceToHost));
      ctrls[g] = c->d_ctrl;
    }
    HIPCHK(hipMemcpy(d_grids, grids.data(), grids.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ctrls, ctrls.data(), ctrls.size() * sizeof(Ctrl *), hipMemcpyHostToDevice));
    std::vector<int> gidx, rows;
    long long strip_rows = 0;
    for (int g = 0; g < n_live; g++) {
      if (parked[live_ids[g]]) continue;  // finished: its work items would only exit
      gidx.push_back(g);
      const sw2d_ctx *c = scen[live_ids[g]];
      strip_rows += (long long)c->tiles_x * c->R;
    }
    const bool fp64 = prec == SW2D_FP64;
    batch_rows = sw2d::batch_seg_rows(strip_rows, sw2d::step_slots(fp64), fp64 ? 64 : 160);
    for (int g : gidx) rows.push_back(rows_for(live_ids[g]));
    std::vector<Tile> work;
    build_tiles(gidx, rows, work);
    rc = grow((void **)&d_work, cap_work, std::max<size_t>(1, work.size()) * sizeof(Tile));
    if (rc) return rc;
    if (!work.empty()) HIPCHK(hipMemcpy(d_work, work.data(), work.size() * sizeof(Tile), hipMemcpyHostToDevice));
    n_work = (int)work.size();
    // one single-wavefront workgroup per work item
    blocks = n_work;
    graph.reset();
    dirty = false;
    return SW2D_OK;
  }

  sw2d::IterLaunch launch_args() const {
    sw2d::IterLaunch a;
    a.grids = d_grids;
    a.work = d_work;
    a.n_work = n_work;
    a.blocks = std::max(1, blocks);
    a.ctrl = d_ctrls;
    a.n_ctrl = n_live;
    a.ev = nullptr;
    return a;
  }
};

extern "C" {

int sw2d_batch_create(int precision, sw2d_batch **out) {
  if (!out || (precision != SW2D_FP64 && precision != SW2D_FP32)) return fail(SW2D_E_INVALID, "batch precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(SW2D_E_DEVICE, "no HIP device available");
  auto *b = new sw2d_batch();
  b->prec = precision;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(SW2D_E_DEVICE, "hipStreamCreate failed");
  }
  *out = b;
  return SW2D_OK;
}

void sw2d_batch_destroy(sw2d_batch *b) { delete b; }

int sw2d_batch_add(sw2d_batch *b, const sw2d_grid *grid, const sw2d_params *params, int share_with, int *id_out) {
  if (!b || !grid || !params || !id_out) return fail(SW2D_E_INVALID, "null argument");
  if (params->precision != b->prec) return fail(SW2D_E_INVALID, "scenario precision differs from the batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  sw2d_ctx *sib = nullptr;
  if (share_with >= 0) {
    if (share_with >= (int)b->scen.size() || !(sib = b->scen[share_with]))
      return fail(SW2D_E_INVALID, "share_with names no live scenario");
    if (sib->R != grid->rows || sib->C != grid->cols || (grid->n_cell != nullptr) != (sib->ncell != nullptr))
      return fail(SW2D_E_INVALID, "share_with scenario has another grid shape");
  }
  sw2d_ctx *c = nullptr;
  int rc = sw2d_create(grid, params, &c);
  if (rc) return rc;
  if (sib) {
    // point the descriptor at the sibling's bed and raster, free our copies
    auto patch = [&](auto &G) -> int {
      using T = std::remove_pointer_t<std::remove_reference_t<decltype(G.h[0])>>;
      HIPCHK(hipMemcpy(&G, c->d_grid, sizeof(G), hipMemcpyDeviceToHost));
      G.z = (const T *)sib->z;
      G.ncell = (const T *)sib->ncell;
      G.datum = sib->datum;
      HIPCHK(hipMemcpy(c->d_grid, &G, sizeof(G), hipMemcpyHostToDevice));
      return SW2D_OK;
    };
    if (b->prec == SW2D_FP64) {
      Grid<double> G;
      rc = patch(G);
    } else {
      Grid<float> G;
      rc = patch(G);
    }
    if (rc) {
      delete c;
      return rc;
    }
    c->release(c->z);
    c->release(c->ncell);
    c->release(c->datum);
    c->z = sib->z;
    c->ncell = sib->ncell;
    c->datum = sib->datum;
  }
  c->owner = b->stream;
  b->scen.push_back(c);
  b->share_of.push_back(sib ? (b->share_of[share_with] >= 0 ? b->share_of[share_with] : share_with) : -1);
  b->rows_override.push_back(0);
  b->parked.push_back(0);
  b->dirty = true;
  *id_out = (int)b->scen.size() - 1;
  return SW2D_OK;
}

int sw2d_batch_scenario(sw2d_batch *b, int id, sw2d_ctx **ctx) {
  if (!b || !ctx || id < 0 || id >= (int)b->scen.size() || !b->scen[id]) return fail(SW2D_E_INVALID, "scenario id");
  HIPCHK(hipStreamSynchronize(b->stream));
  *ctx = b->scen[id];
  return SW2D_OK;
}

int sw2d_batch_touch(sw2d_batch *b) {
  if (!b) return fail(SW2D_E_INVALID, "null batch");
  std::fill(b->parked.begin(), b->parked.end(), 0);
  b->dirty = true;
  return SW2D_OK;
}

int sw2d_batch_remove(sw2d_batch *b, int id) {
  if (!b || id < 0 || id >= (int)b->scen.size() || !b->scen[id]) return fail(SW2D_E_INVALID, "scenario id");
  for (size_t k = 0; k < b->scen.size(); k++)
    if (b->scen[k] && (int)k != id && (b->share_of[k] == id))
      return fail(SW2D_E_STATE, "another scenario shares this scenario's statics");
  HIPCHK(hipStreamSynchronize(b->stream));
  sw2d_ctx *c = b->scen[id];
  if (b->share_of[id] >= 0) {  // statics belong to the sibling: do not free them
    auto drop = [&](void *p) {
      auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
      if (it != c->allocs.end()) c->allocs.erase(it);
    };
    drop(c->z);
    drop(c->ncell);
    drop(c->datum);
  }
  delete c;
  b->scen[id] = nullptr;
  b->dirty = true;
  return SW2D_OK;
}

int sw2d_batch_detach(sw2d_batch *b, int id, sw2d_ctx **ctx_out) {
  if (!b || !ctx_out || id < 0 || id >= (int)b->scen.size() || !b->scen[id])
    return fail(SW2D_E_INVALID, "scenario id");
  if (b->share_of[id] >= 0) return fail(SW2D_E_STATE, "the scenario reads another scenario's statics");
  for (size_t k = 0; k < b->scen.size(); k++)
    if (b->scen[k] && (int)k != id && b->share_of[k] == id)
      return fail(SW2D_E_STATE, "another scenario shares this scenario's statics");
  // launches still in flight read the scenario's arrays through the batch's
  // work list: the context leaves the batch once they are done
  HIPCHK(hipStreamSynchronize(b->stream));
  sw2d_ctx *c = b->scen[id];
  c->owner = nullptr;
  b->scen[id] = nullptr;
  b->dirty = true;
  *ctx_out = c;
  return SW2D_OK;
}

int sw2d_batch_attach(sw2d_batch *b, sw2d_ctx *c, int *id_out) {
  if (!b || !c || !id_out) return fail(SW2D_E_INVALID, "null argument");
  if (c->owner) return fail(SW2D_E_STATE, "the context is in a batch already");
  if (c->prec != b->prec) return fail(SW2D_E_INVALID, "scenario precision differs from the batch");
  // work the context may still have on its own stream (it may have been
  // stepped alone) ends before the batch reads its state
  HIPCHK(hipStreamSynchronize(c->stream));
  // the batch's device lists are rewritten only by the next rebuild, which
  // waits for the batch stream: nothing here touches them
  c->owner = b->stream;
  // the lowest id freed by a detach or remove is taken again (no live
  // scenario can share the statics of a freed id: both calls refuse that),
  // so a batch fed by attach and detach keeps a table as long as its peak
  // number of scenarios
  size_t id = 0;
  while (id < b->scen.size() && b->scen[id]) id++;
  if (id == b->scen.size()) {
    b->scen.push_back(nullptr);
    b->share_of.push_back(-1);
    b->rows_override.push_back(0);
    b->parked.push_back(0);
  }
  b->scen[id] = c;
  b->share_of[id] = -1;
  b->rows_override[id] = 0;
  b->parked[id] = 0;
  b->dirty = true;
  *id_out = (int)id;
  return SW2D_OK;
}

int sw2d_batch_set_tile_rows(sw2d_batch *b, int id, int rows) {
  if (!b || id < 0 || id >= (int)b->scen.size() || !b->scen[id]) return fail(SW2D_E_INVALID, "scenario id");
  if (rows != 0 && (rows < 4 || rows % 4 != 0 || rows > 4096))
    return fail(SW2D_E_INVALID, "segment rows must be a multiple of 4 in [4, 4096], or 0");
  b->rows_override[id] = rows;
  b->dirty = true;
  return SW2D_OK;
}

int sw2d_batch_tile_rows(sw2d_batch *b, int id, int *rows) {
  if (!b || !rows || id < 0 || id >= (int)b->scen.size() || !b->scen[id]) return fail(SW2D_E_INVALID, "scenario id");
  *rows = b->rows_for(id);
  return SW2D_OK;
}

int sw2d_batch_step(sw2d_batch *b, int64_t n_iter) {
  if (!b || n_iter < 0) return fail(SW2D_E_INVALID, "step count");
  if (b->dirty) {
    int rc = b->rebuild();
    if (rc) return rc;
  }
  if (n_iter == 0 || b->n_live == 0) return SW2D_OK;
  int rc = enqueue_iterations(b->prec, b->stream, b->launch_args(), n_iter, &b->graph);
  if (rc) return rc;
  HIPCHK(sw2d::launch_control(b->d_ctrls, b->n_live, 1, b->stream));
  return SW2D_OK;
}

int sw2d_batch_idle(sw2d_batch *b, int *idle) {
  if (!b || !idle) return fail(SW2D_E_INVALID, "null argument");
  const hipError_t e = hipStreamQuery(b->stream);
  if (e == hipSuccess) *idle = 1;
  else if (e == hipErrorNotReady) *idle = 0;
  else return fail(SW2D_E_DEVICE, std::string("batch stream: ") + hipGetErrorString(e));
  return SW2D_OK;
}

int sw2d_batch_sync(sw2d_batch *b) {
  if (!b) return fail(SW2D_E_INVALID, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  return SW2D_OK;
}

int sw2d_batch_status(sw2d_batch *b, int *done, int64_t *iters, int cap, int *n_out) {
  if (!b || !n_out) return fail(SW2D_E_INVALID, "null argument");
  const int n = (int)b->scen.size();
  *n_out = n;
  if (cap < n) return fail(SW2D_E_INVALID, "status buffer too small");
  if (b->dirty) {
    int rc = b->rebuild();
    if (rc) return rc;
  }
  std::vector<int> dn(std::max(1, b->n_live));
  std::vector<long long> it(std::max(1, b->n_live));
  if (b->n_live > 0) {
    HIPCHK(sw2d::launch_status(b->d_ctrls, b->n_live, b->d_done, b->d_iters, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(dn.data(), b->d_done, b->n_live * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(it.data(), b->d_iters, b->n_live * sizeof(long long), hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < n; k++) {
    if (done) done[k] = 1;
    if (iters) iters[k] = 0;
  }
  for (int g = 0; g < b->n_live; g++) {
    if (done) done[b->live_ids[g]] = dn[g];
    if (iters) iters[b->live_ids[g]] = it[g];
    if (dn[g] && !b->parked[b->live_ids[g]]) {  // finished: drop its items at the next step
      b->parked[b->live_ids[g]] = 1;
      b->dirty = true;
    }
  }
  return SW2D_OK;
}

int sw2d_batch_autotune(sw2d_batch *b, int iters) {
  if (!b || iters <= 0) return fail(SW2D_E_INVALID, "autotune iterations");
  if (b->dirty) {
    int rc = b->rebuild();
    if (rc) return rc;
  }
  if (b->n_live == 0) return SW2D_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  // Save every live scenario's control block and volume accumulators, then
  // decide one iteration for all of them without committing it: the step
  // kernel then reads the committed buffers and rewrites the scratch ones
  // the same way on every launch, so candidates are timed on the real mix
  // and restoring the saved blocks leaves the batch bitwise unchanged.
  std::vector<Ctrl> ctrl_save(b->n_live);
  std::vector<size_t> acc_off(b->n_live + 1, 0);
  for (int g = 0; g < b->n_live; g++)
    acc_off[g + 1] = acc_off[g] + (size_t)b->scen[b->live_ids[g]]->n_acc * 2 * sw2d::ACC_N;
  // scratch of the tuning run, released on every return path
  struct Scratch {
    double *acc = nullptr;
    Tile *tw = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch() {
      if (acc) (void)hipFree(acc);
      if (tw) (void)hipFree(tw);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } scr;
  HIPCHK(hipEventCreate(&scr.e0));
  HIPCHK(hipEventCreate(&scr.e1));
  HIPCHK(hipMalloc((void **)&scr.acc, std::max<size_t>(1, acc_off.back()) * sizeof(double)));
  double *d_acc_save = scr.acc;
  for (int g = 0; g < b->n_live; g++) {
    sw2d_ctx *c = b->scen[b->live_ids[g]];
    HIPCHK(hipMemcpy(&ctrl_save[g], c->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(d_acc_save + acc_off[g], c->acc, (acc_off[g + 1] - acc_off[g]) * sizeof(double),
                     hipMemcpyDeviceToDevice));
  }
  HIPCHK(sw2d::launch_control(b->d_ctrls, b->n_live, 0, b->stream));

  // buckets probed largest total first; scenarios with a manual height keep it
  std::map<int, long long> cells;
  std::map<int, int> choice;
  for (int g = 0; g < b->n_live; g++) {
    const int id = b->live_ids[g];
    if (b->rows_override[id]) continue;
    const int bk = sw2d_batch::bucket_of(b->scen[id]);
    cells[bk] += (long long)b->scen[id]->R * b->scen[id]->C;
    if (!choice.count(bk)) choice[bk] = b->rows_for(id);
  }
  std::vector<std::pair<long long, int>> order;
  for (auto &kv : cells) order.push_back({-kv.second, kv.first});
  std::sort(order.begin(), order.end());

  const std::vector<int> cands{16, 32, 64, 128};
  hipEvent_t e0 = scr.e0, e1 = scr.e1;
  Tile *&d_tw = scr.tw;
  size_t cap_tw = 0;
  int rc = SW2D_OK;
  auto time_choice = [&](float &ms) -> int {
    std::vector<int> gidx(b->n_live), rows(b->n_live);
    for (int g = 0; g < b->n_live; g++) {
      const int id = b->live_ids[g];
      gidx[g] = g;
      rows[g] = b->rows_override[id] ? b->rows_override[id] : choice[sw2d_batch::bucket_of(b->scen[id])];
    }
    std::vector<Tile> work;
    b->build_tiles(gidx, rows, work);
    int r = b->grow((void **)&d_tw, cap_tw, work.size() * sizeof(Tile));
    if (r) return r;
    HIPCHK(hipMemcpy(d_tw, work.data(), work.size() * sizeof(Tile), hipMemcpyHostToDevice));
    sw2d::IterLaunch a = b->launch_args();
    a.work = d_tw;
    a.n_work = (int)work.size();
    a.blocks = a.n_work;
    HIPCHK(sw2d::launch_step_only(b->prec, a, b->stream));  // warm
    HIPCHK(hipEventRecord(e0, b->stream));
    for (int k = 0; k < iters; k++) HIPCHK(sw2d::launch_step_only(b->prec, a, b->stream));
    HIPCHK(hipEventRecord(e1, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    return SW2D_OK;
  };
  // The incumbent (the batch default or an earlier choice) is timed first
  // and kept unless a candidate is faster by more than 3 %: one tuning pass
  // sees a few launches of the current state, and a height that wins there
  // by less has lost over whole runs (16-32-row choices made on a freshly
  // wetted ensemble ran 8 % slower than the batch default).
  for (auto &o : order) {
    const int bk = o.second;
    const int inc = choice[bk];
    float inc_ms = 0.0f;
    rc = time_choice(inc_ms);
    if (rc) break;
    int best_rows = inc;
    float best = inc_ms;
    for (int th : cands) {
      if (th == inc) continue;
      choice[bk] = th;
      float ms = 0.0f;
      rc = time_choice(ms);
      if (rc) break;
      if (ms < best && ms < 0.97f * inc_ms) {
        best = ms;
        best_rows = th;
      }
    }
    if (rc) break;
    choice[bk] = best_rows;
  }
  // restore the control blocks and accumulators (also after a failure)
  for (int g = 0; g < b->n_live; g++) {
    sw2d_ctx *c = b->scen[b->live_ids[g]];
    (void)hipMemcpy(c->d_ctrl, &ctrl_save[g], sizeof(Ctrl), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->acc, d_acc_save + acc_off[g], (acc_off[g + 1] - acc_off[g]) * sizeof(double),
                    hipMemcpyDeviceToDevice);
  }
  (void)hipStreamSynchronize(b->stream);
  if (!rc)
    for (auto &kv : choice) b->bucket_rows[kv.first] = kv.second;
  b->dirty = true;
  return rc;
}

}  // extern "C"
