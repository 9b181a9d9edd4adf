// sw2d_io.cpp -- the reference's file surface: case readers, the VTK and
// gauge writers, and the drop-in driver that replaces main().
//
// Reference: SW2D-GPU_linux/src/SW2D-GPU.cu
//   dir.txt L:772-777, setup.dat L:779-809, dem.txt L:815-835,
//   coord_source_sink.dat L:840-856, Q_source_sink.dat L:881-921,
//   rain.dat L:929-960, temperatures/solar_radiation L:963-1016,
//   initial_condition.txt L:1022-1033, coord_out.dat L:1043-1072,
//   time window L:1082-1087, output cadence L:1039-1040 and L:1339-1351,
//   VTK writer L:1358-1472, WaterLevel gauges L:1386-1412,
//   TimeSimu L:1477-1486.
// Numbers are parsed with strtod on whitespace-separated tokens, which is
// what fscanf("%lf") does on these files.  Where the reference reads past
// what a file holds (the last Q row, records missing at the end of a file),
// the value is 0 instead of uninitialised memory.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "sw2d.h"

// message slot shared with the runtime (sw2d_api.hip) behind sw2d_last_error
void sw2d_set_error_text(const std::string &m);

namespace {

int io_fail(int code, const std::string &m) {
  sw2d_set_error_text(m);
  return code;
}

bool read_file(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

std::vector<std::string> tokens(const std::string &s) {
  std::vector<std::string> t;
  size_t i = 0, n = s.size();
  while (i < n) {
    while (i < n && std::isspace((unsigned char)s[i])) i++;
    size_t b = i;
    while (i < n && !std::isspace((unsigned char)s[i])) i++;
    if (i > b) t.emplace_back(s, b, i - b);
  }
  return t;
}

long count_newlines(const std::string &s) {
  long c = 0;
  for (char ch : s)
    if (ch == '\n') c++;
  return c;
}

bool to_double(const std::string &t, double &v) {
  char *end = nullptr;
  v = std::strtod(t.c_str(), &end);
  return end != t.c_str();
}

// C round(): halves away from zero (L:542-543, L:1039-1040, L:1389-1390)
long c_round(double x) { return (long)std::round(x); }

}  // namespace

struct sw2d_case {
  std::string dir;
  sw2d_case_info info{};
  std::vector<double> dem, ic, rain, q, src, outxy, T, Rg, evapo, manning;
  bool has_manning = false;
};

extern "C" {

int sw2d_case_load(const char *case_dir, sw2d_case **out) {
  if (!case_dir || !out) return io_fail(SW2D_E_INVALID, "null argument");
  *out = nullptr;
  auto c = new sw2d_case();
  c->dir = case_dir;
  const std::string in = c->dir + "/input/";
  auto bad = [&](int code, const std::string &m) {
    delete c;
    return io_fail(code, m);
  };
  std::string txt;
  sw2d_case_info &I = c->info;
  // ---- setup.dat, 17 keys in fixed order (L:788-804) ----
  if (!read_file(in + "setup.dat", txt)) return bad(SW2D_E_IO, "unknown file - setup.dat");
  {
    auto t = tokens(txt);
    size_t k = 0;
    auto key = [&](const char *name, int n, double *dst) -> bool {
      if (k >= t.size() || t[k] != std::string(name) + ":") return false;
      k++;
      for (int a = 0; a < n; a++) {
        if (k >= t.size() || !to_double(t[k], dst[a])) return false;
        k++;
      }
      return true;
    };
    double d[6];
    bool ok = key("tday0_thour0_tmin0_tsec0", 4, I.tstart) && key("tday_thour_tmin_tsec", 4, I.tend) &&
              key("dt", 1, &I.dt) && key("dpout", 1, &I.dpout) && key("dkout", 1, &I.dkout) &&
              key("dtoq", 1, &I.dtoq) && key("evaporation_on", 1, &d[0]) && key("EV_WL_min", 1, &I.EV_WL_min) &&
              key("INT", 1, &I.INT) && key("INF", 1, &I.INF) && key("LWL", 1, &I.LWL) &&
              key("manning_coef", 1, &I.manning_coef) && key("out_velocity_x", 1, &d[1]) &&
              key("out_velocity_y", 1, &d[2]) && key("out_elevation", 1, &d[3]) && key("out_depth", 1, &d[4]) &&
              key("out_outlet_on", 1, &d[5]);
    if (!ok) return bad(SW2D_E_IO, "malformed setup.dat near token " + std::to_string(k));
    I.evaporation_on = (int)d[0];
    I.out_velocity_x = (int)d[1];
    I.out_velocity_y = (int)d[2];
    I.out_elevation = (int)d[3];
    I.out_depth = (int)d[4];
    I.out_outlet_on = (int)d[5];
  }
  // ---- dem.txt (L:822-834) ----
  if (!read_file(in + "dem.txt", txt)) return bad(SW2D_E_IO, "unknown file - dem.txt");
  {
    auto t = tokens(txt);
    const char *names[6] = {"ncols", "nrows", "xllcorner", "yllcorner", "cellsize", "NODATA_value"};
    double hv[6];
    if (t.size() < 12) return bad(SW2D_E_IO, "malformed dem.txt header");
    for (int a = 0; a < 6; a++)
      if (t[2 * a] != names[a] || !to_double(t[2 * a + 1], hv[a]))
        return bad(SW2D_E_IO, std::string("malformed dem.txt header at ") + names[a]);
    I.cols = (int)hv[0];
    I.rows = (int)hv[1];
    I.xll = hv[2];
    I.yll = hv[3];
    I.cellsize = hv[4];
    I.nodata = hv[5];
    if (I.rows <= 0 || I.cols <= 0 || !(I.cellsize > 0)) return bad(SW2D_E_IO, "dem.txt: bad dimensions");
    const size_t n = (size_t)I.rows * I.cols;
    if (t.size() < 12 + n) return bad(SW2D_E_IO, "dem.txt: fewer values than ncols*nrows");
    c->dem.resize(n);
    for (size_t a = 0; a < n; a++)
      if (!to_double(t[12 + a], c->dem[a])) return bad(SW2D_E_IO, "dem.txt: bad value");
  }
  const size_t ncell = (size_t)I.rows * I.cols;
  // ---- coord_source_sink.dat (L:848-855) ----
  if (!read_file(in + "coord_source_sink.dat", txt)) return bad(SW2D_E_IO, "unknown file - coord_source_sink.dat");
  {
    auto t = tokens(txt);
    double n;
    if (t.size() < 2 || t[0] != "nst" || !to_double(t[1], n) || n < 0)
      return bad(SW2D_E_IO, "malformed coord_source_sink.dat");
    I.nst = (int)n;
    c->src.assign(2 * (size_t)I.nst, 0.0);
    for (int a = 0; a < 2 * I.nst; a++) {
      if (2 + (size_t)a >= t.size() || !to_double(t[2 + a], c->src[a]))
        return bad(SW2D_E_IO, "coord_source_sink.dat: missing coordinates");
      c->src[a] = c->src[a] / I.cellsize;  // L:853-854
    }
  }
  // ---- Q_source_sink.dat: rows = newline count (L:892-920) ----
  if (!read_file(in + "Q_source_sink.dat", txt)) return bad(SW2D_E_IO, "unknown file - Q_source_sink.dat");
  {
    I.n_q_rows = (int)count_newlines(txt);
    auto t = tokens(txt);
    size_t k = (t.size() >= 2 && t[0] == "flow") ? 2 : 0;
    c->q.assign((size_t)I.n_q_rows * I.nst, 0.0);
    for (size_t a = 0; a < c->q.size() && k < t.size(); a++, k++) to_double(t[k], c->q[a]);
  }
  // ---- rain.dat: dtrain, 4 header tokens, newline count - 1 records ----
  long cont_rain = 0;
  if (!read_file(in + "rain.dat", txt)) return bad(SW2D_E_IO, "unknown file - rain.dat");
  {
    cont_rain = count_newlines(txt);
    auto t = tokens(txt);
    if (t.empty() || !to_double(t[0], I.dtrain)) return bad(SW2D_E_IO, "malformed rain.dat");
    I.n_rain = (int)std::max(0L, cont_rain - 1);
    c->rain.assign(I.n_rain, 0.0);
    for (int a = 0; a < I.n_rain; a++) {
      size_t k = 5 + 2 * (size_t)a + 1;
      if (k < t.size()) to_double(t[k], c->rain[a]);
    }
  }
  // ---- evaporation inputs (L:968-1016) ----
  if (I.evaporation_on == 1) {
    std::string tt, st;
    if (!read_file(in + "temperatures.txt", tt) || !read_file(in + "solar_radiation.txt", st))
      return bad(SW2D_E_IO, "unknown file - evaporation.dat");
    auto t1 = tokens(tt), t2 = tokens(st);
    if (t2.size() < 3 || t2[0] != "albedo:" || !to_double(t2[1], I.albedo))
      return bad(SW2D_E_IO, "malformed solar_radiation.txt");
    c->T.assign(I.n_rain, 0.0);
    c->Rg.assign(I.n_rain, 0.0);
    for (int a = 0; a < I.n_rain; a++) {
      if (1 + (size_t)a < t1.size()) to_double(t1[1 + a], c->T[a]);
      if (3 + (size_t)a < t2.size()) to_double(t2[3 + a], c->Rg[a]);
    }
    c->evapo.assign(I.n_rain, 0.0);
    sw2d_evaporation_table(I.albedo, c->T.data(), c->Rg.data(), I.n_rain, I.dtrain, c->evapo.data());
  }
  // ---- initial_condition.txt (L:1029-1032) ----
  if (!read_file(in + "initial_condition.txt", txt)) return bad(SW2D_E_IO, "unknown file - initial_condition.txt");
  {
    auto t = tokens(txt);
    c->ic.assign(ncell, 0.0);
    for (size_t a = 0; a < ncell && a < t.size(); a++) to_double(t[a], c->ic[a]);
  }
  // ---- coord_out.dat (L:1052-1068), read only with the outlet on ----
  if (I.out_outlet_on == 1) {
    if (!read_file(in + "coord_out.dat", txt)) return bad(SW2D_E_IO, "unknown file - coord_out.dat");
    auto t = tokens(txt);
    double n;
    if (t.size() < 2 || t[0] != "n_out" || !to_double(t[1], n) || n < 0)
      return bad(SW2D_E_IO, "malformed coord_out.dat");
    I.n_out = (int)n;
    c->outxy.assign(2 * (size_t)I.n_out, 0.0);
    for (int a = 0; a < 2 * I.n_out; a++) {
      if (2 + (size_t)a < t.size()) to_double(t[2 + a], c->outxy[a]);
      c->outxy[a] = c->outxy[a] / I.cellsize;
    }
  }
  // ---- optional Manning raster (no reference counterpart) ----
  if (read_file(in + "manning.txt", txt)) {
    auto t = tokens(txt);
    const size_t n = ncell;
    if (t.size() < 12 + n) return bad(SW2D_E_IO, "manning.txt: fewer values than ncols*nrows");
    c->manning.resize(n);
    for (size_t a = 0; a < n; a++) to_double(t[12 + a], c->manning[a]);
    c->has_manning = true;
  }
  // ---- time window (L:1082-1087) ----
  I.time0 = 3600.0 * 24.0 * I.tstart[0] + 3600.0 * I.tstart[1] + 60.0 * I.tstart[2] + I.tstart[3];
  I.timmax = 3600.0 * 24.0 * I.tend[0] + 3600.0 * I.tend[1] + 60.0 * I.tend[2] + I.tend[3];
  if ((I.tend[0] == 0) & (I.tend[1] == 0) & (I.tend[2] == 0) & (I.tend[3] == 0))
    I.timmax = (cont_rain - 2) * I.dtrain;
  *out = c;
  return SW2D_OK;
}

void sw2d_case_free(sw2d_case *c) { delete c; }

int sw2d_case_get_info(const sw2d_case *c, sw2d_case_info *info) {
  if (!c || !info) return io_fail(SW2D_E_INVALID, "null argument");
  *info = c->info;
  return SW2D_OK;
}

int sw2d_case_get_array(const sw2d_case *c, int which, double *out, size_t cap, size_t *n_out) {
  if (!c || !n_out) return io_fail(SW2D_E_INVALID, "null argument");
  const std::vector<double> *v = nullptr;
  switch (which) {
    case SW2D_CASE_DEM: v = &c->dem; break;
    case SW2D_CASE_IC: v = &c->ic; break;
    case SW2D_CASE_RAIN: v = &c->rain; break;
    case SW2D_CASE_Q: v = &c->q; break;
    case SW2D_CASE_SRC_XY: v = &c->src; break;
    case SW2D_CASE_OUT_XY: v = &c->outxy; break;
    case SW2D_CASE_TEMP: v = &c->T; break;
    case SW2D_CASE_RG: v = &c->Rg; break;
    case SW2D_CASE_EVAPO: v = &c->evapo; break;
    default: return io_fail(SW2D_E_INVALID, "unknown case array");
  }
  *n_out = v->size();
  if (out) {
    if (cap < v->size()) return io_fail(SW2D_E_INVALID, "case array buffer too small");
    std::memcpy(out, v->data(), v->size() * sizeof(double));
  }
  return SW2D_OK;
}

int sw2d_case_source_cells(const sw2d_case *c, int64_t *out, int cap) {
  if (!c || (c->info.nst > 0 && !out) || cap < c->info.nst) return io_fail(SW2D_E_INVALID, "source cell buffer");
  const sw2d_case_info &I = c->info;
  const double x0 = I.xll / I.cellsize, y0 = I.yll / I.cellsize;  // L:1278 passes xcoor/resolution
  for (int k = 0; k < I.nst; k++) {
    const long qinx = c_round(std::fabs(x0 - c->src[2 * k]));              // L:542
    const long qiny = I.rows - c_round(std::fabs(y0 - c->src[2 * k + 1])); // L:543
    out[k] = (int64_t)((qiny + 1) * I.cols - (I.cols - (qinx + 1)));       // L:545
  }
  return SW2D_OK;
}

int sw2d_case_create(const sw2d_case *c, const sw2d_params *ov, sw2d_ctx **out) {
  if (!c || !out) return io_fail(SW2D_E_INVALID, "null argument");
  const sw2d_case_info &I = c->info;
  sw2d_params p;
  if (ov) p = *ov;
  else sw2d_params_default(&p);
  p.dt = I.dt;
  p.manning = I.manning_coef;
  p.INT = I.INT;
  p.INF = I.INF;
  p.LWL = I.LWL;
  p.EV_WL_min = I.EV_WL_min;
  p.evaporation_on = I.evaporation_on;
  p.t0 = I.time0;
  if (!ov || ov->t_end <= 0) p.t_end = I.timmax;
  sw2d_grid g;
  g.rows = I.rows;
  g.cols = I.cols;
  g.dx = I.cellsize;  // L:1236-1237
  g.dy = I.cellsize;
  g.nodata = I.nodata;
  g.z = c->dem.data();
  g.h0 = c->ic.data();
  g.n_cell = c->has_manning ? c->manning.data() : nullptr;
  int rc = sw2d_create(&g, &p, out);
  if (rc) return rc;
  rc = sw2d_set_rain(*out, c->rain.data(), I.n_rain, I.dtrain, nullptr, 0);
  if (!rc && I.nst > 0) {
    std::vector<int64_t> cells(I.nst);
    sw2d_case_source_cells(c, cells.data(), I.nst);
    rc = sw2d_set_sources(*out, c->q.data(), I.n_q_rows, I.nst, I.dtoq, cells.data());
  }
  if (!rc && I.evaporation_on == 1) rc = sw2d_set_evaporation(*out, c->evapo.data(), I.n_rain);
  if (rc) {
    sw2d_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

int sw2d_write_vtk(const char *path, const sw2d_case *c, const double *depth, const double *vv,
                   const double *uu, const double *z) {
  if (!path || !c) return io_fail(SW2D_E_INVALID, "null argument");
  const sw2d_case_info &I = c->info;
  FILE *f = std::fopen(path, "w");
  if (!f) return io_fail(SW2D_E_IO, std::string("cannot open ") + path);
  const int rows = I.rows, cols = I.cols;
  std::fprintf(f, "%s\n", "# vtk DataFile Version 2.0");
  std::fprintf(f, "%s\n", "Brazil");
  std::fprintf(f, "%s\n", "ASCII");
  std::fprintf(f, "%s\n", "DATASET STRUCTURED_POINTS");
  std::fprintf(f, "DIMENSIONS %d %d %d\n", cols, rows, 1);
  std::fprintf(f, "ASPECT_RATIO %lf %lf %lf\n", I.cellsize, I.cellsize, 1.0000);
  std::fprintf(f, "ORIGIN %lf %lf %lf\n", I.xll, I.yll, 0.000);
  std::fprintf(f, "POINT_DATA %d\n", cols * rows);
  // rows are written south -> north (L:1380-1382)
  auto block = [&](const char *label, const double *a, double sign) {
    std::fprintf(f, "SCALARS %s float 1\n", label);
    std::fprintf(f, "%s\n", "LOOKUP_TABLE default");
    for (int im = 0; im < rows; im++) {
      const size_t base = (size_t)rows * cols - (size_t)(im + 1) * cols;
      for (int jm = 0; jm < cols; jm++) {
        const double v = a[base + jm];
        std::fprintf(f, "%f\n", sign < 0 ? -v : v);
      }
    }
  };
  if (I.out_depth == 1 && depth) block("depth", depth, 1.0);
  if (I.out_velocity_x == 1 && vv) block("x_velocity", vv, 1.0);
  if (I.out_velocity_y == 1 && uu) block("y_velocity", uu, -1.0);
  if (I.out_elevation == 1 && z) block("Elevations", z, 1.0);
  std::fclose(f);
  return SW2D_OK;
}

int sw2d_run_db(const char *db_root, const sw2d_params *overrides) {
  if (!db_root) return io_fail(SW2D_E_INVALID, "null db root");
  const std::string db = db_root;
  std::string txt;
  if (!read_file(db + "/dir.txt", txt)) return io_fail(SW2D_E_IO, "unknown file - dir.txt");
  auto t = tokens(txt);
  double nd;
  if (t.size() < 2 || t[0] != "dir_number:" || !to_double(t[1], nd)) return io_fail(SW2D_E_IO, "malformed dir.txt");
  const int ncase = (int)nd;
  for (int ci = 0; ci < ncase; ci++) {
    if (2 + (size_t)ci >= t.size()) return io_fail(SW2D_E_IO, "dir.txt lists fewer cases than dir_number");
    const std::string name = t[2 + ci];
    const std::string cdir = db + "/" + name;
    sw2d_case *c = nullptr;
    int rc = sw2d_case_load(cdir.c_str(), &c);
    if (rc) {
      std::printf("%s\n", sw2d_last_error());
      return rc;
    }
    const sw2d_case_info I = c->info;
    const long lpout = c_round(I.dpout / I.dt), lkout = c_round(I.dkout / I.dt);  // L:1039-1040
    if (lpout <= 0 || lkout <= 0) {
      sw2d_case_free(c);
      return io_fail(SW2D_E_INVALID, "dpout/dt and dkout/dt must round to >= 1");
    }
    const std::string outdir = cdir + "/output/";
    FILE *wl = std::fopen((outdir + "WaterLevel.txt").c_str(), "w");
    if (!wl) {
      sw2d_case_free(c);
      return io_fail(SW2D_E_IO, "cannot open " + outdir + "WaterLevel.txt (missing output directory?)");
    }
    if (I.out_outlet_on != 1) std::fprintf(wl, "%s\n", "Outlet off");
    std::printf("%s\n", name.c_str());
    sw2d_ctx *x = nullptr;
    rc = sw2d_case_create(c, overrides, &x);
    if (rc) {
      std::fclose(wl);
      sw2d_case_free(c);
      return rc;
    }
    // gauge raster indices (L:1389-1392), in file order
    std::vector<long> gk(I.n_out);
    for (int g = 0; g < I.n_out; g++) {
      const long outx = c_round(std::fabs(I.xll / I.cellsize - c->outxy[2 * g]));
      const long outy = I.rows - c_round(std::fabs(I.yll / I.cellsize - c->outxy[2 * g + 1]));
      gk[g] = outy * I.cols - (I.cols - (outx + 1));
    }
    const size_t ncell = (size_t)I.rows * I.cols;
    const bool adaptive = overrides && overrides->cfl > 0;
    double time = I.time0;
    long long mstep = 0;
    long out0 = -1;
    std::string tempo;
    long long pending = 0;
    auto t_start = std::chrono::steady_clock::now();
    // An output is a snapshot of the state on the device (sw2d_output_begin)
    // whose copy to host memory runs beside the next iterations; its files
    // are written once the iterations toward the next output are queued, so
    // the host formats output k while the device computes toward k + 1.
    struct Out {
      int slot;
      std::string tempo;
      bool first;  // out0 == 0: depth from the initial condition (L:1415-1416)
    };
    std::vector<Out> outs;  // at most one waiting to be written
    auto write_files = [&](const Out &o) -> int {
      const double *h = nullptr, *vv = nullptr, *uu = nullptr;
      int r = sw2d_output_wait(x, o.slot, &h, &vv, &uu);
      if (r) return r;
      const double *depth = o.first ? c->ic.data() : h;
      // gauges, raster order (L:1380-1412)
      if (I.out_depth == 1 && I.out_outlet_on == 1) {
        int npout = 0;
        for (size_t km = 0; km < ncell; km++)
          for (int g = 0; g < I.n_out; g++)
            if ((long)km == gk[g]) {
              npout++;
              std::fprintf(wl, npout == I.n_out ? " %lf\n" : " %lf", depth[km]);
            }
      }
      const std::string path = outdir + "Results_" + o.tempo + ".vtk";
      return sw2d_write_vtk(path.c_str(), c, depth, vv, uu, c->dem.data());
    };
    auto flush_older = [&]() -> int {  // write every output but the newest
      int r = SW2D_OK;
      while (!r && outs.size() > 1) {
        r = write_files(outs.front());
        outs.erase(outs.begin());
      }
      return r;
    };
    auto take_output = [&]() -> int {
      int r = sw2d_step(x, pending);
      pending = 0;
      int slot = 0;
      if (!r) r = sw2d_output_begin(x, &slot);
      if (r) return r;
      std::ostringstream os;
      if (out0 > 0) os << std::round(mstep * I.dt / I.dkout);  // L:1343
      else os << 0;
      tempo = os.str();
      outs.push_back(Out{slot, tempo, out0 == 0});
      return SW2D_OK;
    };
    if (!adaptive) {
      // the host mirrors the device clock exactly (same FP64 additions)
      while (time + I.dt <= I.timmax) {  // L:1256
        out0 = out0 + 1;
        if (mstep % lpout == 0) std::printf(" %d\n", int(time));
        pending++;
        time = time + I.dt;
        time = time + I.dt;
        mstep += 2;
        if ((mstep % lkout == 0) || (out0 == 0)) {
          rc = take_output();      // queues the iterations up to here and the snapshot
          if (!rc) rc = flush_older();
          if (rc) break;
        }
      }
      if (!rc && pending) rc = sw2d_step(x, pending);
    } else {
      // adaptive step: the device owns the clock and the end test; the host
      // advances to the next iteration whose mstep is a multiple of lkout
      sw2d_stats st;
      bool first = true;
      while (!rc) {
        long long target = mstep / 2 + 1;
        if (!first)
          while ((2 * target) % lkout != 0) target++;
        pending = target - mstep / 2;
        rc = sw2d_step(x, pending);
        pending = 0;
        if (!rc) rc = flush_older();  // the previous output, while the device steps
        if (!rc) rc = sw2d_get_stats(x, &st);
        if (rc) break;
        const long long did = st.iterations - mstep / 2;
        if (did <= 0) break;
        out0 += did;
        mstep = 2 * st.iterations;
        std::printf(" %d\n", int(st.t));
        if ((mstep % lkout == 0) || first) rc = take_output();  // written after the next steps are queued
        first = false;
        if (st.done) break;
      }
    }
    for (const Out &o : outs)
      if (!rc) rc = write_files(o);
    outs.clear();
    if (!rc) rc = sw2d_sync(x);
    const double duration = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::cout << "printf: " << duration << '\n';
    FILE *ts = std::fopen((outdir + "TimeSimu_" + tempo + ".txt").c_str(), "w");
    if (ts) {
      std::fprintf(ts, " %lf\n", duration);
      std::fclose(ts);
    }
    std::fclose(wl);
    sw2d_destroy(x);
    sw2d_case_free(c);
    if (rc) return rc;
  }
  return SW2D_OK;
}

}  // extern "C"
