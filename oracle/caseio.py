"""Restatement of the SW2D-GPU case-file readers and result writers.

TEST INFRASTRUCTURE ONLY: what the tests hold the product's C++ reader/writer to
(sw2d-gpu_amd/csrc/sw2d_io.cpp) and the way tests feed the oracle.

Reference: SW2D-GPU_linux/src/SW2D-GPU.cu
  setup.dat               L:788-804 (17 keys, fixed order)
  dem.txt                 L:822-834 (ESRI ASCII header + rows north->south)
  coord_source_sink.dat   L:848-855
  Q_source_sink.dat       L:881-921 (row count = newline count)
  rain.dat                L:929-960 (record count = newline count - 1)
  temperatures.txt /
  solar_radiation.txt     L:998-1007
  initial_condition.txt   L:1022-1033
  coord_out.dat           L:1060-1067
  time window             L:1082-1087
  VTK writer              L:1364-1472
  WaterLevel gauges       L:1386-1412
"""
import math
import os

import numpy as np

SETUP_KEYS = [
    ("tday0_thour0_tmin0_tsec0", 4), ("tday_thour_tmin_tsec", 4), ("dt", 1),
    ("dpout", 1), ("dkout", 1), ("dtoq", 1), ("evaporation_on", 1),
    ("EV_WL_min", 1), ("INT", 1), ("INF", 1), ("LWL", 1), ("manning_coef", 1),
    ("out_velocity_x", 1), ("out_velocity_y", 1), ("out_elevation", 1),
    ("out_depth", 1), ("out_outlet_on", 1)]
INT_KEYS = {"evaporation_on", "out_velocity_x", "out_velocity_y",
            "out_elevation", "out_depth", "out_outlet_on"}


def _tokens(path):
    with open(path, "rb") as f:
        return f.read().decode("latin-1").split()


def _count_newlines(path):
    with open(path, "rb") as f:
        return f.read().count(b"\n")


def read_setup(path):
    toks = _tokens(path)
    out, k = {}, 0
    for key, n in SETUP_KEYS:
        if k >= len(toks) or toks[k] != key + ":":
            raise ValueError("setup.dat: expected key %s" % key)
        vals = toks[k + 1:k + 1 + n]
        k += 1 + n
        if key in INT_KEYS:
            out[key] = int(vals[0])
        else:
            out[key] = [float(v) for v in vals] if n > 1 else float(vals[0])
    return out


def read_dem(path):
    toks = _tokens(path)
    hdr = {}
    names = ["ncols", "nrows", "xllcorner", "yllcorner", "cellsize", "NODATA_value"]
    for i, nm in enumerate(names):
        if toks[2 * i] != nm:
            raise ValueError("dem.txt: expected %s" % nm)
        hdr[nm] = toks[2 * i + 1]
    cols, rows = int(hdr["ncols"]), int(hdr["nrows"])
    vals = np.array([float(v) for v in toks[12:12 + rows * cols]])
    if vals.size != rows * cols:
        raise ValueError("dem.txt: short raster")
    return dict(cols=cols, rows=rows, xll=float(hdr["xllcorner"]),
                yll=float(hdr["yllcorner"]), res=float(hdr["cellsize"]),
                nodata=float(hdr["NODATA_value"]), z=vals.reshape(rows, cols))


def read_values(path, n):
    toks = _tokens(path)
    vals = np.zeros(n)
    m = min(n, len(toks))
    vals[:m] = [float(v) for v in toks[:m]]
    return vals


def read_coords(path, key):
    toks = _tokens(path)
    if toks[0] != key:
        raise ValueError("%s: expected %s" % (path, key))
    n = int(toks[1])
    xy = np.array([float(v) for v in toks[2:2 + 2 * n]]).reshape(n, 2)
    return xy


def read_qtable(path, nst):
    rows = _count_newlines(path)                       # L:892-899
    toks = _tokens(path)
    vals = [float(v) for v in toks[2:]]
    q = np.zeros(rows * nst)
    m = min(len(vals), rows * nst)
    q[:m] = vals[:m]
    return q.reshape(rows, nst)


def read_rain(path):
    cont = _count_newlines(path)                       # L:939-946
    toks = _tokens(path)
    dtrain = float(toks[0])
    body = toks[5:]                                    # 4 header tokens, L:953
    n = cont - 1
    vals = np.zeros(n)
    for i in range(n):
        if 2 * i + 1 < len(body):
            vals[i] = float(body[2 * i + 1])
    return dtrain, vals, cont


def read_case(case_dir):
    """Parse one case directory db/<case>/ (with input/ inside)."""
    inp = os.path.join(case_dir, "input")
    s = read_setup(os.path.join(inp, "setup.dat"))
    d = read_dem(os.path.join(inp, "dem.txt"))
    src = read_coords(os.path.join(inp, "coord_source_sink.dat"), "nst") / d["res"]
    q = read_qtable(os.path.join(inp, "Q_source_sink.dat"), src.shape[0])
    dtrain, rain, cont_rain = read_rain(os.path.join(inp, "rain.dat"))
    case = dict(setup=s, dem=d, src=src, q=q, dtrain=dtrain, rain=rain,
                cont_rain=cont_rain)
    case["ic"] = read_values(os.path.join(inp, "initial_condition.txt"),
                             d["rows"] * d["cols"]).reshape(d["rows"], d["cols"])
    if s["evaporation_on"] == 1:
        tt = _tokens(os.path.join(inp, "temperatures.txt"))[1:]
        st = _tokens(os.path.join(inp, "solar_radiation.txt"))
        albedo = float(st[1])
        rg = st[3:]
        n = cont_rain - 1
        case["T"] = np.array([float(v) for v in tt[:n]])
        case["Rg"] = np.array([float(v) for v in rg[:n]])
        case["albedo"] = albedo
    if s["out_outlet_on"] == 1:
        case["gauges"] = read_coords(os.path.join(inp, "coord_out.dat"), "n_out") / d["res"]
    t0 = s["tday0_thour0_tmin0_tsec0"]
    t1 = s["tday_thour_tmin_tsec"]
    case["time0"] = 3600.0 * 24.0 * t0[0] + 3600.0 * t0[1] + 60.0 * t0[2] + t0[3]
    tm = 3600.0 * 24.0 * t1[0] + 3600.0 * t1[1] + 60.0 * t1[2] + t1[3]
    if t1 == [0.0, 0.0, 0.0, 0.0]:
        tm = (cont_rain - 2) * dtrain                  # L:1085-1087
    case["timmax"] = tm
    return case


def c_round(x):
    """C round(): halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def source_cells(case):
    """Linear cell index of each source as L:542-545 computes it."""
    d = case["dem"]
    x0, y0 = d["xll"] / d["res"], d["yll"] / d["res"]
    out = []
    for bx, by in case["src"]:
        qinx = int(c_round(abs(x0 - bx)))
        qiny = d["rows"] - int(c_round(abs(y0 - by)))
        out.append((qiny + 1) * d["cols"] - (d["cols"] - (qinx + 1)))
    return out


def gauge_cells(case):
    """Raster index of each gauge as L:1389-1392 computes it."""
    d = case["dem"]
    x0, y0 = d["xll"] / d["res"], d["yll"] / d["res"]
    out = []
    for gx, gy in case.get("gauges", []):
        outx = int(c_round(abs(x0 - gx)))
        outy = d["rows"] - int(c_round(abs(y0 - gy)))
        out.append(outy * d["cols"] - (d["cols"] - (outx + 1)))
    return out


def vtk_text(case, depth, vv, uu, z, flags=None):
    """Bytes of one Results_<k>.vtk (L:1364-1472). depth/vv/uu/z: (rows, cols)
    arrays in DEM order; rows are written south->north."""
    s, d = case["setup"], case["dem"]
    flags = flags or s
    rows, cols = d["rows"], d["cols"]
    out = ["# vtk DataFile Version 2.0", "Brazil", "ASCII",
           "DATASET STRUCTURED_POINTS",
           "DIMENSIONS %d %d %d" % (cols, rows, 1),
           "ASPECT_RATIO %f %f %f" % (d["res"], d["res"], 1.0),
           "ORIGIN %f %f %f" % (d["xll"], d["yll"], 0.0),
           "POINT_DATA %d" % (cols * rows)]

    def block(label, arr, neg=False):
        out.append("SCALARS %s float 1" % label)
        out.append("LOOKUP_TABLE default")
        flipped = arr[::-1].ravel()
        if neg:
            flipped = -flipped
        out.extend("%f" % v for v in flipped)

    if flags["out_depth"] == 1:
        block("depth", depth)
    if flags["out_velocity_x"] == 1:
        block("x_velocity", vv)
    if flags["out_velocity_y"] == 1:
        block("y_velocity", uu, neg=True)
    if flags["out_elevation"] == 1:
        block("Elevations", z)
    return ("\n".join(out) + "\n").encode()
