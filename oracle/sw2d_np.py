"""Literal NumPy transcription of the SW2D-GPU kernels on the reference layout.

TEST INFRASTRUCTURE ONLY: a second, independent restatement used to pin the
C oracle (sw2d_oracle.c).  Only tests and the smoke check import it.

Unlike the C oracle (explicit face arrays, wall semantics written out), this
file keeps the reference's own flat layout from SW2D-GPU_linux/src/SW2D-GPU.cu:
one array of length N = rows*cols + cols/2 per field (L:829), linear index
id = i*cols + j (L:58-59), U faces stored at the id of the cell south of the
face and V faces at the id of the cell east of it, so that the east face of
column cols-1 aliases the west face of the next row (L:232, L:340-342).  Reads
past N return zero (arrays are padded), which is the condition under which
the reference's ghost half row and out-of-bounds south-edge reads (L:172-175,
L:296-299, L:473, L:503, L:654) reduce to closed walls.  Ghost cells start
with h = 0 and bed 0.

Each function vectorises one reference kernel over every id; np.where picks
among branches that are all evaluated with the kernel's own operand order.
"""

import numpy as np

G = 9.80          # L:809
TH = 1.0e-4       # L:1130
EXP = 1.3333330   # L:195, L:312

# The friction power hf^1.3333330: the C oracle's own IEEE-operation
# sequence (orc_pow_fric), which the device FP64 step shares; it is pinned on
# its own against a 60-digit evaluation (tests/test_oracle_pins.py), so this
# transcription takes it from the oracle library rather than restating it.
def _fric_pow(hface):
    import oracle_py
    f = np.vectorize(oracle_py.pow_fric, otypes=[np.float64])
    return f(hface)


def _interp(table, t, dT):
    """The forcing formula shared by stream_flow (L:533-538, L:547-552) and
    evaporation_load (L:592-597), written out as the reference does: a ramp
    a[0]*t up to t = 1 s, then a[it] + (a[it+1] - a[it]) / (dT*(t - dT*it))
    with it = int(t/dT).  Callers keep it + 1 inside the table."""
    if t <= 1.00:
        return table[0] * t
    it = int(t / dT)
    return table[it] + (table[it + 1] - table[it]) / (dT * (t - dT * it))


class RefLayoutSim:
    """State of one case in the reference layout (all float64)."""

    def __init__(self, z, h0, nodata, dx, dt, manning, INT=0.0, INF=0.0,
                 LWL=0.0, EV_WL_min=0.0):
        rows, cols = z.shape
        self.rows, self.cols = rows, cols
        self.N = rows * cols + cols // 2
        self.pad = 2 * cols + 4
        n = self.N + self.pad
        self.dx = self.dy = float(dx)
        self.dt = float(dt)
        self.dt2 = 2.00 * self.dt                   # L:808
        self.rn = float(manning)
        self.INT, self.INF, self.LWL, self.EVmin = INT, INF, LWL, EV_WL_min
        f = lambda: np.zeros(n)
        self.baseo = f()
        self.baseo[:rows * cols] = z.ravel()
        self.h = f()
        self.h[:rows * cols] = h0.ravel()           # L:1151
        for name in ("ho", "hm", "hn", "um", "umo", "uu", "uua", "uu1",
                     "vn", "vno", "vv", "vva", "vv1", "ql"):
            setattr(self, name, f())
        ids = np.arange(self.N)
        self.ids = ids
        self.inj = ids % cols
        self.ini = ids // cols
        # init_inf L:37-44
        self.ho[:self.N] = self.h[:self.N]
        self.inf = np.zeros(n, dtype=np.int64)
        self.inf[:self.N] = np.where(self.baseo[:self.N] != nodata, 1, 0)
        # initiald L:62-87 (infx/infy are redundant with the inf checks)
        self._hm_hn()
        self.time = 0.0
        self.rr = 0.0
        self.Ev = 0.0

    # -- helpers -------------------------------------------------------
    def _at(self, a, off):
        """a[id + off] for every id < N, zero past the padded array."""
        idx = self.ids + off
        out = np.zeros(self.N)
        ok = (idx >= 0) & (idx < a.shape[0])
        out[ok] = a[idx[ok]]
        return out

    def _hm_hn(self):
        """hm_hn L:374-394 (also initiald's hm/hn L:62-87)."""
        h = self.h
        inj, ini, rows, cols = self.inj, self.ini, self.rows, self.cols
        hc = h[:self.N]
        hn = np.where(inj > 0, 0.50 * (hc + self._at(h, -1)), 0.0)
        hm = np.where(ini > 0, 0.50 * (hc + self._at(h, -cols)), 0.0)
        hm = np.where(ini == 0, hc, hm)
        hm = np.where(ini == rows, self._at(h, -cols), hm)
        hn = np.where(inj == 0, hc, hn)
        self.hm[:self.N] = hm
        self.hn[:self.N] = hn

    # -- kernels -------------------------------------------------------
    def flux(self):
        """flux L:98-354."""
        N, cols, rows = self.N, self.cols, self.rows
        inj, ini = self.inj, self.ini
        dt2, dx, dy, gg, rn = self.dt2, self.dx, self.dy, G, self.rn
        h, b, inf = self.h, self.baseo, self.inf
        at = self._at
        with np.errstate(all="ignore"):
            hc = h[:N]
            bc = b[:N]
            # ---- X direction (U faces) ----
            hw, bw = at(h, -cols), at(b, -cols)
            gate = ((ini > 0) & (ini < rows) & (inf[:N] != 0)
                    & (self._ati(inf, -cols) != 0) & ((hw > TH) | (hc > TH)))
            um = self._momentum(hc, bc, hw, bw, self.hm[:N],
                                self.uu1[:N], at(self.uu1, cols), at(self.uu1, -cols),
                                self.umo[:N], at(self.umo, cols), at(self.umo, -cols),
                                at(self.vv1, 1), at(self.vv1, -cols + 1), self.vv1[:N],
                                at(self.vv1, -cols),
                                at(self.umo, 1), at(self.umo, -1),
                                self.uu1[:N], self.vva[:N],
                                cross_zero=(inj == 0) | (inj == cols),
                                d_along=dx, d_cross=dy, along_first=True)
            self.um[:N] = np.where(gate, um, 0.0)
            # ---- Y direction (V faces) ----
            hs, bs = at(h, -1), at(b, -1)
            gate = ((inj > 0) & (inj < cols) & (inf[:N] != 0)
                    & (self._ati(inf, -1) != 0) & ((hs > TH) | (hc > TH)))
            vn = self._momentum(hc, bc, hs, bs, self.hn[:N],
                                self.vv1[:N], at(self.vv1, 1), at(self.vv1, -1),
                                self.vno[:N], at(self.vno, 1), at(self.vno, -1),
                                at(self.uu1, cols), at(self.uu1, cols - 1), self.uu1[:N],
                                at(self.uu1, -1),
                                at(self.vno, cols), at(self.vno, -cols),
                                self.uua[:N], self.vv1[:N],
                                cross_zero=(ini == 0) | (ini == rows),
                                d_along=dy, d_cross=dx, along_first=False)
            self.vn[:N] = np.where(gate, vn, 0.0)

    def _ati(self, a, off):
        idx = self.ids + off
        out = np.zeros(self.N, dtype=a.dtype)
        ok = (idx >= 0) & (idx < a.shape[0])
        out[ok] = a[idx[ok]]
        return out

    def _momentum(self, hc, bc, hp, bp, hface, q1c, q1f, q1b, qoc, qof, qob,
                  x1a, x1b, x1c, x1d, qox, qoy, fa, fb, cross_zero,
                  d_along, d_cross, along_first):
        """Dry-bed rules and momentum update of one direction.

        U direction (L:121-203): hp/bp = cell i-1; q1* = uu1 at this face,
        i+1, i-1; qo* = umo at this face, i+1, i-1; x1a..x1d = vv1 at
        (i,j+1), (i-1,j+1), (i,j), (i-1,j); qox/qoy = umo at (i,j+1),(i,j-1);
        friction velocity pair (uu1, vva).
        V direction (L:237-320): hp/bp = cell j-1; q1* = vv1 at face, j+1,
        j-1; qo* = vno at face, j+1, j-1; x1a..x1d = uu1 at (i+1,j),
        (i+1,j-1), (i,j), (i,j-1); qox/qoy = vno at (i+1,j), (i-1,j);
        friction velocity pair (uua, vv1).
        """
        gg, dt2 = G, self.dt2
        hhe = hc + bc
        hhw = hp + bp
        hhep = hc - TH
        hhwp = hp - TH
        r1a = np.where(hp > TH, 0.5440 * hp * np.sqrt(gg * hp), 0.0)
        r1b = np.where(hc > TH, -0.544 * hc * np.sqrt(gg * hc), 0.0)
        hhan = hhep - hhwp
        sgnm = hhan / np.abs(hhan)
        hh3 = np.fmax((hc + bc), (hp + bp)) - np.fmax(bc, bp)
        r2 = np.where((hc > 0) | (hp > 0),
                      -sgnm * 0.350 * hh3 * np.sqrt(2.00 * gg * hh3), 0.0)
        # gravity
        g13 = gg * hface * (dt2 / d_along) * (hc + bc - hp - bp)
        # along-direction convection
        rr_ = 0.50 * (q1f + q1c)
        ll_ = 0.50 * (q1c + q1b)
        mr = rr_ * (qof + qoc) * 0.50 + np.abs(rr_) * (qoc - qof) * 0.50
        ml = ll_ * (qoc + qob) * 0.50 + np.abs(ll_) * (qob - qoc) * 0.50
        along = (dt2 / d_along) * (mr - ml)
        # cross-direction convection
        cu = 0.50 * (x1a + x1b)
        cd = 0.50 * (x1c + x1d)
        if along_first:
            # u12: umu with umo(i,j+1), umd with umo(i,j-1)  (L:185-189)
            mu = cu * (qox + qoc) * 0.50 + np.abs(cu) * (qoc - qox) * 0.50
            md = cd * (qoy + qoc) * 0.50 + np.abs(cd) * (qoy - qoc) * 0.50
        else:
            # v11: vnr with vno(i+1,j), vnl with vno(i-1,j)  (L:296-300)
            mu = cu * (qox + qoc) * 0.50 + np.abs(cu) * (qoc - qox) * 0.50
            md = cd * (qoc + qoy) * 0.50 + np.abs(cd) * (qoy - qoc) * 0.50
        cross = np.where(cross_zero, 0.0, (dt2 / d_cross) * (mu - md))
        # friction
        sq = np.sqrt(fa * fa + fb * fb)
        ex = _fric_pow(hface)
        ram = gg * self.rn * self.rn * sq / ex
        ram = np.where(hface <= TH, 0.00, ram)
        if along_first:
            upd = ((1.00 - dt2 * ram * 0.50) * qoc + (-along - cross - g13)) \
                / (1.00 + dt2 * ram * 0.50)
        else:
            upd = ((1.00 - dt2 * ram * 0.50) * qoc + (-cross - along - g13)) \
                / (1.00 + dt2 * ram * 0.50)
        out = np.where(hhe < bp, r1a,
                       np.where(hhw < bc, r1b,
                                np.where(hhep * hhwp < 0, r2, upd)))
        return out

    def continuity(self):
        """continuity L:627-660 (uniform rain rr, source array ql)."""
        N, cols = self.N, self.cols
        with np.errstate(all="ignore"):
            ho = self.ho[:N]
            pct = np.where(ho > self.EVmin, 1.00 - self.LWL,
                           1.00 - (self.INT + self.INF))
            ev = np.where(ho > self.EVmin, self.Ev, 0.000)
            h = ho - self.dt2 * ((self._at(self.um, cols) - self.um[:N]) / self.dx
                                 + (self._at(self.vn, 1) - self.vn[:N]) / self.dy
                                 - self.ql[:N] - self.rr * pct + ev)
            h = np.fmax(h, 0.00)
            h = np.where(self.inf[:N] == 0, 0.00, h)
        self.h[:N] = h

    def treat_error(self):
        """treat_error L:704-719 (the zeroing rules commute, so the
        parallel kernel's order does not matter)."""
        N, cols, rows = self.N, self.cols, self.rows
        dry = ((self.inf[:N] == 1) | (self.inf[:N] == 2)) & (self.h[:N] < TH)
        um, vn = self.um, self.vn
        ids = self.ids
        s = ids[dry & (self.ini < rows)] + cols
        s = s[s < um.shape[0]]
        um[s[um[s] > 0]] = 0.0
        c = ids[dry]
        um[c[um[c] < 0.0]] = 0.00
        e = ids[dry & (self.inj < cols)] + 1
        e = e[e < vn.shape[0]]
        vn[e[vn[e] > 0]] = 0.0
        vn[c[vn[c] < 0]] = 0.0

    def uu1_vv1(self):
        """uu1_vv1 L:418-447."""
        N = self.N
        with np.errstate(all="ignore"):
            hm, hn = self.hm[:N], self.hn[:N]
            self.uu1[:N] = np.where(hm >= TH, self.um[:N] / hm, 0.00)
            self.vv1[:N] = np.where(hn >= TH, self.vn[:N] / hn, 0.0)

    def uu_vv(self):
        """uu_vv L:472-479."""
        N, cols = self.N, self.cols
        hc = self.h[:N]
        self.uu[:N] = np.where(hc >= TH, (self._at(self.uu1, cols) + self.uu1[:N]) / 2.00, 0.00)
        self.vv[:N] = np.where(hc >= TH, (self._at(self.vv1, 1) + self.vv1[:N]) / 2.00, 0.00)

    def uua_vva(self):
        """uua_vva L:502-519 (later assignments override earlier ones)."""
        N, cols, rows = self.N, self.cols, self.rows
        inj, ini, at = self.inj, self.ini, self._at
        u1, v1 = self.uu1, self.vv1
        uua = self.uua[:N].copy()
        vva = self.vva[:N].copy()
        uua = np.where(inj > 0, 0.250 * (u1[:N] + at(u1, cols) + at(u1, -1) + at(u1, cols - 1)), uua)
        vva = np.where(ini > 0, 0.250 * (v1[:N] + at(v1, 1) + at(v1, -cols) + at(v1, -cols + 1)), vva)
        uua = np.where(inj == 0, 0.50 * (u1[:N] + at(u1, cols)), uua)
        vva = np.where(ini == 0, 0.50 * (v1[:N] + at(v1, 1)), vva)
        vva = np.where(ini == rows, 0.50 * (v1[:N] + at(v1, 1)), vva)
        self.uua[:N] = uua
        self.vva[:N] = vva

    def forward(self):
        """forward L:679-681."""
        N = self.N
        self.umo[:N] = self.um[:N]
        self.vno[:N] = self.vn[:N]
        self.ho[:N] = self.h[:N]

    def _forcing(self, rain_mm, sources, evap):
        """stream_flow L:529-555 and evaporation_load L:589-601 at self.time.

        sources: (qq, dtoq, cells) with qq[n_rows][nst] in m3/s and the flat
        index of each source cell (L:545); each source assigns its rate to
        h_ql, in source order, so a later source on the same cell wins.
        evap: (evapo_mm, dtrain), the evaporation table of
        gpu_evaporation_calc in mm per dtrain; None leaves Ev untouched
        (evaporation_on != 1, L:1283-1286)."""
        t = self.time
        self.rr = rain_mm(t) if rain_mm else 0.0
        if sources is not None:
            qq, dtoq, cells = sources
            for k, cell in enumerate(cells):
                ql = _interp([row[k] for row in qq], t, dtoq)
                self.ql[cell] = ql / (self.dx * self.dy)          # L:540, L:545
        if evap is not None:
            table, dtr = evap
            self.Ev = _interp(table, t, dtr) / (dtr * 1000.0)    # L:592-599

    def iterate(self, rain_mm=None, dtrain=60.0, sources=None, evap=None):
        """One loop iteration L:1268-1336 with uniform rain given as a
        function of time returning rr in m/s (None = no rain), optional
        point sources and evaporation (see _forcing)."""
        self.flux()
        self._forcing(rain_mm, sources, evap)
        self.continuity()
        self.treat_error()
        self.time = self.time + self.dt
        self._forcing(rain_mm, sources, evap)
        self.continuity()
        self._hm_hn()
        self.uu1_vv1()
        self.uu_vv()
        self.uua_vva()
        self.forward()
        self.time = self.time + self.dt

    # -- views in the explicit-face layout --------------------------------
    def cells(self, a):
        return a[:self.rows * self.cols].reshape(self.rows, self.cols).copy()

    def u_faces(self):
        """U(i,j) for i in 0..rows (row `rows` = wall)."""
        R, C = self.rows, self.cols
        out = np.zeros((R + 1, C))
        out[:R] = self.um[:R * C].reshape(R, C)
        return out

    def v_faces(self):
        """V(i,j) for j in 0..cols (column `cols` = wall)."""
        R, C = self.rows, self.cols
        out = np.zeros((R, C + 1))
        out[:, :C] = self.vn[:R * C].reshape(R, C)
        return out
