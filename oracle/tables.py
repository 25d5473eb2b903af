"""ORACLE (test infrastructure only) -- FP64 numpy restatement of the reference's
physics-table build.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker.  The library never routes through it.

What it restates (all paths relative to the reference tree):
  * per-element tables read from EPDL ENDF-6 (MF23 MT502/504/522, MF27
    MT502/504), as stored by utility_scripts/sql_database_generator.ipynb
    cells 21-29 and read by src/MIDSX/Core/material_data.cpp:145-152;
  * number densities n_i = N_A * rho * w_i / A_i
    (src/MIDSX/Core/material_properties.cpp:66-73);
  * merged energy grids = sorted union with duplicates removed
    (include/MIDSX/Core/material_helpers.h:69-85);
  * weighted sums at the merged nodes, number-density weights for cross
    sections and mass-fraction weights for FF and S, then x1e-24 barn->cm2
    (material_data.cpp:87-114, 44-58);
  * per-element and per-material interpolators: LogLogSpline for the
    coherent / incoherent cross sections, LogLogLinear for the rest
    (material_data.cpp:154-170, 44-58; src/MIDSX/Core/interpolators.cpp);
  * total = sum of the three *interpolated* partials on their merged grid
    (material_data.cpp:75-85, 125-134);
  * majorant = node-wise max over materials on the union of their total grids
    (src/MIDSX/Core/interaction_data.cpp:42-73);
  * RITA tables for the Rayleigh angular distribution
    (material_data.cpp:62-73, src/MIDSX/Core/probability_dist.cpp:62-287).

The B-spline fit (Eigen SplineFitting::Interpolate with knot averaging, used
at interpolators.cpp:3-12) is solved here with scipy's banded collocation
solver on the knot-averaged knot vector (Piegl & Tiller, The NURBS Book,
eq. 9.8), which is independent of the library's own solver.

Deliberate change, documented: ENDF MF27 tables start with an x = 0 row whose
log10 is -inf; those rows are dropped so x below the first positive node
clamps to that node's value (the reference's behaviour there is undefined).
"""
import gzip
import math
import os

import numpy as np
from scipy.interpolate import make_interp_spline

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "midsx_amd", "data")

# include/MIDSX/Core/constants.h:13-23
PI = 3.14159265358979323846
ELECTRON_REST_MASS = 0.51099895000e6
SPEED_OF_LIGHT = 299792458e2
PLANCK_CONSTANT = 4.135667662e-15
AVOGADRO_CONSTANT = 6.022140857e23
ALPHA = ELECTRON_REST_MASS / (math.sqrt(2) * PLANCK_CONSTANT * SPEED_OF_LIGHT * 1e8)
NEUTRON_MASS_U = 1.00866491595  # ENDF AWR is in neutron masses

SECTIONS = {"coh": (23, 502), "incoh": (23, 504), "photo": (23, 522), "ff": (27, 502), "sf": (27, 504)}


# --------------------------------------------------------------------------- ENDF
def endf_float(s):
    s = s.strip()
    if not s:
        return 0.0
    try:
        return float(s)
    except ValueError:
        # ENDF short form "1.2345-5" / "1.2345+5"
        for i in range(len(s) - 1, 0, -1):
            if s[i] in "+-" and s[i - 1] not in "eE":
                return float(s[:i] + "e" + s[i:])
        raise


def read_endf(path):
    """Return (awr, {(mf, mt): (x, y)}) for an ENDF-6 file (.gz or plain)."""
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "rt") as f:
        lines = f.readlines()
    by = {}
    awr = None
    for ln in lines:
        if len(ln) < 75:
            continue
        try:
            mf, mt = int(ln[70:72]), int(ln[72:75])
        except ValueError:
            continue
        if awr is None and mf == 1 and mt == 451:
            awr = endf_float(ln[11:22])
        by.setdefault((mf, mt), []).append(ln)
    out = {}
    for key, ls in by.items():
        if key not in SECTIONS.values():
            continue
        # HEAD, TAB1 control record, NR interpolation pairs, NP (x, y) pairs
        nr = int(ls[1][44:55])
        np_ = int(ls[1][55:66])
        first = 2 + (2 * nr + 5) // 6
        vals = []
        for ln in ls[first:]:
            for k in range(6):
                fld = ln[11 * k:11 * k + 11]
                if fld.strip():
                    vals.append(endf_float(fld))
        vals = vals[:2 * np_]
        out[key] = (np.array(vals[0::2]), np.array(vals[1::2]))
    return awr, out


_ELEMENT_CACHE = {}


def element(z):
    if z not in _ELEMENT_CACHE:
        awr, t = read_endf(os.path.join(DATA, "epdl", "%03d.endf.gz" % z))
        tabs = {}
        for name, key in SECTIONS.items():
            x, y = t[key]
            if name in ("ff", "sf"):
                keep = x > 0.0
                x, y = x[keep], y[keep]
            tabs[name] = (x, y)
        _ELEMENT_CACHE[z] = (awr * NEUTRON_MASS_U, tabs)
    return _ELEMENT_CACHE[z]


# --------------------------------------------------------------------------- materials
def load_registry(path=None):
    path = path or os.path.join(DATA, "materials.csv")
    reg = {}
    with open(path) as f:
        for ln in f:
            if ln.startswith("#") or not ln.strip():
                continue
            a = ln.index('"')
            b = ln.index('"', a + 1)
            mid = int(ln[:a].rstrip(","))
            name = ln[a + 1:b]
            rest = ln[b + 2:].strip().split(",")
            dens = float(rest[0])
            comp = [(int(p.split(":")[0]), float(p.split(":")[1])) for p in rest[1].split(";")]
            reg[mid] = (name, dens, comp)
    return reg


# --------------------------------------------------------------------------- interpolators
class LogLogLinear:
    """interpolators.cpp:37-76: 10^(linear interp of (log10 x, log10 y)),
    clamped outside the node range, bracket by upper_bound."""

    def __init__(self, x, y):
        with np.errstate(divide="ignore"):
            self.lx = np.log10(np.asarray(x, dtype=np.float64))
            self.ly = np.log10(np.asarray(y, dtype=np.float64))
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        lx = np.log10(x)
        lxv, lyv = self.lx, self.ly
        idx = np.searchsorted(lxv, lx, side="right")  # upper_bound
        idx = np.clip(idx, 1, len(lxv) - 1)
        x1, x2 = lxv[idx - 1], lxv[idx]
        y1, y2 = lyv[idx - 1], lyv[idx]
        with np.errstate(invalid="ignore", divide="ignore"):
            slope = (y2 - y1) / (x2 - x1)
            v = y1 + slope * (lx - x1)
        v = np.where(lx <= lxv[0], lyv[0], v)
        v = np.where(lx >= lxv[-1], lyv[-1], v)
        return np.power(10.0, v)


def knot_averaging(u, p):
    """Knot vector of Eigen's KnotAveraging / Piegl & Tiller eq. 9.8."""
    n = len(u)
    t = np.zeros(n + p + 1)
    for j in range(1, n - p):
        t[j + p] = np.mean(u[j:j + p])
    t[n:] = 1.0
    return t


class LogLogSpline:
    """interpolators.cpp:3-31: cubic B-spline through (u_i, log10 y_i),
    u = (log10 x - min)/(max - min), degree min(n-1, 3), knot averaging;
    value 10^spline(u)."""

    def __init__(self, x, y):
        lx = np.log10(np.asarray(x, dtype=np.float64))
        ly = np.log10(np.asarray(y, dtype=np.float64))
        self.xmin, self.xmax = lx.min(), lx.max()
        u = (lx - self.xmin) / (self.xmax - self.xmin)
        self.p = min(len(u) - 1, 3)
        self.t = knot_averaging(u, self.p)
        self.spl = make_interp_spline(u, ly, k=self.p, t=self.t)
        self.coef = np.asarray(self.spl.c, dtype=np.float64)

    def __call__(self, x):
        u = (np.log10(np.asarray(x, dtype=np.float64)) - self.xmin) / (self.xmax - self.xmin)
        return np.power(10.0, self.spl(u, extrapolate=True))


def merge(*grids):
    return np.unique(np.concatenate([np.asarray(g, dtype=np.float64) for g in grids]))


# --------------------------------------------------------------------------- RITA
def _simpson_weights(n):
    w = np.full(n, 2.0)
    w[1::2] = 4.0
    w[0] = w[-1] = 1.0
    return w


class Rita:
    """probability_dist.cpp:62-287 (PENELOPE RITA with adaptive bisection)."""

    def __init__(self, pdf, energies, x_min=-1.0, x_max=1.0, err_thresh=1e-4):
        self.energies = np.asarray(energies, dtype=np.float64)
        self.x_min, self.x_max = x_min, x_max
        self.rows = []  # per energy: (x, y, a, b)
        w1001 = _simpson_weights(1001)
        w51 = _simpson_weights(51)
        for E in self.energies:
            h = (x_max - x_min) / 1000
            xs = x_min + np.arange(1001) * h
            integral = h / 3 * np.sum(w1001 * pdf(xs, E))

            def npdf(x, E=E, integral=integral):
                return pdf(x, E) / integral

            grid = np.linspace(x_min, x_max, 10)
            while True:
                cdf = self._cdf(grid, npdf, w51)
                a, b = self._params(grid, cdf, npdf)
                eps = self._err(grid, cdf, a, b, npdf)
                bad = eps > err_thresh
                if not bad.any():
                    break
                mids = 0.5 * (grid[:-1] + grid[1:])
                new = []
                for i in range(len(grid) - 1):
                    new.append(grid[i])
                    if bad[i]:
                        new.append(mids[i])
                new.append(grid[-1])
                grid = np.array(new)
            cdf = self._cdf(grid, npdf, w51)
            a, b = self._params(grid, cdf, npdf)
            self.rows.append((grid.copy(), cdf, a, b))

    def _cdf(self, grid, npdf, w51):
        x0 = grid[:-1][:, None]
        h = (grid[1:] - grid[:-1]) / 50.0
        xk = x0 + np.arange(51)[None, :] * h[:, None]
        s = np.sum(w51[None, :] * npdf(xk), axis=1)
        inc = h / 3 * s
        cdf = np.empty(len(grid))
        cdf[0] = 0.0
        acc = 0.0
        for j in range(len(inc)):  # sequential, as in the reference
            acc = acc + inc[j]
            cdf[j + 1] = acc
        return cdf

    @staticmethod
    def _params(grid, cdf, npdf):
        x0, x1 = grid[:-1], grid[1:]
        y0, y1 = cdf[:-1], cdf[1:]
        slope = (y1 - y0) / (x1 - x0)
        b = 1 - slope ** 2 * (1 / (npdf(x0) * npdf(x1)))
        a = slope * (1 / npdf(x0)) - b - 1
        return a, b

    @staticmethod
    def pnum(x, x0, x1, y0, y1, a, b):
        tau = (x - x0) / (x1 - x0)
        c = 1 + a + b - a * tau
        with np.errstate(invalid="ignore", divide="ignore"):
            eta = c / (2 * b * tau) * (1 - np.sqrt(1 - 4 * b * tau * tau / c ** 2))
        eta = np.where(tau == 0, 0.0, eta)
        return (1 + a * eta + b * eta * eta) ** 2 / ((1 + a + b) * (1 - b * eta * eta)) * (y1 - y0) / (x1 - x0)

    def _err(self, grid, cdf, a, b, npdf):
        x0, x1 = grid[:-1][:, None], grid[1:][:, None]
        y0, y1 = cdf[:-1][:, None], cdf[1:][:, None]
        A, B = a[:, None], b[:, None]
        d = (x1 - x0) / 51
        xk = x0 + np.arange(1, 51)[None, :] * d
        s = np.sum(np.abs(npdf(xk) - self.pnum(xk, x0, x1, y0, y1, A, B)), axis=1)
        e1 = np.abs(npdf(x1) - self.pnum(x1, x0, x1, y0, y1, A, B))[:, 0]
        e0 = np.abs(npdf(x0) - self.pnum(x0, x0, x1, y0, y1, A, B))[:, 0]
        return d[:, 0] * (s + (e1 + e0) / 2)

    def x_from_y(self, ie, y):
        """probability_dist.cpp:90-119."""
        x, cdf, a, b = self.rows[ie]
        i = np.searchsorted(cdf, y, side="left") - 1  # lower_bound - 1
        if i == len(cdf) - 1:
            return self.x_max
        if i < 0:
            return self.x_min
        nu = y - cdf[i]
        d = cdf[i + 1] - cdf[i]
        return x[i] + (1 + a[i] + b[i]) * d * nu / (d * d + a[i] * d * nu + b[i] * nu * nu) * (x[i + 1] - x[i])

    def sample(self, E, y):
        """probability_dist.cpp:73-88 with the uniform y given."""
        # lower_bound - 1, clamped to the end intervals (the reference reads
        # out of its arrays for energies at or beyond the grid's ends, :76)
        lo = min(max(int(np.searchsorted(self.energies, E, side="left")) - 1, 0), len(self.energies) - 2)
        up = lo + 1
        xl, xu = self.x_from_y(lo, y), self.x_from_y(up, y)
        El, Eu = self.energies[lo], self.energies[up]
        return math.exp(math.log(xl + 1) + (math.log(xu + 1) - math.log(xl + 1))
                        * (math.log(E) - math.log(El)) / (math.log(Eu) - math.log(El))) - 1.0


# --------------------------------------------------------------------------- material tables
class Material:
    def __init__(self, mid, registry=None, rita=True):
        reg = registry or load_registry()
        self.id = mid
        self.name, self.density, comp = reg[mid]
        self.comp = sorted(comp)
        self.elements = {z: element(z) for z, _ in self.comp}
        self.number_density = {z: AVOGADRO_CONSTANT * self.density * w / self.elements[z][0]
                               for z, w in self.comp}
        self.mass_fraction = dict(self.comp)
        self.coh = self._weighted("coh", True, spline=True)
        self.incoh = self._weighted("incoh", True, spline=True)
        self.photo = self._weighted("photo", True, spline=False)
        self.coh_i = LogLogSpline(*self.coh)
        self.incoh_i = LogLogSpline(*self.incoh)
        self.photo_i = LogLogLinear(*self.photo)
        e = merge(self.incoh[0], self.coh[0], self.photo[0])
        self.total = (e, self.incoh_i(e) + self.coh_i(e) + self.photo_i(e))
        self.total_i = LogLogLinear(*self.total)
        self.sf = self._weighted("sf", False, spline=False)
        self.ff = self._weighted("ff", False, spline=False)
        self.sf_i = LogLogLinear(*self.sf)
        self.ff_i = LogLogLinear(*self.ff)
        self.rita = None
        if rita:
            self.rita = Rita(self.rayleigh_pdf, self.incoh[0])

    def _weighted(self, name, macroscopic, spline):
        grids, interps, weights = [], [], []
        for z, _ in self.comp:
            x, y = self.elements[z][1][name]
            grids.append(x)
            interps.append(LogLogSpline(x, y) if spline else LogLogLinear(x, y))
            weights.append(self.number_density[z] if macroscopic else self.mass_fraction[z])
        e = merge(*grids)
        acc = np.zeros_like(e)
        for w, f in zip(weights, interps):
            acc = acc + w * f(e)
        if macroscopic:
            acc = acc * 1e-24
        return e, acc

    def rayleigh_pdf(self, mu, E):
        """material_data.cpp:64-68: (1 + mu^2) FF(x)^2, x = ALPHA k sqrt(1 - mu)."""
        k = E / ELECTRON_REST_MASS
        x = ALPHA * k * np.sqrt(1 - mu)
        # log10(0) -> the clamp returns the first node's value (see module note)
        ff = self.ff_i(np.maximum(x, 0.0))
        return (1 + mu * mu) * ff ** 2


def majorant(materials):
    e = merge(*[m.total[0] for m in materials])
    v = np.max(np.stack([m.total_i(e) for m in materials]), axis=0)
    return e, v
