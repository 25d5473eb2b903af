"""ORACLE -- test infrastructure only.

A CPU restatement of the reference's photon-transport path, used as the
checker by tests/, by __graft_entry__.smoke() and as the timed CPU baseline in
bench.py.  The library (midsx_amd) never imports it.

  tables.py              FP64 numpy restatement of the table build
  transport_oracle.cpp   FP64 C++/OpenMP restatement of runSimulation and the
                         per-step physics and tallies (built by build.py)

Parity status of each part (how it is pinned) is listed in the tests that
check it: tests/test_oracle_pins.py.
"""
import ctypes
import os

import numpy as np

from . import tables as T

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            from . import build
            build.build()
        L = ctypes.CDLL(SO)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_run.restype = ctypes.c_double
        L.orc_rita_sample.restype = ctypes.c_double
        L.orc_interp.restype = ctypes.c_double
        for name in ["orc_destroy", "orc_set_majorant", "orc_add_material", "orc_set_domain", "orc_add_grid",
                     "orc_set_source", "orc_set_tallies", "orc_enable_voxel_dose", "orc_run", "orc_get_tally",
                     "orc_get_voxel_dose", "orc_get_background_dose", "orc_get_counters", "orc_reset",
                     "orc_rita_sample", "orc_interp", "orc_kahn_sample", "orc_rotate", "orc_aacuboid_traversal",
                     "orc_sample_source", "orc_max_threads"]:
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _v3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class _TallyDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("p0", ctypes.c_double * 3), ("p1", ctypes.c_double * 3),
                ("p2", ctypes.c_double * 3), ("normal", ctypes.c_double * 3), ("radius", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("hist_bins", ctypes.c_int),
                ("hist_lo", ctypes.c_double), ("hist_width", ctypes.c_double)]


_MAT_CACHE = {}


def material(mid):
    if mid not in _MAT_CACHE:
        _MAT_CACHE[mid] = T.Material(mid)
    return _MAT_CACHE[mid]


class Simulation:
    """Oracle counterpart of midsx_amd.Context (same call shapes)."""

    def __init__(self, material_ids):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.orc_create())
        self.mats = [material(m) for m in dict.fromkeys(material_ids)]
        maj = T.majorant(self.mats)
        self._keep = []
        x, px = _d(maj[0])
        y, py = _d(maj[1])
        self.L.orc_set_majorant(self.h, ctypes.c_int(len(x)), px, py)
        for m in self.mats:
            self._add(m)
        self.tallies = []
        self.grids = []

    def _add(self, m):
        keep = []

        def d(a):
            arr, p = _d(a)
            keep.append(arr)
            return p

        r = m.rita
        rows_x, rows_y, rows_a, rows_b, off = [], [], [], [], [0]
        for (g, c, a, b) in r.rows:
            rows_x.append(g)
            rows_y.append(c)
            rows_a.append(np.append(a, 0.0))
            rows_b.append(np.append(b, 0.0))
            off.append(off[-1] + len(g))
        off = np.array(off, dtype=np.int32)
        keep.append(off)
        cs, ci = m.coh_i, m.incoh_i
        self.L.orc_add_material(
            self.h, ctypes.c_int(m.id),
            ctypes.c_int(len(m.total[0])), d(m.total[0]), d(m.total[1]),
            ctypes.c_int(len(m.sf[0])), d(m.sf[0]), d(m.sf[1]),
            ctypes.c_int(len(cs.t)), d(cs.t), d(cs.coef), ctypes.c_double(cs.xmin), ctypes.c_double(cs.xmax),
            ctypes.c_int(cs.p),
            ctypes.c_int(len(ci.t)), d(ci.t), d(ci.coef), ctypes.c_double(ci.xmin), ctypes.c_double(ci.xmax),
            ctypes.c_int(ci.p),
            ctypes.c_int(len(r.energies)), d(r.energies), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            d(np.concatenate(rows_x)), d(np.concatenate(rows_y)), d(np.concatenate(rows_a)),
            d(np.concatenate(rows_b)))
        self._keep.append(keep)

    def set_domain(self, dim_space, background_id):
        self.L.orc_set_domain(self.h, _v3(dim_space), ctypes.c_int(background_id))
        self.grids = []

    def add_grid(self, ids, spacing, origin):
        ids = np.asarray(ids, dtype=np.uint8)
        flat = np.ascontiguousarray(ids.transpose(2, 1, 0)).ravel()
        dims = (ctypes.c_int * 3)(*ids.shape)
        self.L.orc_add_grid(self.h, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), dims, _v3(spacing),
                            _v3(origin))
        self.grids.append(ids.shape)

    def set_source(self, direction=0, position=(0, 0, 0), a=(0, 0, 1), b=(0, 0, 0), c=(0, 0, 0), radius=0.0,
                   energy=None, spectrum=None):
        if spectrum is not None:
            e, pe = _d(spectrum[0])
            w, pw = _d(spectrum[1])
            self._keep.append((e, w))
            self.L.orc_set_source(self.h, ctypes.c_int(direction), ctypes.c_int(1), _v3(position), _v3(a), _v3(b),
                                  _v3(c), ctypes.c_double(radius), ctypes.c_double(0.0), ctypes.c_int(len(e)), pe, pw)
        else:
            self.L.orc_set_source(self.h, ctypes.c_int(direction), ctypes.c_int(0), _v3(position), _v3(a), _v3(b),
                                  _v3(c), ctypes.c_double(radius), ctypes.c_double(energy), ctypes.c_int(0), None,
                                  None)

    def set_tallies(self, tallies):
        arr = (_TallyDesc * max(1, len(tallies)))()
        for i, t in enumerate(tallies):
            q = arr[i]
            q.kind = t["kind"]
            q.p0, q.p1, q.p2 = _v3(t.get("p0", (0, 0, 0))), _v3(t.get("p1", (0, 0, 0))), _v3(t.get("p2", (0, 0, 0)))
            q.normal = _v3(t.get("normal", (0, 0, 1)))
            q.radius = t.get("radius", 0.0)
            q.nx, q.ny = t.get("nx", 0), t.get("ny", 0)
            q.hist_bins = t.get("hist_bins", 0)
            q.hist_lo = t.get("hist_lo", 0.0)
            q.hist_width = t.get("hist_width", 0.0)
        self.L.orc_set_tallies(self.h, ctypes.c_int(len(tallies)), arr)
        self.tallies = [dict(t) for t in tallies]

    def enable_voxel_dose(self, on=True):
        self.L.orc_enable_voxel_dose(self.h, ctypes.c_int(1 if on else 0))

    def reset(self):
        self.L.orc_reset(self.h)

    def run(self, n, seed=1, threads=0):
        return self.L.orc_run(self.h, ctypes.c_uint64(n), ctypes.c_uint64(seed), ctypes.c_int(threads))

    def counters(self):
        out = (ctypes.c_uint64 * 12)()
        self.L.orc_get_counters(self.h, out)
        names = ["histories", "steps", "real", "rayleigh", "compton", "photo", "kahn_trials", "kahn_s_tests",
                 "deposits", "surface_hits", "escapes", "reserved"]
        return {k: int(out[i]) for i, k in enumerate(names)}

    def tally(self, i):
        t = self.tallies[i]
        acc = np.zeros(48)
        hb = t.get("hist_bins", 0) if t["kind"] != 0 else 0
        hist = np.zeros(4 * hb) if hb else None
        image = np.zeros(3 * t.get("nx", 0) * t.get("ny", 0)) if t["kind"] == 3 else None
        self.L.orc_get_tally(self.h, ctypes.c_int(i), acc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            None if hist is None else hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            None if image is None else image.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        out = {"acc": acc.reshape(4, 4, 3)}
        if hist is not None:
            out["hist"] = hist.reshape(4, -1)
        if image is not None:
            out["image"] = image.reshape(t["ny"], t["nx"], 3)
        return out

    def voxel_dose(self, g):
        nx, ny, nz = self.grids[g]
        nv = nx * ny * nz
        ss = np.zeros(2 * nv)
        cnt = np.zeros(nv, dtype=np.uint32)
        self.L.orc_get_voxel_dose(self.h, ctypes.c_int(g), ss.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return ss[0::2], ss[1::2], cnt

    def background_dose(self):
        out = (ctypes.c_double * 3)()
        self.L.orc_get_background_dose(self.h, out)
        return tuple(out)

    def rita_sample(self, mid, E, y):
        return self.L.orc_rita_sample(self.h, ctypes.c_int(mid), ctypes.c_double(E), ctypes.c_double(y))

    def kahn_sample(self, mid, E, n, seed=1):
        mu = np.zeros(n)
        ep = np.zeros(n)
        self.L.orc_kahn_sample(self.h, ctypes.c_int(mid), ctypes.c_double(E), ctypes.c_int(n), ctypes.c_uint64(seed),
                               mu.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               ep.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return mu, ep

    def sample_source(self, n, seed=1):
        """n photons of the source as generatePhoton draws them: directions
        (n x 3) and energies."""
        d = np.zeros(3 * n)
        e = np.zeros(n)
        self.L.orc_sample_source(self.h, ctypes.c_int(n), ctypes.c_uint64(seed),
                                 d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return d.reshape(n, 3), e

    def interp(self, mid, which, x):
        return self.L.orc_interp(self.h, ctypes.c_int(mid), ctypes.c_int(which), ctypes.c_double(x))

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass


TRAVERSAL = ("misses", "lands_inside", "passes_through", "starts_inside_exits", "starts_inside_stays")


def aacuboid_traversal(mn, mx, pos, direction, free_path):
    """AACuboidVolumeTally::determineVolumeTraversal for n steps; returns
    indices into TRAVERSAL."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(direction, dtype=np.float64).reshape(-1, 3)
    fp = np.ascontiguousarray(free_path, dtype=np.float64).ravel()
    out = np.zeros(len(fp), dtype=np.int32)
    lib().orc_aacuboid_traversal(_v3(mn), _v3(mx), ctypes.c_int(len(fp)),
                                 pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 fp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out


def max_threads():
    """The OpenMP team size of a run with threads=0 (omp_get_max_threads)."""
    return int(lib().orc_max_threads())


def rotate(d, theta, phi):
    """The reference's quaternion rotation (particle.cpp:7-45), FP64."""
    out = (ctypes.c_double * 3)()
    lib().orc_rotate((ctypes.c_double * 3)(*d), ctypes.c_double(theta), ctypes.c_double(phi), out)
    return np.array(out[:])
