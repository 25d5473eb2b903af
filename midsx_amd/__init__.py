"""Python host mirror of the MI355X photon-transport library.

The reference's surface is its C++ class API (include/MIDSX/Core.h); this
module exposes the same objects -- a domain, interaction tables, a photon
source, volume / surface tallies and runSimulation -- over the C-ABI of
libmidsx_amd.so (include/midsx_amd.h) with ctypes.  Nothing here computes
physics: every number comes from the HIP library, and importing the module
fails loudly when the library has not been built.
"""
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmidsx_amd.so")
DATA_DIR = os.path.join(HERE, "data")

# include/midsx_amd.h
OK, ERR_INVALID, ERR_RANGE, ERR_RUNTIME, ERR_DEVICE, ERR_NOMEM = 0, -1, -2, -3, -4, -5
DIR_BEAM, DIR_ISOTROPIC, DIR_RECT_ISOTROPIC, DIR_DISC_ISOTROPIC = 0, 1, 2, 3
SPEC_MONO, SPEC_POLY = 0, 1
TALLY_AACUBOID, TALLY_DISC, TALLY_RECT, TALLY_IMAGE = 0, 1, 2, 3
XS_TOTAL, XS_COHERENT, XS_INCOHERENT, XS_PHOTOELECTRIC, SF_INCOHERENT, FF_COHERENT = 0, 1, 2, 3, 4, 5
KERNEL_PLAIN, KERNEL_QUEUE = 0, 1
EVENT_QUEUE, EVENT_BANK = 0, 1
BANK_SIZES = (96, 128, 160)  # photons per wave bank of EVENT_BANK
BANK_MAX_HISTORY = 1 << 42  # EVENT_BANK launches: first + n at most this
TUNE_ROUNDS, TUNE_EVENTS = 1, 2
DOSE_F64, DOSE_F32, DOSE_F32C, DOSE_F32C_LDS = 0, 1, 2, 3
NQ, NBUCKET, NFIELD = 4, 4, 3
STAT_NAMES = ["histories", "steps", "real", "rayleigh", "compton", "photo", "kahn_trials", "kahn_s_tests",
              "deposits", "surface_hits", "escapes", "reserved"]
VOLUME_QUANTITIES = ["EnergyDeposition", "IncidentEnergy", "NumberOfPhotons", "NumberOfInteractions"]
SURFACE_QUANTITIES = ["IncidentEnergy", "EntranceCosine", "NumberOfPhotons"]
BUCKETS = ["primary", "single_incoherent", "single_coherent", "multiple"]

EXPORTS = [
    "mdx_last_error", "mdx_create", "mdx_destroy", "mdx_set_data_dir", "mdx_material_id", "mdx_material_name",
    "mdx_build_tables", "mdx_build_tables_for_domain", "mdx_set_domain", "mdx_add_grid", "mdx_load_scene",
    "mdx_num_grids", "mdx_grid_info", "mdx_get_grid_ids", "mdx_set_source", "mdx_set_tallies",
    "mdx_enable_voxel_dose", "mdx_reset_tallies", "mdx_run", "mdx_get_tally", "mdx_get_voxel_dose",
    "mdx_get_background_dose", "mdx_interpolate", "mdx_interpolate_host", "mdx_get_table", "mdx_sample_rayleigh",
    "mdx_sample_rayleigh_host", "mdx_get_rita_row", "mdx_sample_compton", "mdx_voxel_lookup",
    "mdx_select_interaction", "mdx_job_create", "mdx_job_select", "mdx_job_release", "mdx_job_device_bytes",
    "mdx_tables_device_bytes", "mdx_run_jobs", "mdx_set_queue_params", "mdx_get_queue_params",
    "mdx_autotune", "mdx_job_alias_domain", "mdx_set_graphs",
    "mdx_run_jobs_async", "mdx_wait", "mdx_mem_info", "mdx_host_alloc", "mdx_host_free",
    "mdx_set_flights", "mdx_set_rita_builder", "mdx_set_dose_precision", "mdx_get_voxel_dose_f32", "mdx_stage_voxel_dose_f32",
    "mdx_read_staged_dose_f32_async", "mdx_read_done", "mdx_copy_probe", "mdx_atomic_probe", "mdx_air_muen", "mdx_set_air_muen_table",
    "mdx_set_local_majorants", "mdx_get_macro_cells", "mdx_delta_decision", "mdx_philox",
    "mdx_sample_source", "mdx_run_jobs_noalloc", "mdx_roi_stats", "mdx_set_event_kernel", "mdx_get_event_kernel",
    "mdx_rotate",
]


class MdxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (status %d)" % (msg, status))
        self.status = status


class _Source(ctypes.Structure):
    _fields_ = [("direction", ctypes.c_int), ("spectrum", ctypes.c_int), ("position", ctypes.c_double * 3),
                ("a", ctypes.c_double * 3), ("b", ctypes.c_double * 3), ("c", ctypes.c_double * 3),
                ("radius", ctypes.c_double), ("energy", ctypes.c_double), ("n_bins", ctypes.c_int),
                ("bin_energy", ctypes.POINTER(ctypes.c_double)), ("bin_weight", ctypes.POINTER(ctypes.c_double))]


class _Tally(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("p0", ctypes.c_double * 3), ("p1", ctypes.c_double * 3),
                ("p2", ctypes.c_double * 3), ("normal", ctypes.c_double * 3), ("radius", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("hist_bins", ctypes.c_int),
                ("hist_lo", ctypes.c_double), ("hist_width", ctypes.c_double)]


class _Stats(ctypes.Structure):
    _fields_ = [("counters", ctypes.c_uint64 * 12), ("kernel_ms", ctypes.c_double)]


_lib = None


def load_library(path=None):
    """Loads libmidsx_amd.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError("libmidsx_amd.so not built (%s); run python -m midsx_amd.build" % path)
    lib = ctypes.CDLL(path)
    lib.mdx_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        getattr(lib, name)  # every declared entry point must resolve
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _v3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class Context:
    """One device context (device=-1: host-only, tables and registry only)."""

    def __init__(self, device=0, data_dir=None):
        self.lib = load_library()
        self._h = ctypes.c_void_p()
        self._call(self.lib.mdx_create, ctypes.c_int(device), ctypes.byref(self._h))
        self.device = device
        if data_dir:
            self._call(self.lib.mdx_set_data_dir, self._h, data_dir.encode())
        self._job = 0
        self._job_tallies = {0: []}
        # free pinned host buffers by element count, shared by the request
        # loops run on this context (serve.Server), so that a long-running
        # service pins its result buffers once
        self.pinned_pool = {}

    @property
    def tallies(self):
        """Tally descriptions of the selected job."""
        return self._job_tallies.setdefault(self._job, [])

    def _call(self, fn, *args):
        st = fn(*args)
        if st != OK:
            raise MdxError(st, self.lib.mdx_last_error().decode())

    def close(self):
        if self._h:
            self.lib.mdx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- materials / tables
    def material_id(self, name):
        out = ctypes.c_int()
        self._call(self.lib.mdx_material_id, self._h, name.encode(), ctypes.byref(out))
        return out.value

    def material_name(self, mid):
        buf = ctypes.create_string_buffer(256)
        self._call(self.lib.mdx_material_name, self._h, ctypes.c_int(mid), buf, ctypes.c_int(256))
        return buf.value.decode()

    def build_tables(self, material_ids=None):
        if material_ids is None:
            self._call(self.lib.mdx_build_tables_for_domain, self._h)
        else:
            ids = np.ascontiguousarray(material_ids, dtype=np.int32)
            self._call(self.lib.mdx_build_tables, self._h, _ptr(ids, ctypes.c_int), ctypes.c_int(len(ids)))

    def table(self, material_id, which):
        n = ctypes.c_int()
        self._call(self.lib.mdx_get_table, self._h, ctypes.c_int(material_id), ctypes.c_int(which), ctypes.byref(n),
                   None, None)
        x = np.empty(n.value)
        y = np.empty(n.value)
        self._call(self.lib.mdx_get_table, self._h, ctypes.c_int(material_id), ctypes.c_int(which), ctypes.byref(n),
                   _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double))
        return x, y

    def interpolate(self, material_id, which, x):
        x = _f32(x)
        out = np.empty_like(x)
        self._call(self.lib.mdx_interpolate, self._h, ctypes.c_int(material_id), ctypes.c_int(which),
                   ctypes.c_int(x.size), _ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float))
        return out

    def interpolate_host(self, material_id, which, x):
        x = _f64(x)
        out = np.empty_like(x)
        self._call(self.lib.mdx_interpolate_host, self._h, ctypes.c_int(material_id), ctypes.c_int(which),
                   ctypes.c_int(x.size), _ptr(x, ctypes.c_double), _ptr(out, ctypes.c_double))
        return out

    def rita_row(self, material_id, energy_index):
        rows = ctypes.c_int()
        e = ctypes.c_double()
        self._call(self.lib.mdx_get_rita_row, self._h, ctypes.c_int(material_id), ctypes.c_int(energy_index),
                   ctypes.byref(rows), ctypes.byref(e), None, None, None, None)
        x, y, a, b = (np.empty(rows.value) for _ in range(4))
        self._call(self.lib.mdx_get_rita_row, self._h, ctypes.c_int(material_id), ctypes.c_int(energy_index),
                   ctypes.byref(rows), ctypes.byref(e), _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
                   _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double))
        return e.value, x, y, a[:-1], b[:-1]

    def sample_rayleigh(self, material_id, energy, y):
        e, y = _f32(energy), _f32(y)
        out = np.empty_like(e)
        self._call(self.lib.mdx_sample_rayleigh, self._h, ctypes.c_int(material_id), ctypes.c_int(e.size),
                   _ptr(e, ctypes.c_float), _ptr(y, ctypes.c_float), _ptr(out, ctypes.c_float))
        return out

    def sample_rayleigh_host(self, material_id, energy, y):
        e, y = _f64(energy), _f64(y)
        out = np.empty_like(e)
        self._call(self.lib.mdx_sample_rayleigh_host, self._h, ctypes.c_int(material_id), ctypes.c_int(e.size),
                   _ptr(e, ctypes.c_double), _ptr(y, ctypes.c_double), _ptr(out, ctypes.c_double))
        return out

    def sample_compton(self, material_id, energy, seed=1):
        e = _f32(energy)
        mu = np.empty_like(e)
        ep = np.empty_like(e)
        self._call(self.lib.mdx_sample_compton, self._h, ctypes.c_int(material_id), ctypes.c_int(e.size),
                   _ptr(e, ctypes.c_float), ctypes.c_uint64(seed), _ptr(mu, ctypes.c_float), _ptr(ep, ctypes.c_float))
        return mu, ep

    def select_interaction(self, s_coh, s_inc, s_tot, u):
        a, b, c, d = _f32(s_coh), _f32(s_inc), _f32(s_tot), _f32(u)
        out = np.empty(a.size, dtype=np.int32)
        self._call(self.lib.mdx_select_interaction, self._h, ctypes.c_int(a.size), _ptr(a, ctypes.c_float),
                   _ptr(b, ctypes.c_float), _ptr(c, ctypes.c_float), _ptr(d, ctypes.c_float), _ptr(out, ctypes.c_int))
        return out

    def delta_decision(self, s_tot, s_max, u, limited=None):
        """1 where the device takes a real collision (physics_engine.cpp:180-183)."""
        a, b, d = _f32(s_tot), _f32(s_max), _f32(u)
        lim = np.zeros(a.size, np.int32) if limited is None else np.ascontiguousarray(limited, dtype=np.int32)
        out = np.empty(a.size, dtype=np.int32)
        self._call(self.lib.mdx_delta_decision, self._h, ctypes.c_int(a.size), _ptr(a, ctypes.c_float),
                   _ptr(b, ctypes.c_float), _ptr(d, ctypes.c_float), _ptr(lim, ctypes.c_int), _ptr(out, ctypes.c_int))
        return out

    def rotate(self, dirs, mu, phi):
        """The device's direction change (particle.cpp:7-45): unit dirs (n, 3)
        turned by polar cosine mu and azimuth phi."""
        d = _f32(dirs).reshape(-1, 3)
        m, p = _f32(mu), _f32(phi)
        out = np.empty_like(d)
        self._call(self.lib.mdx_rotate, self._h, ctypes.c_int(m.size), _ptr(d, ctypes.c_float),
                   _ptr(m, ctypes.c_float), _ptr(p, ctypes.c_float), _ptr(out, ctypes.c_float))
        return out

    def sample_source(self, n, seed=1):
        """Directions (n x 3) and energies of the selected job's source
        photons for histories 0 .. n-1, drawn on the device."""
        d = np.empty(3 * n, np.float32)
        e = np.empty(n, np.float32)
        self._call(self.lib.mdx_sample_source, self._h, ctypes.c_int(n), ctypes.c_uint64(seed),
                   _ptr(d, ctypes.c_float), _ptr(e, ctypes.c_float))
        return d.reshape(n, 3), e

    def philox(self, counter, key, words=4):
        """Raw Philox-4x32-10 (words 4: counter n x 4, key n x 2) or
        Philox-2x32-10 (words 2: counter n x 2, key n) blocks on the device."""
        c = np.ascontiguousarray(counter, dtype=np.uint32).reshape(-1, words)
        k = np.ascontiguousarray(key, dtype=np.uint32).reshape(c.shape[0], -1)
        out = np.empty_like(c)
        self._call(self.lib.mdx_philox, self._h, ctypes.c_int(c.shape[0]), ctypes.c_int(words),
                   _ptr(c, ctypes.c_uint32), _ptr(k, ctypes.c_uint32), _ptr(out, ctypes.c_uint32))
        return out

    # ---------------------------------------------------------- domain
    def set_domain(self, dim_space, background_id):
        self._call(self.lib.mdx_set_domain, self._h, _v3(dim_space), ctypes.c_int(background_id))

    def add_grid(self, ids, spacing, origin):
        """ids indexed [x, y, z] (stored x-fastest)."""
        ids = np.asarray(ids, dtype=np.uint8)
        dims = (ctypes.c_int * 3)(*ids.shape)
        flat = np.ascontiguousarray(ids.transpose(2, 1, 0)).ravel()
        self._call(self.lib.mdx_add_grid, self._h, _ptr(flat, ctypes.c_uint8), dims, _v3(spacing), _v3(origin))

    def load_scene(self, json_path):
        self._call(self.lib.mdx_load_scene, self._h, os.fspath(json_path).encode())

    def num_grids(self):
        n = ctypes.c_int()
        self._call(self.lib.mdx_num_grids, self._h, ctypes.byref(n))
        return n.value

    def grid_info(self, g):
        dims = (ctypes.c_int * 3)()
        sp = (ctypes.c_double * 3)()
        org = (ctypes.c_double * 3)()
        self._call(self.lib.mdx_grid_info, self._h, ctypes.c_int(g), dims, sp, org)
        return tuple(dims), tuple(sp), tuple(org)

    def grid_ids(self, g):
        dims, _, _ = self.grid_info(g)
        out = np.empty(dims[0] * dims[1] * dims[2], dtype=np.uint8)
        self._call(self.lib.mdx_get_grid_ids, self._h, ctypes.c_int(g), _ptr(out, ctypes.c_uint8))
        return out.reshape(dims[2], dims[1], dims[0]).transpose(2, 1, 0)

    def voxel_lookup(self, xyz):
        p = _f32(xyz).reshape(-1, 3)
        n = p.shape[0]
        g = np.empty(n, dtype=np.int32)
        lin = np.empty(n, dtype=np.int64)
        m = np.empty(n, dtype=np.int32)
        self._call(self.lib.mdx_voxel_lookup, self._h, ctypes.c_int(n), _ptr(p, ctypes.c_float),
                   _ptr(g, ctypes.c_int), _ptr(lin, ctypes.c_longlong), _ptr(m, ctypes.c_int))
        return g, lin, m

    # ---------------------------------------------------------- source / tallies / run
    def set_source(self, direction=DIR_BEAM, position=(0, 0, 0), a=(0, 0, 1), b=(0, 0, 0), c=(0, 0, 0),
                   radius=0.0, energy=None, spectrum=None):
        s = _Source()
        s.direction = direction
        s.position = _v3(position)
        s.a, s.b, s.c = _v3(a), _v3(b), _v3(c)
        s.radius = radius
        keep = []
        if spectrum is not None:
            e = _f64(spectrum[0])
            w = _f64(spectrum[1])
            keep = [e, w]
            s.spectrum = SPEC_POLY
            s.n_bins = e.size
            s.bin_energy = _ptr(e, ctypes.c_double)
            s.bin_weight = _ptr(w, ctypes.c_double)
        else:
            s.spectrum = SPEC_MONO
            s.energy = float(energy)
        self._call(self.lib.mdx_set_source, self._h, ctypes.byref(s))
        del keep

    def set_tallies(self, tallies):
        arr = (_Tally * max(1, len(tallies)))()
        for i, t in enumerate(tallies):
            q = arr[i]
            q.kind = t["kind"]
            q.p0 = _v3(t.get("p0", (0, 0, 0)))
            q.p1 = _v3(t.get("p1", (0, 0, 0)))
            q.p2 = _v3(t.get("p2", (0, 0, 0)))
            q.normal = _v3(t.get("normal", (0, 0, 1)))
            q.radius = t.get("radius", 0.0)
            q.nx, q.ny = t.get("nx", 0), t.get("ny", 0)
            q.hist_bins = t.get("hist_bins", 0)
            q.hist_lo = t.get("hist_lo", 0.0)
            q.hist_width = t.get("hist_width", 0.0)
        self._call(self.lib.mdx_set_tallies, self._h, arr, ctypes.c_int(len(tallies)))
        self._job_tallies[self._job] = [dict(t) for t in tallies]

    def enable_voxel_dose(self, on=True):
        self._call(self.lib.mdx_enable_voxel_dose, self._h, ctypes.c_int(1 if on else 0))

    def reset_tallies(self):
        self._call(self.lib.mdx_reset_tallies, self._h)

    def run(self, n, seed=1, first=0, kernel=KERNEL_PLAIN):
        st = _Stats()
        self._call(self.lib.mdx_run, self._h, ctypes.c_uint64(n), ctypes.c_uint64(seed), ctypes.c_uint64(first),
                   ctypes.c_int(kernel), ctypes.byref(st))
        out = {name: int(st.counters[i]) for i, name in enumerate(STAT_NAMES)}
        out["kernel_ms"] = st.kernel_ms
        return out

    def set_queue_params(self, blocks_per_cu=4, refill_min=32, interact_min=32):
        self._call(self.lib.mdx_set_queue_params, self._h, ctypes.c_int(blocks_per_cu), ctypes.c_int(refill_min),
                   ctypes.c_int(interact_min))

    def autotune(self, probe_histories=1000000, rounds=False, events=True):
        """Tunes the event kernel and its launch parameters on the selected job
        (resets its tallies): rounds=True probes the multi-job kernels that
        run_jobs and graph rounds launch for several jobs; events=False keeps
        to the lane-resident queue kernel."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        flags = (TUNE_ROUNDS if rounds else 0) | (TUNE_EVENTS if events else 0)
        self._call(self.lib.mdx_autotune, self._h, ctypes.c_uint64(probe_histories), ctypes.c_int(flags),
                   ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(ms))
        kind, bank = self.event_kernel()
        return dict(blocks_per_cu=a.value, refill_min=b.value, interact_min=c.value, probe_ms=ms.value,
                    event_kernel="bank" if kind == EVENT_BANK else "queue", bank_size=bank)

    def set_event_kernel(self, kind=EVENT_QUEUE, bank_size=128):
        """The event form of KERNEL_QUEUE runs: EVENT_QUEUE (lane-resident
        photons) or EVENT_BANK (per-wave photon banks of bank_size in LDS, one
        of BANK_SIZES; history indices below BANK_MAX_HISTORY)."""
        self._call(self.lib.mdx_set_event_kernel, self._h, ctypes.c_int(kind), ctypes.c_int(bank_size))

    def event_kernel(self):
        k, b = ctypes.c_int(), ctypes.c_int()
        self._call(self.lib.mdx_get_event_kernel, self._h, ctypes.byref(k), ctypes.byref(b))
        return k.value, b.value

    def queue_params(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._call(self.lib.mdx_get_queue_params, self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    # ---------------------------------------------------------- jobs
    @property
    def job(self):
        """The selected job."""
        return self._job

    def job_create(self):
        j = ctypes.c_int()
        self._call(self.lib.mdx_job_create, self._h, ctypes.byref(j))
        self._job = j.value
        self._job_tallies[j.value] = []
        return j.value

    def job_select(self, j):
        self._call(self.lib.mdx_job_select, self._h, ctypes.c_int(j))
        self._job = j

    def job_release(self, j):
        self._call(self.lib.mdx_job_release, self._h, ctypes.c_int(j))
        self._job_tallies.pop(j, None)
        if self._job == j:
            self._job = 0

    def job_alias_domain(self, src):
        """The selected job uses job src's domain (grids, voxel dose)."""
        self._call(self.lib.mdx_job_alias_domain, self._h, ctypes.c_int(src))

    def set_dose_precision(self, mode):
        """DOSE_F64, DOSE_F32 or DOSE_F32C accumulation of the voxel dose
        (selected job, which must own its domain); resets the dose."""
        self._call(self.lib.mdx_set_dose_precision, self._h, ctypes.c_int(mode))

    def set_rita_builder(self, on_device=True):
        """Build the RITA tables on the device (default) or the host, both FP64."""
        self._call(self.lib.mdx_set_rita_builder, self._h, ctypes.c_int(1 if on_device else 0))

    def set_flights(self, on=True):
        """Background free flights for the selected job (queue kernel)."""
        self._call(self.lib.mdx_set_flights, self._h, ctypes.c_int(1 if on else 0))

    def set_local_majorants(self, on=True):
        """Local majorants on macro cells for the queue kernel (all jobs)."""
        self._call(self.lib.mdx_set_local_majorants, self._h, ctypes.c_int(1 if on else 0))

    def macro_cells(self):
        """The selected job's macro-cell table as (class, radius, uniform)
        arrays of shape (cz, cy, cx) (class 255: none; uniform: the cell holds
        only its class material), or None when it has none."""
        cells = (ctypes.c_int * 3)()
        self._call(self.lib.mdx_get_macro_cells, self._h, cells, None)
        if cells[0] == 0:
            return None
        w = np.zeros(cells[0] * cells[1] * cells[2], np.uint16)
        self._call(self.lib.mdx_get_macro_cells, self._h, cells, _ptr(w, ctypes.c_uint16))
        w = w.reshape(cells[2], cells[1], cells[0])
        low = (w & 255).astype(np.int32)
        cls = np.where(low == 255, 255, low & 0x7F)
        return cls, (w >> 8).astype(np.int32), (low != 255) & ((low & 0x80) != 0)

    def set_graphs(self, on=True):
        self._call(self.lib.mdx_set_graphs, self._h, ctypes.c_int(1 if on else 0))

    def job_device_bytes(self, j):
        b = ctypes.c_uint64()
        self._call(self.lib.mdx_job_device_bytes, self._h, ctypes.c_int(j), ctypes.byref(b))
        return b.value

    def tables_device_bytes(self):
        b = ctypes.c_uint64()
        self._call(self.lib.mdx_tables_device_bytes, self._h, ctypes.byref(b))
        return b.value

    def run_jobs(self, jobs, n, first=None, seeds=None):
        """One queue-kernel launch over several jobs; n[i] histories of job
        jobs[i] starting at history first[i] with seed seeds[i]."""
        k = len(jobs)
        ids = np.ascontiguousarray(jobs, dtype=np.int32)
        nn = np.ascontiguousarray(n, dtype=np.uint64)
        ff = np.ascontiguousarray(first if first is not None else np.zeros(k), dtype=np.uint64)
        ss = np.ascontiguousarray(seeds if seeds is not None else np.ones(k), dtype=np.uint64)
        st = _Stats()
        self._call(self.lib.mdx_run_jobs, self._h, ctypes.c_int(k), _ptr(ids, ctypes.c_int),
                   _ptr(ff, ctypes.c_uint64), _ptr(nn, ctypes.c_uint64), _ptr(ss, ctypes.c_uint64), ctypes.byref(st))
        out = {name: int(st.counters[i]) for i, name in enumerate(STAT_NAMES)}
        out["kernel_ms"] = st.kernel_ms
        return out

    def run_jobs_noalloc(self, jobs, n, first=None, seeds=None):
        """run_jobs without device allocation (mdx_run_jobs_noalloc): the
        launch's records and counters in the library's static scratch; at
        most 256 jobs."""
        k = len(jobs)
        ids = np.ascontiguousarray(jobs, dtype=np.int32)
        nn = np.ascontiguousarray(n, dtype=np.uint64)
        ff = np.ascontiguousarray(first if first is not None else np.zeros(k), dtype=np.uint64)
        ss = np.ascontiguousarray(seeds if seeds is not None else np.ones(k), dtype=np.uint64)
        st = _Stats()
        self._call(self.lib.mdx_run_jobs_noalloc, self._h, ctypes.c_int(k), _ptr(ids, ctypes.c_int),
                   _ptr(ff, ctypes.c_uint64), _ptr(nn, ctypes.c_uint64), _ptr(ss, ctypes.c_uint64), ctypes.byref(st))
        out = {name: int(st.counters[i]) for i, name in enumerate(STAT_NAMES)}
        out["kernel_ms"] = st.kernel_ms
        return out

    def run_jobs_async(self, jobs, n, first=None, seeds=None):
        """Queues one round (as run_jobs) and returns; wait() ends it.  Tally
        and dose reads overlap the round; every other call waits for it."""
        k = len(jobs)
        ids = np.ascontiguousarray(jobs, dtype=np.int32)
        nn = np.ascontiguousarray(n, dtype=np.uint64)
        ff = np.ascontiguousarray(first if first is not None else np.zeros(k), dtype=np.uint64)
        ss = np.ascontiguousarray(seeds if seeds is not None else np.ones(k), dtype=np.uint64)
        self._call(self.lib.mdx_run_jobs_async, self._h, ctypes.c_int(k), _ptr(ids, ctypes.c_int),
                   _ptr(ff, ctypes.c_uint64), _ptr(nn, ctypes.c_uint64), _ptr(ss, ctypes.c_uint64))

    def wait(self):
        st = _Stats()
        self._call(self.lib.mdx_wait, self._h, ctypes.byref(st))
        out = {name: int(st.counters[i]) for i, name in enumerate(STAT_NAMES)}
        out["kernel_ms"] = st.kernel_ms
        return out

    def tally(self, i):
        t = self.tallies[i]
        acc = np.zeros(NQ * NBUCKET * NFIELD)
        hist = np.zeros(4 * t.get("hist_bins", 0)) if t.get("hist_bins", 0) and t["kind"] != TALLY_AACUBOID else None
        image = np.zeros(3 * t.get("nx", 0) * t.get("ny", 0)) if t["kind"] == TALLY_IMAGE else None
        self._call(self.lib.mdx_get_tally, self._h, ctypes.c_int(i), _ptr(acc, ctypes.c_double),
                   None if hist is None else _ptr(hist, ctypes.c_double),
                   None if image is None else _ptr(image, ctypes.c_double))
        out = {"acc": acc.reshape(NQ, NBUCKET, NFIELD)}
        if hist is not None:
            out["hist"] = hist.reshape(4, -1)
        if image is not None:
            out["image"] = image.reshape(t["ny"], t["nx"], 3)
        return out

    def voxel_dose(self, g, out=None, counts=True):
        """(sum, sum^2, count) of the deposits per voxel, x fastest.  out: a
        float64 buffer of 2 * voxels to read (sum, sum^2) pairs into;
        counts=False skips the counts (returned as None)."""
        dims, _, _ = self.grid_info(g)
        nv = dims[0] * dims[1] * dims[2]
        ss = np.zeros(2 * nv) if out is None else out[:2 * nv]
        assert ss.dtype == np.float64 and ss.flags.c_contiguous and ss.size == 2 * nv
        cnt = np.zeros(nv, dtype=np.uint64) if counts else None
        self._call(self.lib.mdx_get_voxel_dose, self._h, ctypes.c_int(g), _ptr(ss, ctypes.c_double),
                   None if cnt is None else _ptr(cnt, ctypes.c_uint64))
        return ss[0::2], ss[1::2], cnt

    def roi_stats(self, jobs, boxes):
        """(n, sum, sum^2) of the deposits in each box (i0, i1, j0, j1, k0, k1,
        half-open voxel index ranges) of each job's grid 0; one row per box."""
        j = np.ascontiguousarray(jobs, dtype=np.int32)
        b = np.ascontiguousarray(boxes, dtype=np.int32).reshape(-1, 6)
        out = np.zeros((len(j), 3))
        if len(j):
            self._call(self.lib.mdx_roi_stats, self._h, ctypes.c_int(len(j)), _ptr(j, ctypes.c_int),
                       _ptr(b, ctypes.c_int), _ptr(out, ctypes.c_double))
        return out

    def voxel_dose_f32(self, g, n_histories, out=None):
        """(mean deposit per history, relative standard error) per voxel as
        float32 arrays (x fastest), computed on the device.  out: a float32
        buffer of 2 * voxels."""
        dims, _, _ = self.grid_info(g)
        nv = dims[0] * dims[1] * dims[2]
        buf = np.empty(2 * nv, np.float32) if out is None else out[:2 * nv]
        assert buf.dtype == np.float32 and buf.flags.c_contiguous and buf.size == 2 * nv
        self._call(self.lib.mdx_get_voxel_dose_f32, self._h, ctypes.c_int(g), ctypes.c_uint64(n_histories),
                   _ptr(buf, ctypes.c_float))
        return buf[0::2], buf[1::2]

    def stage_voxel_dose_f32(self, g, n_histories):
        """Queues the FP32 conversion behind the running round; the next
        voxel_dose_f32(g, n_histories) then only copies."""
        self._call(self.lib.mdx_stage_voxel_dose_f32, self._h, ctypes.c_int(g), ctypes.c_uint64(n_histories))

    def copy_probe(self, nbytes):
        """Milliseconds of one streaming copy of nbytes (read + write)."""
        ms = ctypes.c_double()
        self._call(self.lib.mdx_copy_probe, self._h, ctypes.c_uint64(nbytes), ctypes.byref(ms))
        return ms.value

    def atomic_probe(self, records, deposits, check=False):
        """Milliseconds of `deposits` scattered voxel deposits (three
        device-scope atomics each) into `records` 32-byte voxel records; with
        check, also the deposit count the records hold after the probe's two
        launches (2 x deposits; read back to the host)."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        self._call(self.lib.mdx_atomic_probe, self._h, ctypes.c_uint64(records), ctypes.c_uint64(deposits),
                   ctypes.byref(ms), ctypes.byref(n) if check else None)
        return (ms.value, n.value) if check else ms.value

    def air_muen(self, energy):
        """mu_en / rho of dry air (cm^2/g) at energies in eV."""
        e = _f64(energy)
        out = np.empty_like(e)
        self._call(self.lib.mdx_air_muen, self._h, ctypes.c_int(e.size), _ptr(e, ctypes.c_double),
                   _ptr(out, ctypes.c_double))
        return out

    def set_air_muen_table(self, table):
        """0: NIST grid (default), 1: TG-195 0.5 keV grid."""
        self._call(self.lib.mdx_set_air_muen_table, self._h, ctypes.c_int(table))

    def read_staged_dose_f32_async(self, out):
        """Starts copying the selected job's staged FP32 dose into out (a
        float32 array of 2 * voxels, pinned for a copy that does not block)."""
        assert out.dtype == np.float32 and out.flags.c_contiguous
        self._call(self.lib.mdx_read_staged_dose_f32_async, self._h, _ptr(out, ctypes.c_float))

    def read_done(self, job, wait=False):
        """True once the copy started by read_staged_dose_f32_async for job
        has finished (wait=True blocks until then)."""
        d = ctypes.c_int()
        self._call(self.lib.mdx_read_done, self._h, ctypes.c_int(job), ctypes.c_int(1 if wait else 0),
                   ctypes.byref(d))
        return bool(d.value)

    def mem_info(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        self._call(self.lib.mdx_mem_info, self._h, ctypes.byref(f), ctypes.byref(t))
        return f.value, t.value

    def pinned(self, n, dtype=np.float64):
        """A numpy array of n elements in pinned host memory, freed when the
        array (and every view of it) is garbage collected."""
        nbytes = int(n) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        self._call(self.lib.mdx_host_alloc, self._h, ctypes.c_uint64(nbytes), ctypes.byref(p))
        buf = (ctypes.c_char * max(1, nbytes)).from_address(p.value)
        arr = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dtype)
        lib, addr = self.lib, p.value
        weakref.finalize(buf, lambda: lib.mdx_host_free(ctypes.c_void_p(addr)))
        return arr

    def background_dose(self):
        out = (ctypes.c_double * 3)()
        self._call(self.lib.mdx_get_background_dose, self._h, out)
        return tuple(out)


def vector_stats(acc_qb):
    """VectorValue statistics (quantity.cpp:27-94) from (n, sum, sum^2):
    sum, sample variance, sigma of the sum sqrt(n s^2), mean and its sigma.
    The variance comes from the moments, clamped at 0 (rounding can take the
    moment form below 0 for near-constant values; the reference's two-pass
    form never goes there).  No values: variance and sigma of the sum 0, the
    mean and its sigma NaN (0 / 0); one value: variance NaN (0 / 0), as
    quantity.cpp computes them."""
    n, s, s2 = acc_qb
    nan = float("nan")
    if n == 0:
        return {"count": 0, "sum": s, "variance": 0.0, "sum_std": 0.0, "mean": nan, "mean_std": nan}
    if n == 1:
        return {"count": n, "sum": s, "variance": nan, "sum_std": nan, "mean": s, "mean_std": nan}
    mean = s / n
    var = max((s2 - mean * s) / (n - 1), 0.0)
    return {"count": n, "sum": s, "variance": var, "sum_std": (n * var) ** 0.5, "mean": mean,
            "mean_std": (var / n) ** 0.5}
