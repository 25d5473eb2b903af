"""Continuous-batching request loop for dose jobs on one device.

Jobs (a phantom, a source, tallies and a history count) arrive over time.
The loop keeps the physics tables resident and shared, admits waiting jobs in
arrival order while their device footprint fits the memory budget, and runs
rounds: one multi-job launch of the queue kernel (a replayed hipGraph) in
which every admitted job runs its next slice of histories.  While a round
runs, the results of the jobs that finished in the round before are copied
back and newly admitted jobs are set up (phantom upload, tally zeroing), both
on a second stream beside the kernel; new jobs join the next round, and the
memory of a finished job is freed once its results are back.

A job's histories are keyed by its seed and history index, so its tallies do
not depend on the round structure nor on the other jobs in the mix (the FP64
sums differ only in their order of summation).

A job may ask for a precision instead of a history count: with
"rse_target" set, its "histories" is a cap, and it runs until every dose
voxel of its region of interest (`scenes.dose_job_roi`) has a relative
standard error at or below the target (the reference's sigma of a sum of
deposits, quantity.cpp:37-45), checked on the device after every round
(`Context.roi_stats`); its next slice is sized from the current error
(histories done x (error / target)^2).

Baseline: `run_plain` serves the same stream first-in first-out with the
plain kernel, one launch per job (one simulation per launch).
"""
import time
from collections import deque

import numpy as np

from . import KERNEL_PLAIN
from . import scenes


class JobResult:
    """What a finished job hands back: its voxel dose (sum, sum^2 per voxel,
    optionally reduced by the caller's `on_result`), its image and timing."""

    __slots__ = ("index", "arrival", "admitted", "finished", "histories", "summary", "roi_rse")

    def __init__(self, index, arrival):
        self.index = index
        self.arrival = arrival
        self.admitted = None
        self.finished = None
        self.histories = 0
        self.summary = None
        self.roi_rse = None  # largest relative standard error over the ROI dose voxels

    @property
    def latency(self):
        return self.finished - self.arrival


def roi_summary(job, mean, rel, image, stride=4):
    """From an FP32 dose read (mean per history, relative standard error):
    the largest and median relative standard error over the ROI (voxels
    holding at least 10 % of the peak dose), taken on every stride-th voxel
    along each axis, and the image's total energy fluence."""
    nx, ny, nz = job["shape"]
    sub = (slice(None, None, stride),) * 3
    mean = mean.reshape(nz, ny, nx)[sub]
    rel = rel.reshape(nz, ny, nx)[sub]
    roi = mean >= 0.1 * mean.max() if mean.max() > 0 else np.zeros(mean.shape, bool)
    rse = rel[roi] if roi.any() else np.zeros(1)
    return dict(roi_voxels=int(roi.sum()), roi_max_rse=float(rse.max()), roi_median_rse=float(np.median(rse)),
                image_energy=float(image[..., 0].sum()))


def read_dose(ctx, job, bufs, dose, n_histories=None):
    """The selected job's voxel dose into a reusable (pinned) buffer kept in
    bufs: dose "f32" gives (mean per history, relative standard error) as
    float32 for the n_histories run (default the job's histories), "f64" the
    (sum, sum^2) of the deposits as float64."""
    nv = int(np.prod(job["shape"]))
    key = ("read", nv, dose)
    if key not in bufs:
        dt = np.float32 if dose == "f32" else np.float64
        try:
            bufs[key] = ctx.pinned(2 * nv, dt)
        except Exception:
            bufs[key] = np.empty(2 * nv, dt)
    if dose == "f32":
        return ctx.voxel_dose_f32(0, n_histories or job["histories"], out=bufs[key])
    s, s2, _ = ctx.voxel_dose(0, out=bufs[key], counts=False)
    return s, s2


class _Active:
    __slots__ = ("job", "res", "slot", "done", "bytes", "image", "buf", "need", "rse", "roi")

    def __init__(self, job, res, slot, nbytes, first_slice):
        self.job = job
        self.res = res
        self.slot = slot
        self.done = 0
        self.bytes = nbytes
        self.image = None
        self.buf = None
        self.rse = np.inf
        # precision jobs: the histories planned so far (a first slice, then
        # estimates from the ROI error), within the job's cap
        self.need = min(job["histories"], first_slice) if job.get("rse_target") else job["histories"]
        self.roi = scenes.dose_job_roi(job) if job.get("rse_target") else None

    def finished(self):
        if self.done >= self.job["histories"]:
            return True
        return self.roi is not None and self.rse <= self.job["rse_target"]


class Server:
    """Continuous batching over one Context whose tables cover every job's
    materials.

    budget_bytes: device memory the resident jobs may hold (default: half of
    the free memory at construction).  round_histories: histories per round,
    shared by the admitted jobs in proportion to (phantom voxels)^size_weight
    (at least min_slice each, never more than a job has left); longer rounds
    spend less of the stream in kernel tails, launches and checks between
    rounds, shorter ones stop precision jobs closer to their target (4.5e7
    served a 32-job stream at 1 % ROI error in 4 % less time than 3e7 and
    in less than 2e7 or 6e7).  phantoms: callable job -> uint8 grid (default:
    `scenes.dose_phantom`), so that a client can hand over its own data.
    dose: format of the dose read back ("f32" or "f64", see `read_dose`);
    on_result(job, a, b, image) receives it (the buffers are reused).
    first_slice: the histories of a precision job's first round at most.
    admit_slice: seconds of job set-up per round at most while others run
    (the next round starts only when the loop gets back to it); a set-up is
    not started when its estimated time (voxels / admit_voxels_per_s: the
    phantom's upload and its macro-cell scan dominate) would overrun the
    slice, unless it is the first of the slice.  size_weight: with shares
    growing with the phantom, the jobs whose results take longest to read
    back finish earlier and their reads overlap later rounds instead of
    trailing the last one (0: equal shares; 0.5 cut a 32-job stream's
    makespan by 6 % at the same p50 / p99 latency)."""

    def __init__(self, ctx, budget_bytes=None, round_histories=45_000_000, min_slice=100_000, max_active=512,
                 graphs=True, phantoms=None, on_result=None, dose="f32", admit_slice=0.005,
                 admit_voxels_per_s=1.0e10, first_slice=1_000_000, size_weight=0.5):
        self.ctx = ctx
        if budget_bytes is None:
            free, _ = ctx.mem_info()
            budget_bytes = free // 2
        self.budget = int(budget_bytes)
        self.round_histories = int(round_histories)
        self.min_slice = int(min_slice)
        self.max_active = min(int(max_active), 1024)
        self.graphs = graphs
        self.phantoms = phantoms or (lambda j: scenes.dose_phantom(j["shape"], j["voxel_cm"], j["phantom_seed"]))
        self.on_result = on_result
        self.dose = dose
        self.admit_slice = float(admit_slice)
        self.admit_voxels_per_s = float(admit_voxels_per_s)
        self.first_slice = int(first_slice)
        self.size_weight = float(size_weight)
        self._buf = ctx.pinned_pool  # f64 read buffers by size, kept by the context
        self._free = ctx.pinned_pool  # free pinned float32 buffers by size, kept by the context
        self.rounds = 0
        self.kernel_ms = 0.0
        self.counters = {}
        self.host_s = dict(admit=0.0, read=0.0, wait=0.0, release=0.0, launch=0.0, roi=0.0)
        self.round_log = []  # per round: the launch's event counters and kernel time
        # per loop pass: host seconds in each phase, the voxels of the jobs it
        # admitted and the jobs it finished
        self.pass_log = []

    # ------------------------------------------------------------------ helpers
    def _pinned(self, n):
        free = self._free.get(n)
        if free:
            return free.pop()
        try:
            return self.ctx.pinned(n, np.float32)
        except Exception:
            return np.empty(n, np.float32)

    def _admit(self, job, res, now):
        ctx = self.ctx
        slot = ctx.job_create()
        scenes.setup_dose_job(ctx, job, ids=self.phantoms(job))
        res.admitted = now
        # what the job holds now, plus the FP32 dose it stages for its read
        # (held from its last round until the copy is done)
        staged = 8 * int(np.prod(job["shape"])) if self.dose == "f32" else 0
        return _Active(job, res, slot, ctx.job_device_bytes(slot) + staged, self.first_slice)

    def _read(self, a):
        """Copies a finished job's results back (overlaps a round in flight)."""
        ctx = self.ctx
        ctx.job_select(a.slot)
        d1, d2 = read_dose(ctx, a.job, self._buf, self.dose, a.done)
        image = ctx.tally(0)["image"]
        a.res.summary = self.on_result(a.job, d1, d2, image) if self.on_result else None
        a.res.histories = a.done

    def _check_precision(self, active):
        """The ROI error of every precision job that has run, after a round:
        one device reduction over all their ROI boxes.  A job above its
        target plans done x (error / target)^2 histories (10 % more, at
        least one more slice, at most its cap)."""
        prec = [a for a in active if a.roi is not None and a.done > 0 and not a.finished()]
        if not prec:
            return
        jobs, boxes = [], []
        for a in prec:
            jobs += [a.slot] * len(a.roi)
            boxes += a.roi
        st = self.ctx.roi_stats(jobs, boxes)
        k = 0
        for a in prec:
            rse = scenes.roi_rse(st[k:k + len(a.roi)])
            k += len(a.roi)
            a.rse = float(rse.max())
            a.res.roi_rse = a.rse
            t = a.job["rse_target"]
            if a.rse > t:
                est = 4 * a.done if not np.isfinite(a.rse) else int(a.done * (a.rse / t) ** 2 * 1.1)
                a.need = int(min(a.job["histories"], max(est, a.done + self.min_slice)))

    def _plan(self, active):
        """Histories of each admitted job in the next round: shares of
        round_histories in proportion to (phantom voxels)^size_weight (equal
        shares by default), the share of a job that has fewer left passed on
        to the others (a precision job's planned histories count as its
        total)."""
        left = np.array([a.need - a.done for a in active], dtype=np.int64)
        w = np.ones(len(active)) if self.size_weight == 0 else \
            np.array([float(np.prod(a.job["shape"])) ** self.size_weight for a in active])
        take = np.zeros_like(left)
        budget = max(self.round_histories, self.min_slice * len(active))
        open_ = left > 0
        while budget > 0 and open_.any():
            share = np.maximum(self.min_slice, (budget * w / w[open_].sum()).astype(np.int64))
            add = np.where(open_, np.minimum(left - take, share), 0)
            take += add
            budget -= int(add.sum())
            open_ = (left - take) > 0
            if share[open_].max(initial=0) <= self.min_slice and budget <= 0:
                break
        return take

    # ------------------------------------------------------------------ loop
    def run(self, jobs, arrivals=None):
        """Serves jobs[i] arriving at arrivals[i] seconds after the start
        (all at 0 by default); returns the JobResults in job order."""
        ctx = self.ctx
        ctx.set_graphs(self.graphs)
        n = len(jobs)
        arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, dtype=float)
        for j in jobs:
            if scenes.dose_job_bytes(j, self.dose) > self.budget:
                raise ValueError("a job's footprint exceeds the memory budget")
        order = deque(sorted(range(n), key=lambda i: arrivals[i]))
        results = [JobResult(i, float(arrivals[i])) for i in range(n)]
        waiting = deque()
        active = []
        resident = 0
        to_read = []  # finished in the last round, results not yet copied
        copying = []  # FP32 dose copies in flight (started beside a round)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        while order or waiting or active or to_read or copying:
            h0 = dict(self.host_s)
            now = clock()
            while order and arrivals[order[0]] <= now:
                waiting.append(order.popleft())
            launched = False
            t = clock()
            if active:
                take = self._plan(active)
                ctx.run_jobs_async([a.slot for a in active], take.tolist(), [a.done for a in active],
                                   [a.job["seed"] for a in active])
                launched = True
                for a, t_ in zip(active, take):
                    a.done += int(t_)
            t1 = clock()
            self.host_s["launch"] += t1 - t
            # beside the round in flight: the results of the jobs finished in
            # the round before (FP32: the copies are started and left running),
            # then the set-up of newly admitted jobs
            for a in to_read:
                if self.dose == "f32":
                    ctx.job_select(a.slot)
                    a.buf = self._pinned(2 * int(np.prod(a.job["shape"])))
                    ctx.read_staged_dose_f32_async(a.buf)
                    a.image = ctx.tally(0)["image"]  # beside the round, not between rounds
                    copying.append(a)
                else:
                    self._read(a)
                    a.res.finished = clock()
            t2 = clock()
            self.host_s["read"] += t2 - t1
            admitted = []
            queued = 0  # histories the admitted jobs bring
            while waiting and len(active) + len(admitted) < self.max_active:
                if not launched and queued >= self.round_histories:
                    break  # start the first round; admit the rest beside it
                i = waiting[0]
                est = float(np.prod(jobs[i]["shape"])) / self.admit_voxels_per_s if admitted else 0.0
                if (active or admitted) and clock() - t2 + est > self.admit_slice:
                    break  # leave the rest for the next rounds: the device waits for this loop
                need = scenes.dose_job_bytes(jobs[i], self.dose)
                if resident + need > self.budget and (active or to_read or admitted or copying):
                    break
                waiting.popleft()
                a = self._admit(jobs[i], results[i], clock())
                resident += a.bytes
                queued += a.need
                admitted.append(a)
            t3 = clock()
            self.host_s["admit"] += t3 - t2
            if launched:
                st = ctx.wait()
                self.host_s["wait"] += clock() - t3
                self.rounds += 1
                self.kernel_ms += st["kernel_ms"]
                self.round_log.append(st)
                for k, v in st.items():
                    if k != "kernel_ms":
                        self.counters[k] = self.counters.get(k, 0) + v
                t4 = clock()
                self._check_precision(active)
                self.host_s["roi"] += clock() - t4
            t = clock()
            # finished copies: results out, device memory freed
            block = not launched and not admitted
            still = []
            for a in copying:
                if ctx.read_done(a.slot, wait=block):
                    buf = a.buf
                    if self.on_result:
                        a.res.summary = self.on_result(a.job, buf[0::2], buf[1::2], a.image)
                    a.res.histories = a.done
                    a.res.finished = clock()
                    self._free.setdefault(buf.size, []).append(buf)
                    a.buf = None
                    ctx.job_release(a.slot)
                    resident -= a.bytes
                    block = False
                else:
                    still.append(a)
            copying = still
            if self.dose != "f32":
                for a in to_read:
                    ctx.job_release(a.slot)
                    resident -= a.bytes
            self.host_s["release"] += clock() - t
            idle = not launched and not to_read and not admitted and not copying
            to_read = [a for a in active if a.finished()]
            active = [a for a in active if not a.finished()] + admitted
            if self.dose == "f32":
                for a in to_read:  # conversion queued ahead of the next round; the copies run beside it
                    ctx.job_select(a.slot)
                    ctx.stage_voxel_dose_f32(0, a.done)
            self.pass_log.append(dict({k: self.host_s[k] - h0[k] for k in h0}, launched=launched,
                                      admitted_voxels=[int(np.prod(a.job["shape"])) for a in admitted],
                                      finished=len(to_read)))
            if idle and order:
                time.sleep(max(0.0, min(arrivals[order[0]] - clock(), 0.05)))
        self.makespan = clock()
        return results


def run_plain(ctx, jobs, arrivals=None, phantoms=None, on_result=None, dose="f32"):
    """First-in first-out service with the plain kernel: each job in turn is
    uploaded, run in one launch, read back and freed."""
    phantoms = phantoms or (lambda j: scenes.dose_phantom(j["shape"], j["voxel_cm"], j["phantom_seed"]))
    n = len(jobs)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, dtype=float)
    results = [JobResult(i, float(arrivals[i])) for i in range(n)]
    bufs = ctx.pinned_pool  # read buffers by size, kept by the context
    counters = {}
    kernel_ms = 0.0
    host = dict(admit=0.0, run=0.0, read=0.0, release=0.0)
    t0 = time.perf_counter()
    for i in sorted(range(n), key=lambda k: arrivals[k]):
        wait = arrivals[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        j = jobs[i]
        ta = time.perf_counter()
        slot = ctx.job_create()
        results[i].admitted = ta - t0
        scenes.setup_dose_job(ctx, j, ids=phantoms(j))
        tb = time.perf_counter()
        st = ctx.run(j["histories"], seed=j["seed"], first=0, kernel=KERNEL_PLAIN)
        if j.get("rse_target"):  # the same ROI error the served loop stops on
            boxes = scenes.dose_job_roi(j)
            results[i].roi_rse = float(scenes.roi_rse(ctx.roi_stats([slot] * len(boxes), boxes)).max())
        tc = time.perf_counter()
        kernel_ms += st["kernel_ms"]
        for k, v in st.items():
            if k != "kernel_ms":
                counters[k] = counters.get(k, 0) + v
        d1, d2 = read_dose(ctx, j, bufs, dose)
        image = ctx.tally(0)["image"]
        results[i].summary = on_result(j, d1, d2, image) if on_result else None
        results[i].histories = j["histories"]
        td = time.perf_counter()
        results[i].finished = td - t0
        ctx.job_release(slot)
        te = time.perf_counter()
        host["admit"] += tb - ta
        host["run"] += tc - tb
        host["read"] += td - tc
        host["release"] += te - td
    makespan = time.perf_counter() - t0
    return results, dict(makespan=makespan, kernel_ms=kernel_ms, counters=counters, host_s=host)


def warm_read_buffers(ctx, jobs, dose="f32"):
    """Pins run_plain's read buffers for the jobs' phantom sizes ahead of
    time, as a service that has been running holds them."""
    dt = np.float32 if dose == "f32" else np.float64
    for nv in sorted(set(int(np.prod(j["shape"])) for j in jobs)):
        key = ("read", nv, dose)
        if key not in ctx.pinned_pool:
            try:
                ctx.pinned_pool[key] = ctx.pinned(2 * nv, dt)
            except Exception:
                ctx.pinned_pool[key] = np.empty(2 * nv, dt)


def poisson_arrivals(n, rate, seed):
    """Arrival times (s) of n jobs of a Poisson process of the given rate (1/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n)) - 0.0


def latency_percentiles(results):
    lat = np.array([r.latency for r in results])
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
