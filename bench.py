"""Benchmark of the photon-transport hot path on one MI355X.

Headline workload (the served job stream): 32 dose jobs drawn with seed 5
(cubic phantoms of 64^3 to 512^3 voxels at 0.5-2 mm, monoenergetic 20-120
keV or Kramers 60-120 kVp sources, FP64 voxel dose plus a 128 x 128 image
each), each asking for 1 % relative standard error on its region of
interest: four 1 cm^3 dose voxels on the beam axis, 1-2 cm below the
entrance face (`scenes.dose_job_roi`; the error of a voxel's summed deposits
as the reference computes it, quantity.cpp:37-45).  The continuous-batching
loop runs each job until every ROI voxel is at or below 1 % (checked on the
device after every round); the plain kernel then serves the same jobs with
the same history counts first-in first-out, one launch per job, so both
sides reach the same sigma.  A step is one pass of the whole stream:
--warmup untimed passes, then --steps timed ones; `value` is the served
histories/s over the timed passes with every job of a pass queued at its
start; then p50/p99 latency under Poisson arrivals at 70 % of each
service's own capacity.  Phantoms are built before the clock starts
(they are the clients' data, in pinned host memory); uploads, tally zeroing,
transport, the ROI checks and the read-back of every job's dose and image
are timed.  The round kernel is the event kernel the autotuner picks for the
stream (k_transport_bank_round, photon banks per wave, or
k_transport_queue_round, a photon per lane), refined by a coordinate
search over bank size, blocks per CU and interaction batch on untimed
passes of the stream itself; the
roofline is that kernel's over the timed rounds: algorithmic bytes from its
int64 event counters over its HIP-event time, and its voxel-dose atomics
per second against what the library's scattered-deposit probe sustains on a
512^3 grid of dose records (`atomic_probe`).

Beside it: the radiography workload (256 x 256 x 160 phantom at 1.25 mm,
15-80 keV Kramers cone beam, 256 x 256 image, FP64 voxel dose; 10 steps of
1e7 histories) on the tuned event kernel, the plain kernel, and the event
kernel with background free flights and local majorants off (what batching
alone gives); the CT scan (360 views of a 512 x 512 x 256 phantom in one launch;
plain on every tenth view); the precision study (the slab at 1e9 histories
per accumulation mode); and the FP64 CPU restatement of the reference's
OpenMP path timed on a bounded sample of the served stream's jobs on the
host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RSE_TARGET = 0.01


def algorithmic_bytes(st):
    """Bytes one launch must move at minimum: per Woodcock step the voxel ID
    (1) and the sigma_max and sigma_tot brackets (2 x 16); per real event the
    two scattering cross-section brackets (32); per Rayleigh two RITA rows
    (48); per Kahn S-test two S(x) brackets (32); per deposit an FP64 sum and
    sum^2 read-modify-write (32); per image hit 16; per history 8."""
    return (33 * st["steps"] + 32 * st["real"] + 48 * st["rayleigh"] + 32 * st["kahn_s_tests"]
            + 32 * st["deposits"] + 16 * st["surface_hits"] + 8 * st["histories"])


def progress(msg):
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def profile_value(name, key):
    """A figure recorded in a committed rocprofv3 summary (None if absent)."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.strip().startswith(key + " ="):
                return float(line.split("=")[1])
    except (OSError, ValueError):
        pass
    return None


def copy_probe(ctx, mb):
    """The library's streaming copy of mb MiB (read mb MiB, write mb MiB),
    twice: the reference for scaling the HBM counters of a profiled run, and
    the achievable copy bandwidth."""
    nbytes = mb * 2 ** 20
    ms = ctx.copy_probe(nbytes)
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB_per_s": 2 * nbytes / ms / 1e6}


def atomic_probe(ctx, deposits=200_000_000):
    """The library's scattered-deposit probe (three device-scope atomics per
    deposit at hashed 32-byte voxel records) on grids of the served stream's
    smallest and largest dose arrays: the atomic rate the chip sustains for
    the dose update pattern, against which the transport kernels' own
    deposit rate is put."""
    out = []
    for n in (128, 512):
        records = n ** 3
        ms = ctx.atomic_probe(records, deposits)
        out.append({"grid": "%d^3" % n, "record_bytes": 32 * records, "deposits": deposits, "ms": ms,
                    "atomics_per_s": 3.0 * deposits / (ms / 1e3)})
    return out


def roi_dose(job, mean, var=None, rel=None, n=None):
    """(dose per history, its standard error) of each of the job's four ROI
    dose voxels (scenes.dose_job_roi) from per-voxel means per history (x
    fastest) and either the variances of those means or their relative
    standard errors, the voxels taken as independent; with n, mean and var
    hold the voxels' sums and sums of squares over n histories instead.
    Only the ROI's voxels are read (the arrays may be a whole result)."""
    from midsx_amd import scenes
    nx, ny, nz = job["shape"]
    m3 = np.asarray(mean).reshape(nz, ny, nx)
    e3 = np.asarray(var if var is not None else rel).reshape(nz, ny, nx)
    out = []
    for i0, i1, j0, j1, k0, k1 in scenes.dose_job_roi(job):
        m = m3[k0:k1, j0:j1, i0:i1].astype(np.float64)
        e = e3[k0:k1, j0:j1, i0:i1].astype(np.float64)
        if n is not None:  # sums over n histories: mean and variance of the mean
            m = m / n
            e = np.maximum(e / n - m * m, 0.0) / (n - 1)
        v = e if var is not None else (m * e) ** 2
        out.append((float(m.sum()), float(np.sqrt(v.sum()))))
    return out


def cpu_baseline(jobs, phantoms, seconds, affinity, roi_gpu=None):
    """The FP64 C++/OpenMP restatement of the reference's CPU path on a
    bounded sample of the served stream: one job of each phantom size, each
    run for the histories that take about seconds / 4 on the host's cores
    (the OpenMP team of the environment).  With roi_gpu (job index -> the
    served result's ROI doses), each sampled job's ROI doses from that CPU
    run are compared with the served ones: z = difference over the combined
    standard error, per ROI dose voxel."""
    import oracle
    from midsx_amd import scenes
    threads = oracle.max_threads()
    sim = oracle.Simulation(scenes.DOSE_MATERIALS)
    picks, seen = [], set()
    for j in jobs:
        if j["shape"][0] not in seen:
            seen.add(j["shape"][0])
            picks.append(j)
    total_n, total_t, parts, agreement = 0, 0.0, [], []
    for j in picks:
        progress("CPU baseline: %d^3 job" % j["shape"][0])
        scenes.setup_dose_job(sim, j, ids=phantoms(j))
        n = 2000
        sim.reset()
        dt = sim.run(n, seed=99, threads=0)
        while dt < 0.5 and n < 10 ** 8:
            n *= 4
            sim.reset()
            dt = sim.run(n, seed=99, threads=0)
        n = int(max(n, n * seconds / len(picks) / max(dt, 1e-9)))
        sim.reset()
        dt = sim.run(n, seed=100, threads=0)
        total_n += n
        total_t += dt
        parts.append("%d^3: %d histories in %.1f s" % (j["shape"][0], n, dt))
        idx = next(k for k, jj in enumerate(jobs) if jj is j)
        if roi_gpu and idx in roi_gpu:
            ssum, ssq, _ = sim.voxel_dose(0)
            cpu = roi_dose(j, ssum, var=ssq, n=n)
            z = [(g[0] - c[0]) / np.hypot(g[1], c[1]) for g, c in zip(roi_gpu[idx]["roi"], cpu)]
            agreement.append(dict(job=idx, shape=j["shape"][0], cpu_histories=n,
                                  served_histories=roi_gpu[idx].get("histories"),
                                  z=[round(float(x), 3) for x in z], max_abs_z=float(np.max(np.abs(z)))))
    zs = np.array([z for a in agreement for z in a["z"]])
    chi2 = float(np.sum(zs ** 2)) if zs.size else None
    try:
        from scipy import stats
        p_chi2 = float(stats.chi2.sf(chi2, zs.size)) if zs.size else None
    except ImportError:
        p_chi2 = None
    return dict(value=total_n / total_t, unit="histories/s", cores=threads, kind="port",
                sample="FP64 C++/OpenMP restatement of runSimulation on one dose job of each phantom size of the "
                       "served stream (" + "; ".join(parts) + ")",
                omp_max_threads=threads, nproc=os.cpu_count(), affinity_cpus=affinity,
                omp_proc_bind=os.environ.get("OMP_PROC_BIND"), omp_places=os.environ.get("OMP_PLACES"),
                dose_agreement=dict(jobs=agreement, chi2=chi2, dof=int(zs.size), p=p_chi2))


def timed_steps(ctx, kernel, n, steps, warmup, first):
    for _ in range(warmup):
        ctx.run(n, seed=2, first=first, kernel=kernel)
        first += n
    ctx.reset_tallies()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.run(n, seed=2, first=first, kernel=kernel))
        first += n
    wall = time.perf_counter() - t0
    return stats, wall, first


def roofline(stats_list, traffic=None, source=None):
    kern_ms = np.array([s["kernel_ms"] for s in stats_list])
    byt = np.array([algorithmic_bytes(s) for s in stats_list], dtype=np.float64)
    achieved = byt.sum() / (kern_ms.sum() / 1e3) / 1e9
    dep = float(sum(s["deposits"] for s in stats_list))
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_ms_avg": float(kern_ms.mean()),
           "algorithmic_bytes_per_launch": float(byt.mean()),
           # the voxel-dose atomics (three per deposit) per second of kernel time
           "dose_atomics_per_s": 3.0 * dep / (kern_ms.sum() / 1e3)}
    if source:
        out["traffic_source"] = source
    return out


def event_kernel_name(ctx):
    import midsx_amd as M
    kind, bank = ctx.event_kernel()
    return "bank%d" % bank if kind == M.EVENT_BANK else "queue"


def event_kernel_symbol(ctx):
    """The kernel MDX_KERNEL_QUEUE launches run (without the _round suffix
    of the graph form)."""
    import midsx_amd as M
    return "k_transport_bank" if ctx.event_kernel()[0] == M.EVENT_BANK else "k_transport_queue"


def set_event_kernel(ctx, name):
    """name: queue, bank96, bank128 or bank160."""
    import midsx_amd as M
    if name == "queue":
        ctx.set_event_kernel(M.EVENT_QUEUE)
    else:
        ctx.set_event_kernel(M.EVENT_BANK, int(name[4:]))


def stream_tune(ctx, jobs, ph, tuned):
    """Start-up tuning on the stream itself: the autotuner probes one job,
    while a round mixes every admitted job, so from the autotuner's choice
    the coordinate search of launch_tune runs each candidate for one pass of
    the stream (seconds per history) and keeps the fastest."""
    from midsx_amd.serve import Server
    Server(ctx, phantoms=ph).run(jobs)  # warm: memory pool and pinned buffers

    def one_pass():
        t = time.perf_counter()
        res = Server(ctx, phantoms=ph).run(jobs)
        return (time.perf_counter() - t) / sum(x.histories for x in res)

    out = launch_tune(ctx, one_pass, tuned)
    out["stream_candidates"] = out.pop("launch_candidates")
    return out


def served_stream(ctx, n_jobs, seed, load, cap, steps, warmup, plain=True, params=None):
    """The served job stream at 1 % ROI error: the continuous-batching loop
    against plain FIFO on the same history counts.  A step is one pass of
    the whole stream (every job served to its ROI target): warmup untimed
    passes, then steps timed ones."""
    from midsx_amd import scenes
    from midsx_amd.serve import Server, run_plain, poisson_arrivals, latency_percentiles, warm_read_buffers
    jobs = [dict(j, histories=cap, rse_target=RSE_TARGET) for j in scenes.draw_dose_jobs(n_jobs, seed)]
    cache = {}
    for j in jobs:  # the clients' phantoms, in pinned host memory (x fastest)
        key = (j["shape"], j["voxel_cm"], j["phantom_seed"])
        if key not in cache:
            ids = scenes.dose_phantom(j["shape"], j["voxel_cm"], j["phantom_seed"])
            try:
                buf = ctx.pinned(ids.size, np.uint8).reshape(ids.shape, order="F")
                buf[...] = ids
                ids = buf
            except Exception:
                pass
            cache[key] = ids
    ph = lambda j: cache[(j["shape"], j["voxel_cm"], j["phantom_seed"])]
    # the stream's own launch parameters, tuned before timing on one job of
    # it (a 256^3 phantom if there is one), as a server does at start-up
    keep_job = ctx.job
    probe = next((j for j in jobs if j["shape"][0] == 256), jobs[0])
    pj = ctx.job_create()
    scenes.setup_dose_job(ctx, probe, ids=ph(probe))
    if params:  # fixed: event kernel, blocks per CU, refill, interact
        ev, b, r, i = params.split(",")
        set_event_kernel(ctx, ev)
        ctx.set_queue_params(int(b), int(r), int(i))
        tuned = dict(blocks_per_cu=int(b), refill_min=int(r), interact_min=int(i), event_kernel=ev, fixed=True)
    else:
        # the rounds' multi-job kernels, probed as launches of ~1.5e7
        # histories (a round runs 4.5e7)
        tuned = ctx.autotune(probe_histories=15_000_000, rounds=True)
    ctx.job_release(pj)
    ctx.job_select(keep_job)
    if not params:
        tuned = stream_tune(ctx, jobs, ph, tuned)
    # untimed passes first: the device memory pool and the pinned result
    # buffers are then warm, as in a service that has been running
    progress("served stream: tuned %s; %d warm-up passes" % (tuned, warmup))
    # every pass's rounds and histories in launch order (warm-up, timed,
    # latency pass), so that profiles can attribute their dispatches
    pass_rounds = []
    for _ in range(warmup):
        srv = Server(ctx, phantoms=ph)
        res = srv.run(jobs)
        pass_rounds.append(dict(kind="warmup", rounds=srv.rounds, histories=sum(r.histories for r in res)))
    passes = []  # (server, results, seconds) of each timed pass
    for i in range(steps):
        srv = Server(ctx, phantoms=ph)
        t = time.perf_counter()
        res = srv.run(jobs)
        passes.append((srv, res, time.perf_counter() - t))
        progress("served stream: timed pass %d of %d, %.3f s for %d histories" % (
            i + 1, steps, passes[-1][2], sum(r.histories for r in res)))
    served_s = sum(p[2] for p in passes)
    H = sum(r.histories for p in passes for r in p[1])
    pass_s = served_s / steps
    hist = [r.histories for r in passes[-1][1]]  # the last pass's per-job histories (what plain runs)
    served_rse = [r.roi_rse for p in passes for r in p[1]]
    capped = sum(1 for p in passes for r in p[1] if r.histories >= cap)
    rounds = [st for p in passes for st in p[0].round_log]
    counters = {}
    for p in passes:
        for k, v in p[0].counters.items():
            counters[k] = counters.get(k, 0) + v
    host_s = {}
    for p in passes:
        for k, v in p[0].host_s.items():
            host_s[k] = host_s.get(k, 0.0) + v / steps
    progress("served stream: %d passes, %.3f s for %d histories; latency pass" % (steps, served_s, H))
    # latency under Poisson arrivals at `load` of each service's own capacity
    arr = poisson_arrivals(n_jobs, load * n_jobs / pass_s, seed + 1)
    for srv, res, _ in passes:
        pass_rounds.append(dict(kind="timed", rounds=srv.rounds, histories=sum(r.histories for r in res)))
    roi_gpu = {}  # the latency pass's ROI doses per job, for the comparison with the CPU path

    def keep_roi(job, a, b, image):
        k = next(i for i, jj in enumerate(jobs) if jj is job)
        roi_gpu[k] = {"roi": roi_dose(job, a, rel=b)}

    srv = Server(ctx, phantoms=ph, on_result=keep_roi)
    res = srv.run(jobs, arr)
    for r in res:
        if r.index in roi_gpu:
            roi_gpu[r.index]["histories"] = r.histories
    pass_rounds.append(dict(kind="latency", rounds=srv.rounds, histories=sum(r.histories for r in res)))
    p50, p99 = latency_percentiles(res)
    plain_s = q50 = q99 = None
    plain_rse = [float("nan")]
    pinfo = {"kernel_ms": float("nan"), "counters": {"steps": float("nan")}}
    if plain:
        # plain: the same jobs with the histories each needed when served
        pjobs = [dict(j, histories=h) for j, h in zip(jobs, hist)]
        warm_read_buffers(ctx, pjobs)  # the plain service's result buffers, pinned ahead as well
        progress("served stream: plain FIFO")
        pres, pinfo = run_plain(ctx, pjobs, phantoms=ph)
        plain_s = pinfo["makespan"]
        progress("served stream: plain %.3f s; plain latency pass" % plain_s)
        plain_rse = [r.roi_rse for r in pres]
        parr = poisson_arrivals(n_jobs, load * n_jobs / plain_s, seed + 2)
        plat, _ = run_plain(ctx, pjobs, parr, phantoms=ph)
        q50, q99 = latency_percentiles(plat)
    kernel = event_kernel_symbol(ctx) + "_round"
    prof = "%s_served_dose_stream.md" % kernel
    traffic_ph = profile_value(prof, "traffic_bytes_per_history")
    traffic = traffic_ph * float(np.mean([r["histories"] for r in rounds])) if traffic_ph else None
    Hp = sum(hist)
    return jobs, ph, {
        "_roi_gpu": roi_gpu,
        "workload": "%d dose jobs (seed %d), phantoms %s voxels at 0.5-2 mm, each run to <= %g relative "
                    "standard error on its 4 ROI dose voxels (1 cm^3, on the beam axis 1-2 cm deep); "
                    "histories per job min %.3g, median %.3g, max %.3g, %d per pass; a step is one pass of the "
                    "whole stream" % (
                        n_jobs, seed, sorted(set(j["shape"][0] for j in jobs)), RSE_TARGET, min(hist),
                        float(np.median(hist)), max(hist), Hp),
        "value": H / served_s,
        "plain_value": Hp / plain_s if plain_s else None,
        "speedup_vs_plain": (H / served_s) / (Hp / plain_s) if plain_s else None,
        "jobs_at_cap": capped,
        "passes": steps,
        "warmup_passes": warmup,
        "served_s": served_s,
        "pass_s": [round(p[2], 4) for p in passes],
        "plain_s": plain_s,
        "kernel_s": sum(p[0].kernel_ms for p in passes) / 1e3,
        "plain_kernel_s": pinfo["kernel_ms"] / 1e3,
        "rounds": len(rounds),
        "pass_rounds": pass_rounds,
        "host_s_per_pass": {k: round(v, 4) for k, v in host_s.items()},
        "steps_per_history": counters["steps"] / H,
        "events_per_history": {k: counters[k] / H for k in ("steps", "real", "rayleigh", "compton", "photo",
                                                             "kahn_trials", "deposits", "surface_hits")},
        "plain_steps_per_history": pinfo["counters"]["steps"] / Hp,
        "roi_rse": {"served_max": float(np.max(served_rse)), "served_median": float(np.median(served_rse)),
                    "plain_max": float(np.max(plain_rse)), "plain_median": float(np.median(plain_rse))},
        "latency": {"load": load, "served_arrival_rate_jobs_per_s": load * n_jobs / pass_s, "p50_s": p50,
                    "p99_s": p99, "plain_arrival_rate_jobs_per_s": load * n_jobs / plain_s if plain_s else None,
                    "plain_fifo_p50_s": q50, "plain_fifo_p99_s": q99},
        "roofline": dict(kernel=kernel, **roofline(
            rounds, traffic, "rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 per history of the served stream "
                             "(factors from the k_copy16 streaming copy), profiles/" + prof if traffic else None)),
        "event_kernel": event_kernel_name(ctx),
        "queue_params": tuned,
    }


def radiography(ctx, args):
    """Config 2 on the tuned event kernel (autotuned), the plain kernel, and
    the event kernel with free flights and local majorants off."""
    import midsx_amd as M
    from midsx_amd import scenes
    kw = dict(shape=(256, 256, 160), voxel_cm=0.125, seed=2, image=(256, 256))
    job = ctx.job_create()
    scenes.setup_radiography(ctx, **kw)
    n = args.histories_per_step
    tuned = None
    if args.queue_params:
        set_event_kernel(ctx, args.event_kernel)
        ctx.set_queue_params(*[int(v) for v in args.queue_params.split(",")])
    elif not args.no_autotune:
        tuned = ctx.autotune(probe_histories=4000000)  # launches of ~3 ms: stable timings
        # refined on launches of the timed size (history indices above the
        # timed ones; the faster of two launches per candidate: the
        # candidates differ by ~1 %)
        tuned = launch_tune(ctx, lambda: min(ctx.run(n, seed=2, first=(1 << 41) + k * n, kernel=M.KERNEL_QUEUE)["kernel_ms"]
                                             for k in range(2)), tuned)
    progress("radiography: queue kernel")
    steps, warm = args.radiography_steps, args.radiography_warmup
    stats, wall, first = timed_steps(ctx, M.KERNEL_QUEUE, n, steps, warm, 0)
    total = n * steps
    out = {
        "workload": "radiography: 256x256x160 uint8 phantom at 1.25 mm, 15-80 keV Kramers cone beam, "
                    "256x256 image + FP64 voxel dose, %d steps of %d histories" % (steps, n),
        "value": total / wall, "unit": "histories/s", "ms_per_step": wall / steps * 1e3,
        "queue_params": dict(zip(("blocks_per_cu", "refill_min", "interact_min"), ctx.queue_params())),
        "event_kernel": event_kernel_name(ctx),
    }
    agg = {k: sum(s[k] for s in stats) for k in M.STAT_NAMES}
    out["events_per_history"] = {k: agg[k] / total for k in ("steps", "real", "rayleigh", "compton", "photo",
                                                              "kahn_trials", "deposits", "surface_hits")}
    kernel = event_kernel_symbol(ctx)
    prof = "%s_radiography_256x256x160.md" % kernel
    tph = profile_value(prof, "traffic_bytes_per_history")
    out["roofline"] = dict(kernel=kernel, **roofline(
        stats, tph * n if tph else None, "rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 per launch (factors "
                                         "from the k_copy16 streaming copy), profiles/" + prof if tph else None))
    if tuned:
        out["autotune_probe_ms"] = tuned["probe_ms"]
    if not args.no_plain:
        progress("radiography: plain kernel")
        pstats, pwall, first = timed_steps(ctx, M.KERNEL_PLAIN, n, steps, warm, first)
        out["plain"] = {"value": total / pwall, "ms_per_step": pwall / steps * 1e3,
                        "steps_per_history": sum(s["steps"] for s in pstats) / total}
        out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
        # what batching and regrouping alone give: majorant steps everywhere
        progress("radiography: ablation")
        ctx.set_flights(False)
        ctx.set_local_majorants(False)
        astats, awall, first = timed_steps(ctx, M.KERNEL_QUEUE, n, steps, warm, first)
        ctx.set_flights(True)
        ctx.set_local_majorants(True)
        out["ablation_majorant_steps"] = {
            "what": "the tuned event kernel with background free flights and local majorants off (the plain "
                    "kernel's Woodcock steps, event-regrouped and batched), same launch parameters",
            "value": total / awall, "ms_per_step": awall / steps * 1e3,
            "steps_per_history": sum(s["steps"] for s in astats) / total,
            "speedup_vs_plain": (total / awall) / out["plain"]["value"]}
    ctx.job_release(job)
    return out


def launch_tune(ctx, run, tuned):
    """From the autotuner's choice, a coordinate search over the bank size,
    the blocks per CU (2-4) and the interaction batch (32, 48, 64) on
    run() (a reduced launch of the workload itself; returns its kernel
    milliseconds), sweeping until a sweep changes nothing (at most three
    sweeps); leaves the fastest set."""
    import midsx_amd as M
    kind, bank = ctx.event_kernel()
    b, r, i = ctx.queue_params()
    timed = {}

    def cost(c):
        if c not in timed:
            k, bs, bb, rr, ii = c
            ctx.set_event_kernel(k, bs)
            ctx.set_queue_params(bb, rr, ii)
            timed[c] = run()
        return timed[c]

    best = (kind, bank, b, r, i)
    axes = [(1, M.BANK_SIZES if kind == M.EVENT_BANK else (bank,)), (2, (2, 3, 4)), (4, (32, 48, 64))]
    for _ in range(3):
        start = best
        for pos, values in axes:
            for v in values:
                c = best[:pos] + (v,) + best[pos + 1:]
                if cost(c) < cost(best):
                    best = c
        if best == start:
            break
    k, bs, bb, rr, ii = best
    ctx.set_event_kernel(k, bs)
    ctx.set_queue_params(bb, rr, ii)
    return dict(tuned, blocks_per_cu=bb, refill_min=rr, interact_min=ii,
                event_kernel="bank" if k == M.EVENT_BANK else "queue", bank_size=bs,
                autotuner=dict(event_kernel=tuned["event_kernel"], bank_size=tuned.get("bank_size"),
                               blocks_per_cu=tuned["blocks_per_cu"], interact_min=tuned["interact_min"]),
                launch_candidates=len(timed))


def ct_scan(ctx, n_views, scale, plain_every, seed=4, params=None):
    """CT: 512 x 512 x 256 phantom at 0.8 mm (water, lung, adipose, bone),
    n_views views about its z axis, 120 kVp, a 512 x 64 detector per view,
    per-view histories U[2e6, 2e7] x scale.  All views in one multi-job launch
    of the queue kernel (views alias the phantom job; the phantom's voxel dose
    sums over views) versus the plain kernel one view per launch, on every
    plain_every-th view (histories/s of that subsample)."""
    import midsx_amd as M
    from midsx_amd import scenes
    owner = ctx.job_create()
    geo, views, angles = scenes.setup_ct(ctx, n_views=n_views, seed=seed)
    n = [max(1, int(h * scale)) for h in scenes.ct_histories(n_views, seed)]
    seeds = [10_000 + v for v in range(n_views)]
    H = sum(n)
    ctx.job_select(views[0])
    if params:  # fixed: event kernel, blocks per CU, refill, interact
        ev, b, r, i = params.split(",")
        set_event_kernel(ctx, ev)
        ctx.set_queue_params(int(b), int(r), int(i))
        tuned = dict(blocks_per_cu=int(b), refill_min=int(r), interact_min=int(i), event_kernel=ev, fixed=True)
    else:
        # one view, probed as the multi-job launch the views run in, then
        # refined on every view at a tenth of its histories (the probes of
        # one view do not separate banks of 128 and 160 reliably)
        tuned = ctx.autotune(probe_histories=4000000, rounds=True)
        small = [max(1, h // 10) for h in n]
        tuned = launch_tune(ctx, lambda: ctx.run_jobs(views, small, [1 << 41] * n_views, seeds)["kernel_ms"], tuned)
    ctx.set_graphs(True)
    ctx.run_jobs(views[:8], [100000] * 8, [1 << 40] * 8, seeds[:8])  # warm (capture)
    for j in [owner] + views:
        ctx.job_select(j)
        ctx.reset_tallies()
    t = time.perf_counter()
    progress("CT: one launch of %d views" % n_views)
    st = ctx.run_jobs(views, n, [0] * n_views, seeds)
    q_s = time.perf_counter() - t
    progress("CT: %.3f s; plain on every %d-th view" % (q_s, plain_every))
    ctx.set_graphs(False)
    for j in [owner] + views:
        ctx.job_select(j)
        ctx.reset_tallies()
    sub = list(range(0, n_views, plain_every))
    t = time.perf_counter()
    pk, psteps, pH = 0.0, 0, 0
    for v in sub:
        ctx.job_select(views[v])
        ps = ctx.run(n[v], seed=seeds[v], first=0, kernel=M.KERNEL_PLAIN)
        pk += ps["kernel_ms"]
        psteps += ps["steps"]
        pH += n[v]
    p_s = time.perf_counter() - t
    for j in views + [owner]:
        ctx.job_release(j)
    return {
        "workload": "CT: 512x512x256 phantom at 0.8 mm, %d views, 120 kVp, 512x64 detector per view, "
                    "histories per view U[2e6,2e7] x %g (total %d)" % (n_views, scale, H),
        "value": H / q_s,
        "unit": "histories/s",
        "plain_value": pH / p_s,
        "plain_sample": "every %d-th view (%d views, %d histories)" % (plain_every, len(sub), pH),
        "speedup_vs_plain": (H / q_s) / (pH / p_s),
        "queue_s": q_s,
        "queue_kernel_ms": st["kernel_ms"],
        "plain_s": p_s,
        "plain_kernel_ms": pk,
        "launches": {"queue": 1, "plain": len(sub)},
        "steps_per_history": st["steps"] / H,
        "plain_steps_per_history": psteps / pH,
        "queue_params": tuned,
    }


def precision_study(ctx, n_hist, step=10 ** 8):
    """Config 3: the slab phantom (40 x 40 x 10 water voxels, 60 keV pencil
    beam), n_hist histories on the queue kernel with the voxel dose
    accumulated in FP64, FP32, compensated FP32 and compensated FP32 summed
    per block in LDS first, on identical RNG streams; per-voxel relative
    error of each against the FP64 run."""
    import midsx_amd as M
    from midsx_amd import scenes
    job = ctx.job_create()
    scenes.setup_slab(ctx)
    keep, keep_ev = ctx.queue_params(), ctx.event_kernel()
    # the slab's own launch parameters on the lane-resident kernel, which
    # every mode (the per-block LDS one included) runs: identical RNG streams
    tuned = ctx.autotune(probe_histories=4000000, events=False)
    out = {}
    res = {}
    for name, mode in (("fp64", M.DOSE_F64), ("fp32", M.DOSE_F32), ("fp32_compensated", M.DOSE_F32C),
                       ("fp32_compensated_lds", M.DOSE_F32C_LDS)):
        progress("precision study: %s" % name)
        ctx.set_dose_precision(mode)
        ctx.reset_tallies()
        ms = 0.0
        done = 0
        while done < n_hist:
            k = min(step, n_hist - done)
            ms += ctx.run(k, seed=33, first=done, kernel=M.KERNEL_QUEUE)["kernel_ms"]
            done += k
        s, s2, cnt = ctx.voxel_dose(0)
        out[name] = (s, s2, cnt)
        res[name] = {"kernel_s": ms / 1e3, "histories_per_s": n_hist / (ms / 1e3)}
    s64, q64, c64 = out["fp64"]
    roi = c64 >= 1000
    for name in ("fp32", "fp32_compensated", "fp32_compensated_lds"):
        s, s2, cnt = out[name]
        e = np.abs(s[roi] / s64[roi] - 1)
        e2 = np.abs(s2[roi] / q64[roi] - 1)
        res[name].update({"rel_err_sum_max": float(e.max()), "rel_err_sum_median": float(np.median(e)),
                          "rel_err_sum2_max": float(e2.max()), "same_deposits": bool(np.array_equal(cnt, c64))})
    ctx.job_release(job)
    ctx.set_queue_params(*keep)
    ctx.set_event_kernel(*keep_ev)
    return {"workload": "slab 40x40x10 water voxels, 60 keV beam, %d histories per mode, identical RNG streams; "
                        "errors over the %d voxels with >= 1000 deposits" % (n_hist, int(roi.sum())),
            "queue_params": {k: tuned[k] for k in ("blocks_per_cu", "refill_min", "interact_min")},
            "modes": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed passes of the served stream (a step: every job served to its ROI target)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed passes of the served stream before them")
    ap.add_argument("--radiography-steps", type=int, default=10, help="radiography steps of --histories-per-step")
    ap.add_argument("--radiography-warmup", type=int, default=2)
    ap.add_argument("--histories-per-step", type=int, default=10 ** 7)
    ap.add_argument("--served-jobs", type=int, default=32)
    ap.add_argument("--served-seed", type=int, default=5, help="seed of the served job stream (5: the stream the "
                    "served-stream configuration names)")
    ap.add_argument("--served-load", type=float, default=0.7)
    ap.add_argument("--served-cap", type=int, default=2 * 10 ** 9, help="histories cap of a served job")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-radiography", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--queue-params", default=None,
                    help="blocks_per_cu,refill_min,interact_min for the radiography (skips its autotuner)")
    ap.add_argument("--event-kernel", default="queue", choices=("queue", "bank96", "bank128", "bank160"),
                    help="event kernel of the radiography with --queue-params")
    ap.add_argument("--served-params", default=None,
                    help="event_kernel,blocks_per_cu,refill_min,interact_min of the served stream, e.g. "
                         "bank128,3,64,48 (skips its autotuner)")
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-ct", action="store_true")
    ap.add_argument("--ct-views", type=int, default=360)
    ap.add_argument("--ct-scale", type=float, default=1.0,
                    help="fraction of the CT workload's per-view histories (1.0: the full ~4e9)")
    ap.add_argument("--ct-plain-every", type=int, default=10)
    ap.add_argument("--ct-params", default=None,
                    help="event_kernel,blocks_per_cu,refill_min,interact_min of the CT launch (skips its autotuner)")
    ap.add_argument("--copy-probe-mb", type=int, default=0)
    ap.add_argument("--no-atomic-probe", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-histories", type=int, default=10 ** 9)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: sharding histories over several GPUs is not built in this "
              "tree; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    affinity = len(os.sched_getaffinity(0))  # before OpenMP binds this thread
    # the CPU baseline's OpenMP team: threads on cores, kept close
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")

    import midsx_amd as M
    from midsx_amd import scenes

    ctx = M.Context(0)
    ctx.build_tables(sorted(set(scenes.C2_MATERIALS) | set(scenes.DOSE_MATERIALS)))
    out = {
        "metric": "photon histories/sec vs plain GPU at matched tally σ; p50/p99 request latency",
        "value": None, "unit": "histories/s", "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded ellipsoid phantoms; EPDL2017 tables for air, water, lung, adipose, cortical bone)",
    }
    jobs = ph = roi_gpu = None
    if not args.no_served:
        if args.steps < 1 or args.warmup < 0:
            ap.error("--steps must be at least 1 and --warmup at least 0")
        jobs, ph, sv = served_stream(ctx, args.served_jobs, args.served_seed, args.served_load, args.served_cap,
                                     args.steps, args.warmup, plain=not args.no_plain, params=args.served_params)
        out.update(value=sv["value"], steps=args.steps, warmup=args.warmup,
                   ms_per_step=sv["served_s"] / args.steps * 1e3,
                   speedup_vs_plain=sv["speedup_vs_plain"], latency=sv["latency"], roofline=sv.pop("roofline"))
        out["config"] = {"workload": "served dose-job stream: " + sv["workload"],
                         "kernel": "served = continuous batching, " + out["roofline"]["kernel"] + " per round (hipGraph); "
                                   "plain = k_transport_plain, one launch per job, FIFO; same histories per job, "
                                   "hence matched sigma",
                         "step": "one pass of the served stream: all %d jobs served to their ROI target"
                                 % args.served_jobs}
        roi_gpu = sv.pop("_roi_gpu")
        out["served"] = sv
    if not args.no_radiography:
        out["radiography"] = radiography(ctx, args)
    if not args.no_ct:
        out["ct"] = ct_scan(ctx, args.ct_views, args.ct_scale, args.ct_plain_every, params=args.ct_params)
    if not args.no_precision:
        out["precision"] = precision_study(ctx, args.precision_histories)
    if args.copy_probe_mb:
        out["copy_probe"] = copy_probe(ctx, args.copy_probe_mb)
    if not args.no_atomic_probe:
        out["atomic_probe"] = atomic_probe(ctx)
        if out.get("roofline"):  # the served kernel's deposit atomics against the probe on a 512^3 grid
            peak = out["atomic_probe"][-1]["atomics_per_s"]
            out["roofline"]["dose_atomics_probe_per_s"] = peak
            out["roofline"]["dose_atomics_frac"] = out["roofline"]["dose_atomics_per_s"] / peak
    if not args.no_cpu_baseline and jobs is not None:
        out["cpu_baseline"] = cpu_baseline(jobs, ph, args.cpu_seconds, affinity, roi_gpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
