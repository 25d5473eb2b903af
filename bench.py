"""Benchmark of the photon-transport hot path on one MI355X.

Workload (radiography): 256 x 256 x 160 uint8 phantom of seeded ellipsoids at
1.25 mm (water, cortical bone, adipose) in air, point source 60 cm above the
phantom, cone beam over its top face, Kramers 15-80 keV spectrum in 1 keV
bins, a 256 x 256 energy-fluence image at the far face and FP64 per-voxel
deposited energy.  One step transports one batch of histories; the default
10 steps of 1e7 histories are the 1e8 histories of the workload.  Inputs
(tables, phantom, tallies) are resident in HBM before timing starts.

Prints one JSON line: throughput, the roofline of the transport kernel
(algorithmic bytes from its int64 event counters over its HIP-event
duration), and the FP64 CPU restatement timed on a bounded sample of the same
workload on the host cores.

Served job stream (the "served" object): dose jobs drawn with a fixed seed
(cubic phantoms of 64^3 to 512^3 voxels at 0.5-2 mm, monoenergetic 20-120 keV
or Kramers 60-120 kVp sources, 1e6-1e8 histories log-uniform, FP64 voxel dose
plus a 128 x 128 image each), served by the continuous-batching loop and,
for comparison, first-in first-out by the plain kernel one job per launch:
throughput with every job queued at the start, then p50/p99 latency under
Poisson arrivals at a fraction of the served capacity.  Phantoms are built
before the clock starts (they are the clients' data, handed over in pinned
host memory); uploads, tally zeroing, transport and the read-back of every
job's dose and image are timed; the ROI statistics come from one more
served pass outside the timed runs.
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


def algorithmic_bytes(st):
    """Bytes one launch must move at minimum: per Woodcock step the voxel ID
    (1) and the sigma_max and sigma_tot brackets (2 x 16); per real event the
    two scattering cross-section brackets (32); per Rayleigh two RITA rows
    (48); per Kahn S-test two S(x) brackets (32); per deposit an FP64 sum and
    sum^2 read-modify-write (32); per image hit 16; per history 8."""
    return (33 * st["steps"] + 32 * st["real"] + 48 * st["rayleigh"] + 32 * st["kahn_s_tests"]
            + 32 * st["deposits"] + 16 * st["surface_hits"] + 8 * st["histories"])


def measured_traffic_per_history():
    """HBM bytes per history of the queue kernel on this workload from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", "k_transport_queue_radiography_256x256x160.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_history"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_baseline(kw, seconds=12.0, threads=16):
    import oracle
    from midsx_amd import scenes
    sim = oracle.Simulation(scenes.C2_MATERIALS)
    scenes.setup_radiography(sim, **kw)
    n = 2000
    dt = sim.run(n, seed=99, threads=threads)
    while dt < 1.0 and n < 10 ** 7:
        n *= 4
        sim.reset()
        dt = sim.run(n, seed=99, threads=threads)
    n = int(max(n, n * seconds / max(dt, 1e-9)))
    sim.reset()
    dt = sim.run(n, seed=100, threads=threads)
    return dict(value=n / dt, unit="histories/s", cores=threads, kind="port",
                sample="%d histories of the radiography workload (full phantom, image and voxel-dose tallies), "
                       "FP64 C++/OpenMP restatement, %.1f s wall" % (n, dt))


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


def copy_probe(ctx, mb):
    """The library's streaming copy of mb MiB (read mb MiB, write mb MiB),
    twice: the reference for scaling the HBM counters of a profiled run, and
    the achievable copy bandwidth."""
    nbytes = mb * 2 ** 20
    ms = ctx.copy_probe(nbytes)
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB_per_s": 2 * nbytes / ms / 1e6}


def served_stream(ctx, n_jobs, seed, load):
    """The served job stream: saturated throughput of the continuous-batching
    loop and of plain FIFO service, then latency under Poisson arrivals."""
    from midsx_amd import scenes
    from midsx_amd.serve import Server, run_plain, poisson_arrivals, latency_percentiles, roi_summary, \
        warm_read_buffers
    jobs = scenes.draw_dose_jobs(n_jobs, seed)
    cache = {}
    for j in jobs:  # the clients' phantoms, in pinned host memory (x fastest)
        ids = scenes.dose_phantom(j["shape"], j["voxel_cm"], j["phantom_seed"])
        try:
            buf = ctx.pinned(ids.size, np.uint8).reshape(ids.shape, order="F")
            buf[...] = ids
            ids = buf
        except Exception:
            pass
        cache[id(j)] = ids
    ph = lambda j: cache[id(j)]
    H = sum(j["histories"] for j in jobs)
    # the stream's own launch parameters, tuned before timing on one job of
    # it (a 256^3 phantom if there is one), as a server does at start-up
    keep_params = ctx.queue_params()
    keep_job = ctx.job
    probe = next((j for j in jobs if j["shape"][0] == 256), jobs[0])
    pj = ctx.job_create()
    scenes.setup_dose_job(ctx, probe, ids=ph(probe))
    tuned = ctx.autotune(probe_histories=4000000)
    ctx.job_release(pj)
    ctx.job_select(keep_job)
    # an untimed pass of the same stream first, which also yields the dose
    # statistics: the device memory pool and the pinned result buffers are
    # then warm, as in a service that has been running
    summaries = {}

    def keep(job, mean, rel, image):
        summaries[id(job)] = roi_summary(job, mean, rel, image)

    Server(ctx, phantoms=ph, on_result=keep).run(jobs)
    rse = [summaries[id(j)]["roi_median_rse"] for j in jobs]
    srv = Server(ctx, phantoms=ph)
    t = time.perf_counter()
    srv.run(jobs)
    served_s = time.perf_counter() - t
    warm_read_buffers(ctx, jobs)  # the plain service's result buffers, pinned ahead as well
    _, pinfo = run_plain(ctx, jobs, phantoms=ph)
    plain_s = pinfo["makespan"]
    cap = n_jobs / served_s
    arr = poisson_arrivals(n_jobs, load * cap, seed + 1)
    srv2 = Server(ctx, phantoms=ph)
    lat = srv2.run(jobs, arr)
    p50, p99 = latency_percentiles(lat)
    plat, _ = run_plain(ctx, jobs, arr, phantoms=ph)
    q50, q99 = latency_percentiles(plat)
    ctx.set_queue_params(*keep_params)
    return {
        "workload": "%d dose jobs (seed %d): phantoms %s, histories 1e6-1e8 log-uniform, total %d" % (
            n_jobs, seed, sorted(set(j["shape"][0] for j in jobs)), H),
        "value": H / served_s,
        "unit": "histories/s",
        "plain_value": H / plain_s,
        "speedup_vs_plain": plain_s / served_s,
        "served_s": served_s,
        "plain_s": plain_s,
        "kernel_s": srv.kernel_ms / 1e3,
        "plain_kernel_s": pinfo["kernel_ms"] / 1e3,
        "rounds": srv.rounds,
        "host_s": {k: round(v, 4) for k, v in srv.host_s.items()},
        "steps_per_history": srv.counters["steps"] / H,
        "plain_steps_per_history": pinfo["counters"]["steps"] / H,
        "latency": {"arrival_rate_jobs_per_s": load * cap, "load": load, "p50_s": p50, "p99_s": p99,
                    "plain_fifo_p50_s": q50, "plain_fifo_p99_s": q99},
        "roi_median_rse": {"median_over_jobs": float(np.median(rse)), "max_over_jobs": float(np.max(rse))},
        "queue_params": tuned,
    }


def ct_scan(ctx, n_views, scale, seed=4):
    """CT: 512 x 512 x 256 phantom at 0.8 mm (water, lung, adipose, bone),
    n_views views about its z axis, 120 kVp, a 512 x 64 detector per view,
    per-view histories U[2e6, 2e7] x scale.  All views in one multi-job launch
    of the queue kernel (views alias the phantom job; the phantom's voxel dose
    sums over views) versus the plain kernel one view per launch."""
    import midsx_amd as M
    from midsx_amd import scenes
    owner = ctx.job_create()
    geo, views, angles = scenes.setup_ct(ctx, n_views=n_views, seed=seed)
    n = [max(1, int(h * scale)) for h in scenes.ct_histories(n_views, seed)]
    seeds = [10_000 + v for v in range(n_views)]
    H = sum(n)
    ctx.job_select(views[0])
    tuned = ctx.autotune(probe_histories=4000000)  # one view: launches of ~15 ms time stably
    ctx.set_graphs(True)
    ctx.run_jobs(views[:8], [100000] * 8, [1 << 40] * 8, seeds[:8])  # warm (capture)
    for j in [owner] + views:
        ctx.job_select(j)
        ctx.reset_tallies()
    t = time.perf_counter()
    st = ctx.run_jobs(views, n, [0] * n_views, seeds)
    q_s = time.perf_counter() - t
    for j in [owner] + views:
        ctx.job_select(j)
        ctx.reset_tallies()
    t = time.perf_counter()
    pk = 0.0
    psteps = 0
    for v, nv, sv in zip(views, n, seeds):
        ctx.job_select(v)
        ps = ctx.run(nv, seed=sv, first=0, kernel=M.KERNEL_PLAIN)
        pk += ps["kernel_ms"]
        psteps += ps["steps"]
    p_s = time.perf_counter() - t
    ctx.set_graphs(False)
    for j in views + [owner]:
        ctx.job_release(j)
    return {
        "workload": "CT: 512x512x256 phantom at 0.8 mm, %d views, 120 kVp, 512x64 detector per view, "
                    "histories per view U[2e6,2e7] x %g (total %d)" % (n_views, scale, H),
        "value": H / q_s,
        "unit": "histories/s",
        "plain_value": H / p_s,
        "speedup_vs_plain": p_s / q_s,
        "queue_s": q_s,
        "queue_kernel_ms": st["kernel_ms"],
        "plain_s": p_s,
        "plain_kernel_ms": pk,
        "launches": {"queue": 1, "plain": n_views},
        "steps_per_history": st["steps"] / H,
        "plain_steps_per_history": psteps / H,
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
    keep = ctx.queue_params()
    tuned = ctx.autotune(probe_histories=4000000)  # the slab's own launch parameters
    out = {}
    res = {}
    for name, mode in (("fp64", M.DOSE_F64), ("fp32", M.DOSE_F32), ("fp32_compensated", M.DOSE_F32C),
                       ("fp32_compensated_lds", M.DOSE_F32C_LDS)):
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
    return {"workload": "slab 40x40x10 water voxels, 60 keV beam, %d histories per mode, identical RNG streams; "
                        "errors over the %d voxels with >= 1000 deposits" % (n_hist, int(roi.sum())),
            "queue_params": {k: tuned[k] for k in ("blocks_per_cu", "refill_min", "interact_min")},
            "modes": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--histories-per-step", type=int, default=10 ** 7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--queue-params", default=None,
                    help="blocks_per_cu,refill_min,interact_min for the queue kernel (skips the autotuner)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-probe-mb", type=int, default=0)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=32)
    ap.add_argument("--served-seed", type=int, default=2026)
    ap.add_argument("--served-load", type=float, default=0.7)
    ap.add_argument("--no-ct", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-histories", type=int, default=10 ** 9)
    ap.add_argument("--ct-views", type=int, default=360)
    ap.add_argument("--ct-scale", type=float, default=0.05,
                    help="fraction of the CT workload's per-view histories (1.0 = the full scan, ~4e9 histories)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: sharding histories over several GPUs is not built in this "
              "tree; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import midsx_amd as M
    from midsx_amd import scenes

    kw = dict(shape=(256, 256, 160), voxel_cm=0.125, seed=2, image=(256, 256))
    ctx = M.Context(0)
    ctx.build_tables(sorted(set(scenes.C2_MATERIALS) | set(scenes.DOSE_MATERIALS)))
    scenes.setup_radiography(ctx, **kw)
    n = args.histories_per_step
    tuned = None
    if args.queue_params:
        ctx.set_queue_params(*[int(v) for v in args.queue_params.split(",")])
    elif not args.no_autotune:
        tuned = ctx.autotune(probe_histories=4000000)  # launches of ~3 ms: stable timings
    stats, wall, first = timed_steps(ctx, M.KERNEL_QUEUE, n, args.steps, args.warmup, 0)
    plain = None
    if not args.no_plain:
        pstats, pwall, _ = timed_steps(ctx, M.KERNEL_PLAIN, n, args.steps, args.warmup, first)
        plain = {"value": n * args.steps / pwall, "ms_per_step": pwall / args.steps * 1e3,
                 "kernel_ms_avg": float(np.mean([s["kernel_ms"] for s in pstats]))}
    total = n * args.steps
    kern_ms = np.array([s["kernel_ms"] for s in stats])
    bytes_per_launch = np.mean([algorithmic_bytes(s) for s in stats])
    achieved = bytes_per_launch / (kern_ms.mean() / 1e3) / 1e9
    agg = {k: sum(s[k] for s in stats) for k in M.STAT_NAMES}
    value = total / wall

    out = {
        "metric": "photon histories/sec vs plain GPU at matched tally \u03c3; p50/p99 request latency",
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded ellipsoid phantom; EPDL2017 tables for air, water, cortical bone, adipose)",
        "config": {
            "workload": "radiography: 256x256x160 uint8 phantom at 1.25 mm, 15-80 keV Kramers cone beam, "
                        "256x256 image + FP64 voxel dose, 1e8 histories",
            "histories_per_step": n,
            "kernel": "queue (event-regrouped, background free flights); plain = one thread per history, "
                      "majorant steps everywhere, every step replayed for the tallies; equal histories, "
                      "hence equal tally sigma",
            "queue_params": dict(zip(("blocks_per_cu", "refill_min", "interact_min"), ctx.queue_params())),
            "steps_per_history": agg["steps"] / total,
        },
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (measured_traffic_per_history() or 0.0) * n or None,
            "traffic_source": "rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 per launch (factors from the "
                              "k_copy16 streaming copy), profiles/k_transport_queue_radiography_256x256x160.md",
            "kernel_ms_avg": float(kern_ms.mean()),
            "algorithmic_bytes_per_launch": float(bytes_per_launch),
        },
        "events_per_history": {k: agg[k] / total for k in ("steps", "real", "rayleigh", "compton", "photo",
                                                            "kahn_trials", "deposits", "surface_hits")},
    }
    if tuned:
        out["config"]["autotune_probe_ms"] = tuned["probe_ms"]
    if not args.no_served:
        out["served"] = served_stream(ctx, args.served_jobs, args.served_seed, args.served_load)
    if not args.no_ct:
        out["ct"] = ct_scan(ctx, args.ct_views, args.ct_scale)
    if not args.no_precision:
        out["precision"] = precision_study(ctx, args.precision_histories)
    if args.copy_probe_mb:
        out["copy_probe"] = copy_probe(ctx, args.copy_probe_mb)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(kw, seconds=args.cpu_seconds, threads=args.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
