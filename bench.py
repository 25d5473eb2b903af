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


def copy_probe(mb):
    """A plain device-to-device copy of known size, for scaling HBM counters."""
    import torch
    a = torch.empty(mb * 2 ** 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    return 2 * mb * 2 ** 20


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--histories-per-step", type=int, default=10 ** 7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-probe-mb", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: sharding histories over several GPUs is not built in this "
              "tree; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import midsx_amd as M
    from midsx_amd import scenes

    kw = dict(shape=(256, 256, 160), voxel_cm=0.125, seed=2, image=(256, 256))
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    scenes.setup_radiography(ctx, **kw)
    n = args.histories_per_step
    tuned = None if args.no_autotune else ctx.autotune(probe_histories=1000000)
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
        "metric": "photon histories/sec (radiography workload; queue kernel vs the plain one-simulation-per-launch "
                  "kernel at equal histories, hence equal tally sigma)",
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
            "kernel": "queue",
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
            "traffic": None,
            "kernel_ms_avg": float(kern_ms.mean()),
            "algorithmic_bytes_per_launch": float(bytes_per_launch),
        },
        "events_per_history": {k: agg[k] / total for k in ("steps", "real", "rayleigh", "compton", "photo",
                                                            "kahn_trials", "deposits", "surface_hits")},
    }
    if tuned:
        out["config"]["autotune_probe_ms"] = tuned["probe_ms"]
    if args.copy_probe_mb:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe_mb)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(kw, seconds=args.cpu_seconds, threads=args.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
