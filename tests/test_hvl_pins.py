"""Half / quarter value layer runs against the reference's own recorded
outputs (cpp_simulations/half_value_layer/data/Al_*_AK_*.txt, ten runs each).

Scene and source follow cpp_simulations/half_value_layer/hvl_30keV.cpp:9-80
and hvl.json: a 4 x 4 x 100 cm air domain, the Al layer grid at z = 10 cm,
a point source at (2, 2, 0) aimed by DiscIsotropicDirectionality at a
0.1 cm disc around (2, 2, 10), and a 1 cm disc surface tally at z = 100.
For a monoenergetic beam the recorded primary air kerma is the primary photon
count at the tally times a constant, so AK(layer) / AK(0) is the ratio of
primary counts with and without the layer.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import midsx_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SCENES = os.path.join(GOLDEN, "scene")


def recorded_ratio(energy_kev, layer):
    a0 = np.loadtxt(os.path.join(GOLDEN, "hvl", "Al_%dkeV_AK_0.txt" % energy_kev))
    a1 = np.loadtxt(os.path.join(GOLDEN, "hvl", "Al_%dkeV_AK_%s.txt" % (energy_kev, layer)))
    r = a1.mean() / a0.mean()
    s = r * math.hypot(a1.std(ddof=1) / a1.mean() / math.sqrt(len(a1)), a0.std(ddof=1) / a0.mean() / math.sqrt(len(a0)))
    return r, s


def configure(sim, energy):
    sim.set_source(direction=M.DIR_DISC_ISOTROPIC, position=(2, 2, 0), a=(2, 2, 10), b=(0, 0, 1), radius=0.1,
                   energy=energy)
    sim.set_tallies([dict(kind=M.TALLY_DISC, p0=(2, 2, 100), normal=(0, 0, 1), radius=1.0)])


def scene_into(sim, json_name):
    """Reads the scene with the library's host-only reader and hands the
    grids to sim (a Context or the oracle)."""
    h = M.Context(-1)
    h.load_scene(os.path.join(SCENES, json_name))
    sim.set_domain((4.0, 4.0, 100.0), 3)
    for g in range(h.num_grids()):
        dims, sp, org = h.grid_info(g)
        sim.add_grid(h.grid_ids(g), sp, org)
    h.close()


def primary_fraction(sim, runner, n):
    runner(n)
    return sim.tally(0)["acc"][2, 0, 0] / n


def test_recorded_ratios_are_consistent():
    r, s = recorded_ratio(30, "hvl")
    assert abs(r - 0.5) < 1e-3 and s < 1e-4
    r, s = recorded_ratio(30, "qvl")
    assert abs(r - 0.25) < 1e-3


def test_oracle_hvl_30kev_matches_recorded():
    import oracle
    n = 2000000
    out = {}
    for name in ("air_only.json", "hvl_30keV.json"):
        sim = oracle.Simulation([3, 61])
        scene_into(sim, name)
        configure(sim, 30e3)
        out[name] = primary_fraction(sim, lambda k: sim.run(k, seed=3, threads=0), n)
    p0, p1 = out["air_only.json"], out["hvl_30keV.json"]
    ratio = p1 / p0
    sig = ratio * math.sqrt((1 - p1) / (p1 * n) + (1 - p0) / (p0 * n))
    ref, ref_s = recorded_ratio(30, "hvl")
    assert abs(ratio - ref) < 4 * math.hypot(sig, ref_s), (ratio, ref, sig)


@pytest.mark.gpu
@pytest.mark.parametrize("energy_kev,layer", [(30, "hvl"), (30, "qvl"), (100, "hvl"), (100, "qvl")])
def test_gpu_layer_ratio_matches_recorded(energy_kev, layer):
    n = 40000000
    fr = {}
    for name in ("air_only.json", "%s_%dkeV.json" % (layer, energy_kev)):
        ctx = M.Context(0)
        scene_into(ctx, name)
        ctx.build_tables([3, 61])
        configure(ctx, energy_kev * 1e3)
        fr[name] = primary_fraction(ctx, lambda k: ctx.run(k, seed=energy_kev), n)
        ctx.close()
    p0 = fr["air_only.json"]
    p1 = fr["%s_%dkeV.json" % (layer, energy_kev)]
    ratio = p1 / p0
    sig = ratio * math.sqrt((1 - p1) / (p1 * n) + (1 - p0) / (p0 * n))
    ref, ref_s = recorded_ratio(energy_kev, layer)
    assert abs(ratio - ref) < 4 * math.hypot(sig, ref_s), (ratio, ref, sig, ref_s)


def build_cpp_driver(tmpdir):
    exe = os.path.join(str(tmpdir), "hvl_slab")
    lib = os.path.join(ROOT, "midsx_amd", "_lib")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "midsx_amd", "cpp"),
           os.path.join(ROOT, "midsx_amd", "cpp", "examples", "hvl_slab.cpp"), "-o", exe, "-L", lib,
           "-lmidsx_amd", "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return exe


def test_cpp_driver_compiles(tmp_path):
    """A driver written against the reference's C++ API compiles and links
    against the library (C++ host API over the C-ABI)."""
    assert os.path.exists(build_cpp_driver(tmp_path))


@pytest.mark.gpu
def test_cpp_driver_hvl_ratio(tmp_path):
    exe = build_cpp_driver(tmp_path)
    n = 20000000
    res = {}
    for name in ("air_only.json", "hvl_30keV.json"):
        r = subprocess.run([exe, os.path.join(SCENES, name), str(n), "30000"], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        kv = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
        res[name] = int(kv["primary_photons"]) / n
    ratio = res["hvl_30keV.json"] / res["air_only.json"]
    ref, ref_s = recorded_ratio(30, "hvl")
    sig = ratio * math.sqrt(2.2 / n)
    assert abs(ratio - ref) < 4 * math.hypot(sig, ref_s), (ratio, ref)
