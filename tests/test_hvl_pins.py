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


def test_recorded_kvp_ratios_are_consistent():
    """The recorded kVp runs: HVL halves and QVL quarters the air kerma."""
    for kvp in (30, 100):
        r, s = recorded_ratio_kvp(kvp, "hvl")
        assert abs(r - 0.5) < 5e-3 and s < 1e-3, (kvp, r, s)
        r, s = recorded_ratio_kvp(kvp, "qvl")
        assert abs(r - 0.25) < 5e-3, (kvp, r)


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


KVP_SPECTRA = {30: "RQRM3_MO_MO_30kVp_E_spectrum.csv", 100: "RQR8_W_AL_100kVp_E_spectrum.csv"}


def kvp_spectrum(kvp):
    """The reference's spectrum file (data/source_distributions/, keV and
    weight per 0.5 keV bin centre), energies in eV as the drivers convert
    them (hvl_30kVp.cpp:21-25)."""
    d = np.loadtxt(os.path.join(GOLDEN, "spectra", KVP_SPECTRA[kvp]), delimiter=",")
    return d[:, 0] * 1e3, d[:, 1]


def recorded_ratio_kvp(kvp, layer):
    a0 = np.loadtxt(os.path.join(GOLDEN, "hvl", "Al_%dkVp_AK_0.txt" % kvp))
    a1 = np.loadtxt(os.path.join(GOLDEN, "hvl", "Al_%dkVp_AK_%s.txt" % (kvp, layer)))
    r = a1.mean() / a0.mean()
    s = r * math.hypot(a1.std(ddof=1) / a1.mean() / math.sqrt(len(a1)), a0.std(ddof=1) / a0.mean() / math.sqrt(len(a0)))
    return r, s


@pytest.mark.gpu
@pytest.mark.parametrize("kvp,layer", [(30, "hvl"), (30, "qvl"), (100, "hvl"), (100, "qvl")])
def test_gpu_kvp_air_kerma_ratio_matches_recorded(kvp, layer):
    """Polyenergetic beams (hvl_30kVp.cpp, hvl_100kVp.cpp): the primary air
    kerma at the disc, sum over the spectrum's bins of E * fluence * mu_en/rho
    of air (DerivedQuantity::getPrimarySpectrumAirKerma), with and without the
    Al layer; the ratio within 4 combined sigma (40 batches each) of the
    reference's recorded outputs."""
    from midsx_amd import derived
    e, w = kvp_spectrum(kvp)
    k, n = 40, 1000000
    ak = {}
    for name in ("air_only.json", "%s_%dkVp.json" % (layer, kvp)):
        ctx = M.Context(0)
        scene_into(ctx, name)
        ctx.build_tables([3, 61])
        ctx.set_source(direction=M.DIR_DISC_ISOTROPIC, position=(2, 2, 0), a=(2, 2, 10), b=(0, 0, 1), radius=0.1,
                       spectrum=(e, w))
        ctx.set_tallies([derived.spectrum_disc_tally((2, 2, 100), (0, 0, 1), 1.0, e)])
        vals = []
        for b in range(k):
            ctx.reset_tallies()
            ctx.run(n, seed=kvp, first=b * n, kernel=M.KERNEL_QUEUE)
            vals.append(derived.primary_spectrum_air_kerma(ctx, 0, e) / n)
        ak[name] = np.array(vals)
        ctx.close()
    a0, a1 = ak["air_only.json"], ak["%s_%dkVp.json" % (layer, kvp)]
    ratio = a1.mean() / a0.mean()
    sig = ratio * math.hypot(a1.std(ddof=1) / a1.mean(), a0.std(ddof=1) / a0.mean()) / math.sqrt(k)
    ref, ref_s = recorded_ratio_kvp(kvp, layer)
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


def build_cpp_kvp_driver(tmpdir):
    exe = os.path.join(str(tmpdir), "hvl_kvp")
    lib = os.path.join(ROOT, "midsx_amd", "_lib")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "midsx_amd", "cpp"),
           os.path.join(ROOT, "midsx_amd", "cpp", "examples", "hvl_kvp.cpp"), "-o", exe, "-L", lib,
           "-lmidsx_amd", "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return exe


def test_cpp_kvp_driver_compiles(tmp_path):
    assert os.path.exists(build_cpp_kvp_driver(tmp_path))


@pytest.mark.gpu
def test_cpp_kvp_driver_air_kerma_ratio(tmp_path):
    """The reference's 30 kVp driver, ported to this API, reproduces the
    recorded AK(HVL) / AK(0)."""
    exe = build_cpp_kvp_driver(tmp_path)
    spec = os.path.join(GOLDEN, "spectra", KVP_SPECTRA[30])
    n = 20000000
    ak = {}
    for name in ("air_only.json", "hvl_30kVp.json"):
        r = subprocess.run([exe, os.path.join(SCENES, name), spec, str(n)], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        kv = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
        ak[name] = float(kv["primary_air_kerma"])
    ratio = ak["hvl_30kVp.json"] / ak["air_only.json"]
    ref, ref_s = recorded_ratio_kvp(30, "hvl")
    sig = ratio * math.sqrt(3.0 / n)
    assert abs(ratio - ref) < 4 * math.hypot(sig, ref_s), (ratio, ref)


def test_cpp_matrix_call_forms(tmp_path):
    """The reference drivers' call forms that pass or receive Eigen types
    (readCSV into a matrix, PolyenergeticSpectrum from it, getDimSpace into a
    3-vector, the quantity-name helpers) compile and run against the header,
    with matrix and vector types that have only the members those drivers
    use (host only)."""
    exe = os.path.join(str(tmp_path), "matrix_call_forms")
    lib = os.path.join(ROOT, "midsx_amd", "_lib")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "midsx_amd", "cpp"),
           os.path.join(ROOT, "tests", "cpp", "matrix_call_forms.cpp"), "-o", exe, "-L", lib, "-lmidsx_amd",
           "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    csv = tmp_path / "spectrum.csv"
    csv.write_text("10000,0.25\n20000,0.5\n30000,0.25\n")
    r = subprocess.run([exe, str(csv)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0 and "matrix call forms ok" in r.stdout, r.stdout
