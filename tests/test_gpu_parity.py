"""Transport parity: the HIP kernel against the FP64 oracle on the same
scenes, tallies compared within combined standard errors (from independent
batches on each side) and with per-tally chi^2 tests at p > 0.01; plus
invariants the physics fixes exactly (energy conservation, primary
transmission) and determinism of the counter-based RNG."""
import math

import numpy as np
import pytest

import midsx_amd as M
import oracle
from midsx_amd import scenes
from helpers import batch_mean_sigma, chi2_pvalue, count_chi2, hotelling, welch_normal_scores

pytestmark = pytest.mark.gpu


def gpu_batches(ctx, n, k, seed, extract, kernel=M.KERNEL_PLAIN):
    out = []
    for b in range(k):
        ctx.reset_tallies()
        st = ctx.run(n, seed=seed, first=b * n, kernel=kernel)
        out.append(extract(ctx, st, n))
    return out


def orc_batches(sim, n, k, seed, extract):
    out = []
    for b in range(k):
        sim.reset()
        sim.run(n, seed=seed + 7919 * b, threads=8)
        out.append(extract(sim, sim.counters(), n))
    return out


def slab_extract(src, st, n):
    v = src.tally(0)["acc"]
    s = src.tally(1)
    dose = src.voxel_dose(0)[0].reshape(10, 40, 40).sum(axis=(1, 2))  # per z slice
    return np.concatenate([
        v[:, :, 1].ravel() / n,          # volume quantity sums per bucket
        s["acc"][:, :, 1].ravel() / n,   # surface quantity sums per bucket
        dose / n,
        [st["steps"] / n, st["real"] / n, st["rayleigh"] / n, st["compton"] / n, st["photo"] / n],
    ])


@pytest.fixture(scope="module")
def slab():
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C1_MATERIALS))
    scenes.setup_slab(ctx)
    sim = oracle.Simulation(scenes.C1_MATERIALS)
    scenes.setup_slab(sim)
    yield ctx, sim
    ctx.close()


def test_slab_tallies_match_oracle(slab):
    """Volume- and surface-tally quantities per bucket, the depth-dose
    profile and the event counts: every bin within 3.5 combined standard
    errors (Welch) and the whole vector by a two-sample Hotelling T^2 test on
    equal-size batches (the bins are correlated; several are proportional)."""
    ctx, sim = slab
    k, n = 48, 50000
    g = np.array(gpu_batches(ctx, n, k, 11, slab_extract))
    o = np.array(orc_batches(sim, n, k, 101, slab_extract))
    gm, gs = batch_mean_sigma(g)
    om, os_ = batch_mean_sigma(o)
    keep = (gs > 0) & (os_ > 0)
    z = welch_normal_scores(gm, gs, k, om, os_, k)[keep]
    assert keep.sum() > 25
    assert np.all(np.abs(z) < 3.5), (np.max(np.abs(z)), np.nonzero(keep)[0][np.argmax(np.abs(z))])
    p, t2, rank = hotelling(g[:, keep], o[:, keep])
    assert p > 0.01, (p, t2, rank)


def test_slab_spectrum_chi2(slab):
    ctx, sim = slab
    ctx.reset_tallies()
    ng = 2000000
    ctx.run(ng, seed=21)
    gh = ctx.tally(1)["hist"]
    sim.reset()
    no = 600000
    sim.run(no, seed=23, threads=0)
    oh = sim.tally(1)["hist"]
    p, chi2, dof = count_chi2(gh, ng, oh, no)
    assert dof >= 5 and p > 0.01, (p, chi2, dof)
    # primary photons all sit in the 60 keV bin
    assert gh[0, 60] == gh[0].sum() and oh[0, 60] == oh[0].sum()


def test_slab_primary_transmission_closed_form(slab):
    ctx, _ = slab
    ctx.reset_tallies()
    n = 2000000
    ctx.run(n, seed=31)
    got = ctx.tally(1)["acc"][2, 0, 0] / n
    w, a = oracle.material(48), oracle.material(3)
    p = math.exp(-w.total_i(60e3) * 1.0 - a.total_i(60e3) * 99.0)
    assert abs(got - p) < 4 * math.sqrt(p * (1 - p) / n), (got, p)


def six_faces(dim):
    X, Y, Z = dim
    return [
        dict(kind=2, p0=(0, 0, 0), p1=(0, Y, 0), p2=(X, 0, 0)),
        dict(kind=2, p0=(0, 0, Z), p1=(X, 0, 0), p2=(0, Y, 0)),
        dict(kind=2, p0=(0, 0, 0), p1=(0, 0, Z), p2=(0, Y, 0)),
        dict(kind=2, p0=(X, 0, 0), p1=(0, Y, 0), p2=(0, 0, Z)),
        dict(kind=2, p0=(0, 0, 0), p1=(X, 0, 0), p2=(0, 0, Z)),
        dict(kind=2, p0=(0, Y, 0), p1=(0, 0, Z), p2=(X, 0, 0)),
    ]


def test_energy_conservation_gpu():
    ctx = M.Context(0)
    ctx.build_tables([3, 48, 169])
    dim = (6.0, 6.0, 20.0)
    ctx.set_domain(dim, 3)
    ids = np.full((30, 30, 20), 48, np.uint8)
    ids[10:20, 10:20, 5:15] = 169
    ctx.add_grid(ids, (0.2, 0.2, 0.2), (0, 0, 8))
    ctx.set_source(direction=M.DIR_ISOTROPIC, position=(3, 3, 10), energy=50e3)
    ctx.set_tallies(six_faces(dim))
    ctx.enable_voxel_dose(True)
    n = 1000000
    st = ctx.run(n, seed=41)
    dep = ctx.voxel_dose(0)[0].sum() + ctx.background_dose()[0]
    out = sum(ctx.tally(i)["acc"][0, :, 1].sum() for i in range(6))
    assert st["histories"] == n
    assert abs(dep + out - n * 50e3) / (n * 50e3) < 2e-6, (dep, out)
    assert st["photo"] + st["escapes"] == n
    ctx.close()


def test_rng_streams_are_per_history(slab):
    """Histories are keyed by (seed, index): a run split in two gives the same
    event counts as one run, and a repeat gives the same tallies up to FP64
    summation order."""
    ctx, _ = slab
    ctx.reset_tallies()
    a = ctx.run(300000, seed=7, first=0)
    ta = ctx.tally(0)["acc"].copy()
    ctx.reset_tallies()
    b1 = ctx.run(100000, seed=7, first=0)
    b2 = ctx.run(200000, seed=7, first=100000)
    tb = ctx.tally(0)["acc"].copy()
    for k in ("steps", "real", "rayleigh", "compton", "photo", "kahn_trials", "deposits", "surface_hits"):
        assert a[k] == b1[k] + b2[k], k
    assert np.allclose(ta, tb, rtol=1e-12, atol=0)
    ctx.reset_tallies()
    c = ctx.run(300000, seed=8, first=0)
    assert c["steps"] != a["steps"]


def radiography_extract(src, st, n):
    img = src.tally(0)["image"]
    coarse = img[:, :, 0].reshape(4, 8, 4, 8).sum(axis=(1, 3)).ravel()  # 4 x 4 blocks of energy fluence
    dose = src.voxel_dose(0)[0].reshape(4, 16, 4, 16, 4, 16).sum(axis=(1, 3, 5)).ravel()  # 4^3 blocks
    return np.concatenate([coarse / n, dose / n, [src.background_dose()[0] / n],
                           [st["steps"] / n, st["real"] / n, st["compton"] / n, st["photo"] / n]])


@pytest.mark.parametrize("kernel", [M.KERNEL_PLAIN, M.KERNEL_QUEUE])
def test_radiography_small_matches_oracle(kernel):
    """Polyenergetic Kramers source, cone beam, heterogeneous phantom
    (water / bone / adipose), image and per-voxel dose tallies: each coarse
    bin within 3.5 combined standard errors and Hotelling T^2 on the whole.
    The queue kernel runs with background free flights (image and dose
    tallies only)."""
    kw = dict(shape=(64, 64, 64), voxel_cm=0.5, seed=3, image=(32, 32))
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    scenes.setup_radiography(ctx, **kw)
    sim = oracle.Simulation(scenes.C2_MATERIALS)
    scenes.setup_radiography(sim, **kw)
    k, n = 160, 10000
    g = np.array(gpu_batches(ctx, n, k, 51, radiography_extract, kernel))
    o = np.array(orc_batches(sim, n, k, 151, radiography_extract))
    if kernel == M.KERNEL_QUEUE:  # free flights take fewer (rejected) steps: drop the step count
        g, o = np.delete(g, -4, axis=1), np.delete(o, -4, axis=1)
        assert ctx.run(20000, seed=5, first=10 ** 9, kernel=kernel)["steps"] < 0.7 * ctx.run(
            20000, seed=5, first=10 ** 9)["steps"]
    gm, gs = batch_mean_sigma(g)
    om, os_ = batch_mean_sigma(o)
    keep = (gs > 0) & (os_ > 0) & (np.abs(om) > 3 * os_)
    z = welch_normal_scores(gm, gs, k, om, os_, k)[keep]
    assert keep.sum() > 40
    assert np.all(np.abs(z) < 3.5), np.max(np.abs(z))
    p, t2, rank = hotelling(g[:, keep], o[:, keep])
    assert p > 0.01, (p, t2, rank)
    ctx.close()
