"""Source energy sampling on the GPU (probability_dist.cpp:18-28, lower_bound
of u in the spectrum's CDF, which the kernels search from a guide table):
primary photons crossing a thin air domain are counted per energy by an
IncidentEnergy histogram of one bin per spectrum bin, and the counts must
follow the bins' probabilities (chi^2, p > 0.01).  The spectra have many
uneven bins, so that the guide's bins hold several CDF values each, and bins
of zero weight, which must never be drawn."""
import numpy as np
import pytest
from scipy import stats

import midsx_amd as M
from midsx_amd import scenes

pytestmark = pytest.mark.gpu


def sampled_counts(ctx, energies, weights, n, kernel):
    ctx.set_domain((4.0, 4.0, 1.0), scenes.AIR)
    ctx.set_source(direction=M.DIR_BEAM, position=(2.0, 2.0, 0.0), a=(2.0, 2.0, 1.0),
                   spectrum=(energies, weights))
    nb = len(energies)
    ctx.set_tallies([dict(kind=M.TALLY_DISC, p0=(2.0, 2.0, 1.0), normal=(0.0, 0.0, 1.0), radius=1.0,
                          hist_bins=nb, hist_lo=energies[0] - 500.0, hist_width=1000.0)])
    ctx.reset_tallies()
    ctx.run(n, seed=17, kernel=kernel)
    return ctx.tally(0)["hist"][0]  # primaries


@pytest.mark.parametrize("kernel", [M.KERNEL_PLAIN, M.KERNEL_QUEUE])
@pytest.mark.parametrize("nb", [65, 300])
def test_spectrum_sampling_chi2(kernel, nb):
    rng = np.random.default_rng(nb)
    energies = 1000.0 * (np.arange(nb) + 20.0)  # histogram bin k holds energy 20 + k keV
    weights = rng.uniform(0.0, 1.0, nb) ** 3
    weights[rng.choice(nb, nb // 10, replace=False)] = 0.0
    ctx = M.Context(0)
    ctx.build_tables([scenes.AIR])
    n = 2_000_000
    h = sampled_counts(ctx, energies, weights, n, kernel)
    ctx.close()
    # 1 cm of air removes at most ~0.1 % of the primaries (at 20 keV), far
    # below the counts' statistical error per bin
    p = weights / weights.sum()
    tot = h.sum()
    assert tot > 0.99 * n
    assert np.all(h[p == 0.0] == 0)
    keep = p * tot >= 10
    exp = p[keep] * tot
    chi2 = float(np.sum((h[keep] - exp) ** 2 / exp))
    pval = stats.chi2.sf(chi2, df=int(keep.sum()) - 1)
    assert pval > 0.01, (pval, chi2, keep.sum())


def test_device_source_directions_closed_forms():
    """The device's generate() (source.cpp:64-115) against the closed forms
    the oracle's samplers are pinned to: isotropic theta ~ U[0, pi] and phi
    ~ U[0, 2 pi] (Q4); disc aim point r ~ U[0, R] and angle uniform (Q5);
    rectangle aim point a1, a2 ~ U[0, 1]; beam along the pass-through
    point.  Directions are unit vectors to FP32 rounding."""
    from helpers import disc_basis, uniform_chi2_p
    ctx = M.Context(0)
    ctx.build_tables([scenes.AIR])
    ctx.set_domain((40.0, 40.0, 40.0), scenes.AIR)
    n = 400000
    # isotropic
    ctx.set_source(direction=M.DIR_ISOTROPIC, position=(20, 20, 20), energy=50e3)
    d, e = ctx.sample_source(n, seed=3)
    d = d.astype(np.float64)
    assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1) < 1e-6) and np.all(e == np.float32(50e3))
    assert uniform_chi2_p(np.arccos(np.clip(d[:, 2], -1, 1)), 0, np.pi) > 1e-3
    assert uniform_chi2_p(np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi), 0, 2 * np.pi) > 1e-3
    # disc
    P, C, nrm, R = np.array([20.0, 20.0, 1.0]), np.array([20.5, 19.5, 11.0]), np.array([0.0, 0.0, 1.0]), 0.7
    ctx.set_source(direction=M.DIR_DISC_ISOTROPIC, position=P, a=C, b=nrm, radius=R, energy=30e3)
    d, _ = ctx.sample_source(n, seed=4)
    d = d.astype(np.float64)
    u, v = disc_basis(nrm)
    X = P + (((C - P) @ nrm) / (d @ nrm))[:, None] * d
    r = np.linalg.norm(X - C, axis=1)
    assert np.all(r <= R * (1 + 1e-5))
    assert uniform_chi2_p(np.minimum(r, R), 0, R) > 1e-3
    assert uniform_chi2_p(np.mod(np.arctan2((X - C) @ v, (X - C) @ u), 2 * np.pi), 0, 2 * np.pi) > 1e-3
    # rectangle
    P = np.array([21.0, 22.0, 30.0])
    corner, e1, e2 = np.array([18.0, 18.0, 1.0]), np.array([4.0, 0.0, 0.0]), np.array([0.0, 3.0, 0.0])
    ctx.set_source(direction=M.DIR_RECT_ISOTROPIC, position=P, a=corner, b=e1, c=e2, energy=30e3)
    d, _ = ctx.sample_source(n, seed=5)
    d = d.astype(np.float64)
    X = P + ((corner[2] - P[2]) / d[:, 2])[:, None] * d
    a1, a2 = (X - corner) @ e1 / 16.0, (X - corner) @ e2 / 9.0
    assert uniform_chi2_p(np.clip(a1, 0, 1), 0, 1) > 1e-3 and uniform_chi2_p(np.clip(a2, 0, 1), 0, 1) > 1e-3
    # beam
    ctx.set_source(direction=M.DIR_BEAM, position=(20, 20, 0), a=(23, 24, 20), energy=30e3)
    d, _ = ctx.sample_source(1000, seed=6)
    want = np.array([3.0, 4.0, 20.0]) / np.linalg.norm([3.0, 4.0, 20.0])
    assert np.allclose(d, want, atol=1e-6)
    ctx.close()
