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
