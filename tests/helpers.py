"""Statistics helpers shared by the parity tests, and the brute-force check
of the macro-cell table shared by the host and device tests."""
import math

import numpy as np
import pytest
from scipy import stats

import midsx_amd as M

# the event kernels MDX_KERNEL_QUEUE runs: (kind, bank size) for
# Context.set_event_kernel
EVENT_FORMS = [pytest.param((M.EVENT_QUEUE, 128), id="queue"),
               pytest.param((M.EVENT_BANK, 96), id="bank96"),
               pytest.param((M.EVENT_BANK, 128), id="bank128"),
               pytest.param((M.EVENT_BANK, 160), id="bank160")]


def batch_mean_sigma(samples):
    """Mean of per-batch values and the standard error of that mean."""
    a = np.asarray(samples, dtype=np.float64)
    k = a.shape[0]
    return a.mean(axis=0), a.std(axis=0, ddof=1) / math.sqrt(k)


def zscore(m1, s1, m2, s2):
    s = np.sqrt(np.asarray(s1) ** 2 + np.asarray(s2) ** 2)
    return np.where(s > 0, (np.asarray(m1) - np.asarray(m2)) / np.where(s > 0, s, 1), 0.0)


def chi2_pvalue(z):
    z = np.asarray(z, dtype=np.float64).ravel()
    return float(stats.chi2.sf(np.sum(z * z), df=len(z))), float(np.sum(z * z)), len(z)


def count_chi2(c1, n1, c2, n2, min_count=10):
    """Two-sample chi^2 of histograms of counts (Poisson), scaled to the
    number of histories of each sample; returns (p, chi2, dof)."""
    c1 = np.asarray(c1, dtype=np.float64).ravel()
    c2 = np.asarray(c2, dtype=np.float64).ravel()
    keep = (c1 + c2) >= min_count
    c1, c2 = c1[keep], c2[keep]
    r = n1 / n2
    z = (c1 - r * c2) / np.sqrt(c1 + r * r * c2)
    return chi2_pvalue(z)


def welch_normal_scores(m1, s1, k1, m2, s2, k2):
    """Per-bin Welch t statistics of two batch means (k1, k2 batches), mapped
    to standard-normal scores through their two-sided p-values so that a
    chi^2 over bins is not inflated by the noise of the estimated sigmas."""
    m1, s1, m2, s2 = (np.asarray(a, dtype=np.float64) for a in (m1, s1, m2, s2))
    v1, v2 = s1 ** 2, s2 ** 2
    se = np.sqrt(v1 + v2)
    t = np.where(se > 0, (m1 - m2) / np.where(se > 0, se, 1), 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        df = (v1 + v2) ** 2 / (v1 ** 2 / (k1 - 1) + v2 ** 2 / (k2 - 1))
    df = np.where(np.isfinite(df) & (df > 0), df, k1 + k2 - 2)
    p = 2 * stats.t.sf(np.abs(t), df)
    return np.sign(t) * stats.norm.isf(np.clip(p, 1e-300, 1) / 2)


def hotelling(a, b, rel_eig=1e-9):
    """Two-sample Hotelling T^2 test of equal mean vectors for per-batch
    samples a (k1 x d) and b (k2 x d) with a pooled covariance; correlated
    and redundant components are handled by whitening on the non-null
    eigenvectors.  Returns (p, T2, rank)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    k1, k2 = len(a), len(b)
    scale = np.maximum(np.abs(np.concatenate([a, b])).mean(axis=0), 1e-300)
    a, b = a / scale, b / scale
    diff = a.mean(axis=0) - b.mean(axis=0)
    S = ((k1 - 1) * np.cov(a, rowvar=False) + (k2 - 1) * np.cov(b, rowvar=False)) / (k1 + k2 - 2)
    S = np.atleast_2d(S) * (1.0 / k1 + 1.0 / k2)
    w, V = np.linalg.eigh(S)
    keep = w > rel_eig * w.max()
    p_dim = int(keep.sum())
    proj = V[:, keep].T @ diff
    t2 = float(np.sum(proj * proj / w[keep]))
    dof2 = k1 + k2 - p_dim - 1
    f = dof2 / (p_dim * (k1 + k2 - 2)) * t2
    return float(stats.f.sf(f, p_dim, dof2)), t2, p_dim


def check_macro_table(ctx, ids, materials, background):
    """The macro-cell table against a brute-force restatement: each cell's
    class is a material present in it that bounds the total cross section of
    every material within its radius (host FP64 tables, 400 energies over
    1-500 keV), cells outside the grid hold the background material, and the
    radius cannot grow by one (up to its cap of 255); a cell is flagged
    uniform exactly when it holds one material, which is then its class."""
    uniq = list(dict.fromkeys(materials))  # slot order of build_tables
    energies = np.geomspace(1e3, 5e5, 400)
    sig = np.array([ctx.interpolate_host(m, M.XS_TOTAL, energies) for m in uniq])
    bounds = np.all(sig[:, None, :] >= sig[None, :, :] * (1 + 1e-6), axis=2) | np.eye(len(uniq), dtype=bool)
    slot_of = {m: s for s, m in enumerate(uniq)}
    cls, rad, uni = ctx.macro_cells()
    cz, cy, cx = cls.shape
    dims = np.array([cz, cy, cx])
    e = 8
    present = np.zeros((cz, cy, cx, len(uniq)), bool)
    for k in range(cz):
        for j in range(cy):
            for i in range(cx):
                for m in np.unique(ids[i * e:(i + 1) * e, j * e:(j + 1) * e, k * e:(k + 1) * e]):
                    present[k, j, i, slot_of[int(m)]] = True

    def held(lo, hi):  # materials in the cells [lo, hi), the background outside the grid
        a, b = np.maximum(lo, 0), np.minimum(hi, dims)
        m = present[a[0]:b[0], a[1]:b[1], a[2]:b[2]].reshape(-1, len(uniq)).any(axis=0)
        if (lo < 0).any() or (hi > dims).any():
            m[slot_of[background]] = True
        return m

    local = 0
    for k in range(cz):
        for j in range(cy):
            for i in range(cx):
                c, r = cls[k, j, i], rad[k, j, i]
                here = present[k, j, i]
                only = np.nonzero(here)[0]
                assert uni[k, j, i] == (len(only) == 1 and c != 255), (k, j, i)
                if uni[k, j, i]:
                    assert only[0] == c
                if c == 255:
                    assert not any(bounds[s][here].all() for s in only)
                    continue
                assert here[c] and bounds[c][here].all()
                if r == 0:
                    continue
                local += 1
                lo = np.array([k, j, i]) - r
                hi = np.array([k, j, i]) + r + 1
                assert bounds[c][held(lo, hi)].all(), (k, j, i, r)
                if r < 255:  # one more cell would hold a material c does not bound
                    assert not bounds[c][held(lo - 1, hi + 1)].all(), (k, j, i, r)
    return local


def uniform_chi2_p(x, lo, hi, bins=40):
    """p-value of a chi^2 test that x is uniform on [lo, hi] (bins equal bins)."""
    counts, _ = np.histogram(x, np.linspace(lo, hi, bins + 1))
    e = len(x) / bins
    return float(stats.chi2.sf(np.sum((counts - e) ** 2 / e), bins - 1))


def disc_basis(normal):
    """DiscIsotropicDirectionality::setUAndV (source.cpp:117-128): u = n x x^
    (n x y^ when n is perpendicular to x^), v = n x u, both normalised."""
    n = np.asarray(normal, float)
    x, y = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    u = np.cross(n, y) if abs(np.dot(n, x)) < 1e-9 else np.cross(n, x)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    return u, v / np.linalg.norm(v)


def binwise(g, o, events, min_events=200, screen_alpha=1e-3):
    """Every bin of one tally, two codes compared: g (k_g x n) and o (k_o x n)
    per-batch values, events the events behind each bin summed over both
    sides' batches (a selection symmetric in the two samples).  Bins with at
    least min_events stand alone; the others are summed into one pooled bin,
    so that every bin takes part.  Per bin a Welch t of the two batch means, mapped to a
    normal score (welch_normal_scores); the tally's chi^2 = sum of the
    squared scores over its bins, with as many degrees of freedom (the bins
    of a voxel-dose, image or spectrum tally are close to independent: a
    history scores in few of them).  The screen at 3 combined standard
    errors takes the number of bins: the count of bins beyond 3 sigma may
    not exceed the (1 - screen_alpha) quantile of Binomial(n, P(|Z| > 3)),
    and no bin may lie beyond the Sidak bound of family-wise level
    screen_alpha.  Returns a dict; `ok` holds when p > 0.01 and the screen
    passes."""
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    alone = np.asarray(events, dtype=np.float64).ravel() >= min_events
    G, O = [g[:, alone]], [o[:, alone]]
    rest = ~alone
    if rest.any():
        G.append(g[:, rest].sum(axis=1, keepdims=True))
        O.append(o[:, rest].sum(axis=1, keepdims=True))
    G, O = np.concatenate(G, axis=1), np.concatenate(O, axis=1)
    gm, gs = batch_mean_sigma(G)
    om, os_ = batch_mean_sigma(O)
    live = (gs > 0) | (os_ > 0)  # a bin empty on both sides carries no test
    z = welch_normal_scores(gm, gs, len(G), om, os_, len(O))[live]
    n = len(z)
    chi2 = float(np.sum(z * z))
    p = float(stats.chi2.sf(chi2, n))
    over3 = int(np.sum(np.abs(z) > 3.0))
    allow3 = int(stats.binom.ppf(1 - screen_alpha, n, 2 * stats.norm.sf(3.0)))
    zmax = float(np.max(np.abs(z))) if n else 0.0
    zcap = float(stats.norm.isf(-np.expm1(np.log1p(-screen_alpha) / n) / 2)) if n else np.inf
    return dict(ok=bool(p > 0.01 and over3 <= allow3 and zmax <= zcap), p=p, chi2=chi2, bins=n,
                standalone=int(alone.sum()), pooled=int(rest.sum()), over3=over3, allow3=allow3, zmax=zmax,
                zcap=zcap)
