"""Device building blocks against FP64 and FP32 references, through the C-ABI.

* voxel indexing, material lookup and interaction-type selection: bit-exact
  against the reference formulas evaluated in IEEE FP32 (numpy) on the same
  inputs;
* cross-section / scattering-function interpolation: within 2 FP32 ulp of the
  FP64 evaluation of the same interpolant;
* RITA inverse-CDF sampling: within 2 FP32 units (at max(|mu|, 1 + mu)) of
  the FP64 sampler on the same uniforms, every sample;
* the delta-scatter decision: bit-exact against IEEE FP32;
* the direction change of a scattering event: against the reference's FP64
  quaternion rotation within the hardware sine / cosine error;
* Kahn + S(x) Compton sampling: chi^2 against Klein-Nishina x S(x).
"""
import numpy as np
import pytest
from scipy import integrate, stats

import midsx_amd as M
import oracle
from oracle import tables as T

pytestmark = pytest.mark.gpu
MATS = (3, 48, 169, 168)


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    c.build_tables(list(MATS))
    yield c
    c.close()


def ulp_err(dev, ref):
    ref = np.asarray(ref, dtype=np.float64)
    spacing = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(np.asarray(dev, dtype=np.float64) - ref) / spacing


def energies(ctx, mid, which, lo=1.5e3, hi=1.5e5):
    E = list(np.geomspace(lo, hi, 20011))
    x, _ = ctx.table(mid, which) if mid >= 0 else ctx.table(-1, M.XS_TOTAL)
    x = x[(x > lo) & (x < hi)]
    E += list(x) + list(np.nextafter(x.astype(np.float32), np.float32(0))) + list(0.5 * (x[1:] + x[:-1]))
    return np.unique(np.asarray(E, dtype=np.float32))


@pytest.mark.parametrize("mid", MATS)
@pytest.mark.parametrize("which", [M.XS_TOTAL, M.XS_COHERENT, M.XS_INCOHERENT, M.XS_PHOTOELECTRIC])
def test_cross_sections_within_2_ulp(ctx, mid, which):
    E = energies(ctx, mid, which)
    dev = ctx.interpolate(mid, which, E)
    ref = ctx.interpolate_host(mid, which, E.astype(np.float64))
    u = ulp_err(dev, ref)
    assert u.max() <= 2.0, (u.max(), E[np.argmax(u)])


def test_majorant_within_2_ulp(ctx):
    E = energies(ctx, -1, M.XS_TOTAL)
    u = ulp_err(ctx.interpolate(-1, M.XS_TOTAL, E), ctx.interpolate_host(-1, M.XS_TOTAL, E.astype(np.float64)))
    assert u.max() <= 2.0


@pytest.mark.parametrize("mid", MATS)
def test_scattering_function_within_2_ulp(ctx, mid):
    x = np.unique(np.geomspace(1e-4, 12.0, 20011).astype(np.float32))
    u = ulp_err(ctx.interpolate(mid, M.SF_INCOHERENT, x),
                ctx.interpolate_host(mid, M.SF_INCOHERENT, x.astype(np.float64)))
    assert u.max() <= 2.0


def test_rita_tables_built_on_device_match_host(ctx):
    """The FP64 RITA tables built on the device (one wavefront per energy,
    the fixture's context) against the host build of the same tables: the
    same adaptive grids (grid points are dyadic, hence exact) at every energy
    and CDF / parameters within 1e-12."""
    host = M.Context(-1)
    host.build_tables(list(MATS))
    for mid in MATS:
        ie, same = 0, 0
        while True:
            try:
                e_d, x_d, y_d, a_d, b_d = ctx.rita_row(mid, ie)
            except M.MdxError:
                break
            e_h, x_h, y_h, a_h, b_h = host.rita_row(mid, ie)
            assert e_d == e_h
            if x_d.size == x_h.size and np.array_equal(x_d, x_h):
                same += 1
                assert np.allclose(y_d, y_h, rtol=1e-12, atol=1e-15), (mid, ie)
                assert np.allclose(a_d, a_h, rtol=1e-9, atol=1e-12), (mid, ie)
                assert np.allclose(b_d, b_h, rtol=1e-9, atol=1e-12), (mid, ie)
            ie += 1
        assert ie > 300 and same >= ie - 2, (mid, same, ie)
    host.close()


def rita_units(dev, ref):
    """|dev - ref| in FP32 spacings at max(|mu|, 1 + mu): the reference forms
    mu as w - 1 with w = 1 + mu the log-log interpolant (probability_dist.cpp:
    86), so 2 ulp of mu times that subtraction's condition number
    max(|mu|, 1 + mu) / |mu| is 2 ulp of its larger operand."""
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.maximum(np.abs(ref), 1.0 + ref).astype(np.float32)
    return np.abs(np.asarray(dev, dtype=np.float64) - ref) / np.spacing(scale).astype(np.float64)


@pytest.mark.parametrize("mid", (48, 169, 3, 168))
def test_rita_device_within_2_ulp_of_fp64(ctx, mid):
    """ContinuousInversion::sample (probability_dist.cpp:73-119) on the device
    against the FP64 evaluation of the same tables on the same (E, y), every
    sample within 2 FP32 units (rita_units): 2e5 random pairs in 15-150 keV,
    plus, on every third row of the grid in 1-500 keV, y on every CDF node
    (the FP32 value nearest to it and its neighbours: the bracket must be the
    FP64 lower_bound's) at the row's energy and its FP32 neighbours; and
    finite values on rows up to 100 MeV, whose CDF widths reach 1e-26."""
    rng = np.random.default_rng(4 + mid)
    E = list(rng.uniform(15e3, 150e3, 200000))
    y = list(rng.random(200000))
    for ie in range(0, 1000, 3):
        try:
            e, x, yy, a, b = ctx.rita_row(mid, ie)
        except M.MdxError:
            break
        if not 1e3 <= e <= 5e5:  # the transport range (tables packed over 1-500 keV)
            continue
        e32 = np.float32(e)
        for ee in (np.nextafter(e32, np.float32(0)), e32, np.nextafter(e32, np.float32(1e9))):
            for yk in yy[1:]:
                y32 = np.float32(yk)
                for yv in (np.nextafter(y32, np.float32(0)), y32, np.nextafter(y32, np.float32(2))):
                    if 0 < yv < 1:
                        E.append(float(ee))
                        y.append(float(yv))
    E = np.asarray(E, dtype=np.float32)
    y = np.asarray(y, dtype=np.float32)
    dev = ctx.sample_rayleigh(mid, E, y).astype(np.float64)
    ref = ctx.sample_rayleigh_host(mid, E.astype(np.float64), y.astype(np.float64))
    u = rita_units(dev, ref)
    k = int(np.argmax(u))
    assert u.max() <= 2.0, (u.max(), float(E[k]), float(y[k]), dev[k], ref[k])
    assert np.all(dev >= -1.0) and np.all(dev <= 1.0)
    hiE, hiy = [], []  # high-energy rows: tiny CDF widths must not underflow
    for ie in range(0, 2000, 50):
        try:
            e, x, yy, a, b = ctx.rita_row(mid, ie)
        except M.MdxError:
            break
        hiE += [e] * (len(yy) - 1)
        hiy += list(np.float32(yy[1:]))
    dev = ctx.sample_rayleigh(mid, np.asarray(hiE, np.float32), np.clip(np.asarray(hiy, np.float32), 1e-30, 1))
    assert np.all(np.isfinite(dev)) and np.all(np.abs(dev) <= 1.0)


def test_voxel_lookup_bit_exact(ctx):
    """computational_domain.cpp:23-31, voxel_grid.cpp:55-59,119-128 on FP32
    inputs: truncation of the IEEE quotient, first grid wins, background
    outside; the far face maps to the last voxel."""
    ctx.set_domain((10.0, 10.0, 10.0), 3)
    rng = np.random.default_rng(5)
    g0 = rng.choice(np.array([48, 169, 168], np.uint8), size=(13, 7, 11))
    g1 = rng.choice(np.array([48, 169], np.uint8), size=(5, 9, 4))
    grids = [(g0, (0.3, 0.7, 0.11), (1.0, 2.0, 3.0)), (g1, (0.5, 0.25, 0.9), (2.0, 1.0, 2.5))]
    for ids, sp, org in grids:
        ctx.add_grid(ids, sp, org)
    ctx.build_tables(list(MATS))
    pts = [rng.uniform(0, 10, size=(50000, 3))]
    for ids, sp, org in grids:  # faces, corners and neighbours of every face
        dims = np.array(ids.shape, dtype=np.float64) * np.array(sp)
        corner = np.array(org) + dims
        for base in (np.array(org), corner):
            q = np.tile(base, (2000, 1)) + rng.uniform(-1, 1, size=(2000, 3)) * 0.5
            ax = rng.integers(0, 3, 2000)
            q[np.arange(2000), ax] = base[ax]
            pts.append(q)
            pts.append(np.nextafter(q.astype(np.float32), np.float32(np.inf)))
            pts.append(np.nextafter(q.astype(np.float32), np.float32(-np.inf)))
        k = np.arange(ids.shape[0] + 1)[:, None] * np.array([sp[0], 0, 0]) + np.array(org)
        pts.append(k + np.array([0, sp[1] * 0.5, sp[2] * 0.5]))
    P = np.concatenate(pts).astype(np.float32)
    g, lin, mat = ctx.voxel_lookup(P)
    # FP32 restatement of the same formula
    eg = np.full(len(P), -1)
    elin = np.full(len(P), -1, dtype=np.int64)
    emat = np.full(len(P), 3)
    for gi in reversed(range(len(grids))):
        ids, sp, org = grids[gi]
        n = np.array(ids.shape)
        dimf = (n * np.array(sp)).astype(np.float32)
        q = P - np.array(org, dtype=np.float32)
        inside = np.all((q >= 0) & (q <= dimf), axis=1)
        idx = np.trunc(q / np.array(sp, dtype=np.float32)).astype(np.int64)
        idx = np.minimum(idx, n - 1)
        l = idx[:, 0] + n[0] * (idx[:, 1] + n[1] * idx[:, 2])
        eg[inside] = gi
        elin[inside] = l[inside]
        flat = np.ascontiguousarray(ids.transpose(2, 1, 0)).ravel()
        emat[inside] = flat[np.clip(l[inside], 0, flat.size - 1)]
    assert np.array_equal(g, eg)
    assert np.array_equal(lin, elin)
    assert np.array_equal(mat, emat)


def test_interaction_type_bit_exact(ctx):
    """physics_engine.cpp:82-109 on FP32 inputs, including u equal to the
    thresholds."""
    rng = np.random.default_rng(6)
    n = 400000
    st = rng.uniform(0.01, 10.0, n).astype(np.float32)
    sc = (st * rng.uniform(0, 0.3, n)).astype(np.float32)
    si = (st * rng.uniform(0, 0.7, n)).astype(np.float32)
    u = rng.random(n).astype(np.float32)
    pc = sc / st
    pi = si / st
    u[: n // 4] = pc[: n // 4]
    u[n // 4: n // 2] = (pc + pi)[n // 4: n // 2]
    u[n // 2: n // 2 + 1000] = np.nextafter(pc[n // 2: n // 2 + 1000], np.float32(0))
    expect = np.where(u < pc, 1, np.where(u < pc + pi, 2, 0))
    got = ctx.select_interaction(sc, si, st, u)
    assert np.array_equal(got, expect)


def test_delta_decision_bit_exact(ctx):
    """physics_engine.cpp:180-183 as both transport kernels take it: real
    collision unless u < 1 - sigma_tot / sigma_max, the quotient and the
    difference in IEEE FP32, with the uniforms the device draws (24-bit
    grid points in (0, 1]) and u placed exactly on the threshold and one
    FP32 step either side of it; a step cut at a local majorant's bound
    takes no collision test."""
    rng = np.random.default_rng(9)
    n = 400000
    sm = np.exp(rng.uniform(np.log(0.05), np.log(200.0), n)).astype(np.float32)
    st = (sm * rng.uniform(0.0, 1.0, n)).astype(np.float32)
    st[:1000] = sm[:1000]  # sigma_tot equal to the majorant: never a delta collision
    u = ((rng.integers(0, 1 << 24, n) + 1) / float(1 << 24)).astype(np.float32)
    thr = np.float32(1.0) - st / sm  # numpy float32: correctly rounded quotient and difference
    assert thr.dtype == np.float32
    q = n // 8
    u[q:2 * q] = thr[q:2 * q]
    u[2 * q:3 * q] = np.nextafter(thr[2 * q:3 * q], np.float32(0))
    u[3 * q:4 * q] = np.nextafter(thr[3 * q:4 * q], np.float32(2))
    limited = np.zeros(n, np.int32)
    limited[4 * q:5 * q] = 1
    expect = ((~(u < thr)) & (limited == 0)).astype(np.int32)
    got = ctx.delta_decision(st, sm, u, limited)
    assert np.array_equal(got, expect), np.nonzero(got != expect)[0][:10]
    assert got[q:2 * q].all() and not got[2 * q:3 * q].any()  # u == thr is real; one step below is delta
    assert not got[4 * q:5 * q].any() and got[:1000].all()
    # against the FP64 decision on the same FP32 inputs: they differ only
    # where u lies within the FP32 rounding of the threshold
    thr64 = 1.0 - st.astype(np.float64) / sm.astype(np.float64)
    rnd = slice(4 * q + q, n)
    real64 = ~(u[rnd].astype(np.float64) < thr64[rnd])
    differ = real64 != got[rnd].astype(bool)
    assert np.all(np.abs(u[rnd][differ] - thr64[rnd][differ]) <= 2 * np.spacing(np.float32(1.0)))


def test_rotation_matches_quaternion_form(ctx):
    """particle.cpp:7-45 on the device (the closed form of the reference's
    quaternion product, hardware sine and cosine for the azimuth) against the
    oracle's FP64 quaternion rotation of the same FP32 inputs: random
    directions, axis-aligned ones and ties between |d_i| (the rotation axis
    comes from the smallest component, x before y before z), polar cosines
    including +-1.  Every component within 4e-6 (the hardware sine and cosine
    are accurate to a few 1e-7 over [0, 2 pi)); |d'| = 1 and d.d' = mu."""
    rng = np.random.default_rng(17)
    d = rng.normal(size=(3000, 3))
    special = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, 0, -1), (1, 1, 1), (1, 1, 0), (0, 1, 1),
               (1, 0, 1), (-1, 1, 1), (1, -1, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2), (3, -3, 1)]
    d[:len(special)] = special
    d /= np.linalg.norm(d, axis=1)[:, None]
    d = d.astype(np.float32)
    n = len(d)
    mu = rng.uniform(-1, 1, n).astype(np.float32)
    mu[::7] = 1.0
    mu[3::7] = -1.0
    phi = rng.uniform(0, 2 * np.pi, n).astype(np.float32)
    phi[5::11] = 0.0
    got = ctx.rotate(d, mu, phi).astype(np.float64)
    ref = np.array([oracle.rotate(d[i].astype(np.float64), float(np.arccos(np.float64(mu[i]))), float(phi[i]))
                    for i in range(n)])
    err = np.abs(got - ref).max(axis=1)
    assert err.max() < 4e-6, (err.max(), np.argmax(err))
    assert np.abs(np.linalg.norm(got, axis=1) - 1).max() < 1e-6
    assert np.abs(np.sum(got * d.astype(np.float64), axis=1) - mu).max() < 4e-6


def kn_s(m, E, mu):
    k = E / T.ELECTRON_REST_MASS
    r = 1.0 / (1 + k * (1 - mu))
    x = T.ALPHA * k * np.sqrt(np.maximum(1 - mu, 0))
    return r * r * (r + 1 / r - (1 - mu * mu)) * m.sf_i(np.maximum(x, 1e-12))


@pytest.mark.parametrize("E", [20e3, 60e3, 120e3])
def test_compton_sampling_chi2(ctx, E):
    m = oracle.material(48)
    n = 400000
    mu, ep = ctx.sample_compton(48, np.full(n, E, np.float32), seed=int(E))
    k = E / T.ELECTRON_REST_MASS
    assert np.allclose(ep, E / (1 + k * (1 - mu.astype(np.float64))), rtol=2e-6)
    edges = np.linspace(-1, 1, 51)
    counts, _ = np.histogram(mu, edges)
    f = lambda q: kn_s(m, E, np.atleast_1d(q))[0]
    norm = integrate.quad(f, -1, 1, limit=400)[0]
    expect = np.array([integrate.quad(f, edges[i], edges[i + 1])[0] for i in range(50)]) / norm * n
    chi2 = np.sum((counts - expect) ** 2 / expect)
    assert stats.chi2.sf(chi2, 49) > 0.01, chi2


def test_atomic_probe_counts_every_deposit(ctx):
    """The atomic-rate probe: both of its launches land every deposit (the
    records' counts sum to 2 x deposits) and it is timed; bad sizes fail."""
    for records, deposits in ((1, 1000), (4096, 300_000), (1 << 20, 1_000_003)):
        ms, n = ctx.atomic_probe(records, deposits, check=True)
        assert n == 2 * deposits, (records, deposits, n)
        assert ms > 0.0
    with pytest.raises(M.MdxError):
        ctx.atomic_probe(0, 10)
