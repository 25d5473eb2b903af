"""The library's host FP64 table builder (the tables the device tables are
packed from) against the oracle's independent restatement (numpy + scipy's
banded B-spline solver).  CPU only: a host-only context (device -1)."""
import numpy as np
import pytest

import midsx_amd as M
from oracle import tables as T

MATERIALS = (3, 48, 169, 168)
KINDS = [(M.XS_TOTAL, "total"), (M.XS_COHERENT, "coh"), (M.XS_INCOHERENT, "incoh"),
         (M.XS_PHOTOELECTRIC, "photo"), (M.SF_INCOHERENT, "sf"), (M.FF_COHERENT, "ff")]


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(-1)
    c.build_tables(list(MATERIALS))
    yield c
    c.close()


@pytest.fixture(scope="module")
def omats():
    return {m: T.Material(m, rita=True) for m in MATERIALS}


@pytest.mark.parametrize("mid", MATERIALS)
def test_table_nodes_match_oracle(ctx, omats, mid):
    o = omats[mid]
    for which, name in KINDS:
        x, y = ctx.table(mid, which)
        ox, oy = getattr(o, name)
        assert np.array_equal(x, ox), name
        tol = 1e-11 if name in ("coh", "incoh", "total") else 1e-13
        assert np.max(np.abs(y / oy - 1)) < tol, name


@pytest.mark.parametrize("mid", MATERIALS)
def test_interpolants_match_oracle(ctx, omats, mid):
    o = omats[mid]
    E = np.geomspace(1.0, 1e11, 5001)
    for which, f in [(M.XS_TOTAL, o.total_i), (M.XS_COHERENT, o.coh_i), (M.XS_INCOHERENT, o.incoh_i),
                     (M.XS_PHOTOELECTRIC, o.photo_i)]:
        assert np.max(np.abs(ctx.interpolate_host(mid, which, E) / f(E) - 1)) < 1e-10
    x = np.geomspace(1e-6, 50.0, 2001)
    assert np.max(np.abs(ctx.interpolate_host(mid, M.SF_INCOHERENT, x) / o.sf_i(x) - 1)) < 1e-13
    assert np.max(np.abs(ctx.interpolate_host(mid, M.FF_COHERENT, x) / o.ff_i(x) - 1)) < 1e-13


def test_majorant_matches_oracle(ctx, omats):
    e, v = T.majorant([omats[m] for m in MATERIALS])
    x, y = ctx.table(-1, M.XS_TOTAL)
    assert np.array_equal(x, e)
    assert np.max(np.abs(y / v - 1)) < 1e-11


@pytest.mark.parametrize("mid", (48, 169))
def test_rita_rows_match_oracle(ctx, omats, mid):
    """Same adaptive grid at (almost) every energy; rows agree to rounding."""
    r = omats[mid].rita
    same = 0
    for ie in range(len(r.energies)):
        e, x, y, a, b = ctx.rita_row(mid, ie)
        ox, oy, oa, ob = r.rows[ie]
        assert e == r.energies[ie]
        if len(x) != len(ox):
            continue
        same += 1
        assert np.allclose(x, ox, atol=1e-15)
        assert np.allclose(y, oy, rtol=1e-11, atol=1e-14)
        assert np.allclose(a, oa, rtol=1e-8, atol=1e-9)
        assert np.allclose(b, ob, rtol=1e-8, atol=1e-9)
    assert same >= len(r.energies) - 2


def test_rita_sampling_matches_oracle(ctx, omats):
    r = omats[48].rita
    rng = np.random.default_rng(1)
    E = rng.uniform(15e3, 150e3, 3000)
    y = rng.random(3000)
    got = ctx.sample_rayleigh_host(48, E, y)
    ref = np.array([r.sample(e, u) for e, u in zip(E, y)])
    assert np.max(np.abs(got - ref)) < 1e-9


def test_packing_is_exact_on_host():
    """The FP32 device records are packed so that a*exp(c1 t + ...) reproduces
    the FP64 interpolant at the rounded nodes; checked here through the
    registry path of the library (no device)."""
    c = M.Context(-1)
    assert c.material_id("Water, Liquid") == 48
    assert c.material_name(169) == "Cortical Bone - C5"
    with pytest.raises(M.MdxError):
        c.material_id("Unobtainium")
    c.close()


def test_rita_energy_bracket_clamped_at_grid_ends():
    """An energy at or below the RITA grid's first energy (or above its last)
    has no lower_bound - 1 bracket: the reference reads outside its arrays
    there (probability_dist.cpp:76); the host sampler, the device and the
    oracle extend the end interval instead.  At the first grid energy the
    end interval's lower row is taken exactly (f = 0)."""
    import midsx_amd as M
    ctx = M.Context(-1)
    ctx.build_tables([48])
    e0, x0, y0, a0, b0 = ctx.rita_row(48, 0)
    y = np.array([0.1, 0.5, 0.9])
    first = ctx.sample_rayleigh_host(48, np.full(3, e0), y)
    below = ctx.sample_rayleigh_host(48, np.full(3, e0 * 0.999), y)
    above = ctx.sample_rayleigh_host(48, np.full(3, 1e12), y)
    assert np.all(np.isfinite(first)) and np.all(np.isfinite(below)) and np.all(np.isfinite(above))
    r = T.Material(48).rita
    assert np.allclose(first, [r.x_from_y(0, v) for v in y], rtol=0, atol=1e-12)
    assert np.allclose(below, [r.sample(e0 * 0.999, v) for v in y], rtol=0, atol=1e-9)
    ctx.close()
