"""The C-ABI boundary on the CPU: the library loads, exports every entry
point include/midsx_amd.h declares, maps the reference's exception classes to
status codes, and reads scenes (JSON + NIfTI-1) like ComputationalDomain."""
import os
import re

import numpy as np
import pytest

import midsx_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "scene")


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "midsx_amd.h")).read()
    return sorted(set(re.findall(r"\b(mdx_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported(lib):
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(M.EXPORTS)


def test_status_codes():
    c = M.Context(-1)
    with pytest.raises(M.MdxError) as e:
        c.build_tables([250])  # not in the registry: runtime_error
    assert e.value.status == M.ERR_RUNTIME
    with pytest.raises(M.MdxError) as e:
        c.run(10)  # host-only context
    assert e.value.status == M.ERR_INVALID
    with pytest.raises(M.MdxError) as e:
        c.set_source(direction=M.DIR_RECT_ISOTROPIC, b=(1, 0, 0), c=(1, 1, 0), energy=50e3)
    assert e.value.status == M.ERR_INVALID and "orthogonal" in str(e.value)
    with pytest.raises(M.MdxError) as e:
        c.interpolate_host(48, M.XS_TOTAL, [1e4])  # tables not built
    assert e.value.status == M.ERR_INVALID
    with pytest.raises(M.MdxError):
        M.Context(-1).load_scene(os.path.join(GOLDEN, "missing.json"))
    c.close()


def test_load_scene_hvl():
    """cpp_simulations/half_value_layer/hvl.json with the 30 keV HVL Al slab
    (data/voxels/al_hvl_30keV.nii.gz: 1 x 1 x 2273 voxels of 40 x 40 x 0.001 mm,
    material 61)."""
    c = M.Context(-1)
    c.load_scene(os.path.join(GOLDEN, "hvl_30keV.json"))
    assert c.num_grids() == 1
    dims, sp, org = c.grid_info(0)
    assert dims == (1, 1, 2273)
    assert np.allclose(sp, (4.0, 4.0, 1e-4), rtol=1e-6)
    assert org == (0.0, 0.0, 10.0)
    ids = c.grid_ids(0)
    assert ids.shape == (1, 1, 2273) and np.all(ids == 61)
    c.build_tables()  # background + grid materials (computational_domain.cpp:109-123)
    x, _ = c.table(61, M.XS_TOTAL)
    assert len(x) > 100
    c.close()


def test_load_scene_first_grid_order():
    c = M.Context(-1)
    c.load_scene(os.path.join(GOLDEN, "lead_blocks.json"))
    assert c.num_grids() == 2
    assert c.grid_info(1)[2] == (0.2, 5.2, 0.0)
    assert np.allclose(c.grid_info(0)[1], (8.2, 2.5, 8.2), rtol=1e-6)
    assert np.all(c.grid_ids(0) == 130)
    c.close()


def test_scene_rejects_bad_files(tmp_path):
    p = tmp_path / "scene.json"
    p.write_text('{"dim_space": [1, 2], "background_material_name": "Air, Dry (near sea level)"}')
    with pytest.raises(M.MdxError):
        M.Context(-1).load_scene(str(p))
    p.write_text('{"dim_space": [1, 2, 3], "background_material_name": "Air, Dry (near sea level)",'
                 ' "voxel_grids": [{"file_path": "x.raw", "origin": [0, 0, 0]}]}')
    with pytest.raises(M.MdxError) as e:
        M.Context(-1).load_scene(str(p))
    assert "NIFTI" in str(e.value)
    q = tmp_path / "scene.txt"
    q.write_text("{}")
    with pytest.raises(M.MdxError) as e:
        M.Context(-1).load_scene(str(q))
    assert "JSON" in str(e.value)


def test_grid_layout_x_fastest():
    c = M.Context(-1)
    c.set_domain((1, 1, 1), 3)
    ids = np.zeros((3, 4, 5), np.uint8)
    ids[2, 1, 0] = 48
    ids[0, 3, 4] = 169
    c.add_grid(ids, (0.1, 0.1, 0.1), (0, 0, 0))
    back = c.grid_ids(0)
    assert np.array_equal(back, ids)
    c.close()


def test_job_alias_domain_host():
    """A job aliasing another's domain sees its grids; it cannot add its own."""
    c = M.Context(-1)
    c.set_domain((4.0, 4.0, 4.0), 3)
    ids = np.arange(2 * 3 * 4, dtype=np.uint8).reshape(2, 3, 4) % 3 + 48
    c.add_grid(ids, (0.5, 0.5, 0.5), (1.0, 1.0, 1.0))
    b = c.job_create()
    assert c.num_grids() == 0
    c.job_alias_domain(0)
    assert c.num_grids() == 1
    assert np.array_equal(c.grid_ids(0), ids)
    with pytest.raises(M.MdxError):
        c.add_grid(ids, (0.5, 0.5, 0.5), (1.0, 1.0, 1.0))
    with pytest.raises(M.MdxError):
        c.enable_voxel_dose(True)
    d = c.job_create()
    with pytest.raises(M.MdxError):
        c.job_alias_domain(d)  # itself
    c.job_alias_domain(b)  # alias of an alias resolves to the owner
    assert c.num_grids() == 1
    c.job_select(b)
    c.set_domain((2.0, 2.0, 2.0), 3)  # a fresh domain of its own
    assert c.num_grids() == 0
    c.job_select(d)
    assert c.num_grids() == 1
    with pytest.raises(M.MdxError):
        c.job_release(0)  # d still uses it
    c.job_release(d)
    c.job_release(0)
    c.close()


def test_ct_scene_host():
    """The CT set-up on a host-only context: view jobs alias the phantom job,
    each view's detector faces its source, and views cover the circle."""
    from midsx_amd import scenes
    c = M.Context(-1)
    geo, views, angles = scenes.setup_ct(c, n_views=4, shape=(16, 16, 8), voxel_cm=0.5, seed=1)
    assert len(views) == 4 and angles == [0.0, 90.0, 180.0, 270.0]
    dim = geo["dim"]
    for v, a in zip(views, angles):
        c.job_select(v)
        assert c.num_grids() == 1  # the phantom job's grid
        t = c.tallies[0]
        n = np.cross(t["p1"], t["p2"])
        th = np.radians(a)
        travel = -np.array([np.cos(th), np.sin(th), 0.0])
        assert np.dot(n, travel) > 0.99 * np.linalg.norm(n)
        # source and detector inside the domain
        src = np.array(geo["axis"]) - geo["sod"] * travel
        assert np.all(src > 0) and np.all(src < dim)
        corner = np.array(t["p0"])
        for q in (corner, corner + t["p1"], corner + t["p2"], corner + t["p1"] + t["p2"]):
            assert np.all(q > 0) and np.all(q < dim)
    h = scenes.ct_histories(360)
    assert h.min() >= 2e6 and h.max() <= 2e7 and 3e9 < h.sum() < 5e9
    c.close()


def test_air_muen_table_and_air_kerma_arithmetic():
    """The air coefficients at the table's nodes are the TG-195 values (the
    CSV is extracted verbatim), log-log between nodes; the air kerma sums
    E * fluence * mu_en over the spectrum's bins."""
    from midsx_amd import derived
    c = M.Context(-1)
    nist = np.loadtxt(os.path.join(os.path.dirname(M.__file__), "data", "air_muen_nist.csv"), delimiter=",",
                      skiprows=1)
    keep = np.r_[np.diff(nist[:, 0]) > 0, True]  # at an edge's energy the value above the edge
    got = c.air_muen(nist[keep, 0] * 1e3)
    assert np.allclose(got[1:], nist[keep, 1][1:], rtol=1e-14)
    # log-log between nodes: exact for a power law through two nodes
    e60, e80 = 60e3, 80e3
    v60, v80 = c.air_muen([e60, e80])
    mid = c.air_muen([70e3])[0]
    assert abs(mid - v60 * (v80 / v60) ** (np.log(70 / 60) / np.log(80 / 60))) < 1e-12 * mid
    assert c.air_muen([3.25e3])[0] == pytest.approx(146.0 * (76.36 / 146.0) ** (np.log(3.25 / 3.2029) / np.log(4 / 3.2029)))
    c.set_air_muen_table(1)
    tab = np.loadtxt(os.path.join(os.path.dirname(M.__file__), "data", "air_muen_tg195.csv"), delimiter=",",
                     skiprows=1)
    e = tab[:, 0] * 1e3
    assert np.array_equal(c.air_muen(e), tab[:, 1])
    c.set_air_muen_table(0)
    energies = np.arange(0.25, 30.0, 0.5) * 1e3
    t = derived.spectrum_disc_tally((2, 2, 100), (0, 0, 1), 1.0, energies)
    assert t["hist_bins"] == energies.size and t["hist_lo"] == 0.0 and t["hist_width"] == 500.0

    class Fake:
        tallies = [t]

        def tally(self, i):
            h = np.zeros((4, energies.size))
            h[0, 20] = 1000.0  # primaries in the 10.25 keV bin
            h[1, 21] = 77.0    # scattered ones do not count
            return {"hist": h}

        def air_muen(self, x):
            return c.air_muen(x)

    ak = derived.primary_spectrum_air_kerma(Fake(), 0, energies)
    assert abs(ak - 10.25e3 * 1000.0 / np.pi * c.air_muen([10.25e3])[0]) < 1e-9 * ak
    c.close()


def test_macro_table_host():
    """The macro-cell table (local majorants) built on a host-only context
    against the brute-force restatement the GPU tests also apply: class,
    uniform flag and safe radius of every cell of a radiography phantom in
    air, whose surface cells' class (water) bounds the background and so
    reaches past the grid's faces."""
    from helpers import check_macro_table
    from midsx_amd import scenes
    c = M.Context(-1)
    materials = list(scenes.C2_MATERIALS)
    c.build_tables(materials)
    shape = (64, 64, 48)
    scenes.setup_radiography(c, shape=shape, voxel_cm=0.5, seed=4, image=(16, 16))
    ids = scenes.ellipsoid_phantom(shape, 4, voxel_cm=0.5)
    assert check_macro_table(c, ids, materials, scenes.AIR) > 10
    cls, rad, uni = c.macro_cells()
    assert uni.mean() > 0.5  # most cells hold water only
    water = materials.index(scenes.WATER)
    corner = (cls[0, 0, 0], rad[0, 0, 0], uni[0, 0, 0])
    assert corner[0] == water and corner[2] and corner[1] > 0, corner  # not stopped by the grid's faces
    c.set_local_majorants(False)
    assert c.macro_cells() is None
    c.close()


def test_event_kernel_choices_host():
    """The event forms and bank sizes the library takes, on a host-only
    context (no device work): banks of 96, 128 or 160 photons; any other
    size or an unknown form is refused and leaves the setting unchanged."""
    c = M.Context(-1)
    for bs in M.BANK_SIZES:
        c.set_event_kernel(M.EVENT_BANK, bs)
        assert c.event_kernel() == (M.EVENT_BANK, bs)
    for bad in ((M.EVENT_BANK, 152), (M.EVENT_BANK, 64), (7, 128)):
        with pytest.raises(M.MdxError) as e:
            c.set_event_kernel(*bad)
        assert e.value.status == M.ERR_INVALID
        assert c.event_kernel() == (M.EVENT_BANK, M.BANK_SIZES[-1])
    assert M.BANK_MAX_HISTORY == 1 << 42
    c.close()


def test_queue_params_limits_host():
    """Launch parameters on a host-only context: blocks per CU, refill batch
    and interaction batch each in [1, 64] (a batch is at most one wave's 64
    lanes); a value outside is refused and leaves the setting unchanged."""
    c = M.Context(-1)
    for p in ((1, 1, 1), (64, 64, 64), (4, 64, 48)):
        c.set_queue_params(*p)
        assert c.queue_params() == p
    for bad in ((0, 64, 48), (65, 64, 48), (4, 0, 48), (4, 65, 48), (4, 64, 0), (4, 64, 65)):
        with pytest.raises(M.MdxError) as e:
            c.set_queue_params(*bad)
        assert e.value.status == M.ERR_INVALID
        assert c.queue_params() == (4, 64, 48)
    c.close()
