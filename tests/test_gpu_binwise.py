"""Every tally bin against the oracle (the FP64 CPU restatement of the
reference's path), per tally: per-voxel deposited energy, per-pixel image
energy and counts, spectrum bins, each quantity x bucket of the scalar
tallies.  Both codes run independent seeded batches of the same scene; per
bin the batch means are compared by a Welch t mapped to a normal score.

* Voxel-dose, image and spectrum tallies (bins close to independent): a
  chi^2 over every bin (bins with fewer than 200 events on both sides
  together are summed into one pooled bin) must give p > 0.01, and the 3
  sigma screen must pass with its multiplicity rule: no more bins beyond 3
  combined standard errors than the 99.9th percentile of Binomial(bins,
  0.0027), none beyond the Sidak bound at family-wise level 1e-3
  (helpers.binwise).
* Quantity x bucket bins of a scalar tally are exact functions of each
  other in places (a monoenergetic primary's IncidentEnergy is E0 times its
  count), so their chi^2 is taken after whitening: Hotelling's T^2, p >
  0.01; the same 3 sigma screen applies to each bin.

Scenes: the slab of the plain method (40x40x10 water voxels, 60 keV pencil
beam), with the plain kernel and every event kernel of the queue launch;
the other scenes on every event kernel; the radiography workload at
32x32x24 voxels (cone beam, Kramers 15-80 keV, water / bone / adipose); one
CT view of a 32x32x16 phantom; a served dose job (16^3 phantom, 80 kVp)
through the request loop.
"""
import numpy as np
import pytest

import midsx_amd as M
import oracle
from midsx_amd import scenes
from midsx_amd.serve import Server
from helpers import EVENT_FORMS, binwise, hotelling

pytestmark = pytest.mark.gpu


def gpu_batches(ctx, k, n, seed, read, kernel=M.KERNEL_QUEUE):
    out = []
    for b in range(k):
        ctx.reset_tallies()
        ctx.run(n, seed=seed, first=b * n, kernel=kernel)
        out.append(read(ctx))
    return out


def orc_batches(sim, k, n, seed, read):
    out = []
    for b in range(k):
        sim.reset()
        sim.run(n, seed=seed + 7919 * b, threads=8)
        out.append(read(sim))
    return out


def stack(batches, key):
    return np.array([b[key] for b in batches])


def assert_binwise(name, g, o, key, events_key, min_standalone=20):
    r = binwise(stack(g, key), stack(o, key), stack(g, events_key).sum(0) + stack(o, events_key).sum(0))
    assert r["standalone"] >= min_standalone, (name, r)
    assert r["ok"], (name, r)
    return r


def assert_scalar_tally(name, g, o, key):
    """Quantity x bucket sums of one scalar tally: Hotelling T^2 and the 3
    sigma screen of binwise on each bin (no pooling: min_events 0)."""
    G, O = stack(g, key), stack(o, key)
    live = (G.std(axis=0) > 0) & (O.std(axis=0) > 0)
    p, t2, rank = hotelling(G[:, live], O[:, live])
    assert p > 0.01, (name, p, t2, rank)
    r = binwise(G[:, live], O[:, live], np.full(int(live.sum()), 1e9))
    assert r["over3"] <= r["allow3"] and r["zmax"] <= r["zcap"], (name, r)


# --------------------------------------------------------------------- slab (C1)
def slab_read(src):
    s, _, cnt = src.voxel_dose(0)
    v, d = src.tally(0), src.tally(1)
    return dict(dose=s.copy(), dose_n=cnt.astype(np.float64), vol=v["acc"][:, :, 1].ravel().copy(),
                vol_n=v["acc"][:, :, 0].ravel().copy(), surf=d["acc"][:, :, 1].ravel().copy(),
                spec=d["hist"].ravel().copy())


@pytest.fixture(scope="module")
def slab_oracle():
    sim = oracle.Simulation(scenes.C1_MATERIALS)
    scenes.setup_slab(sim)
    return orc_batches(sim, 40, 250000, 301, slab_read)


@pytest.mark.parametrize("kernel, events", [pytest.param(M.KERNEL_PLAIN, None, id="plain")] +
                         [pytest.param(M.KERNEL_QUEUE, e.values[0], id=e.id) for e in EVENT_FORMS])
def test_slab_every_bin(slab_oracle, kernel, events):
    ctx = M.Context(0)
    if events:
        ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.C1_MATERIALS))
    scenes.setup_slab(ctx)
    g = gpu_batches(ctx, 40, 250000, 71 + kernel + 10 * (events[0] if events else 0), slab_read, kernel)
    o = slab_oracle
    r = assert_binwise("slab voxel dose", g, o, "dose", "dose_n")
    assert r["standalone"] > 500  # the pencil beam's dose is on and near its axis
    # a 60 keV beam: the primaries fill one bin per bucket, single Compton
    # scatters a few around 50-60 keV
    assert_binwise("slab spectrum", g, o, "spec", "spec", min_standalone=5)
    assert_scalar_tally("slab volume tally", g, o, "vol")
    assert_scalar_tally("slab disc tally", g, o, "surf")
    ctx.close()


# --------------------------------------------------------------------- radiography (C2, reduced)
RAD = dict(shape=(32, 32, 24), voxel_cm=0.5, seed=7, image=(16, 16))


def image_read(src, t=0):
    im = src.tally(t)["image"]
    return dict(img_e=im[..., 0].ravel().copy(), img_n=im[..., 2].ravel().copy())


def rad_read(src):
    s, _, cnt = src.voxel_dose(0)
    out = dict(dose=s.copy(), dose_n=cnt.astype(np.float64))
    out.update(image_read(src))
    return out


@pytest.fixture(scope="module")
def rad_oracle():
    sim = oracle.Simulation(scenes.C2_MATERIALS)
    scenes.setup_radiography(sim, **RAD)
    return orc_batches(sim, 20, 100000, 401, rad_read)


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_radiography_every_bin(rad_oracle, events):
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    scenes.setup_radiography(ctx, **RAD)
    g = gpu_batches(ctx, 20, 100000, 81, rad_read)
    o = rad_oracle
    r = assert_binwise("radiography voxel dose", g, o, "dose", "dose_n")
    assert r["standalone"] > 3000
    assert_binwise("radiography image energy", g, o, "img_e", "img_n")
    assert_binwise("radiography image counts", g, o, "img_n", "img_n")
    ctx.close()


# --------------------------------------------------------------------- CT view (C4, reduced)
CT_SHAPE, CT_VOX = (32, 32, 16), 0.5


def ct_view_setup(sim):
    geo = scenes.ct_geometry(CT_SHAPE, CT_VOX)
    sim.set_domain(geo["dim"], scenes.AIR)
    sim.add_grid(scenes.ct_phantom(CT_SHAPE, CT_VOX, 4), (CT_VOX,) * 3, geo["origin"])
    sim.enable_voxel_dose(True)
    scenes.setup_ct_view(sim, geo, 30.0, pixels=(32, 8))


@pytest.fixture(scope="module")
def ct_oracle():
    sim = oracle.Simulation(scenes.CT_MATERIALS)
    ct_view_setup(sim)
    return orc_batches(sim, 20, 60000, 501, rad_read)


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_ct_view_every_bin(ct_oracle, events):
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.CT_MATERIALS))
    ct_view_setup(ctx)
    g = gpu_batches(ctx, 20, 60000, 91, rad_read)
    o = ct_oracle
    r = assert_binwise("CT voxel dose", g, o, "dose", "dose_n")
    assert r["standalone"] > 200
    assert_binwise("CT image energy", g, o, "img_e", "img_n")
    assert_binwise("CT image counts", g, o, "img_n", "img_n")
    ctx.close()


# --------------------------------------------------------------------- served dose job (C5, reduced)
@pytest.mark.parametrize("events", EVENT_FORMS)
def test_served_dose_job_every_bin(events):
    """24 copies of one dose job (different seeds) served through the
    request loop (FP64 dose read) against 24 oracle batches of the same job;
    the pooling uses the events of a separate GPU pilot run of the job."""
    job = dict(shape=(16, 16, 16), voxel_cm=0.25, phantom_seed=11, histories=100000, seed=1, image=(8, 8),
               kvp=80.0)
    k = 24
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.DOSE_MATERIALS))
    keep = lambda j, s, s2, image: (s.copy(), image.copy())
    srv = Server(ctx, round_histories=400_000, min_slice=20_000, on_result=keep, dose="f64")
    res = srv.run([dict(job, seed=1000 + b) for b in range(k)])
    g = [dict(dose=r.summary[0], img_e=r.summary[1][..., 0].ravel(), img_n=r.summary[1][..., 2].ravel())
         for r in res]
    sim = oracle.Simulation(scenes.DOSE_MATERIALS)
    scenes.setup_dose_job(sim, job)
    o = orc_batches(sim, k, job["histories"], 601,
                    lambda s: dict(dose=s.voxel_dose(0)[0].copy(), **image_read(s)))
    pilot = ctx.job_create()  # events per bin from an independent run (4x a batch per side)
    scenes.setup_dose_job(ctx, job)
    ctx.run(8 * job["histories"], seed=99, kernel=M.KERNEL_QUEUE)
    ev_dose = ctx.voxel_dose(0)[2].astype(np.float64) * k / 4
    ev_img = ctx.tally(0)["image"][..., 2].ravel() * k / 4
    ctx.job_release(pilot)
    r = binwise(stack(g, "dose"), stack(o, "dose"), ev_dose)
    assert r["ok"] and r["standalone"] > 500, ("served voxel dose", r)
    r = binwise(stack(g, "img_e"), stack(o, "img_e"), ev_img)
    assert r["ok"] and r["standalone"] >= 20, ("served image energy", r)
    ctx.close()
