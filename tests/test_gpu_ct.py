"""CT scan: views as jobs that share one phantom (domain aliasing), all views
in one multi-job launch.

* every view's image is the same in the one launch of all views as run
  alone, and the shared voxel dose is the sum over views (FP64 summation
  order only, 1e-12 / 1e-9 relative);
* one view agrees with the oracle (the same source, detector and phantom set
  up directly): Welch scores and Hotelling T^2 over 24 batches of detector
  blocks and phantom dose blocks.
"""
import numpy as np
import pytest

import midsx_amd as M
import oracle
from midsx_amd import scenes
from helpers import EVENT_FORMS, batch_mean_sigma, hotelling, welch_normal_scores

pytestmark = pytest.mark.gpu

SHAPE, VOX = (32, 32, 16), 0.5


@pytest.fixture(scope="module")
def ct():
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.CT_MATERIALS))
    geo, views, angles = scenes.setup_ct(ctx, n_views=12, shape=SHAPE, voxel_cm=VOX, seed=4, pixels=(32, 8))
    yield ctx, geo, views, angles
    ctx.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_views_same_alone_and_in_one_launch(ct, events):
    ctx, geo, views, angles = ct
    ctx.set_event_kernel(*events)
    n = [20000 + 1000 * i for i in range(len(views))]
    seeds = [100 + i for i in range(len(views))]
    ctx.set_graphs(True)
    ctx.job_select(0)
    ctx.reset_tallies()
    for v in views:
        ctx.job_select(v)
        ctx.reset_tallies()
    st = ctx.run_jobs(views, n, [0] * len(views), seeds)
    assert st["histories"] == sum(n)
    together = []
    for v in views:
        ctx.job_select(v)
        together.append(ctx.tally(0)["image"].copy())
    ctx.job_select(0)
    dose_all = ctx.voxel_dose(0)[0].copy()
    dose_sum = np.zeros_like(dose_all)
    for i, v in enumerate(views):
        ctx.job_select(0)
        ctx.reset_tallies()
        ctx.job_select(v)
        ctx.reset_tallies()
        ctx.run_jobs([v], [n[i]], [0], [seeds[i]])
        im = ctx.tally(0)["image"]
        assert np.count_nonzero(im[..., 2]) > 20
        assert np.allclose(im, together[i], rtol=1e-12, atol=1e-9), i
        ctx.job_select(0)
        dose_sum += ctx.voxel_dose(0)[0]
    assert np.allclose(dose_all, dose_sum, rtol=1e-9, atol=1e-6)
    ctx.set_graphs(False)
    ctx.set_event_kernel(M.EVENT_QUEUE)


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_view_matches_oracle(ct, events):
    ctx, geo, views, angles = ct
    ctx.set_event_kernel(*events)
    v = views[1]  # 30 degrees
    k, n = 24, 6000
    ids = scenes.ct_phantom(SHAPE, VOX, 4)

    def extract(src, im, dose):
        blocks = im[..., 0].reshape(8, 4, 2, 4).sum(axis=(1, 3)).ravel()  # 8 column blocks x 2 row blocks
        d = dose.reshape(SHAPE[2], SHAPE[1], SHAPE[0]).reshape(2, 8, 2, 16, 2, 16).sum(axis=(1, 3, 5)).ravel()
        return np.concatenate([blocks, d]) / n

    g = []
    for b in range(k):
        ctx.job_select(0)
        ctx.reset_tallies()
        ctx.job_select(v)
        ctx.reset_tallies()
        ctx.run(n, seed=7, first=b * n, kernel=M.KERNEL_QUEUE)
        im = ctx.tally(0)["image"]
        ctx.job_select(0)
        g.append(extract(ctx, im, ctx.voxel_dose(0)[0]))
    sim = oracle.Simulation(scenes.CT_MATERIALS)
    sim.set_domain(geo["dim"], scenes.AIR)
    sim.add_grid(ids, (VOX,) * 3, geo["origin"])
    sim.enable_voxel_dose(True)
    scenes.setup_ct_view(sim, geo, angles[1], pixels=(32, 8))
    o = []
    for b in range(k):
        sim.reset()
        sim.run(n, seed=900 + b, threads=8)
        o.append(extract(sim, sim.tally(0)["image"], sim.voxel_dose(0)[0]))
    g, o = np.array(g), np.array(o)
    gm, gs = batch_mean_sigma(g)
    om, os_ = batch_mean_sigma(o)
    keep = (gs > 0) & (os_ > 0) & (np.abs(om) > 3 * os_)
    assert keep.sum() > 8
    z = welch_normal_scores(gm, gs, k, om, os_, k)[keep]
    assert np.all(np.abs(z) < 4.0), np.max(np.abs(z))
    p, t2, rank = hotelling(g[:, keep], o[:, keep])
    assert p > 0.01, (p, t2, rank)
    ctx.set_event_kernel(M.EVENT_QUEUE)
