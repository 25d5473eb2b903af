"""The continuous-batching loop on the device.

* each job's voxel dose and image are the same served in a mixed stream
  (several rounds, a memory budget that makes jobs wait, other jobs in every
  round) as served alone: its histories are keyed by (seed, history), so only
  the FP64 order of summation may differ (checked to 1e-12 relative);
* served dose jobs agree with the oracle: 24 copies of one dose job with
  different seeds run as one served stream are 24 independent batches; their
  block-summed doses and image quadrants are compared with 24 oracle batches
  by Welch scores and a two-sample Hotelling T^2 test.
"""
import numpy as np
import pytest

import midsx_amd as M
import oracle
from midsx_amd import scenes
from midsx_amd.serve import Server, poisson_arrivals, run_plain
from helpers import EVENT_FORMS, batch_mean_sigma, hotelling, welch_normal_scores

pytestmark = pytest.mark.gpu


def keep_arrays(job, s, s2, image):
    return (s.copy(), s2.copy(), image.copy())


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    c.build_tables(list(scenes.DOSE_MATERIALS))
    yield c
    c.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_served_same_alone_and_in_mix(ctx, events):
    ctx.set_event_kernel(*events)
    jobs = scenes.draw_dose_jobs(10, 5, shapes=(16, 24, 32), h_lo=5e4, h_hi=4e5)
    budget = 3 * max(scenes.dose_job_bytes(j) for j in jobs)
    srv = Server(ctx, budget_bytes=budget, round_histories=300_000, min_slice=20_000, on_result=keep_arrays,
                 dose="f64")
    mixed = srv.run(jobs, poisson_arrivals(len(jobs), 50.0, 9))
    assert srv.rounds > 3
    for i in (0, 3, 7):
        alone = Server(ctx, budget_bytes=budget, round_histories=10 ** 9, on_result=keep_arrays,
                       dose="f64").run([jobs[i]])[0]
        for a, m in zip(alone.summary, mixed[i].summary):
            assert np.count_nonzero(a) > 10
            assert np.allclose(m, a, rtol=1e-12, atol=1e-12 * np.abs(a).max()), (i, np.max(np.abs(m - a)))
    ctx.job_select(0)
    ctx.set_event_kernel(M.EVENT_QUEUE)


def test_plain_fifo_serves_same_jobs(ctx):
    jobs = scenes.draw_dose_jobs(4, 6, shapes=(16,), h_lo=5e4, h_hi=1e5)
    res, info = run_plain(ctx, jobs, on_result=keep_arrays, dose="f64")
    assert info["counters"]["histories"] == sum(j["histories"] for j in jobs)
    for r in res:
        assert r.finished >= r.admitted and np.count_nonzero(r.summary[0]) > 10


def block_sums(s, shape, b=2):
    nx, ny, nz = shape
    d = s.reshape(nz, ny, nx)
    return d.reshape(b, nz // b, b, ny // b, b, nx // b).sum(axis=(1, 3, 5)).ravel()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_served_dose_matches_oracle(ctx, events):
    ctx.set_event_kernel(*events)
    job = dict(shape=(16, 16, 16), voxel_cm=0.25, phantom_seed=11, histories=20000, seed=1, image=(8, 8),
               kvp=80.0)
    k = 24
    jobs = [dict(job, seed=1000 + b) for b in range(k)]
    srv = Server(ctx, round_histories=100_000, min_slice=5_000, on_result=keep_arrays, dose="f64")
    res = srv.run(jobs)

    def extract(s, image):
        im = image[..., 0]
        quads = im.reshape(2, 4, 2, 4).sum(axis=(1, 3)).ravel()
        return np.concatenate([block_sums(s, job["shape"]), quads]) / job["histories"]

    g = np.array([extract(r.summary[0], r.summary[2]) for r in res])
    sim = oracle.Simulation(scenes.DOSE_MATERIALS)
    scenes.setup_dose_job(sim, job)
    o = []
    for b in range(k):
        sim.reset()
        sim.run(job["histories"], seed=500 + b, threads=8)
        o.append(extract(sim.voxel_dose(0)[0], sim.tally(0)["image"]))
    o = np.array(o)
    gm, gs = batch_mean_sigma(g)
    om, os_ = batch_mean_sigma(o)
    keep = (gs > 0) & (os_ > 0)
    z = welch_normal_scores(gm, gs, k, om, os_, k)[keep]
    assert np.all(np.abs(z) < 4.0), np.max(np.abs(z))
    p, t2, rank = hotelling(g[:, keep], o[:, keep])
    assert p > 0.01, (p, t2, rank)
    ctx.set_event_kernel(M.EVENT_QUEUE)


def test_dose_f32_read_matches_f64(ctx):
    job = dict(shape=(20, 16, 12), voxel_cm=0.3, phantom_seed=3, histories=200000, seed=77, image=(8, 8),
               energy=50e3)
    slot = ctx.job_create()
    scenes.setup_dose_job(ctx, job)
    ctx.run(job["histories"], seed=job["seed"], kernel=M.KERNEL_QUEUE)
    s, s2, _ = ctx.voxel_dose(0)
    mean, rel = ctx.voxel_dose_f32(0, job["histories"])
    n = job["histories"]
    m = s / n
    r = np.where(m > 0, np.sqrt(np.maximum(s2 / n - m * m, 0) / (n - 1)) / np.where(m > 0, m, 1), 0)
    assert np.count_nonzero(m) > 100
    assert np.allclose(mean, m, rtol=1e-6, atol=0)
    assert np.allclose(rel, r, rtol=1e-5, atol=1e-7)
    ctx.stage_voxel_dose_f32(0, n)  # queued conversion, then a plain copy
    mean2, rel2 = ctx.voxel_dose_f32(0, n)
    assert np.array_equal(mean2, mean) and np.array_equal(rel2, rel)
    ctx.job_release(slot)


def test_served_f32_reads_match_f64(ctx):
    """The FP32 (mean, relative error) results, converted on the device and
    copied beside later rounds, equal the FP64 sums read synchronously."""
    jobs = scenes.draw_dose_jobs(6, 8, shapes=(16, 24), h_lo=5e4, h_hi=2e5)
    budget = 2 * max(scenes.dose_job_bytes(j) for j in jobs)
    a = Server(ctx, budget_bytes=budget, round_histories=150_000, min_slice=20_000, on_result=keep_arrays,
               dose="f32").run(jobs)
    b = Server(ctx, budget_bytes=budget, round_histories=150_000, min_slice=20_000, on_result=keep_arrays,
               dose="f64").run(jobs)
    for ra, rb, j in zip(a, b, jobs):
        n = j["histories"]
        mean, rel, im_a = ra.summary
        s, s2, im_b = rb.summary
        m = s / n
        r = np.where(m > 0, np.sqrt(np.maximum(s2 / n - m * m, 0) / (n - 1)) / np.where(m > 0, m, 1), 0)
        assert np.count_nonzero(m) > 100
        assert np.allclose(mean, m, rtol=1e-6, atol=0)
        assert np.allclose(rel, r, rtol=1e-4, atol=1e-6)
        assert np.allclose(im_a, im_b, rtol=1e-12, atol=1e-9)
