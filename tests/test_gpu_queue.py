"""The persistent event kernels (lane-resident queue, per-wave photon banks)
and multi-job launches, each test on every event kernel.

* parity of the queue kernel against the oracle on the slab workload (same
  Hotelling T^2 test as the plain kernel);
* a job's results are the same alone and batched with other jobs: its RNG
  streams are keyed by (seed, job, history), so only the FP64 summation order
  of its tallies may differ (checked to 1e-12 relative).
"""
import numpy as np
import pytest

import midsx_amd as M
import oracle
from midsx_amd import scenes
from helpers import EVENT_FORMS, batch_mean_sigma, check_macro_table, hotelling, welch_normal_scores
from test_gpu_parity import orc_batches, slab_extract

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_queue_kernel_slab_matches_oracle(events):
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.C1_MATERIALS))
    scenes.setup_slab(ctx)
    k, n = 48, 50000
    g = []
    for b in range(k):
        ctx.reset_tallies()
        st = ctx.run(n, seed=13, first=b * n, kernel=M.KERNEL_QUEUE)
        assert st["histories"] == n
        g.append(slab_extract(ctx, st, n))
    g = np.array(g)
    sim = oracle.Simulation(scenes.C1_MATERIALS)
    scenes.setup_slab(sim)
    o = np.array(orc_batches(sim, n, k, 303, slab_extract))
    gm, gs = batch_mean_sigma(g)
    om, os_ = batch_mean_sigma(o)
    keep = (gs > 0) & (os_ > 0)
    z = welch_normal_scores(gm, gs, k, om, os_, k)[keep]
    assert np.all(np.abs(z) < 3.5), np.max(np.abs(z))
    p, t2, rank = hotelling(g[:, keep], o[:, keep])
    assert p > 0.01, (p, t2, rank)
    ctx.close()


def make_jobs(ctx):
    """Three scenes of different shapes sharing one table set."""
    jobs = []
    # job A: slab (job 0)
    scenes.setup_slab(ctx)
    jobs.append(0)
    # job B: small radiography with image + voxel dose
    b = ctx.job_create()
    scenes.setup_radiography(ctx, shape=(32, 32, 24), voxel_cm=0.5, seed=7, image=(16, 16))
    jobs.append(b)
    # job C: bone block, isotropic source, AABB tally
    c = ctx.job_create()
    ctx.set_domain((10.0, 10.0, 10.0), 3)
    ids = np.full((20, 20, 20), scenes.BONE, np.uint8)
    ids[5:15, 5:15, 5:15] = scenes.WATER
    ctx.add_grid(ids, (0.25, 0.25, 0.25), (2.5, 2.5, 2.5))
    ctx.set_source(direction=M.DIR_ISOTROPIC, position=(5, 5, 5), energy=40e3)
    ctx.set_tallies([dict(kind=M.TALLY_AACUBOID, p0=(2.5, 2.5, 2.5), p1=(7.5, 7.5, 7.5))])
    ctx.enable_voxel_dose(True)
    jobs.append(c)
    return jobs


def snapshot(ctx, job):
    ctx.job_select(job)
    out = [ctx.tally(i)["acc"].ravel() for i in range(len(ctx.tallies))]
    if ctx.num_grids():
        s, s2, n = ctx.voxel_dose(0)
        out += [s, s2, n.astype(np.float64)]
    out.append(np.array(ctx.background_dose()))
    return np.concatenate(out)


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_job_same_alone_and_in_mix(events):
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(set(scenes.C1_MATERIALS) | set(scenes.C2_MATERIALS)))
    jobs = make_jobs(ctx)
    n = [300000, 200000, 100000]
    seeds = [21, 22, 23]
    alone = []
    for j, nj, sj in zip(jobs, n, seeds):
        for jj in jobs:
            ctx.job_select(jj)
            ctx.reset_tallies()
        st = ctx.run_jobs([j], [nj], [0], [sj])
        assert st["histories"] == nj
        alone.append(snapshot(ctx, j))
        if j == jobs[1]:
            dose_b = ctx.voxel_dose(0)[0].copy()
    for jj in jobs:
        ctx.job_select(jj)
        ctx.reset_tallies()
    st = ctx.run_jobs(jobs, n, [0, 0, 0], seeds)
    assert st["histories"] == sum(n)
    for j, a in zip(jobs, alone):
        m = snapshot(ctx, j)
        assert a.shape == m.shape
        assert np.count_nonzero(a) > 10
        assert np.allclose(m, a, rtol=1e-12, atol=1e-9), (j, np.max(np.abs(m - a)))
    # and a job split over two launches equals one launch
    ctx.job_select(jobs[1])
    ctx.reset_tallies()
    ctx.run_jobs([jobs[1]], [120000], [0], [22])
    ctx.run_jobs([jobs[1], jobs[2]], [80000, 1000], [120000, 0], [22, 99])
    ctx.job_select(jobs[1])
    s = ctx.voxel_dose(0)[0]
    assert np.allclose(s, dose_b, rtol=1e-12, atol=1e-9)
    ctx.close()


def test_job_footprint_and_release():
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    j = ctx.job_create()
    scenes.setup_radiography(ctx, shape=(64, 64, 32), voxel_cm=0.5, seed=1, image=(8, 8))
    b = ctx.job_device_bytes(j)
    assert b >= 64 * 64 * 32 * (1 + 16 + 4)
    assert ctx.tables_device_bytes() > 0
    ctx.job_release(j)
    with pytest.raises(M.MdxError):
        ctx.job_select(j)
    ctx.close()


def play_rounds(ctx, jobs, rounds):
    for jj in jobs:
        ctx.job_select(jj)
        ctx.reset_tallies()
    counts = []
    for r in rounds:
        st = ctx.run_jobs(*r)
        counts.append({k: v for k, v in st.items() if k != "kernel_ms"})
    return [snapshot(ctx, j) for j in jobs], counts


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_graph_rounds_match_eager(events):
    """Rounds replayed from one captured hipGraph (round-buffer copy, counter
    zeroing kernels, queue kernel) give the eager launches' results; eager
    work and synchronisations run between replays, and a change of queue
    parameters forces a new capture."""
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(set(scenes.C1_MATERIALS) | set(scenes.C2_MATERIALS)))
    jobs = make_jobs(ctx)
    rounds = [(jobs, [100000, 60000, 30000], [0, 0, 0], [5, 6, 7]),
              ([jobs[1]], [50000], [60000], [6]),
              ([jobs[2], jobs[0]], [20000, 40000], [30000, 100000], [7, 5])]
    eager, ce = play_rounds(ctx, jobs, rounds)
    ctx.set_graphs(True)
    for rep in range(4):
        if rep == 2:
            ctx.set_queue_params(4, 4, 8)
        g, cg = play_rounds(ctx, jobs, rounds)
        assert cg == ce, rep  # integer event counters: exact
        for j, a, m in zip(jobs, eager, g):
            assert np.count_nonzero(a) > 10
            assert np.allclose(m, a, rtol=1e-12, atol=1e-9), (rep, j, np.max(np.abs(m - a)))
        # eager work between replays
        ctx.job_select(jobs[0])
        ctx.interpolate(scenes.WATER, M.XS_TOTAL, np.array([3e4], np.float32))
    ctx.set_graphs(False)
    again, ca = play_rounds(ctx, jobs, rounds)
    assert ca == ce
    for a, m in zip(eager, again):
        assert np.allclose(m, a, rtol=1e-12, atol=1e-9)
    ctx.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_alias_domain_matches_own_copy(events):
    """A job that aliases another job's phantom tallies what a job with its
    own copy of the phantom tallies; the shared per-voxel dose holds both
    jobs' deposits."""
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    kw = dict(shape=(32, 32, 24), voxel_cm=0.5, seed=7, image=(16, 16))
    scenes.setup_radiography(ctx, **kw)  # job 0 owns the phantom
    own = ctx.job_create()
    scenes.setup_radiography(ctx, **kw)
    alias = ctx.job_create()
    ctx.job_alias_domain(0)  # job 0's phantom and dose, its own source and tallies
    scenes.setup_radiography(ctx, **kw, with_domain=False)
    for j in (0, own, alias):
        ctx.job_select(j)
        ctx.reset_tallies()
    ctx.run_jobs([0], [100000], [0], [41])
    ctx.job_select(0)
    d0 = ctx.voxel_dose(0)[0].copy()
    ctx.run_jobs([own, alias], [80000, 80000], [0, 0], [42, 42])
    ctx.job_select(own)
    d_own = ctx.voxel_dose(0)[0]
    im_own = ctx.tally(0)["image"]
    ctx.job_select(alias)
    im_alias = ctx.tally(0)["image"]
    assert np.count_nonzero(im_own) > 10
    assert np.allclose(im_alias, im_own, rtol=1e-12, atol=1e-9)
    d_shared = ctx.voxel_dose(0)[0]  # job 0's grid, read through the alias
    assert np.allclose(d_shared, d0 + d_own, rtol=1e-9, atol=1e-6)
    with pytest.raises(M.MdxError):
        ctx.job_release(0)
    ctx.job_release(alias)
    ctx.job_release(0)
    ctx.close()


@pytest.mark.parametrize("rounds", [False, True])
def test_autotune_on_aliased_job_clears_shared_dose(rounds):
    """The autotuner's probes on a job that uses another job's phantom
    deposit in that phantom's voxels and background voxel: after it, the
    owner's voxel dose (sums and counts) and background dose read zero, as
    do the tuned job's own tallies; in both probe forms (the job alone, or
    two slices of it in one multi-job launch), and the event kernel it keeps
    is one of those it tried."""
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    kw = dict(shape=(32, 32, 24), voxel_cm=0.5, seed=7, image=(16, 16))
    scenes.setup_radiography(ctx, **kw)  # job 0 owns the phantom
    view = ctx.job_create()
    ctx.job_alias_domain(0)
    scenes.setup_radiography(ctx, **kw, with_domain=False)
    ctx.run_jobs([view], [50000], [0], [3])
    ctx.job_select(0)
    assert ctx.voxel_dose(0)[0].sum() > 0
    ctx.job_select(view)
    tuned = ctx.autotune(probe_histories=20000, rounds=rounds)
    assert tuned["event_kernel"] in ("queue", "bank") and tuned["bank_size"] in M.BANK_SIZES
    assert ctx.event_kernel() == (M.EVENT_BANK if tuned["event_kernel"] == "bank" else M.EVENT_QUEUE,
                                  tuned["bank_size"])
    assert ctx.queue_params() == (tuned["blocks_per_cu"], tuned["refill_min"], tuned["interact_min"])
    assert not ctx.tally(0)["image"].any()
    s, s2, n = ctx.voxel_dose(0)
    assert not s.any() and not s2.any() and not n.any()
    assert ctx.background_dose() == (0.0, 0.0, 0.0)
    ctx.job_select(0)
    assert not ctx.voxel_dose(0)[0].any()
    ctx.job_release(view)
    ctx.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_free_flights_match_majorant_steps(events):
    """Background free flights change only the number of rejected steps: the
    image, dose and interaction counts of a radiography job agree with the
    same job run with majorant steps everywhere (Welch scores and Hotelling
    T^2 over 40 batches each), in far fewer steps.  A job with a volume
    tally never takes them."""
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.C2_MATERIALS))
    scenes.setup_radiography(ctx, shape=(48, 48, 40), voxel_cm=0.5, seed=4, image=(16, 16))
    k, n = 40, 40000

    def batches(seed):
        out, steps = [], 0
        for b in range(k):
            ctx.reset_tallies()
            st = ctx.run(n, seed=seed, first=b * n, kernel=M.KERNEL_QUEUE)
            steps += st["steps"]
            im = ctx.tally(0)["image"][..., 0].reshape(4, 4, 4, 4).sum(axis=(1, 3)).ravel()
            dose = ctx.voxel_dose(0)[0].reshape(4, 10, 4, 12, 4, 12).sum(axis=(1, 3, 5)).ravel()
            out.append(np.concatenate([im, dose, [ctx.background_dose()[0], st["real"], st["compton"],
                                                  st["photo"], st["escapes"]]]) / n)
        return np.array(out), steps

    ctx.set_flights(True)
    f, sf = batches(61)
    ctx.set_flights(False)
    m, sm = batches(62)
    assert sf < 0.5 * sm, (sf, sm)
    fm, fs = batch_mean_sigma(f)
    mm, ms = batch_mean_sigma(m)
    keep = (fs > 0) & (ms > 0) & (np.abs(mm) > 3 * ms)
    assert keep.sum() > 30
    z = welch_normal_scores(fm, fs, k, mm, ms, k)[keep]
    assert np.all(np.abs(z) < 3.5), np.max(np.abs(z))
    p, t2, rank = hotelling(f[:, keep], m[:, keep])
    assert p > 0.01, (p, t2, rank)
    # with a volume tally the job keeps majorant steps: same counts as flights off
    ctx.set_tallies([dict(kind=M.TALLY_AACUBOID, p0=(1, 1, 1), p1=(20, 20, 10))])
    ctx.set_flights(True)
    a = ctx.run(20000, seed=9, first=0, kernel=M.KERNEL_QUEUE)
    ctx.set_flights(False)
    b = ctx.run(20000, seed=9, first=0, kernel=M.KERNEL_QUEUE)
    assert a["steps"] == b["steps"] and a["real"] == b["real"]
    ctx.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_setup_during_round_clears_reused_blocks(events):
    """A job set up while a graph round is in flight has its dose and tally
    buffers cleared by the next round's graph (the set-up calls do not wait
    for the round).  Its buffers here come back from the pool holding a
    released job's sums, so a lost clear would show in its results, which
    must equal those of the same job set up with the device idle."""
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(scenes.DOSE_MATERIALS))
    ctx.set_graphs(True)
    small = dict(shape=(32, 32, 32), voxel_cm=0.2, phantom_seed=2, image=(16, 16), energy=50e3)
    big = dict(shape=(64, 64, 64), voxel_cm=0.2, phantom_seed=1, image=(32, 32), energy=80e3)

    def results(j):
        ctx.job_select(j)
        s, s2, n = ctx.voxel_dose(0)
        return s.copy(), n.copy(), ctx.tally(0)["image"].copy(), ctx.background_dose()

    ref = ctx.job_create()
    scenes.setup_dose_job(ctx, small)
    ref_st = ctx.run_jobs([ref], [200000], [0], [5])
    want = results(ref)
    ctx.job_release(ref)
    dirty = ctx.job_create()  # same shapes: its blocks return to the pool with sums in them
    scenes.setup_dose_job(ctx, small)
    ctx.run_jobs([dirty], [200000], [0], [77])
    ctx.job_release(dirty)
    b = ctx.job_create()
    scenes.setup_dose_job(ctx, big)
    ctx.run_jobs_async([b], [4_000_000], [0], [3])
    j = ctx.job_create()
    scenes.setup_dose_job(ctx, small)  # while the round runs
    ctx.wait()
    st = ctx.run_jobs([j], [200000], [0], [5])
    got = results(j)
    for k in ("steps", "real", "deposits", "surface_hits"):
        assert st[k] == ref_st[k], k
    for a, m in zip(want[:3], got[:3]):
        assert np.count_nonzero(a) > 10
        assert np.allclose(m, a, rtol=1e-12, atol=1e-12 * np.abs(a).max()), np.max(np.abs(m - a))
    assert np.allclose(got[3], want[3], rtol=1e-12, atol=0)
    ctx.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_local_majorants_match_global_majorant(events):
    """Local majorants (macro cells with a bounding material and a safe
    radius) change only the number of rejected steps: a radiography job's
    image, dose and interaction counts agree with the same job run with the
    global majorant (Welch scores below 4.0 on each of the ~85 bins, a
    family-wise level of ~0.5 %, and Hotelling T^2 over 40 batches), in fewer
    steps; the macro-cell table holds what it claims."""
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    materials = list(scenes.C2_MATERIALS)
    ctx.build_tables(materials)
    shape = (64, 64, 48)
    scenes.setup_radiography(ctx, shape=shape, voxel_cm=0.5, seed=4, image=(16, 16))
    ids = scenes.ellipsoid_phantom(shape, 4, voxel_cm=0.5)
    assert check_macro_table(ctx, ids, materials, scenes.AIR) > 10
    k, n = 40, 40000

    def batches(seed):
        out, steps = [], 0
        for b in range(k):
            ctx.reset_tallies()
            st = ctx.run(n, seed=seed, first=b * n, kernel=M.KERNEL_QUEUE)
            steps += st["steps"]
            im = ctx.tally(0)["image"][..., 0].reshape(4, 4, 4, 4).sum(axis=(1, 3)).ravel()
            dose = ctx.voxel_dose(0)[0].reshape(4, 12, 4, 16, 4, 16).sum(axis=(1, 3, 5)).ravel()
            out.append(np.concatenate([im, dose, [ctx.background_dose()[0], st["real"], st["compton"],
                                                  st["photo"], st["escapes"]]]) / n)
        return np.array(out), steps

    ctx.set_local_majorants(True)
    f, sf = batches(71)
    ctx.set_local_majorants(False)
    assert ctx.macro_cells() is None
    m, sm = batches(72)
    assert sf < 0.98 * sm, (sf, sm)
    fm, fs = batch_mean_sigma(f)
    mm, ms = batch_mean_sigma(m)
    keep = (fs > 0) & (ms > 0) & (np.abs(mm) > 3 * ms)
    assert keep.sum() > 30
    z = welch_normal_scores(fm, fs, k, mm, ms, k)[keep]
    assert np.all(np.abs(z) < 4.0), np.max(np.abs(z))
    p, t2, rank = hotelling(f[:, keep], m[:, keep])
    assert p > 0.01, (p, t2, rank)
    ctx.close()
