"""mdx_run_jobs_noalloc, the entry point that allocates nothing, against
mdx_run_jobs on the same inputs: two calls of it in a row, the second with
fewer jobs and fewer histories (the static scratch keeps the first call's
contents), each equal to mdx_run_jobs's results -- event counters and
per-voxel deposit counts exactly, FP64 sums to summation order -- with the
device's free memory unchanged across each call; more jobs than the scratch
holds is refused."""
import numpy as np
import pytest

import midsx_amd as M
from midsx_amd import scenes
from helpers import EVENT_FORMS

pytestmark = pytest.mark.gpu

KW = dict(shape=(32, 32, 24), voxel_cm=0.5, image=(16, 16))


def make_jobs(ctx):
    ctx.build_tables(list(scenes.C2_MATERIALS))
    a = ctx.job
    scenes.setup_radiography(ctx, seed=7, **KW)
    b = ctx.job_create()
    scenes.setup_radiography(ctx, seed=8, kvp=60.0, **KW)
    c = ctx.job_create()  # a view of a's phantom with its own source and image
    ctx.job_alias_domain(a)
    scenes.setup_radiography(ctx, seed=7, kvp=100.0, with_domain=False, **KW)
    return [a, b, c]


def reset(ctx, jobs):
    for j in jobs:
        ctx.job_select(j)
        ctx.reset_tallies()


def results(ctx, jobs):
    out = []
    for j in jobs:
        ctx.job_select(j)
        t = ctx.tally(0)
        s, s2, n = ctx.voxel_dose(0)
        out.append(dict(acc=t["acc"].copy(), image=t["image"].copy(), s=s.copy(), s2=s2.copy(), n=n.copy(),
                        bg=np.array(ctx.background_dose())))
    return out


def same(ref, got):
    for r, g in zip(ref, got):
        assert np.array_equal(r["n"], g["n"])  # deposit counts: integers, exact
        assert np.array_equal(r["image"][..., 2], g["image"][..., 2])  # pixel hit counts
        for k in ("acc", "image", "s", "s2", "bg"):
            assert np.allclose(r[k], g[k], rtol=1e-12, atol=1e-9), k


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_noalloc_matches_run_jobs_twice(events):
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    jobs = make_jobs(ctx)
    calls = [(jobs, [200000, 150000, 100000], [0, 0, 0], [11, 12, 13]),
             (jobs[:2], [50000, 30000], [1000, 7], [21, 22])]  # smaller: fewer jobs, fewer histories
    refs = []
    for ids, n, first, seeds in calls:
        reset(ctx, jobs)
        st = ctx.run_jobs(ids, n, first, seeds)
        refs.append((st, results(ctx, jobs)))
    for (ids, n, first, seeds), (st_ref, ref) in zip(calls, refs):
        reset(ctx, jobs)
        free0 = ctx.mem_info()[0]
        st = ctx.run_jobs_noalloc(ids, n, first, seeds)
        assert ctx.mem_info()[0] == free0
        assert st["histories"] == sum(n)
        for k in M.STAT_NAMES:
            assert st[k] == st_ref[k], k
        same(ref, results(ctx, jobs))
    with pytest.raises(M.MdxError) as e:
        ctx.run_jobs_noalloc([jobs[0]] * 257, [10] * 257, [0] * 257, list(range(257)))
    assert e.value.status == M.ERR_INVALID and "256" in str(e.value)
    ctx.job_release(jobs[2])
    ctx.close()
