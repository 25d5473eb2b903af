"""Edge cases of the launch paths, on the plain kernel and every event form:

* empty and tiny launches (0, 1, 63, 64, 65 histories: fewer histories than
  a wave has lanes, or a bank has photons): the history count is exact,
  every history ends in photoelectric absorption or escape, and an empty
  launch leaves every tally at zero;
* history indices that cross 2^32 (the event kernels keep a history's index
  in two 32-bit halves): one launch over [2^32 - 300, 2^32 + 300) equals the
  two launches split at 2^32, event counters exactly and tallies to FP64
  summation order;
* a multi-job launch in which one job has an empty slice: the other jobs'
  results are those of their launches alone, and the empty job's tallies
  stay zero.
"""
import numpy as np
import pytest

import midsx_amd as M
from midsx_amd import scenes
from helpers import EVENT_FORMS
from test_gpu_queue import make_jobs, snapshot

pytestmark = pytest.mark.gpu

KERNELS = [pytest.param(("plain", None), id="plain")] + [
    pytest.param(("events", p.values[0]), id=p.id) for p in EVENT_FORMS]


def slab_ctx(kernel):
    ctx = M.Context(0)
    if kernel[0] == "events":
        ctx.set_event_kernel(*kernel[1])
    ctx.build_tables(list(scenes.C1_MATERIALS))
    scenes.setup_slab(ctx)
    return ctx, (M.KERNEL_PLAIN if kernel[0] == "plain" else M.KERNEL_QUEUE)


COUNTERS = ("histories", "steps", "real", "rayleigh", "compton", "photo", "kahn_trials", "kahn_s_tests",
            "deposits", "surface_hits", "escapes")


@pytest.mark.parametrize("kernel", KERNELS)
def test_empty_and_tiny_launches(kernel):
    ctx, K = slab_ctx(kernel)
    for n in (0, 1, 63, 64, 65, 1000):
        ctx.reset_tallies()
        st = ctx.run(n, seed=5, first=0, kernel=K)
        assert st["histories"] == n, (n, st)
        assert st["photo"] + st["escapes"] == n, (n, st)
        snap = snapshot(ctx, 0)
        if n == 0:
            assert st["steps"] == 0 and not np.any(snap), n
        else:
            assert np.any(snap), n
    ctx.close()


@pytest.mark.parametrize("kernel", KERNELS)
def test_history_indices_across_2_pow_32(kernel):
    ctx, K = slab_ctx(kernel)
    lo = (1 << 32) - 300
    ctx.reset_tallies()
    one = ctx.run(600, seed=9, first=lo, kernel=K)
    whole = snapshot(ctx, 0)
    ctx.reset_tallies()
    a = ctx.run(300, seed=9, first=lo, kernel=K)
    b = ctx.run(300, seed=9, first=1 << 32, kernel=K)
    split = snapshot(ctx, 0)
    for c in COUNTERS:
        assert one[c] == a[c] + b[c], (c, one[c], a[c], b[c])
    assert one["histories"] == 600
    assert np.allclose(whole, split, rtol=1e-12, atol=1e-9), np.max(np.abs(whole - split))
    # the upper half is a different stream from the same low indices
    ctx.reset_tallies()
    low = ctx.run(300, seed=9, first=0, kernel=K)
    assert any(low[c] != b[c] for c in ("steps", "real", "deposits"))
    ctx.close()


@pytest.mark.parametrize("events", EVENT_FORMS)
def test_multi_job_launch_with_an_empty_slice(events):
    ctx = M.Context(0)
    ctx.set_event_kernel(*events)
    ctx.build_tables(list(set(scenes.C1_MATERIALS) | set(scenes.C2_MATERIALS)))
    jobs = make_jobs(ctx)
    n = [40000, 0, 20000]
    seeds = [31, 32, 33]
    alone = {}
    for j, nj, sj in zip(jobs, n, seeds):
        if nj == 0:
            continue
        for jj in jobs:
            ctx.job_select(jj)
            ctx.reset_tallies()
        ctx.run_jobs([j], [nj], [0], [sj])
        alone[j] = snapshot(ctx, j)
    for jj in jobs:
        ctx.job_select(jj)
        ctx.reset_tallies()
    st = ctx.run_jobs(jobs, n, [0, 0, 0], seeds)
    assert st["histories"] == sum(n)
    for j, a in alone.items():
        m = snapshot(ctx, j)
        assert np.allclose(m, a, rtol=1e-12, atol=1e-9), (j, np.max(np.abs(m - a)))
    assert not np.any(snapshot(ctx, jobs[1]))
    ctx.close()


@pytest.mark.parametrize("events", [p for p in EVENT_FORMS if p.values[0][0] == M.EVENT_BANK])
def test_bank_history_index_limit(events):
    """The bank kernel keeps 42 bits of a history index: a launch that ends
    at 2^42 runs, one past it is refused before anything runs (the queue
    kernel has no such limit)."""
    ctx, K = slab_ctx(("events", events))
    top = M.BANK_MAX_HISTORY
    ctx.reset_tallies()
    st = ctx.run(300, seed=9, first=top - 300, kernel=K)
    assert st["histories"] == 300 and st["photo"] + st["escapes"] == 300
    with pytest.raises(M.MdxError) as e:
        ctx.run(300, seed=9, first=top - 299, kernel=K)
    assert e.value.status == M.ERR_RANGE
    with pytest.raises(M.MdxError):
        ctx.run_jobs([ctx.job], [300], [top], [9])
    ctx.set_event_kernel(M.EVENT_QUEUE)
    assert ctx.run(300, seed=9, first=top, kernel=K)["histories"] == 300
    ctx.close()
