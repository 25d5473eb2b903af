"""Host logic of the benchmark's launch tuning (no device): the coordinate
search over bank size, blocks per CU and interaction batch finds the
fastest setting of a separable cost, keeps the refill batch, times each
candidate once and leaves the chosen setting on the context."""
import midsx_amd as M

import bench


class FakeCtx:
    def __init__(self, kind, bank, blocks, refill, interact):
        self.ev = (kind, bank)
        self.qp = (blocks, refill, interact)

    def event_kernel(self):
        return self.ev

    def queue_params(self):
        return self.qp

    def set_event_kernel(self, kind, bank):
        self.ev = (kind, bank)

    def set_queue_params(self, b, r, i):
        self.qp = (b, r, i)


def test_launch_tune_finds_the_separable_minimum():
    ctx = FakeCtx(M.EVENT_BANK, 128, 3, 64, 48)
    runs = []
    bank_cost = {96: 3.0, 128: 2.0, 160: 1.0}
    block_cost = {2: 2.5, 3: 1.5, 4: 0.5}
    inter_cost = {32: 0.3, 48: 0.2, 64: 0.1}

    def run():
        (kind, bank), (b, r, i) = ctx.ev, ctx.qp
        assert r == 64  # the refill batch is not searched
        runs.append((bank, b, i))
        return bank_cost[bank] + block_cost[b] + inter_cost[i]

    tuned = {"event_kernel": "bank", "bank_size": 128, "blocks_per_cu": 3, "interact_min": 48, "probe_ms": 1.0}
    out = bench.launch_tune(ctx, run, tuned)
    assert ctx.ev == (M.EVENT_BANK, 160) and ctx.qp == (4, 64, 64)
    assert (out["bank_size"], out["blocks_per_cu"], out["interact_min"]) == (160, 4, 64)
    assert out["autotuner"]["bank_size"] == 128 and out["probe_ms"] == 1.0
    assert len(runs) == len(set(runs)) == out["launch_candidates"]  # each candidate timed once


def test_launch_tune_on_the_queue_kernel_keeps_its_bank():
    ctx = FakeCtx(M.EVENT_QUEUE, 128, 4, 32, 32)
    out = bench.launch_tune(ctx, lambda: float(ctx.qp[0]), {"event_kernel": "queue", "bank_size": 128,
                                                            "blocks_per_cu": 4, "interact_min": 32})
    assert ctx.ev == (M.EVENT_QUEUE, 128)
    assert ctx.qp[0] == 2 and out["event_kernel"] == "queue"


def test_roi_dose_sums_and_errors():
    """ROI doses from per-voxel means with relative errors equal those from
    the same means with variances, and from sums and sums of squares over n
    histories; each ROI dose voxel covers the 1 cm cube of scenes.dose_job_roi."""
    import numpy as np
    from midsx_amd import scenes
    job = dict(shape=(40, 40, 40), voxel_cm=0.5)
    rng = np.random.default_rng(3)
    n = 5000
    s = rng.random(40 ** 3) * n
    s2 = s * s / n * (1 + rng.random(40 ** 3))
    m = s / n
    var = np.maximum(s2 / n - m * m, 0) / (n - 1)
    a = bench.roi_dose(job, s, var=s2, n=n)
    b = bench.roi_dose(job, m, var=var)
    c = bench.roi_dose(job, m.astype(np.float32), rel=(np.sqrt(var) / m).astype(np.float32))
    assert np.allclose(a, b, rtol=1e-12) and np.allclose(b, c, rtol=1e-5)
    m3 = m.reshape(40, 40, 40)
    for (i0, i1, j0, j1, k0, k1), (d, _) in zip(scenes.dose_job_roi(job), b):
        assert (i1 - i0, j1 - j0, k1 - k0) == (2, 2, 2)
        assert abs(d - m3[k0:k1, j0:j1, i0:i1].sum()) < 1e-9 * d
