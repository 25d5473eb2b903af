"""Voxel-dose accumulation precision (FP64, FP32, FP32 with compensated
adds) on identical RNG streams: the deposits are the same events in every
mode (counts identical), so the sums differ only by accumulation rounding;
the compensated FP32 sums of deposits match FP64, their sums of squares
within ~1 FP32 ulp, and plain FP32 sums are far worse."""
import numpy as np
import pytest

import midsx_amd as M
from midsx_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [M.KERNEL_PLAIN, M.KERNEL_QUEUE])
def test_dose_precision_modes(kernel):
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C1_MATERIALS))
    scenes.setup_slab(ctx)
    n = 2_000_000
    out = {}
    for mode in (M.DOSE_F64, M.DOSE_F32, M.DOSE_F32C, M.DOSE_F32C_LDS):
        ctx.set_dose_precision(mode)
        ctx.reset_tallies()
        ctx.run(n // 2, seed=5, first=0, kernel=kernel)
        ctx.run(n // 2, seed=5, first=n // 2, kernel=kernel)
        out[mode] = ctx.voxel_dose(0)
    s64, q64, c64 = out[M.DOSE_F64]
    hot = c64 >= 20
    assert hot.sum() > 40
    for mode in (M.DOSE_F32, M.DOSE_F32C, M.DOSE_F32C_LDS):
        assert np.array_equal(out[mode][2], c64)  # the same deposits
    # block sums in LDS (queue kernel; the plain kernel takes the global form)
    e32l = np.abs(out[M.DOSE_F32C_LDS][0][hot] / s64[hot] - 1)
    assert e32l.max() < 1e-9, e32l.max()
    e32 = np.abs(out[M.DOSE_F32][0][hot] / s64[hot] - 1)
    q32 = np.abs(out[M.DOSE_F32][1][hot] / q64[hot] - 1)
    e32c = np.abs(out[M.DOSE_F32C][0][hot] / s64[hot] - 1)
    q32c = np.abs(out[M.DOSE_F32C][1][hot] / q64[hot] - 1)
    # the sum of deposits carries its exact rounding error: FP64-level;
    # the sum of squares spans a wider range, and its error term is itself
    # an FP32 sum: about one FP32 ulp
    assert e32c.max() < 1e-9, e32c.max()
    assert q32c.max() < 1e-6, q32c.max()
    assert e32.max() < 1e-2 and q32.max() < 1e-2
    assert e32.max() > 10 * max(e32c.max(), 1e-12) and q32.max() > 10 * q32c.max()
    with pytest.raises(M.MdxError):
        ctx.voxel_dose_f32(0, n)  # FP32 (mean, error) reads need FP64 sums
    ctx.close()


def test_lds_dose_rejects_large_grids():
    ctx = M.Context(0)
    ctx.build_tables(list(scenes.C1_MATERIALS))
    ctx.set_domain((10.0, 10.0, 10.0), 3)
    ctx.add_grid(np.full((40, 40, 40), 48, np.uint8), (0.25, 0.25, 0.25), (0, 0, 0))
    ctx.enable_voxel_dose(True)
    with pytest.raises(M.MdxError):
        ctx.set_dose_precision(M.DOSE_F32C_LDS)  # 64 000 voxels do not fit in one block's LDS
    ctx.close()
