"""fvk_residual_l2_noalloc (the allocation-free residual norm) against
fvk_residual_l2, through the C ABI.

The two entry points run the same kernels (per-block partial sums, then
their compensated sum), one through the context's buffers, the other through
a static device array, so their norms are equal bit for bit.  Checked on two calls in a row on different contexts, the
second with a smaller mesh (the static array keeps the first call's larger
contents beyond what the second reads), in both precisions, and on a mesh
above the array's 65,536-block limit, which is refused.
"""
import numpy as np
import pytest

from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, FvkError, Solver

pytestmark = pytest.mark.gpu


def _solver(grid, case, steps, **ov):
    s = Solver(grid, case, **ov)
    s.init_freestream()
    s.step(steps)
    return s


@pytest.mark.parametrize("precision", [0, FP32_MIXED])
def test_noalloc_norm_equals_residual_l2_larger_then_smaller(precision):
    u, v = inputs.mach_velocity(2.5)
    big = _solver(gg.ramp(512, 256), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                                 dt_type="LOCAL", CFL=0.8), 12,
                  precision=precision)
    u, v = inputs.mach_velocity(0.5)
    small = _solver(gg.bump(64, 32), inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL", CFL=0.5),
                    17, precision=precision)
    print(f"[gpu] noalloc norm precision={precision}", flush=True)
    for s in (big, small):
        want = s.residual_l2()
        got = s.residual_l2_noalloc()
        assert np.all(want > 0) and np.isfinite(want).all()
        assert np.array_equal(got, want), (got, want)
    # a second round on the larger one after the smaller one
    assert np.array_equal(big.residual_l2_noalloc(), big.residual_l2())


def test_noalloc_norm_refuses_more_than_65536_blocks():
    # 4.3M cells in 64-cell blocks: more blocks than the static array holds
    u, v = inputs.mach_velocity(2.5)
    grid = gg.ramp(2048, 1400)
    s = Solver(grid, inputs.Case(u=u, v=v, poly=0, dt_type="LOCAL", CFL=0.5), block_size=64)
    assert grid.num_cells > 64 * 65536
    s.init_freestream()
    s.step(1)
    with pytest.raises(FvkError) as e:
        s.residual_l2_noalloc()
    assert e.value.code == -1
    assert np.isfinite(s.residual_l2()).all()
