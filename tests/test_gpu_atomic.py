"""Atomic-scatter residual accumulation (accumulate = 1) against the owner
gather (accumulate = 0, the reference's bits).

The atomic form scatters each face's terms into LDS accumulators of the
block's cells with LDS atomics, in run-time order, so its sums differ from the
gather's by summation order only.  Where the scheme is well conditioned (no limiter, or first order) the
two FP64 solutions must agree to 1e-10 relative after 100 steps, the bound
stated for summation-order differences; with the SQUEEZE limiter, whose
lagged ratio of differences amplifies rounding, only a loose bound holds.
"""
import numpy as np
import pytest

from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, FvkError, Solver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _pair(grid, case, steps, **ov):
    out = []
    for acc in (0, 1):
        s = Solver(grid, case, accumulate=acc, **ov)
        s.init_freestream()
        s.step(steps)
        out.append(s.state_and_grad())
        assert s.launch_options()["accumulate"] == acc
        s.close()
    return out


def _cases():
    u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
    yield "bump o2 NONE LOCAL", gg.bump(), inputs.Case(
        u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.8)
    yield "bump o1 GLOBAL", gg.bump(), inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL", CFL=0.5)
    u, v = inputs.mach_velocity(0.2)
    yield "ogrid NS o2 NONE GLOBAL", gg.ogrid(128, 64), inputs.Case(
        u=u, v=v, poly=1, limiter="NONE", eqn="NS", mu=1.225 * u / 100.0, BC1="NO_SLIP_WALL",
        dt_type="GLOBAL", CFL=0.5)
    u, v = inputs.mach_velocity(0.6)
    yield "box mixed o2 NONE", gg.box(40, 30, jitter=0.2, split_fraction=0.5,
                                      markers=(0, -1, -2, -1)), inputs.Case(
        u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.5, BC1="SLIP_WALL",
        BC2="FREESTREAM")


@pytest.mark.parametrize("name,grid,case", list(_cases()), ids=lambda x: x if isinstance(x, str) else "")
def test_atomic_fp64_matches_gather_to_summation_order(name, grid, case):
    print(f"[gpu] atomic vs gather FP64: {name}", flush=True)
    (q0, g0), (q1, g1) = _pair(grid, case, 100)
    assert np.isfinite(q1).all()
    assert _rel(q1, q0) <= 1e-10, (name, _rel(q1, q0))


def test_atomic_fp32_matches_gather():
    u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
    case = inputs.Case(u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.8)
    print("[gpu] atomic vs gather FP32-mixed", flush=True)
    (q0, _), (q1, _) = _pair(gg.bump(), case, 100, precision=FP32_MIXED)
    assert _rel(q1, q0) <= 1e-5, _rel(q1, q0)


def test_atomic_squeeze_stays_close():
    """SQUEEZE on the ramp: the rounding-order change is amplified by the
    limiter, as any perturbation of 1 ulp is; the states stay close."""
    u, v = inputs.mach_velocity(2.5)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8)
    print("[gpu] atomic vs gather SQUEEZE ramp", flush=True)
    (q0, _), (q1, _) = _pair(gg.ramp(96, 48), case, 30)
    assert np.isfinite(q1).all()
    assert np.linalg.norm(q1[0] - q0[0]) / np.linalg.norm(q0[0]) <= 1e-3


def test_atomic_batch_with_retirement():
    """Batch, graph replay and retirement with the atomic form: every
    instance within the summation-order bound of its gather run."""
    cases = []
    for mach, ang in [(0.5, 2.0), (0.7, 0.0), (0.35, 4.0)]:
        u, v = inputs.mach_velocity(mach, angle_deg=ang)
        cases.append((gg.bump(48 + int(20 * mach), 24), inputs.Case(
            u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.7)))
    cases.append(list(_cases())[3][1:])
    print("[gpu] atomic batch", flush=True)
    out = []
    for acc in (0, 1):
        b = Batch(cases, accumulate=acc)
        b.init()
        b.step(40)
        flags = np.ones(len(cases), dtype=np.uint8)
        flags[1] = 0
        b.set_active(flags)
        b.step(60)
        out.append([b.state(i) for i in range(len(cases))])
        b.close()
    for i in range(len(cases)):
        assert _rel(out[1][i], out[0][i]) <= 1e-10, (i, _rel(out[1][i], out[0][i]))


def test_atomic_needs_the_lds_shape():
    g = gg.bump(16, 8)
    with pytest.raises(FvkError):
        Solver(g, inputs.Case(poly=1), accumulate=1, kernel_variant=1)
    with pytest.raises(FvkError):
        Solver(g, inputs.Case(poly=1), accumulate=2)
