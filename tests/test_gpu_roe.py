"""The Roe flux (flux_solver ROE, an extension) on the device.

The reference reads ROE but never built its Roe flux (res.cpp:66-68 leaves the
flux at -101), so the oracle's Roe flux, pinned by mathematics in
tests/test_oracle_roe.py, is the reference here:
  * FP64: state, stage-1 limited gradients and residual equal the oracle's bit
    for bit after 1, 2 and 40 steps, on Euler and NS, first and second order,
    every BC kind and cells of up to nine sides; in batches as single runs;
  * FP32-mixed: within 1e-5 (relative density L2) of FP64 Roe on cases
    without a limiter, and within 5 x the FP64 sensitivity floor + 1e-6 with
    SQUEEZE (the rule of tests/test_gpu_fp32.py);
  * the options it needs (the LDS-staged gather pass B) are checked.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, FvkError, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu

FVK_ERR_OPTION = -3


def roe_case(kind):
    if kind == "bump_o1_global":
        u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
        return gg.bump(48, 24), inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL", CFL=0.5,
                                            flux="ROE")
    if kind == "bump_o2_squeeze":
        u, v = inputs.mach_velocity(0.7, angle_deg=1.0)
        return gg.bump(64, 32), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                            dt_type="LOCAL", CFL=0.8, flux="ROE")
    if kind == "ramp_o2_squeeze":
        u, v = inputs.mach_velocity(2.5)
        return gg.ramp(96, 48), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                            dt_type="LOCAL", CFL=0.8, flux="ROE")
    if kind == "box_all_bc":
        # BC1 slip wall, BC2 freestream, extrapolation and no-slip sides
        u, v = inputs.mach_velocity(0.6, angle_deg=10.0)
        return (gg.box(14, 10, jitter=0.2, split_fraction=0.5, markers=(0, -1, -2, -3)),
                inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.5,
                            BC1="SLIP_WALL", BC2="FREESTREAM", flux="ROE"))
    if kind == "ogrid_ns_global":
        u, v = inputs.mach_velocity(0.2)
        return gg.ogrid(64, 32), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS",
                                             mu=1.225 * u / 100.0, BC1="NO_SLIP_WALL",
                                             dt_type="GLOBAL", CFL=0.5, flux="ROE")
    if kind == "airfoil_dual":
        u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
        return gg.airfoil_dual(20000), inputs.Case(u=u, v=v, poly=1, limiter="NONE",
                                                   dt_type="LOCAL", CFL=0.5, flux="ROE")
    raise ValueError(kind)


KINDS = ["bump_o1_global", "ramp_o2_squeeze", "box_all_bc", "ogrid_ns_global", "airfoil_dual"]


@pytest.mark.parametrize("kind", KINDS)
def test_roe_matches_oracle(kind):
    grid, case = roe_case(kind)
    print(f"[gpu] roe {kind} cells={grid.num_cells}", flush=True)
    s = Solver(grid, case)
    orc = O.Oracle(grid, case)
    s.init_freestream()
    done = 0
    for n in (1, 2, 40):
        s.step(n - done)
        orc.step(n - done)
        done = n
        q, gr = s.state_and_grad()
        assert np.isfinite(q).all()
        assert same(q, orc.state()), f"state step {n}: {mismatch(q, orc.state())}"
        assert same(gr, orc.grad()), f"grad step {n}: {mismatch(gr, orc.grad())}"
        assert same(s.residual(), orc.residual()), f"residual step {n}"
    # the Roe flux differs from the Rusanov flux on the same run
    r = Solver(grid, inputs.Case(**{**case.__dict__, "flux": "RUSANOV"}))
    r.init_freestream()
    r.step(40)
    assert not np.array_equal(r.state(), s.state())


def test_roe_batch_equals_single_runs():
    cases = [roe_case(k) for k in ("ramp_o2_squeeze", "box_all_bc")]
    b = Batch(cases)
    b.init()
    b.step(30)
    for i, (g, c) in enumerate(cases):
        s = Solver(g, c)
        s.init_freestream()
        s.step(30)
        assert same(b.state(i), s.state()), f"instance {i}"


def _rel(a, b):
    return float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))


# (not the M 2.5 ramp: from the perturbed freestream of the floor run its
# impulsive start breaks down in FP64 too, by step 60; Roe's flux does not
# keep pressures positive, and the Rusanov flux's larger dissipation does here)
@pytest.mark.parametrize("kind", ["bump_o1_global", "airfoil_dual", "bump_o2_squeeze",
                                  "ogrid_ns_global"])
def test_roe_fp32_mixed_against_fp64(kind):
    grid, case = roe_case(kind)
    steps = 200
    out = {}
    for prec in (0, FP32_MIXED):
        s = Solver(grid, case, precision=prec)
        s.init_freestream()
        s.step(steps)
        out[prec] = s.state()
    err = _rel(out[FP32_MIXED], out[0])
    if case.limiter == "NONE":
        assert err <= 1e-5, err
        return
    # SQUEEZE: the lagged limiter amplifies rounding-level differences; the
    # FP64 run from a freestream perturbed by 1e-7 sets the floor
    rng = np.random.default_rng(7)
    s = Solver(grid, case)
    u, v = case.u, case.v
    qinf = np.array([case.rho, case.rho * u, case.rho * v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (u * u + v * v)])
    s.set_state(qinf[:, None] * (1 + 1e-7 * rng.uniform(-1, 1, (4, grid.num_cells))))
    s.step(steps)
    floor = _rel(s.state(), out[0])
    print(f"[gpu] roe fp32 {kind}: err {err:.3e} floor {floor:.3e}", flush=True)
    assert err <= 5 * floor + 1e-6, (err, floor)


@pytest.mark.parametrize("over", [dict(accumulate=1), dict(kernel_variant=1),
                                  dict(kernel_variant=2)])
def test_roe_needs_the_lds_gather_shape(over):
    grid, case = roe_case("bump_o1_global")
    with pytest.raises(FvkError) as e:
        Solver(grid, case, **over)
    assert e.value.code == FVK_ERR_OPTION
