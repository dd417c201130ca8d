"""The Barth-Jespersen limiter (limiter 3, an extension) on the device.

The reference maps MINMOD to a no-op and never built its minmod.cpp draft, so
the oracle's Barth-Jespersen limiter, pinned in tests/test_oracle_bj.py by its
defining bound, is the reference here:
  * FP64: state, stage-1 limited gradients and residual equal the oracle's bit
    for bit after 1, 2 and 40 steps (Euler and NS, every BC kind, polygonal
    cells, the three launch shapes, with the Roe flux and the two-stage
    SSP-RK2 too); a batch mixing limiters per instance equals single runs;
  * FP32-mixed: within 5 x the FP64 sensitivity floor + 1e-6 of FP64.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu

BJ = "BARTH_JESPERSEN"


def bj_case(kind):
    if kind == "ramp":
        u, v = inputs.mach_velocity(2.5)
        return gg.ramp(96, 48), inputs.Case(u=u, v=v, poly=1, limiter=BJ, dt_type="LOCAL",
                                            CFL=0.8)
    if kind == "bump":
        u, v = inputs.mach_velocity(0.7, angle_deg=1.0)
        return gg.bump(64, 32), inputs.Case(u=u, v=v, poly=1, limiter=BJ, dt_type="GLOBAL",
                                            CFL=0.5)
    if kind == "box_all_bc":
        u, v = inputs.mach_velocity(0.6, angle_deg=10.0)
        return (gg.box(14, 10, jitter=0.2, split_fraction=0.5, markers=(0, -1, -2, -3)),
                inputs.Case(u=u, v=v, poly=1, limiter=BJ, dt_type="LOCAL", CFL=0.5,
                            BC1="SLIP_WALL", BC2="FREESTREAM"))
    if kind == "ogrid_ns":
        u, v = inputs.mach_velocity(0.2)
        return gg.ogrid(64, 32), inputs.Case(u=u, v=v, poly=1, limiter=BJ, eqn="NS",
                                             mu=1.225 * u / 100.0, BC1="NO_SLIP_WALL",
                                             dt_type="GLOBAL", CFL=0.5)
    if kind == "airfoil_dual":
        u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
        return gg.airfoil_dual(20000), inputs.Case(u=u, v=v, poly=1, limiter=BJ,
                                                   dt_type="LOCAL", CFL=0.5)
    if kind == "ramp_roe_rk2":
        u, v = inputs.mach_velocity(2.5)
        return gg.ramp(96, 48), inputs.Case(u=u, v=v, poly=1, limiter=BJ, dt_type="LOCAL",
                                            CFL=0.5, flux="ROE", scheme="SSP_RK2_TWO_STAGE")
    raise ValueError(kind)


@pytest.mark.parametrize("kind,variant", [
    ("ramp", 3), ("bump", 3), ("box_all_bc", 3), ("ogrid_ns", 3), ("airfoil_dual", 3),
    ("ramp_roe_rk2", 3), ("ramp", 1), ("ogrid_ns", 2), ("box_all_bc", 1)])
def test_bj_matches_oracle(kind, variant):
    grid, case = bj_case(kind)
    print(f"[gpu] bj {kind} variant={variant} cells={grid.num_cells}", flush=True)
    s = Solver(grid, case, kernel_variant=variant)
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
    # a different limiter from SQUEEZE on the same run
    r = Solver(grid, inputs.Case(**{**case.__dict__, "limiter": "SQUEEZE"}))
    r.init_freestream()
    r.step(40)
    assert not np.array_equal(r.state(), s.state())


def test_bj_batch_mixing_limiters_equals_single_runs():
    g, c = bj_case("ramp")
    cases = [(g, c), (g, inputs.Case(**{**c.__dict__, "limiter": "SQUEEZE"})),
             (gg.bump(48, 24), inputs.Case(**{**bj_case("bump")[1].__dict__, "dt_type": "LOCAL"})),
             (gg.bump(40, 20), inputs.Case(**{**bj_case("bump")[1].__dict__, "dt_type": "LOCAL",
                                              "limiter": "NONE"}))]
    b = Batch(cases)
    b.init()
    b.step(30)
    for i, (gi, ci) in enumerate(cases):
        s = Solver(gi, ci)
        s.init_freestream()
        s.step(30)
        assert same(b.state(i), s.state()), f"instance {i}"


def _rel(a, b):
    return float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))


@pytest.mark.parametrize("kind", ["bump", "ogrid_ns", "airfoil_dual"])
def test_bj_fp32_mixed_against_fp64(kind):
    grid, case = bj_case(kind)
    steps = 200
    out = {}
    for prec in (0, FP32_MIXED):
        s = Solver(grid, case, precision=prec)
        s.init_freestream()
        s.step(steps)
        out[prec] = s.state()
    err = _rel(out[FP32_MIXED], out[0])
    rng = np.random.default_rng(7)
    s = Solver(grid, case)
    u, v = case.u, case.v
    qinf = np.array([case.rho, case.rho * u, case.rho * v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (u * u + v * v)])
    s.set_state(qinf[:, None] * (1 + 1e-7 * rng.uniform(-1, 1, (4, grid.num_cells))))
    s.step(steps)
    floor = _rel(s.state(), out[0])
    print(f"[gpu] bj fp32 {kind}: err {err:.3e} floor {floor:.3e}", flush=True)
    assert err <= 5 * floor + 1e-6, (err, floor)
