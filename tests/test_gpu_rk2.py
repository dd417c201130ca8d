"""The two-stage SSP-RK2 (time_scheme 3, an extension) on the device.

The reference runs SSP-RK3 for SSP_RK2 (main.cpp:123-126) and never built its
ssp_rk2.cpp draft, so the oracle's two-stage scheme, pinned by mathematics in
tests/test_oracle_rk2.py (second order in time, conservation, freestream), is
the reference here:
  * FP64: state, stage-1 limited gradients and residual equal the oracle's bit
    for bit after 1, 2 and 40 steps (Euler and NS, first and second order,
    GLOBAL, LOCAL and uniform dt, polygonal cells, every launch shape, with the
    Roe flux too); graph replay, kernel timing and checkpoints leave runs
    unchanged; batches equal single runs;
  * FP32-mixed: within 1e-5 (relative density L2) of FP64 without a limiter,
    within 5 x the FP64 sensitivity floor + 1e-6 with SQUEEZE.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, FvkError, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu

RK2 = "SSP_RK2_TWO_STAGE"


def rk2_case(kind):
    if kind == "bump_o1_global":
        u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
        return gg.bump(48, 24), inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL", CFL=0.5,
                                            scheme=RK2)
    if kind == "ramp_o2_squeeze":
        u, v = inputs.mach_velocity(2.5)
        return gg.ramp(96, 48), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                            dt_type="LOCAL", CFL=0.8, scheme=RK2)
    if kind == "bump_o2_squeeze":
        u, v = inputs.mach_velocity(0.7, angle_deg=1.0)
        return gg.bump(64, 32), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                            dt_type="LOCAL", CFL=0.8, scheme=RK2)
    if kind == "ogrid_ns_global":
        u, v = inputs.mach_velocity(0.2)
        return gg.ogrid(64, 32), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS",
                                             mu=1.225 * u / 100.0, BC1="NO_SLIP_WALL",
                                             dt_type="GLOBAL", CFL=0.5, scheme=RK2)
    if kind == "box_uniform":
        u, v = inputs.mach_velocity(0.4, angle_deg=30)
        return (gg.box(20, 16, jitter=0.25, split_fraction=0.5, markers=(0, 0, 0, 0)),
                inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="UNIFORM", CFL=0.5,
                            BC1="SLIP_WALL", BC2="FREESTREAM", scheme=RK2))
    if kind == "airfoil_dual":
        u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
        return gg.airfoil_dual(20000), inputs.Case(u=u, v=v, poly=1, limiter="NONE",
                                                   dt_type="LOCAL", CFL=0.5, scheme=RK2)
    if kind == "ramp_roe":
        u, v = inputs.mach_velocity(2.5)
        return gg.ramp(96, 48), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                            dt_type="LOCAL", CFL=0.5, scheme=RK2, flux="ROE")
    raise ValueError(kind)


@pytest.mark.parametrize("kind,variant", [
    ("bump_o1_global", 3), ("ramp_o2_squeeze", 3), ("ogrid_ns_global", 3), ("box_uniform", 3),
    ("airfoil_dual", 3), ("ramp_roe", 3), ("ramp_o2_squeeze", 1), ("ogrid_ns_global", 2)])
def test_rk2_matches_oracle(kind, variant):
    grid, case = rk2_case(kind)
    print(f"[gpu] rk2 {kind} variant={variant} cells={grid.num_cells}", flush=True)
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
    # a different scheme from SSP-RK3 on the same run
    r = Solver(grid, inputs.Case(**{**case.__dict__, "scheme": "SSP_RK3"}))
    r.init_freestream()
    r.step(40)
    assert not np.array_equal(r.state(), s.state())


@pytest.mark.parametrize("kind", ["ramp_o2_squeeze", "ogrid_ns_global"])
def test_rk2_graph_timing_and_checkpoint_leave_runs_unchanged(kind):
    grid, case = rk2_case(kind)
    ref = Solver(grid, case)
    ref.use_graph(False)
    ref.init_freestream()
    ref.step(40)
    # graph replay over the whole run
    a = Solver(grid, case)
    a.init_freestream()
    a.step(40)
    assert same(a.state(), ref.state())
    # kernel timing in the middle (it writes buffers the two-stage scheme
    # does not use), then a checkpoint taken and resumed in a new context
    b = Solver(grid, case)
    b.init_freestream()
    b.step(15)
    b.time_kernels(5)
    b.step(10)
    ck = b.checkpoint()
    c = Solver(grid, case)
    c.restore(ck)
    c.step(15)
    qa, ga = ref.state_and_grad()
    qc, gc = c.state_and_grad()
    assert same(qc, qa) and same(gc, ga)
    assert same(c.residual(), ref.residual())


def test_rk2_batch_equals_single_runs_and_schemes_must_agree():
    cases = [rk2_case(k) for k in ("ramp_o2_squeeze", "bump_o2_squeeze")]
    b = Batch(cases)
    b.init()
    b.step(30)
    for i, (g, c) in enumerate(cases):
        s = Solver(g, c)
        s.init_freestream()
        s.step(30)
        assert same(b.state(i), s.state()), f"instance {i}"
    g, c = cases[1]
    with pytest.raises(FvkError):
        Batch([cases[0], (g, inputs.Case(**{**c.__dict__, "scheme": "SSP_RK3"}))])


def _rel(a, b):
    return float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))


@pytest.mark.parametrize("kind", ["bump_o1_global", "airfoil_dual", "bump_o2_squeeze",
                                  "ogrid_ns_global"])
def test_rk2_fp32_mixed_against_fp64(kind):
    grid, case = rk2_case(kind)
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
    rng = np.random.default_rng(7)
    s = Solver(grid, case)
    u, v = case.u, case.v
    qinf = np.array([case.rho, case.rho * u, case.rho * v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (u * u + v * v)])
    s.set_state(qinf[:, None] * (1 + 1e-7 * rng.uniform(-1, 1, (4, grid.num_cells))))
    s.step(steps)
    floor = _rel(s.state(), out[0])
    print(f"[gpu] rk2 fp32 {kind}: err {err:.3e} floor {floor:.3e}", flush=True)
    assert err <= 5 * floor + 1e-6, (err, floor)
