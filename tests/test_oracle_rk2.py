"""The two-stage SSP-RK2 (time_scheme 3, an extension) of the oracle, pinned by
what the mathematics fixes.

The reference reads SSP_RK2 but runs SSP-RK3 for it (main.cpp:123-126); its
ssp_rk2.cpp is a draft outside its solver library.  time_scheme 1 keeps that
behaviour; time_scheme 3 runs the two-stage scheme of the draft's stage
structure.  Pinned by:
  * order in time: with a fixed dt on a smooth unsteady flow (no limiter), the
    differences between runs at dt, dt/2 and dt/4 to the same time shrink by
    4x per halving (2nd order), and the SSP-RK3's by 8x (3rd order);
  * conservation in a closed box under the uniform time step, and
    freestream preservation;
  * time_scheme 1 is SSP-RK3, as in the reference.
The device path is held to the oracle bit for bit in tests/test_gpu_rk2.py.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _pulse_case(scheme):
    g = gg.box(24, 24, jitter=0.15, split_fraction=0.3, markers=(-1, -1, -1, -1))
    u, v = inputs.mach_velocity(0.3, angle_deg=20)
    case = inputs.Case(u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.5,
                       BC1="FREESTREAM", BC2="FREESTREAM", scheme=scheme)
    return g, case


def _pulse_state(o, case):
    """freestream with a smooth pressure and density bump in the middle"""
    cx, cy = o.get("geom/cell_centerx"), o.get("geom/cell_centery")
    x0, y0 = 0.5 * (cx.min() + cx.max()), 0.5 * (cy.min() + cy.max())
    w = 0.15 * (cx.max() - cx.min())
    b = np.exp(-((cx - x0) ** 2 + (cy - y0) ** 2) / w ** 2)
    rho = case.rho * (1 + 0.05 * b)
    P = case.P * (1 + 0.07 * b)
    u, v = case.u, case.v
    return np.array([rho, rho * u, rho * v, P / (case.gamma - 1) + 0.5 * rho * (u * u + v * v)])


def _run(scheme, dt, steps):
    g, case = _pulse_case(scheme)
    o = O.Oracle(g, case)
    q0 = _pulse_state(o, case)
    o.set_state(q0)
    o.set_fixed_dt(dt)
    o.step(steps)
    return o.state()


def _dt0():
    g, case = _pulse_case("SSP_RK3")
    o = O.Oracle(g, case)
    o.set_state(_pulse_state(o, case))
    o.step(1)
    return 0.5 * o.get("dt").min()


@pytest.mark.parametrize("scheme,order", [("SSP_RK2_TWO_STAGE", 2), ("SSP_RK3", 3)])
def test_order_in_time(scheme, order):
    dt = _dt0()
    q1, q2, q4 = (_run(scheme, dt / m, 16 * m) for m in (1, 2, 4))
    e1 = np.linalg.norm(q1 - q2)
    e2 = np.linalg.norm(q2 - q4)
    ratio = e1 / e2
    print(f"{scheme}: e(dt)={e1:.3e} e(dt/2)={e2:.3e} ratio {ratio:.2f}")
    assert abs(np.log2(ratio) - order) < 0.25, ratio


def test_rk2_conserves_in_a_closed_box():
    g = gg.box(20, 16, jitter=0.25, split_fraction=0.5, markers=(0, 0, 0, 0))
    u, v = inputs.mach_velocity(0.4, angle_deg=30)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="UNIFORM", CFL=0.5,
                       BC1="SLIP_WALL", BC2="FREESTREAM", scheme="SSP_RK2_TWO_STAGE")
    o = O.Oracle(g, case)
    vol = o.get("geom/cell_vol")
    q0 = o.state()
    o.step(100)
    q = o.state()
    for k in (0, 3):
        assert abs((vol * q[k]).sum() / (vol * q0[k]).sum() - 1) < 1e-13
    assert np.abs(q[0] - q0[0]).max() > 1e-3


def test_rk2_preserves_the_freestream():
    g = gg.box(16, 12, jitter=0.3, split_fraction=0.5, markers=(0, -1, 0, -1))
    u, v = inputs.mach_velocity(0.7, angle_deg=20)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8,
                       BC1="FREESTREAM", BC2="FREESTREAM", scheme="SSP_RK2_TWO_STAGE")
    o = O.Oracle(g, case)
    q0 = o.state()
    o.step(50)
    assert np.abs(o.state() - q0).max() <= 1e-12 * np.abs(q0).max()


def test_scheme_1_runs_ssp_rk3_as_the_reference():
    g, case = _pulse_case("SSP_RK2")
    a = O.Oracle(g, case)
    b = O.Oracle(g, inputs.Case(**{**case.__dict__, "scheme": "SSP_RK3"}))
    for o in (a, b):
        o.set_state(_pulse_state(o, case))
        o.step(10)
    assert np.array_equal(a.state(), b.state())
