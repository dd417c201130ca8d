"""The FP32-mixed precision form (FVK_FP32_MIXED) against the FP64 path.

The FP64 path is bit-identical to the reference (test_gpu_parity.py), so it
is the FP64 oracle here.  The FP32-mixed path stores dQ = Q - Q_inf in FP32
and evaluates face states, fluxes and the stage update in FP32 perturbation
form; its results differ from FP64 by rounding, so every check states its
tolerance.

Where the reference scheme is well conditioned (first order, or second order
without a limiter) the FP32-mixed density and pressure relative L2 errors
must be <= 1e-5.  With the SQUEEZE limiter the reference scheme itself is not:
its lagged limiter (squeeze.cpp:68-121) divides two differences that are both
rounding noise wherever a cell's state has not yet changed but a neighbour's
has, so the FP64 solution moves by ~1e-3 (L2) in response to a 1e-7 relative
perturbation of its initial state, and does so at convergence too.  There the
FP32-mixed error must stay within a small multiple of that FP64 sensitivity,
measured in the same test ("floor").

Exact properties: a uniform freestream with freestream/extrapolation
boundaries stays exactly uniform (dQ = 0 is a fixed point); the state round
trip is Q_inf + float(Q - Q_inf); the results do not depend on the launch
shape, cell renumbering or block size (bit for bit: per-cell sums keep the
face order in every shape).
"""
import numpy as np
import pytest

from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Solver
from fvtest import Golden, golden_names, same

pytestmark = pytest.mark.gpu

RHO_L2_TOL = 1e-5        # well-conditioned cases: density/pressure relative L2
LINF_TOL = 1e-4
FLOOR_FACTOR = 5.0       # SQUEEZE cases: error <= FLOOR_FACTOR * floor + 1e-6
PERTURBATION = 1e-7      # relative size of the floor run's initial perturbation


def pressure(q, gamma):
    rho, mx, my, E = q
    return (E - 0.5 * (mx * mx + my * my) / rho) * (gamma - 1)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_linf(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def errors(q32, q64, gamma):
    p32, p64 = pressure(q32, gamma), pressure(q64, gamma)
    return {"rho_l2": rel_l2(q32[0], q64[0]), "rho_linf": rel_linf(q32[0], q64[0]),
            "p_l2": rel_l2(p32, p64), "p_linf": rel_linf(p32, p64)}


def check(err, l2=RHO_L2_TOL, linf=LINF_TOL):
    assert err["rho_l2"] <= l2 and err["p_l2"] <= l2, err
    assert err["rho_linf"] <= linf and err["p_linf"] <= linf, err


def qinf_of(case):
    return np.array([case.rho, case.rho * case.u, case.rho * case.v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (case.u ** 2 + case.v ** 2)])


def floor_error(grid, case, steps, q0=None):
    """FP64 sensitivity: FP64 from q0 * (1 + 1e-7 u), u uniform in [-1, 1],
    against FP64 from q0 (both started through fvk_set_state)."""
    if q0 is None:
        q0 = np.broadcast_to(qinf_of(case)[:, None], (4, grid.num_cells)).copy()
    u = np.random.default_rng(7).uniform(-1.0, 1.0, q0.shape)
    out = []
    for q in (q0, q0 * (1.0 + PERTURBATION * u)):
        s = Solver(grid, case)
        s.set_state(q)
        s.step(steps)
        out.append(s.state())
    return errors(out[1], out[0], case.gamma)


def check_vs_floor(err, floor):
    for k in ("rho_l2", "p_l2"):
        assert err[k] <= FLOOR_FACTOR * floor[k] + 1e-6, (k, err, floor)


def well_conditioned(case):
    return case.poly == 0 or case.limiter != "SQUEEZE"


def _start(s, g):
    if g.meta["state"]:
        s.set_state(g.a["init/state"])
    else:
        s.init_freestream()


@pytest.mark.parametrize("name", [n for n in golden_names() if not Golden(n).meta["res_only"]])
def test_fp32_error_on_fixture(name):
    g = Golden(name)
    s = Solver(g.grid, g.case, precision=FP32_MIXED)
    _start(s, g)
    n = max(g.meta["dumps"])
    s.step(n)
    q = s.state()
    assert np.isfinite(q).all()
    err = errors(q, g.state(n), g.case.gamma)
    if well_conditioned(g.case):
        print(f"[gpu] fp32 fixture {name} step {n}: {err}", flush=True)
        check(err)
    else:
        floor = floor_error(g.grid, g.case, n, g.a["init/state"] if g.meta["state"] else None)
        print(f"[gpu] fp32 fixture {name} step {n}: {err} floor {floor}", flush=True)
        check_vs_floor(err, floor)


def _bench_case(kind):
    if kind == "bump":
        u, v = inputs.mach_velocity(0.5)
        return gg.bump(64, 32), inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL", CFL=0.5)
    if kind == "ramp":
        u, v = inputs.mach_velocity(2.5)
        return gg.ramp(512, 256), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                              dt_type="LOCAL", CFL=0.8)
    if kind == "ogrid":
        u, v = inputs.mach_velocity(0.2)
        return gg.ogrid(1024, 512), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS",
                                                mu=1.225 * u / 100.0, BC1="NO_SLIP_WALL",
                                                dt_type="GLOBAL", CFL=0.5)
    if kind == "airfoil":
        # the polygonal Delaunay-dual workload at 1M points (979k cells of 4 to
        # 9 sides), its options unchanged
        u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
        return gg.airfoil_dual(1000000), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                                     dt_type="LOCAL", CFL=0.5)
    raise ValueError(kind)


# bump: the 2000-iteration first-order channel case, and the same channel at
# second order without a limiter; ramp and ogrid: the second-order SQUEEZE
# benchmark meshes at full size; airfoil: the polygonal dual at 1M points
@pytest.mark.parametrize("kind,steps", [("bump", 2000), ("bump_o2", 2000), ("ramp", 500),
                                        ("ogrid", 200), ("airfoil", 200)])
def test_fp32_error_vs_fp64(kind, steps):
    if kind == "bump_o2":
        grid, case = _bench_case("bump")
        case = inputs.Case(**{**case.__dict__, "poly": 1, "limiter": "NONE", "dt_type": "LOCAL",
                              "CFL": 0.8})
    else:
        grid, case = _bench_case(kind)
    print(f"[gpu] fp32 vs fp64 {kind} cells={grid.num_cells} steps={steps}", flush=True)
    out = {}
    for prec in (0, FP32_MIXED):
        s = Solver(grid, case, precision=prec)
        s.init_freestream()
        s.step(steps)
        out[prec] = s.state()
    assert np.isfinite(out[FP32_MIXED]).all()
    err = errors(out[FP32_MIXED], out[0], case.gamma)
    if well_conditioned(case):
        print(f"[gpu] fp32 vs fp64 {kind} step {steps}: {err}", flush=True)
        check(err)
    else:
        floor = floor_error(grid, case, steps)
        print(f"[gpu] fp32 vs fp64 {kind} step {steps}: {err} floor {floor}", flush=True)
        check_vs_floor(err, floor)


def test_fp32_freestream_is_a_fixed_point():
    grid = gg.box(12, 9, jitter=0.2, split_fraction=0.5, markers=(0, -1, -2, -1))
    u, v = inputs.mach_velocity(0.7, angle_deg=3.0)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8,
                       BC1="FREESTREAM", BC2="FREESTREAM")
    s = Solver(grid, case, precision=FP32_MIXED)
    s.init_freestream()
    s.step(50)
    q = s.state()
    assert np.array_equal(q, np.broadcast_to(qinf_of(case)[:, None], q.shape))
    assert not s.residual().any()


def test_fp32_state_round_trip():
    g = Golden("bump_restart_state")
    s = Solver(g.grid, g.case, precision=FP32_MIXED)
    q0 = g.a["init/state"]
    s.set_state(q0)
    qinf = qinf_of(g.case)
    expect = qinf[:, None] + (q0 - qinf[:, None]).astype(np.float32).astype(np.float64)
    assert same(s.state(), expect)


def test_fp32_residual_norm():
    grid, case = _bench_case("bump")
    s = Solver(grid, case, precision=FP32_MIXED)
    s.init_freestream()
    s.step(20)
    r = s.residual()
    ref = np.sqrt((r ** 2).sum(axis=1))
    assert np.allclose(s.residual_l2(), ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("overrides", [{"kernel_variant": 1}, {"kernel_variant": 2},
                                       {"renumber": 0}, {"renumber": 2}, {"block_size": 128},
                                       {"xcd_remap": 0}, {"xcd_remap": 8},
                                       {"block_size": 128, "xcd_remap": 5}])
def test_fp32_layout_invariance(overrides):
    grid = gg.ramp(96, 48)
    u, v = inputs.mach_velocity(2.5)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8)
    out = []
    for ov in ({}, overrides):
        s = Solver(grid, case, precision=FP32_MIXED, **ov)
        s.init_freestream()
        s.step(30)
        out.append(s.state_and_grad())
    assert same(out[0][0], out[1][0]) and same(out[0][1], out[1][1]), overrides


def test_fp32_ns_layout_invariance():
    grid = gg.ogrid(64, 32)
    u, v = inputs.mach_velocity(0.2)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS", mu=1.225 * u / 100.0,
                       BC1="NO_SLIP_WALL", dt_type="GLOBAL", CFL=0.5)
    out = []
    for ov in ({}, {"kernel_variant": 1}):
        s = Solver(grid, case, precision=FP32_MIXED, **ov)
        s.init_freestream()
        s.step(30)
        out.append(s.state())
    assert same(out[0], out[1])


def test_fp32_lsq_inverses_in_fp32(monkeypatch):
    """On a well-conditioned mesh the FP32-mixed form reads the LSQ inverses
    rounded to FP32 (the products stay FP64); FVK_LSQ_FP64=1 keeps them FP64.
    Both stay within FP32 rounding of each other and of the FP64 path."""
    u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
    grid = gg.bump(64, 32)
    case = inputs.Case(u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.8)
    out = {}
    for force in ("0", "1"):
        monkeypatch.setenv("FVK_LSQ_FP64", force)
        s = Solver(grid, case, precision=FP32_MIXED)
        s.init_freestream()
        s.step(200)
        out[force] = s.state()
    s = Solver(grid, case)
    s.init_freestream()
    s.step(200)
    q64 = s.state()
    rel = lambda a, b: float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))
    assert 0 < rel(out["0"], out["1"]) < 1e-6
    assert rel(out["0"], q64) < 1e-6 and rel(out["1"], q64) < 1e-6
