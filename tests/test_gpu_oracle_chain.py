"""The batched form, the atomic-scatter accumulation and the 256-instance batch
workload against the oracle, and closed-form checks on polygonal cells.

* A batch of downsized batch-workload instances (bump channels and ramps of
  mixed size, Mach number and flow angle; second order, SQUEEZE, LOCAL or
  GLOBAL dt), stepped 100 steps with two instances (a channel and a ramp)
  retired after 40: every instance's state and stage-1 limited gradients
  equal the oracle's after its own step count, bit for bit, and its residual
  norm agrees to 1e-12.  The
  oracle is the reference restated (oracle/fv_oracle.c, pinned bit for bit to
  the reference's own outputs), so this closes batch = reference.
* The atomic-scatter accumulation (accumulate = 1) in FP64, single and
  batched: within 1e-10 relative of the oracle after 100 steps where the
  scheme is well conditioned (first order, or no limiter): its sums differ
  from the reference's by summation order only.
* The full 256-instance batch workload (11.3M cells), FP64 and FP32-mixed,
  50 steps: every state finite; three instances (smallest, median, largest)
  equal single-context runs bit for bit, and the smallest and median equal
  the oracle; every instance's FP32-mixed density and pressure errors against
  FP64 within the stated rule (5 x the FP64 sensitivity floor + 1e-6).
* Polygonal cells (the Delaunay dual, 4 to 9 sides): a uniform flow with
  freestream boundaries everywhere (the body included) stays uniform, and the
  least-squares gradient of a linear field is exact on interior cells.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs
from fv2d import workloads as W
from fv2d.lib import FP32_MIXED, Batch, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu


def _workload_instances(n=8, dt_type="LOCAL"):
    """Downsized instances of the batch workload: same generators, Mach and
    angle draws and options, 600 to 3000 cells each."""
    plan = [W.build_instance(i) for i in W.batch_plan(n, lo=600, hi=3000)]
    out = []
    for i in plan:
        case = i.case if dt_type == "LOCAL" else inputs.Case(**{**i.case.__dict__,
                                                               "dt_type": dt_type})
        out.append((i.grid, case))
    return out


@pytest.mark.parametrize("dt_type", ["LOCAL", "GLOBAL"])
def test_batch_fp64_equals_oracle_with_retirement(dt_type):
    cases = _workload_instances(dt_type=dt_type)
    kinds = sorted({c.u > 400 for _, c in cases})
    assert kinds == [False, True], "both channel and ramp instances"
    print(f"[gpu] batch vs oracle {dt_type}: cells {[g.num_cells for g, _ in cases]}", flush=True)
    b = Batch(cases)
    b.init()
    b.step(40)
    active = np.ones(len(cases), dtype=np.uint8)
    active[[1, 6]] = 0     # a channel and a ramp instance
    b.set_active(active)
    b.step(60)
    r = b.residual_l2()
    for i, (g, c) in enumerate(cases):
        n = 100 if active[i] else 40
        orc = O.Oracle(g, c)
        orc.step(n)
        q, gr = b.state_and_grad(i)
        assert same(q, orc.state()), f"instance {i} ({n} steps): {mismatch(q, orc.state())}"
        assert same(gr, orc.grad()), f"instance {i} grad: {mismatch(gr, orc.grad())}"
        assert np.allclose(r[i], orc.residual_l2(), rtol=1e-12, atol=0), (i, r[i])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _well_conditioned_cases():
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


@pytest.mark.parametrize("name,grid,case", list(_well_conditioned_cases()),
                         ids=lambda x: x if isinstance(x, str) else "")
def test_atomic_fp64_within_summation_order_of_oracle(name, grid, case):
    print(f"[gpu] atomic FP64 vs oracle: {name}", flush=True)
    s = Solver(grid, case, accumulate=1)
    assert s.launch_options()["accumulate"] == 1
    s.init_freestream()
    s.step(100)
    orc = O.Oracle(grid, case)
    orc.step(100)
    q = s.state()
    assert np.isfinite(q).all()
    assert _rel(q, orc.state()) <= 1e-10, (name, _rel(q, orc.state()))


def test_atomic_fp64_batch_within_summation_order_of_oracle():
    cases = [(g, c) for _, g, c in _well_conditioned_cases() if c.eqn == "EULER" and c.poly == 1]
    assert len(cases) == 2
    b = Batch(cases, accumulate=1)
    b.init()
    b.step(100)
    for i, (g, c) in enumerate(cases):
        orc = O.Oracle(g, c)
        orc.step(100)
        assert _rel(b.state(i), orc.state()) <= 1e-10, (i, _rel(b.state(i), orc.state()))


# ---------------------------------------------------------------------------
# the 256-instance batch workload at full size
# ---------------------------------------------------------------------------
FLOOR_FACTOR = 5.0
PERTURBATION = 1e-7
FULL_STEPS = 50


def _pressure(q, gamma):
    rho, mx, my, E = q
    return (E - 0.5 * (mx * mx + my * my) / rho) * (gamma - 1)


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def batch256():
    name, plan = W.batch(256)
    return [(i.grid, i.case) for i in plan]


def test_batch256_fp64_properties(batch256):
    cases = batch256
    nc = np.array([g.num_cells for g, _ in cases])
    print(f"[gpu] batch256 FP64: {int(nc.sum())} cells, {FULL_STEPS} steps", flush=True)
    b = Batch(cases)
    b.init()
    b.step(FULL_STEPS)
    r = b.residual_l2()
    assert np.isfinite(r).all()
    order = np.argsort(nc)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for i in range(len(cases)):
        assert np.isfinite(b.state(i)).all(), i
    for j, i in enumerate(pick):
        g, c = cases[i]
        q, gr = b.state_and_grad(i)
        s = Solver(g, c)
        s.init_freestream()
        s.step(FULL_STEPS)
        q1, g1 = s.state_and_grad()
        assert same(q, q1) and same(gr, g1), f"instance {i} ({g.num_cells} cells)"
        s.close()
        if j < 2:
            orc = O.Oracle(g, c)
            orc.step(FULL_STEPS)
            assert same(q, orc.state()), f"instance {i} vs oracle: {mismatch(q, orc.state())}"
            assert same(gr, orc.grad()), f"instance {i} grad vs oracle"
    b.close()


def test_batch256_fp32_mixed_error_rule(batch256):
    cases = batch256
    print(f"[gpu] batch256 FP32-mixed vs FP64, {FULL_STEPS} steps", flush=True)
    runs = {}
    for prec in (0, FP32_MIXED):
        b = Batch(cases, precision=prec)
        b.init()
        b.step(FULL_STEPS)
        runs[prec] = [b.state(i) for i in range(len(cases))]
        b.close()
    # the FP64 scheme's own response to a 1e-7 relative perturbation of its start
    rng = np.random.default_rng(11)
    starts = []
    for g, c in cases:
        qinf = np.array([c.rho, c.rho * c.u, c.rho * c.v,
                         c.P / (c.gamma - 1) + 0.5 * c.rho * (c.u ** 2 + c.v ** 2)])
        starts.append(qinf[:, None] * (1.0 + PERTURBATION * rng.uniform(-1, 1, (4, g.num_cells))))
    b = Batch(cases)
    b.init(starts)
    b.step(FULL_STEPS)
    floor_q = [b.state(i) for i in range(len(cases))]
    b.close()
    bad = []
    errs = []
    for i, (g, c) in enumerate(cases):
        q32, q64, qf = runs[FP32_MIXED][i], runs[0][i], floor_q[i]
        assert np.isfinite(q32).all(), i
        e = {"rho": _rel_l2(q32[0], q64[0]),
             "p": _rel_l2(_pressure(q32, c.gamma), _pressure(q64, c.gamma))}
        f = {"rho": _rel_l2(qf[0], q64[0]),
             "p": _rel_l2(_pressure(qf, c.gamma), _pressure(q64, c.gamma))}
        errs.append(e["rho"])
        for k in ("rho", "p"):
            if e[k] > FLOOR_FACTOR * f[k] + 1e-6:
                bad.append((i, k, e[k], f[k]))
    print(f"[gpu] batch256 FP32 rho L2 error median {np.median(errs):.3e} max {max(errs):.3e}",
          flush=True)
    assert not bad, bad[:10]


# ---------------------------------------------------------------------------
# polygonal cells (4 to 9 sides): closed forms
# ---------------------------------------------------------------------------
def _dual_case(**kw):
    u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
    base = dict(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.5,
                BC1="FREESTREAM", BC2="FREESTREAM")
    base.update(kw)
    return gg.airfoil_dual(40000), inputs.Case(**base)


@pytest.mark.parametrize("precision", [0, FP32_MIXED])
def test_dual_freestream_preserved_on_polygons(precision):
    """Uniform flow, freestream boundaries on the body and outside: the
    residual of every cell is a sum of identical face fluxes over a closed
    polygon, zero up to the rounding of the normals (FP64), and the FP32-mixed
    perturbation state stays exactly zero."""
    grid, case = _dual_case()
    deg = np.bincount(grid.face_c1, minlength=grid.num_cells)
    deg += np.bincount(grid.face_c2[grid.face_c2 >= 0], minlength=grid.num_cells)
    assert (deg > 6).any()
    s = Solver(grid, case, precision=precision)
    s.init_freestream()
    s.step(50)
    q = s.state()
    qinf = np.array([case.rho, case.rho * case.u, case.rho * case.v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (case.u ** 2 + case.v ** 2)])
    if precision == FP32_MIXED:
        assert np.array_equal(q, np.broadcast_to(qinf[:, None], q.shape))
    else:
        assert np.allclose(q, qinf[:, None], rtol=1e-11, atol=0)


def test_dual_lsq_gradient_exact_for_linear_field():
    """The unweighted least-squares gradient (cell_gradient.cpp:9-131)
    reproduces a linear field's gradient exactly on every interior cell,
    whatever its number of sides: the stage-1 gradient returned after one
    step from a linear state (no limiter) equals the field's gradient."""
    grid, case = _dual_case(limiter="NONE", BC1="SLIP_WALL")
    s = Solver(grid, case)
    cx, cy = s.geometry("cell_centerx"), s.geometry("cell_centery")
    qinf = np.array([case.rho, case.rho * case.u, case.rho * case.v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (case.u ** 2 + case.v ** 2)])
    a = np.array([0.02, -0.03, 0.05, 0.01])
    bcoef = np.array([-0.04, 0.02, 0.03, -0.02])
    q = qinf[:, None] * (1.0 + a[:, None] * cx + bcoef[:, None] * cy)
    s.set_state(q)
    s.step(1)
    _, gr = s.state_and_grad()
    interior = np.ones(grid.num_cells, bool)
    interior[grid.face_c1[grid.face_c2 < 0]] = False
    deg = np.bincount(grid.face_c1, minlength=grid.num_cells)
    deg += np.bincount(grid.face_c2[grid.face_c2 >= 0], minlength=grid.num_cells)
    assert (deg[interior] > 6).any()
    gx_true = qinf * a
    gy_true = qinf * bcoef
    for k in range(4):
        assert np.allclose(gr[k][interior], gx_true[k], rtol=1e-9, atol=0), k
        assert np.allclose(gr[4 + k][interior], gy_true[k], rtol=1e-9, atol=0), k
