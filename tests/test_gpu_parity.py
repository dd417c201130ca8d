"""FP64 parity of the HIP path (through the C ABI) with the reference.

* Every golden fixture: the device state, stage-1 limited gradients and
  stage-3 residual after each dumped step equal the reference's outputs as
  IEEE values (bit-identical up to the sign of zero).
* The benchmark-sized meshes (bump 64x32, ramp 512x256, O-grid 1024x512,
  100 steps each) and a 196k-cell polygonal airfoil dual against the oracle
  on the same inputs.
* Size-independent properties: results do not depend on cell renumbering,
  block size, the XCD block order or the launch shape (one workgroup per
  block, or the persistent grid of kernel_variant 4); a uniform freestream with only
  freestream/extrapolation boundaries stays uniform; conservation.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, FvkError, Solver
from fvtest import Golden, golden_names, mismatch, same

pytestmark = pytest.mark.gpu


def _start(solver, orc, g: Golden):
    if g.meta["state"]:
        solver.set_state(g.a["init/state"])
        if orc is not None:
            orc.set_state(g.a["init/state"])
    else:
        solver.init_freestream()


@pytest.mark.parametrize("variant", [0, 4])
@pytest.mark.parametrize("name", golden_names())
def test_matches_reference_fixture(name, variant):
    g = Golden(name)
    print(f"[gpu] fixture {name} cells={g.grid.num_cells} variant={variant}", flush=True)
    s = Solver(g.grid, g.case, kernel_variant=variant)
    if g.meta["res_only"]:
        # the residual operator is pinned through one step from the fixture
        # state, against the oracle (itself pinned to the reference's res())
        orc = O.Oracle(g.grid, g.case)
        _start(s, orc, g)
        s.step(1)
        orc.step(1)
        assert same(s.state(), orc.state()), mismatch(s.state(), orc.state())
        assert same(s.residual(), orc.residual()), mismatch(s.residual(), orc.residual())
        return
    _start(s, None, g)
    done = 0
    for n in sorted(g.meta["dumps"]):
        s.step(n - done)
        done = n
        q, gr = s.state_and_grad()
        assert same(q, g.state(n)), f"state step {n}: {mismatch(q, g.state(n))}"
        assert same(gr, g.grad(n)), f"grad step {n}: {mismatch(gr, g.grad(n))}"
        r = s.residual()
        assert same(r, g.residual(n)), f"residual step {n}: {mismatch(r, g.residual(n))}"
        l2 = s.residual_l2()
        ref = np.sqrt((g.residual(n) ** 2).sum(axis=1))
        assert np.allclose(l2, ref, rtol=1e-12, atol=0, equal_nan=True), (l2, ref)


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
        # polygonal cells of 4 to 9 sides (the reference itself stops at 6;
        # the oracle keeps its face loops for any count)
        u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
        return gg.airfoil_dual(200000), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                                    dt_type="LOCAL", CFL=0.5)
    raise ValueError(kind)


_ORACLE_RUNS = {}


def _oracle_run(kind, steps):
    """The oracle's state, gradients and residual after `steps` steps (one
    run per workload, shared by the launch shapes)."""
    key = (kind, steps)
    if key not in _ORACLE_RUNS:
        _ORACLE_RUNS.clear()
        grid, case = _bench_case(kind)
        orc = O.Oracle(grid, case)
        orc.step(steps)
        _ORACLE_RUNS[key] = (grid, case, orc.state(), orc.grad(), orc.residual())
    return _ORACLE_RUNS[key]


# the oracle's 100 steps of the 524k-cell NS O-grid take about 90 s on one core
@pytest.mark.timeout(900)
@pytest.mark.parametrize("variant", [3, 4, 1])
@pytest.mark.parametrize("kind,steps", [("bump", 100), ("ramp", 100), ("ogrid", 100),
                                        ("airfoil", 100)])
def test_full_size_against_oracle(kind, steps, variant):
    grid, case, q0, g0, r0 = _oracle_run(kind, steps)
    print(f"[gpu] full size {kind} cells={grid.num_cells} steps={steps} variant={variant}", flush=True)
    s = Solver(grid, case, kernel_variant=variant)
    s.init_freestream()
    s.step(steps)
    q, gr = s.state_and_grad()
    assert same(q, q0), mismatch(q, q0)
    assert same(gr, g0), mismatch(gr, g0)
    assert same(s.residual(), r0), mismatch(s.residual(), r0)


@pytest.mark.parametrize("over", [dict(renumber=0), dict(renumber=2), dict(kernel_variant=1),
                                  dict(kernel_variant=2), dict(kernel_variant=1, block_size=64),
                                  dict(kernel_variant=3, block_size=256),
                                  dict(kernel_variant=3, block_size=64, renumber=2),
                                  dict(xcd_remap=0), dict(xcd_remap=3),
                                  dict(kernel_variant=2, xcd_remap=16),
                                  dict(kernel_variant=4), dict(kernel_variant=4, block_size=64)])
def test_layout_invariance(over):
    g = Golden("ogrid_ns_squeeze_global")
    a = Solver(g.grid, g.case)
    b = Solver(g.grid, g.case, **{"kernel_variant": 1, **over})
    a.init_freestream()
    b.init_freestream()
    a.step(7)
    b.step(7)
    qa, ga = a.state_and_grad()
    qb, gb = b.state_and_grad()
    assert same(qa, qb) and same(ga, gb), over


def test_automatic_launch_options():
    """block_size 0 and xcd_remap -1 resolve by precision and size; explicit
    values are kept; out-of-range orders are refused."""
    g = Golden("ramp_o2_squeeze")
    auto = {"block_size": 128, "xcd_remap": 1, "kernel_variant": 3, "accumulate": 0}
    assert Solver(g.grid, g.case).launch_options() == auto
    assert Solver(g.grid, g.case, precision=FP32_MIXED).launch_options() == {**auto,
                                                                             "block_size": 256}
    got = Solver(g.grid, g.case, block_size=192, xcd_remap=5).launch_options()
    assert got == {"block_size": 192, "xcd_remap": 5, "kernel_variant": 3, "accumulate": 0}
    with pytest.raises(FvkError):
        Solver(g.grid, g.case, xcd_remap=-2)


def test_restart_semantics_match_oracle():
    """fvk_set_state re-arms the -101 sentinel and the freestream wave speeds,
    as a restart of the reference does (main.cpp:82-118)."""
    g = Golden("ramp_o2_squeeze")
    s = Solver(g.grid, g.case)
    orc = O.Oracle(g.grid, g.case)
    s.init_freestream()
    s.step(5)
    q = s.state()
    orc.set_state(q)
    s.set_state(q)
    s.step(6)
    orc.step(6)
    assert same(s.state(), orc.state()), mismatch(s.state(), orc.state())


def test_freestream_preserved():
    """Uniform flow with only freestream/extrapolation boundaries has a zero
    residual to round-off on any closed mesh."""
    grid = gg.box(12, 9, jitter=0.3, split_fraction=0.4, markers=(-1, -1, -2, -1))
    u, v = inputs.mach_velocity(0.6, angle_deg=20.0)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", BC1="EXTRAPOLATION",
                       BC2="FREESTREAM", dt_type="LOCAL", CFL=0.5)
    s = Solver(grid, case)
    s.init_freestream()
    s.step(20)
    q = s.state()
    qinf = np.array([case.rho, case.rho * u, case.rho * v,
                     case.P / 0.4 + 0.5 * case.rho * (u * u + v * v)])
    assert np.allclose(q, qinf[:, None], rtol=1e-9, atol=1e-6 * np.abs(qinf).max())


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["ramp_o2_squeeze", "ogrid_ns_squeeze_global"])
def test_checkpoint_resumes_exactly(name, precision):
    """A checkpoint taken mid-run and loaded into a new context continues bit
    for bit as if the run had never stopped (lagged limiter and dt included)."""
    g = Golden(name)
    a = Solver(g.grid, g.case, precision=precision)
    a.init_freestream()
    a.step(25)
    b = Solver(g.grid, g.case, precision=precision)
    b.init_freestream()
    b.step(12)
    ck = b.checkpoint()
    b.close()
    c = Solver(g.grid, g.case, precision=precision)
    c.restore(ck)
    c.step(13)
    qa, ga = a.state_and_grad()
    qc, gc = c.state_and_grad()
    assert same(qa, qc) and same(ga, gc)
    assert same(a.residual(), c.residual())


def test_checkpoint_rejects_another_layout():
    g = Golden("ramp_o2_squeeze")
    a = Solver(g.grid, g.case, block_size=128)
    a.init_freestream()
    ck = a.checkpoint()
    b = Solver(g.grid, g.case, block_size=256)
    with pytest.raises(FvkError):
        b.restore(ck)


@pytest.mark.parametrize("kind,steps", [("bump", 75), ("ogrid_small", 21), ("ramp", 19),
                                        ("ogrid", 7)])
def test_graph_replay_equals_eager(kind, steps):
    """fvk_step replays a captured graph of 32, 8 or 2 steps (by mesh size)
    after the first steps of a start; the results equal step-by-step
    launches bit for bit, also after a restart reuses the graph."""
    if kind == "ogrid_small":
        u, v = inputs.mach_velocity(0.2)
        grid = gg.ogrid(256, 128)
        case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS", mu=1.225 * u / 100.0,
                           BC1="NO_SLIP_WALL", dt_type="GLOBAL", CFL=0.5)
    else:
        grid, case = _bench_case(kind)
    print(f"[gpu] graph vs eager {kind} cells={grid.num_cells} steps={steps}", flush=True)
    out = []
    for graph in (True, False):
        s = Solver(grid, case)
        s.use_graph(graph)
        s.init_freestream()
        s.step(steps)
        q1, g1 = s.state_and_grad()
        r1 = s.residual()
        # a restart from freestream replays the same graph
        s.init_freestream()
        s.step(steps)
        q2, g2 = s.state_and_grad()
        assert same(q1, q2) and same(g1, g2), kind
        out.append((q1, g1, r1))
    for a, b in zip(out[0], out[1]):
        assert same(a, b), f"{kind}: {mismatch(a, b)}"


@pytest.mark.parametrize("precision", [0, 1])
def test_conservation_sums(precision):
    """fvk_conservation_sums = sum_i V_i Q_k,i of the current state, with
    compensated summation: within 4 ulp of the sum of |V Q| of an exactly
    rounded sum (math.fsum) of the same products."""
    import math
    g = Golden("ramp_o2_squeeze")
    s = Solver(g.grid, g.case, precision=precision)
    s.init_freestream()
    s.step(30)
    q = s.state()
    vol = s.geometry("cell_vol")
    got = s.conservation_sums()
    for k in range(4):
        prod = vol * q[k]
        ref = math.fsum(prod.tolist())
        bound = 4 * np.finfo(np.float64).eps * float(np.abs(prod).sum())
        assert abs(got[k] - ref) <= bound, (k, got[k], ref)


def test_conservation_sums_closed_box_uniform_dt():
    """In a closed box of slip walls with the uniform time step, mass and
    energy sums stay constant to round-off (the face fluxes cancel pairwise
    and the walls carry no mass or energy flux)."""
    grid = gg.box(24, 18, jitter=0.25, split_fraction=0.5, markers=(0, 0, 0, 0))
    u, v = inputs.mach_velocity(0.4, angle_deg=30)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="UNIFORM", CFL=0.5,
                       BC1="SLIP_WALL", BC2="FREESTREAM")
    s = Solver(grid, case)
    s.init_freestream()
    c0 = s.conservation_sums()
    s.step(100)
    c1 = s.conservation_sums()
    for k in (0, 3):
        assert abs(c1[k] - c0[k]) <= 1e-13 * abs(c0[k]), (k, c0[k], c1[k])


def test_checkpoint_rejects_other_flow_options():
    """A checkpoint records the flow options it was written under: loading it
    into a context of the same mesh and layout but another freestream (the
    FP32-mixed state is stored relative to it) is refused."""
    g = Golden("ramp_o2_squeeze")
    a = Solver(g.grid, g.case, precision=1)
    a.init_freestream()
    a.step(3)
    ck = a.checkpoint()
    other = inputs.Case(**{**g.case.__dict__, "u": g.case.u * 1.01})
    b = Solver(g.grid, other, precision=1)
    with pytest.raises(FvkError):
        b.restore(ck)
    c = Solver(g.grid, g.case, precision=1)
    c.restore(ck)
