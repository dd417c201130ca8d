"""The uniform time step (time_step_type 3, an extension with no reference
counterpart: one dt = 2 CFL / max_i(Sigma_i / V_i) for every cell, so that a
run is time-accurate) on the GPU.

The FP64 kernels must equal the oracle's restatement of the mode bit for bit,
state, limited gradients and physical time; the oracle's mode is itself
pinned by conservation in a closed box and by its dt (test_oracle_golden.py).
Also covered: every launch shape, the atomic-scatter accumulation, the
FP32-mixed form, batches with retirement, checkpoints, and the refusal of a
physical time for GLOBAL/LOCAL runs.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, FvkError, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu


def _cases():
    u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
    yield "bump o2 SQUEEZE", gg.bump(), inputs.Case(
        u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="UNIFORM", CFL=0.8), 30
    u, v = inputs.mach_velocity(0.4, angle_deg=30)
    yield "closed box o2 SQUEEZE", gg.box(24, 18, jitter=0.25, split_fraction=0.5,
                                          markers=(0, 0, 0, 0)), inputs.Case(
        u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="UNIFORM", CFL=0.5), 30
    # the uniform step takes the convective spectral sums only, as the
    # reference's steps do: at Re 100 with 1e-3 wall spacing the viscous
    # limit needs CFL 0.2 here (0.3 diverges)
    u, v = inputs.mach_velocity(0.2)
    yield "ogrid NS o2 SQUEEZE", gg.ogrid(128, 64), inputs.Case(
        u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS", mu=1.225 * u / 100.0,
        BC1="NO_SLIP_WALL", dt_type="UNIFORM", CFL=0.2), 10


CASES = list(_cases())


@pytest.mark.parametrize("name,grid,case,steps", CASES, ids=[c[0] for c in CASES])
def test_uniform_fp64_equals_oracle(name, grid, case, steps):
    print(f"[gpu] uniform dt vs oracle: {name}", flush=True)
    s = Solver(grid, case)
    s.init_freestream()
    o = O.Oracle(grid, case)
    s.step(steps)
    o.step(steps)
    q, g = s.state_and_grad()
    assert same(q, o.state()), mismatch(q, o.state())
    assert same(g, o.grad()), mismatch(g, o.grad())
    assert s.time() == o.get("time")[0]


@pytest.mark.parametrize("over", [dict(kernel_variant=1), dict(kernel_variant=2),
                                  dict(block_size=256, xcd_remap=4)])
def test_uniform_layout_invariance(over):
    name, grid, case, steps = CASES[1]
    print(f"[gpu] uniform dt layouts {over}", flush=True)
    out = []
    for ov in ({}, over):
        s = Solver(grid, case, **ov)
        s.init_freestream()
        s.step(steps)
        out.append((s.state_and_grad(), s.time()))
    assert same(out[0][0][0], out[1][0][0]) and same(out[0][0][1], out[1][0][1]), over
    assert out[0][1] == out[1][1]


def test_uniform_atomic_accumulation_conserves():
    """The atomic scatter with one dt: mass and energy of the closed box kept
    to rounding (the sums only change order)."""
    name, grid, case, _ = CASES[1]
    s = Solver(grid, case, accumulate=1)
    s.init_freestream()
    vol = s.geometry("cell_vol")
    q0 = s.state()
    s.step(50)
    q = s.state()
    for k in (0, 3):
        assert abs((vol * q[k]).sum() / (vol * q0[k]).sum() - 1) < 1e-13


def test_uniform_fp32_mixed():
    """FP32-mixed with one dt: the physical time (an FP64 sum of FP64 dt
    values computed from FP32 spectral sums) stays within FP32 rounding of
    the FP64 run's, and the closed box keeps its mass to FP32 rounding."""
    name, grid, case, _ = CASES[1]
    out = {}
    for prec in (0, FP32_MIXED):
        s = Solver(grid, case, precision=prec)
        s.init_freestream()
        vol = s.geometry("cell_vol")
        m0 = (vol * s.state()[0]).sum()
        s.step(50)
        out[prec] = (s.time(), abs((vol * s.state()[0]).sum() / m0 - 1))
    assert abs(out[FP32_MIXED][0] / out[0][0] - 1) < 1e-5, out
    assert out[FP32_MIXED][1] < 1e-6, out


def test_uniform_batch_time_and_retirement():
    cases = [(g, c) for _, g, c, _ in CASES[:2]]
    print("[gpu] uniform dt batch", flush=True)
    b = Batch(cases)
    b.init()
    b.step(10)
    flags = np.ones(len(cases), dtype=np.uint8)
    flags[1] = 0
    b.set_active(flags)
    t_retired = b.time()[1]
    b.step(7)
    t = b.time()
    assert t[1] == t_retired
    for i, (g, c) in enumerate(cases):
        s = Solver(g, c)
        s.init_freestream()
        s.step(17 if flags[i] else 10)
        assert same(b.state(i), s.state()), f"instance {i}: {mismatch(b.state(i), s.state())}"
        assert t[i] == s.time(), i


def test_uniform_checkpoint_keeps_time():
    name, grid, case, _ = CASES[0]
    a = Solver(grid, case)
    a.init_freestream()
    a.step(6)
    ck = a.checkpoint()
    a.step(5)
    b = Solver(grid, case)
    b.init_freestream()
    b.restore(ck)
    b.step(5)
    assert b.time() == a.time() and same(b.state(), a.state())


def test_time_needs_the_uniform_step():
    u, v = inputs.mach_velocity(0.5)
    s = Solver(gg.bump(16, 8), inputs.Case(u=u, v=v, dt_type="GLOBAL"))
    s.init_freestream()
    with pytest.raises(FvkError):
        s.time()
