"""The persistent launch of the LDS-staged kernels (kernel_variant 4).

A grid of as many workgroups as the device holds at once walks the mesh
blocks, each workgroup a stride of the blocks its XCD would run in the
one-workgroup-per-block launch; each block's work is the LDS-staged kernels'
own, so every result must equal kernel_variant 3's bit for bit:
single solves in both precisions, batches with graph replay and retirement,
the Roe flux, and (within the summation-order bound of the atomic form) the
atomic-scatter accumulation.  Meshes of one block (seven idle XCDs) and of
thousands of blocks per workgroup are covered here and by the fixture and
full-size tests of test_gpu_parity.py, which run variant 4 too.
"""
import numpy as np
import pytest

from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, FvkError, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu


def _run(grid, case, steps, **ov):
    s = Solver(grid, case, **ov)
    s.init_freestream()
    s.step(steps)
    out = s.state_and_grad(), s.residual()
    s.close()
    return out


def _cases():
    u, v = inputs.mach_velocity(2.5)
    yield "ramp o2 SQUEEZE LOCAL", gg.ramp(96, 48), inputs.Case(
        u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8)
    u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
    yield "bump 8x4 (one block)", gg.bump(8, 4), inputs.Case(
        u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="GLOBAL", CFL=0.5)
    yield "bump o1 GLOBAL", gg.bump(64, 32), inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL",
                                                        CFL=0.5)
    u, v = inputs.mach_velocity(0.2)
    yield "ogrid NS SQUEEZE GLOBAL", gg.ogrid(128, 64), inputs.Case(
        u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS", mu=1.225 * u / 100.0,
        BC1="NO_SLIP_WALL", dt_type="GLOBAL", CFL=0.5)


@pytest.mark.parametrize("prec", [0, FP32_MIXED], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name,grid,case", list(_cases()),
                         ids=lambda x: x if isinstance(x, str) else "")
def test_persistent_equals_one_workgroup_per_block(name, grid, case, prec):
    print(f"[gpu] persistent vs per-block: {name} precision={prec}", flush=True)
    ((q3, g3), r3) = _run(grid, case, 40, precision=prec, kernel_variant=3)
    ((q4, g4), r4) = _run(grid, case, 40, precision=prec, kernel_variant=4)
    assert same(q4, q3), mismatch(q4, q3)
    assert same(g4, g3), mismatch(g4, g3)
    assert same(r4, r3), mismatch(r4, r3)


@pytest.mark.parametrize("over", [dict(block_size=256, xcd_remap=0), dict(xcd_remap=1),
                                  dict(xcd_remap=3, renumber=2)])
def test_persistent_launch_shapes(over):
    name, grid, case = list(_cases())[0]
    print(f"[gpu] persistent shapes {over}", flush=True)
    ((q3, _), _) = _run(grid, case, 25, kernel_variant=3, **over)
    ((q4, _), _) = _run(grid, case, 25, kernel_variant=4, **over)
    assert same(q4, q3), (over, mismatch(q4, q3))


def _batch_cases():
    out = []
    for i, (nx, ny, mach, ang, lim) in enumerate([(32, 16, 0.5, 0.0, "SQUEEZE"),
                                                  (48, 24, 0.7, 2.0, "NONE"),
                                                  (40, 20, 0.3, 4.5, "SQUEEZE")]):
        u, v = inputs.mach_velocity(mach, angle_deg=ang)
        out.append((gg.bump(nx, ny), inputs.Case(u=u, v=v, poly=1, limiter=lim, dt_type="LOCAL",
                                                 CFL=0.8 - 0.1 * i)))
    u, v = inputs.mach_velocity(2.5, angle_deg=1.0)
    out.append((gg.ramp(64, 32), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                             dt_type="LOCAL", CFL=0.7)))
    return out


@pytest.mark.parametrize("prec", [0, FP32_MIXED], ids=["fp64", "fp32"])
def test_persistent_batch_with_retirement(prec):
    """Graph replay across a change of the active-block list: the persistent
    grid is sized again for the new block count."""
    cases = _batch_cases()
    print(f"[gpu] persistent batch precision={prec}", flush=True)
    out = []
    for variant in (3, 4):
        b = Batch(cases, precision=prec, kernel_variant=variant)
        assert b.launch_options()["kernel_variant"] == variant
        b.init()
        b.step(30)
        flags = np.ones(len(cases), dtype=np.uint8)
        flags[2] = 0
        b.set_active(flags)
        b.step(50)
        out.append([b.state(i) for i in range(len(cases))])
        b.close()
    for i in range(len(cases)):
        assert same(out[1][i], out[0][i]), (i, mismatch(out[1][i], out[0][i]))


def test_persistent_roe():
    u, v = inputs.mach_velocity(0.7, angle_deg=1.0)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8,
                       flux="ROE")
    print("[gpu] persistent Roe", flush=True)
    ((q3, _), _) = _run(gg.bump(64, 32), case, 30, kernel_variant=3)
    ((q4, _), _) = _run(gg.bump(64, 32), case, 30, kernel_variant=4)
    assert same(q4, q3), mismatch(q4, q3)


def test_persistent_atomic_scatter():
    """The atomic-scatter pass B on the persistent grid: within the
    summation-order bound (1e-10 relative after 100 steps, FP64, no limiter)
    of the gather."""
    u, v = inputs.mach_velocity(0.5, angle_deg=2.0)
    case = inputs.Case(u=u, v=v, poly=1, limiter="NONE", dt_type="LOCAL", CFL=0.8)
    print("[gpu] persistent atomic scatter", flush=True)
    ((q3, _), _) = _run(gg.bump(), case, 100, kernel_variant=3)
    ((q4, _), _) = _run(gg.bump(), case, 100, kernel_variant=4, accumulate=1)
    rel = float(np.max(np.abs(q4 - q3)) / np.max(np.abs(q3)))
    assert rel <= 1e-10, rel


def test_persistent_engines_interleaved_and_timed():
    """Two engines stepped in turn, and the per-launch timing, which
    launches each kernel back to back."""
    name, grid, case = list(_cases())[0]
    print("[gpu] persistent engines interleaved", flush=True)
    a = Solver(grid, case, kernel_variant=4)
    b = Solver(gg.bump(64, 32), list(_cases())[2][2], kernel_variant=4)
    a.init_freestream()
    b.init_freestream()
    for _ in range(5):
        a.step(3)
        b.step(2)
    qa = a.state()
    ((q3, _), _) = _run(grid, case, 15, kernel_variant=3)
    assert same(qa, q3), mismatch(qa, q3)
    t = a.time_kernels(5)
    assert t["pass_a"] > 0 and t["pass_b"] > 0
    a.step(1)
    assert np.isfinite(a.state()).all()


def test_variant_out_of_range():
    name, grid, case = list(_cases())[1]
    with pytest.raises(FvkError):
        Solver(grid, case, kernel_variant=5)
