"""The batched form (fvk_batch_*) against single-instance runs.

A batch concatenates its instances' device layouts and advances all of them
with one launch per kernel and stage; each instance takes its own freestream,
CFL, BC kinds, limiter switch and (GLOBAL dt) time-step maximum.  Every
instance's results must equal, bit for bit, a single fvk_ctx run of the same
mesh and options -- which for FP64 is the reference's own result
(test_gpu_parity.py).  Covered: mixed meshes, Mach numbers, angles, limiter
switches and BC sets in one batch; GLOBAL and LOCAL dt; Euler and NS; FP64
and FP32-mixed; graph replay against step-by-step launches; retirement
(retired instances keep their state, the others go on unchanged);
run-to-residual-drop; stall retirement; the FP64 finish of stalled
FP32-mixed instances; option checks.
"""
import numpy as np
import pytest

from fv2d import gridgen as gg
from fv2d import inputs
from fv2d.lib import FP32_MIXED, Batch, FvkError, Solver
from fvtest import mismatch, same

pytestmark = pytest.mark.gpu


def euler_cases(dt_type="LOCAL"):
    out = []
    for i, (nx, ny, mach, ang, lim) in enumerate([(32, 16, 0.5, 0.0, "SQUEEZE"),
                                                  (48, 24, 0.7, 2.0, "NONE"),
                                                  (40, 20, 0.3, 4.5, "SQUEEZE")]):
        u, v = inputs.mach_velocity(mach, angle_deg=ang)
        out.append((gg.bump(nx, ny), inputs.Case(u=u, v=v, poly=1, limiter=lim, dt_type=dt_type,
                                                 CFL=0.8 - 0.1 * i)))
    u, v = inputs.mach_velocity(2.5, angle_deg=1.0)
    out.append((gg.ramp(64, 32), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                             dt_type=dt_type, CFL=0.7)))
    u, v = inputs.mach_velocity(0.6)
    # slip walls, freestream and extrapolation sides on a jittered mixed mesh
    out.append((gg.box(12, 9, jitter=0.2, split_fraction=0.5, markers=(0, -1, -2, -1)),
                inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type=dt_type, CFL=0.5,
                            BC1="SLIP_WALL", BC2="FREESTREAM")))
    return out


def ns_cases():
    out = []
    for nr, mach in [(16, 0.2), (24, 0.3)]:
        u, v = inputs.mach_velocity(mach)
        out.append((gg.ogrid(2 * nr, nr), inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE",
                                                      eqn="NS", mu=1.225 * u / 100.0,
                                                      BC1="NO_SLIP_WALL", dt_type="GLOBAL",
                                                      CFL=0.5)))
    return out


def single(grid, case, steps, **ov):
    s = Solver(grid, case, **ov)
    s.init_freestream()
    s.step(steps)
    return s


@pytest.mark.parametrize("precision", [0, FP32_MIXED])
@pytest.mark.parametrize("dt_type", ["LOCAL", "GLOBAL"])
def test_batch_equals_single_runs(precision, dt_type):
    cases = euler_cases(dt_type)
    steps = 25   # two eager steps, graph pairs, one eager step at the end
    print(f"[gpu] batch euler {dt_type} precision={precision} n={len(cases)}", flush=True)
    b = Batch(cases, precision=precision)
    b.init()
    b.step(steps)
    r = b.residual_l2()
    for i, (g, c) in enumerate(cases):
        s = single(g, c, steps, precision=precision)
        q, gr = b.state_and_grad(i)
        q1, gr1 = s.state_and_grad()
        assert same(q, q1), f"instance {i}: {mismatch(q, q1)}"
        assert same(gr, gr1), f"instance {i} grad: {mismatch(gr, gr1)}"
        assert np.isfinite(q).all(), f"instance {i}"
        assert np.allclose(r[i], s.residual_l2(), rtol=1e-12, atol=0)


def test_batch_ns_global_dt():
    cases = ns_cases()
    b = Batch(cases)
    b.init()
    b.step(12)
    for i, (g, c) in enumerate(cases):
        q1 = single(g, c, 12).state()
        assert same(b.state(i), q1), f"instance {i}: {mismatch(b.state(i), q1)}"


def test_batch_graph_equals_eager():
    cases = euler_cases("GLOBAL")
    out = []
    for graph in (True, False):
        b = Batch(cases, precision=FP32_MIXED)
        b.use_graph(graph)
        b.init()
        b.step(17)
        out.append([b.state_and_grad(i) for i in range(len(cases))])
    for i in range(len(cases)):
        assert same(out[0][i][0], out[1][i][0]) and same(out[0][i][1], out[1][i][1]), i


def test_batch_retirement():
    cases = euler_cases("GLOBAL")
    b = Batch(cases)
    b.init()
    b.step(10)
    flags = np.ones(len(cases), dtype=np.uint8)
    flags[1] = 0
    flags[3] = 0
    b.set_active(flags)
    assert b.info()["active_instances"] == len(cases) - 2
    b.step(11)
    for i, (g, c) in enumerate(cases):
        n = 10 if flags[i] == 0 else 21
        q1 = single(g, c, n).state()
        assert same(b.state(i), q1), f"instance {i} ({n} steps): {mismatch(b.state(i), q1)}"


@pytest.mark.parametrize("opts", [dict(block_size=128, xcd_remap=0),
                                  dict(block_size=128, xcd_remap=4), dict(xcd_remap=2)])
def test_batch_launch_orders(opts):
    """Block size and XCD block order change no bit, with retired instances
    left out of the active-block list."""
    cases = euler_cases("GLOBAL")
    print(f"[gpu] batch launch order {opts}", flush=True)
    flags = np.ones(len(cases), dtype=np.uint8)
    flags[2] = 0
    out = []
    for ov in ({}, opts):
        b = Batch(cases, precision=FP32_MIXED, **ov)
        b.init()
        b.step(6)
        b.set_active(flags)
        b.step(9)
        out.append([b.state_and_grad(i) for i in range(len(cases))])
        b.close()
    for i in range(len(cases)):
        assert same(out[0][i][0], out[1][i][0]) and same(out[0][i][1], out[1][i][1]), i


def test_batch_run_to_drop():
    cases = []
    for nx, mach in [(24, 0.4), (32, 0.5), (40, 0.6)]:
        u, v = inputs.mach_velocity(mach)
        cases.append((gg.bump(nx, nx // 2), inputs.Case(u=u, v=v, poly=1, limiter="NONE",
                                                        dt_type="LOCAL", CFL=0.8)))
    b = Batch(cases)
    b.init()
    steps, r0, r1 = b.run_to_drop(1e-2, 4000, 50)
    print(f"[gpu] run_to_drop steps={steps.tolist()} drop={(r1 / r0).tolist()}", flush=True)
    assert b.info()["active_instances"] == 0
    assert (r1 <= 1e-2 * r0).all() and (steps < 4000).all()
    assert len(set(steps.tolist())) > 1     # instances retire at different steps
    for i, (g, c) in enumerate(cases):
        s = single(g, c, int(steps[i]))
        assert same(b.state(i), s.state()), f"instance {i}"
        assert np.isclose(s.residual_l2()[0], r1[i], rtol=1e-12, atol=0)


def test_batch_rejects_mixed_kernel_options():
    cases = euler_cases()
    with pytest.raises(FvkError, match="OPTION"):
        Batch([cases[0], (cases[1][0], inputs.Case(**{**cases[1][1].__dict__, "eqn": "NS"}))])
    with pytest.raises(FvkError, match="OPTION"):
        Batch(cases[:2], kernel_variant=1)


def test_batch_checkpoint_resumes_exactly():
    cases = euler_cases("GLOBAL")
    a = Batch(cases, precision=FP32_MIXED)
    a.init()
    a.step(9)
    flags = np.ones(len(cases), dtype=np.uint8)
    flags[2] = 0
    a.set_active(flags)
    a.step(4)
    ck = a.checkpoint()
    a.step(8)
    b = Batch(cases, precision=FP32_MIXED)
    b.restore(ck)
    assert b.info()["active_instances"] == len(cases) - 1
    b.step(8)
    for i in range(len(cases)):
        assert same(a.state(i), b.state(i)), i


def test_batch_conservation_sums_equal_single():
    cases = euler_cases("LOCAL")
    b = Batch(cases)
    b.init()
    b.step(9)
    sums = b.conservation_sums()
    for i, (g, c) in enumerate(cases):
        s = single(g, c, 9)
        assert np.array_equal(sums[i], s.conservation_sums()), i


def test_batch_set_active_checks_length():
    cases = euler_cases("LOCAL")
    b = Batch(cases)
    with pytest.raises(ValueError):
        b.set_active(np.ones(len(cases) - 1, dtype=np.uint8))


def test_batch_graph_survives_restart():
    """fvk_batch_init keeps the captured graph (it reads everything from
    device memory); a restarted batch replays it and gives the results of a
    fresh batch."""
    cases = euler_cases("GLOBAL")
    b = Batch(cases)
    b.init()
    b.step(9)
    b.init()
    b.step(13)
    f = Batch(cases)
    f.init()
    f.step(13)
    for i in range(len(cases)):
        assert same(b.state(i), f.state(i)), i


def test_batch_stall_retirement_and_error_at_convergence():
    """First-order channel and ramp instances run to an unreachable drop with
    stall retirement on: both precisions retire every instance by stall (FP64
    near 1e-14 of the first-step residual, FP32-mixed where its state's
    rounding stops the residual), and the converged FP32-mixed density is
    within 1e-5 (relative L2) of the converged FP64 density."""
    from fv2d.lib import RETIRED_STALL
    u, v = inputs.mach_velocity(0.5, angle_deg=1.0)
    u2, v2 = inputs.mach_velocity(2.2, angle_deg=2.0)
    cases = [(gg.bump(64, 32), inputs.Case(u=u, v=v, poly=0, dt_type="LOCAL", CFL=0.5)),
             (gg.ramp(48, 24), inputs.Case(u=u2, v=v2, poly=0, dt_type="LOCAL", CFL=0.5))]
    final = {}
    for prec in (0, FP32_MIXED):
        b = Batch(cases, precision=prec)
        b.set_stall(5, 0.5)
        b.init()
        steps, r0, r1 = b.run_to_drop(1e-30, 40000, 200)
        why = b.retired()
        print(f"[gpu] stall precision={prec}: steps {steps.tolist()}, drop {(r1 / r0).tolist()}",
              flush=True)
        assert (why == RETIRED_STALL).all(), why
        assert (steps < 40000).all(), steps
        final[prec] = [b.state(i) for i in range(len(cases))]
    for i in range(len(cases)):
        q32, q64 = final[FP32_MIXED][i], final[0][i]
        err = float(np.linalg.norm(q32[0] - q64[0]) / np.linalg.norm(q64[0]))
        assert err <= 1e-5, (i, err)


def test_finish_stalled_in_fp64():
    """The FP32-mixed batch retires first-order instances by a stall above an
    FP64-reachable drop; finish_stalled continues exactly those instances in
    an FP64 batch from their FP32-mixed states, measured from the FP32-mixed
    run's reference norms, to the drop.  Each continued instance equals a
    single FP64 context started from the same state and run the same steps
    (bit for bit), its norm is within the drop of the FP32-mixed reference,
    and its density is within 1e-5 (relative L2) of the FP64 batch's own
    converged density."""
    from fv2d.lib import RETIRED_CALLER, RETIRED_DROP, RETIRED_STALL, finish_stalled
    u, v = inputs.mach_velocity(0.5, angle_deg=1.0)
    u2, v2 = inputs.mach_velocity(2.2, angle_deg=2.0)
    cases = [(gg.bump(64, 32), inputs.Case(u=u, v=v, poly=0, dt_type="LOCAL", CFL=0.5)),
             (gg.ramp(48, 24), inputs.Case(u=u2, v=v2, poly=0, dt_type="LOCAL", CFL=0.5)),
             (gg.bump(32, 16), inputs.Case(u=u, v=v, poly=0, dt_type="LOCAL", CFL=0.5))]
    drop = 1e-9
    b32 = Batch(cases, precision=FP32_MIXED)
    b32.set_stall(5, 0.5)
    b32.init()
    b32.set_active([1, 1, 0])           # the third instance stays out (retired by the caller)
    s32, r0, r1 = b32.run_to_drop(drop, 40000, 200)
    why32 = b32.retired()
    print(f"[gpu] finish: FP32-mixed steps {s32.tolist()} drop {(r1 / r0).tolist()} "
          f"why {why32.tolist()}", flush=True)
    assert (why32[:2] == RETIRED_STALL).all() and why32[2] == RETIRED_CALLER, why32
    q32 = [b32.state(i) for i in range(2)]

    b64 = Batch(cases)
    b64.set_stall(5, 0.5)
    finished, steps, res_end, why = finish_stalled(b32, b64, r0, drop, 40000, 200)
    print(f"[gpu] finish: FP64 steps {steps.tolist()} drop {(res_end / r0).tolist()} "
          f"why {why.tolist()}", flush=True)
    assert finished.tolist() == [True, True, False]
    assert (why[:2] == RETIRED_DROP).all() and why[2] == RETIRED_CALLER, why
    assert (res_end[:2] <= drop * r0[:2]).all()
    assert steps[2] == 0
    ref = Batch(cases[:2])
    ref.init()
    ref.run_to_drop(drop, 40000, 200)
    for i in range(2):
        s = Solver(*cases[i])
        s.set_state(q32[i])
        s.step(int(steps[i]))
        assert same(b64.state(i), s.state()), f"instance {i}"
        assert np.isclose(s.residual_l2()[0], res_end[i], rtol=1e-12, atol=0)
        q64 = ref.state(i)
        err = float(np.linalg.norm(b64.state(i)[0] - q64[0]) / np.linalg.norm(q64[0]))
        assert err <= 1e-5, (i, err)
    # the third instance was not continued: freestream state, untouched
    s = Solver(*cases[2])
    s.init_freestream()
    assert same(b64.state(2), s.state())
    # a reference norm > 0 replaces the first step's; <= 0 or None restores it
    b64.set_reference([5.0, 0.0, -1.0])
    b64.init()
    _, r0b, r1b = b64.run_to_drop(1e-2, 1, 1)
    assert r0b[0] == 5.0 and r0b[1] == r1b[1] and r0b[2] == r1b[2]
    b64.set_reference(None)
    b64.init()
    _, r0c, r1c = b64.run_to_drop(1e-2, 1, 1)
    assert (r0c == r1c).all()
