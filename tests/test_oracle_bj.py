"""The Barth-Jespersen limiter (limiter 3, an extension) of the oracle, pinned
by what defines it.

The reference maps MINMOD to a no-op (res.cpp:36-38) and never builds its
minmod.cpp draft; limiter 1 keeps that.  Limiter 3 is the multidimensional
minmod-type limiter of Barth and Jespersen on the current reconstruction.
Pinned by:
  * its defining bound: every limited face value Q_i + phi grad Q_i . d_f lies
    within the min/max of cell i and its face neighbours, on random data;
  * linear data on a uniform mesh pass unlimited in the interior;
  * a local extremum gets a zero gradient;
  * in full runs: conservation in a closed box under the uniform time step,
    freestream preservation, and a supersonic ramp that stays positive.
The device path is held to the oracle bit for bit in tests/test_gpu_bj.py.
"""
import numpy as np
import pytest

import oracle as O
from fv2d import gridgen as gg
from fv2d import inputs


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _case(limiter="BARTH_JESPERSEN", **kw):
    u, v = inputs.mach_velocity(0.5, angle_deg=15)
    base = dict(u=u, v=v, poly=1, limiter=limiter, dt_type="LOCAL", CFL=0.5,
                BC1="SLIP_WALL", BC2="FREESTREAM")
    base.update(kw)
    return inputs.Case(**base)


def _qinf(case):
    u, v = case.u, case.v
    return np.array([case.rho, case.rho * u, case.rho * v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (u * u + v * v)])


def _bounds(g, q):
    """min/max of each cell and its face neighbours, per component"""
    qmin, qmax = q.copy(), q.copy()
    c1, c2 = g.face_c1, g.face_c2
    inner = c2 >= 0
    for k in range(4):
        np.minimum.at(qmin[k], c1[inner], q[k][c2[inner]])
        np.maximum.at(qmax[k], c1[inner], q[k][c2[inner]])
        np.minimum.at(qmin[k], c2[inner], q[k][c1[inner]])
        np.maximum.at(qmax[k], c2[inner], q[k][c1[inner]])
    return qmin, qmax


def test_limited_face_values_stay_within_the_neighbour_bounds():
    g = gg.box(20, 16, jitter=0.3, split_fraction=0.5, markers=(0, -1, 0, -1))
    case = _case()
    rng = np.random.default_rng(3)
    o = O.Oracle(g, case)
    cx, cy = o.get("geom/cell_centerx"), o.get("geom/cell_centery")
    q = _qinf(case)[:, None] * (1 + 0.05 * np.sin(3 * cx + 2 * cy)[None, :]
                                + 0.02 * rng.uniform(-1, 1, (4, g.num_cells)))
    o.set_state(q)
    o.res()
    qmin, qmax = _bounds(g, q)
    c1, c2 = g.face_c1, g.face_c2
    inner = c2 >= 0
    scale = np.abs(q).max(axis=1)
    for k in range(4):
        fl = o.get(f"fL{k + 1}")
        fr = o.get(f"fR{k + 1}")
        tol = 1e-12 * scale[k]
        assert np.all(fl >= qmin[k][c1] - tol) and np.all(fl <= qmax[k][c1] + tol), k
        assert np.all(fr[inner] >= qmin[k][c2[inner]] - tol), k
        assert np.all(fr[inner] <= qmax[k][c2[inner]] + tol), k
    # the limiter is active on this data: the unlimited gradient differs
    u = O.Oracle(g, _case(limiter="NONE"))
    u.set_state(q)
    u.res()
    lim, unl = o.grad(), u.grad()
    ratio = np.abs(lim).sum() / np.abs(unl).sum()
    assert 0.1 < ratio < 0.99, ratio


def test_linear_data_pass_unlimited_in_the_interior():
    g = gg.box(16, 12, jitter=0.0, split_fraction=0.0, markers=(0, -1, 0, -1))
    case = _case()
    o = O.Oracle(g, case)
    cx, cy = o.get("geom/cell_centerx"), o.get("geom/cell_centery")
    q = _qinf(case)[:, None] * (1 + 0.01 * (cx + 2 * cy))[None, :]
    o.set_state(q)
    o.res()
    u = O.Oracle(g, _case(limiter="NONE"))
    u.set_state(q)
    u.res()
    boundary = np.zeros(g.num_cells, bool)
    boundary[g.face_c1[g.face_c2 < 0]] = True
    lim, unl = o.grad(), u.grad()
    assert np.array_equal(lim[:, ~boundary], unl[:, ~boundary])
    # the exact gradient of the linear density (rows: d/dx of the four
    # components, then d/dy)
    assert np.allclose(unl[0, ~boundary], 0.01 * case.rho, rtol=1e-9)
    assert np.allclose(unl[4, ~boundary], 0.02 * case.rho, rtol=1e-9)


def test_a_local_maximum_gets_a_zero_gradient():
    g = gg.box(16, 12, jitter=0.2, split_fraction=0.5, markers=(0, -1, 0, -1))
    case = _case()
    o = O.Oracle(g, case)
    cx, cy = o.get("geom/cell_centerx"), o.get("geom/cell_centery")
    q = _qinf(case)[:, None] * (1 + 0.01 * (cx + 2 * cy))[None, :]
    boundary = np.zeros(g.num_cells, bool)
    boundary[g.face_c1[g.face_c2 < 0]] = True
    i = int(np.flatnonzero(~boundary)[len(np.flatnonzero(~boundary)) // 2])
    q[0, i] *= 1.5                       # a density spike
    o.set_state(q)
    o.res()
    gr = o.grad()
    assert gr[0, i] == 0 and gr[4, i] == 0
    u = O.Oracle(g, _case(limiter="NONE"))
    u.set_state(q)
    u.res()
    assert u.grad()[0, i] != 0 or u.grad()[4, i] != 0


def test_bj_runs_conserve_in_a_closed_box():
    g = gg.box(20, 16, jitter=0.25, split_fraction=0.5, markers=(0, 0, 0, 0))
    u, v = inputs.mach_velocity(0.4, angle_deg=30)
    case = inputs.Case(u=u, v=v, poly=1, limiter="BARTH_JESPERSEN", dt_type="UNIFORM", CFL=0.5,
                       BC1="SLIP_WALL", BC2="FREESTREAM")
    o = O.Oracle(g, case)
    vol = o.get("geom/cell_vol")
    q0 = o.state()
    o.step(100)
    q = o.state()
    for k in (0, 3):
        assert abs((vol * q[k]).sum() / (vol * q0[k]).sum() - 1) < 1e-13
    assert np.abs(q[0] - q0[0]).max() > 1e-3


def test_bj_preserves_the_freestream_and_keeps_a_ramp_positive():
    g = gg.box(16, 12, jitter=0.3, split_fraction=0.5, markers=(0, -1, 0, -1))
    u, v = inputs.mach_velocity(0.7, angle_deg=20)
    case = inputs.Case(u=u, v=v, poly=1, limiter="BARTH_JESPERSEN", dt_type="LOCAL", CFL=0.8,
                       BC1="FREESTREAM", BC2="FREESTREAM")
    o = O.Oracle(g, case)
    q0 = o.state()
    o.step(50)
    assert np.abs(o.state() - q0).max() <= 1e-12 * np.abs(q0).max()
    u, v = inputs.mach_velocity(2.5)
    ramp = inputs.Case(u=u, v=v, poly=1, limiter="BARTH_JESPERSEN", dt_type="LOCAL", CFL=0.8)
    r = O.Oracle(gg.ramp(64, 32), ramp)
    r.step(300)
    q = r.state()
    P = (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]) * (ramp.gamma - 1)
    assert np.isfinite(q).all() and q[0].min() > 0 and P.min() > 0
