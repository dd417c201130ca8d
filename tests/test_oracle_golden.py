"""The oracle (oracle/fv_oracle.c) against the reference's own outputs.

The fixtures were written by the reference solver compiled from its own
sources (tests/golden/make_golden.py).  Every dumped array must agree as IEEE
values: geometry, state, primitives, limited gradients, residual, face
states, face wave speeds and (NS) viscous face terms.
"""
import numpy as np
import pytest

import oracle as O
from fvtest import Golden, golden_names, mismatch, same

FACE_BOUNDARY_UNUSED = ("geom/face_dist", "geom/face_lx", "geom/face_ly")


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


@pytest.mark.parametrize("name", golden_names())
def test_geometry(name):
    g = Golden(name)
    orc = O.Oracle(g.grid, g.case)
    interior = g.grid.face_c2 >= 0
    for k, v in g.a.items():
        if not k.startswith("geom/") or k == "geom/face_cell2":
            continue
        a = orc.get(k)
        if k in FACE_BOUNDARY_UNUSED:  # boundary entries are never read by the solver
            a, v = a[interior], v[interior]
        assert same(a, v), f"{name} {k}: {mismatch(a, v)}"
    assert np.array_equal(orc.get("geom/face_cell2"), g.a["geom/face_cell2"])


@pytest.mark.parametrize("name", golden_names())
def test_trajectory(name):
    g = Golden(name)
    orc = O.Oracle(g.grid, g.case)
    if g.meta["state"]:
        orc.set_state(g.a["init/state"])
    if g.meta["res_only"]:
        orc.res()
        keys = [k for k in g.a if k.startswith("res/")]
        assert keys
        for k in keys:
            a = orc.get(k.split("/", 1)[1])
            assert same(a, g.a[k]), f"{name} {k}: {mismatch(a, g.a[k])}"
        return
    done = 0
    for s in sorted(g.meta["dumps"]):
        orc.step(s - done)
        done = s
        for k, v in g.a.items():
            if not k.startswith(f"step{s}/"):
                continue
            a = orc.get(k.split("/", 1)[1])
            assert same(a, v), f"{name} {k}: {mismatch(a, v)}"


def test_residual_norm_matches_sequential_sum():
    g = Golden("bump_o2_squeeze_local")
    orc = O.Oracle(g.grid, g.case)
    orc.step(2)
    r = g.residual(2)
    expect = np.array([np.sqrt(sum(float(x) * float(x) for x in r[k])) for k in range(4)])
    assert same(orc.residual_l2(), expect)


def _closed_box_case(dt_type, poly=1, limiter="SQUEEZE"):
    from fv2d import gridgen as gg
    from fv2d import inputs
    g = gg.box(24, 18, jitter=0.25, split_fraction=0.5, markers=(0, 0, 0, 0))
    u, v = inputs.mach_velocity(0.4, angle_deg=30)
    return g, inputs.Case(u=u, v=v, poly=poly, limiter=limiter, dt_type=dt_type, CFL=0.5,
                          BC1="SLIP_WALL", BC2="FREESTREAM")


@pytest.mark.parametrize("poly,limiter", [(1, "SQUEEZE"), (0, "NONE")])
def test_uniform_dt_conserves_and_keeps_time(poly, limiter):
    """The uniform time step (time_step_type 3, an extension with no reference
    counterpart) is pinned by what it is for.  One dt for every cell makes the
    face fluxes cancel pairwise, so a closed box of slip walls keeps its mass
    and energy to rounding; the reference's GLOBAL and LOCAL steps (dt_i
    proportional to V_i, ssp_rk3.cpp:38-47) do not.  Every cell gets the same
    dt, the largest the CFL bound allows (dt * max(Sigma/V) = 2 CFL), and the
    physical time is the sum of the steps' dt."""
    drift = {}
    for dt_type in ("UNIFORM", "GLOBAL"):
        g, case = _closed_box_case(dt_type, poly, limiter)
        o = O.Oracle(g, case)
        vol = o.get("geom/cell_vol")
        m0, e0 = (vol * o.state()[0]).sum(), (vol * o.state()[3]).sum()
        t = 0.0
        for _ in range(100):
            o.step(1)
            if dt_type == "UNIFORM":
                dt = o.get("dt")
                assert np.all(dt == dt[0])
                t += dt[0]
                assert o.get("time")[0] == t
        q = o.state()
        drift[dt_type] = max(abs((vol * q[0]).sum() / m0 - 1), abs((vol * q[3]).sum() / e0 - 1))
    assert drift["UNIFORM"] < 1e-13, drift
    assert drift["GLOBAL"] > 1e-6, drift


def test_uniform_dt_is_the_smallest_local_dt():
    """With the first step's spectral sums (|V_inf| + c_inf on every face,
    class_q.hpp:61-62), the uniform dt is the minimum over cells of the LOCAL
    dt, 2 CFL V_i / Sigma_i (ssp_rk3.cpp:43-47), taken in the same arithmetic."""
    g, case = _closed_box_case("UNIFORM")
    o = O.Oracle(g, case)
    o.step(1)
    gl, cl = _closed_box_case("LOCAL")
    ol = O.Oracle(gl, cl)
    ol.step(1)
    dl = ol.get("dt")
    vol = ol.get("geom/cell_vol")
    # dt_local = 2 CFL V / Sigma, so Sigma / V = 2 CFL / dt_local up to rounding
    assert np.isclose(o.get("dt")[0], dl.min(), rtol=1e-14, atol=0)
    assert np.all(o.get("dt")[0] <= dl * (1 + 1e-14))
    assert vol.min() > 0
