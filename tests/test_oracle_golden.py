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
