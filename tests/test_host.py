"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared symbol, the host geometry is bit-identical to the reference reader,
mesh/option validation, the grid generators and the grid/input file formats."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from fv2d import gridgen as gg
from fv2d import inputs
from fv2d import lib as F
from fvtest import Golden, golden_names, mismatch, same

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "fvk.h")).read()
    declared = set(re.findall(r"\b(fvk_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(F.EXPORTS)
    L = ctypes.CDLL(F.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}$", out, re.M), name


def test_struct_layout_matches_header(tmp_path):
    # offsets the C compiler gives the header structs
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "fvk.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(fvk_mesh), offsetof(fvk_mesh, face_cell2),
         sizeof(fvk_options), offsetof(fvk_options, kernel_variant));
  return 0;
}
"""
    d = str(tmp_path)
    with open(os.path.join(d, "abi.c"), "w") as f:
        f.write(src)
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o",
                           os.path.join(d, "abi"), os.path.join(d, "abi.c")])
    got = subprocess.run([os.path.join(d, "abi")], capture_output=True, text=True).stdout.split()
    assert [int(v) for v in got] == [ctypes.sizeof(F.FvkMesh), F.FvkMesh.face_cell2.offset,
                                     ctypes.sizeof(F.FvkOptions), F.FvkOptions.kernel_variant.offset]


GEOM_FACE = ["face_area", "face_nx", "face_ny", "face_centerx", "face_centery"]
GEOM_CELL = ["cell_vol", "cell_centerx", "cell_centery", "Ixx", "Iyy", "Ixy"]


@pytest.mark.parametrize("name", golden_names())
def test_host_geometry_bit_identical_to_reference(name):
    g = Golden(name)
    for key in GEOM_FACE + GEOM_CELL:
        a = F.host_geometry(g.grid, g.case, key)
        ref = g.a[f"geom/{key}"]
        assert same(a, ref), f"{key}: {mismatch(a, ref)}"
        assert np.array_equal(a.view(np.uint64), ref.view(np.uint64)), key


def test_parallel_preparation_is_bit_identical(monkeypatch):
    """The mesh preparation's threads change no bit of the geometry (ramp
    512x256: large enough for every parallel loop to split)."""
    g = gg.ramp()
    case = inputs.Case(poly=1, limiter="SQUEEZE", BC1="SLIP_WALL", BC2="FREESTREAM")
    out = {}
    for t in ("1", "8"):
        monkeypatch.setenv("FVK_PREP_THREADS", t)
        out[t] = {k: F.host_geometry(g, case, k) for k in GEOM_FACE + GEOM_CELL}
    for k in GEOM_FACE + GEOM_CELL:
        assert np.array_equal(out["1"][k].view(np.uint64), out["8"][k].view(np.uint64)), k


def _mesh_err(grid, case):
    with pytest.raises(F.FvkError) as e:
        F.host_geometry(grid, case, "cell_vol")
    return e.value


def test_mesh_validation():
    case = inputs.Case()
    g = gg.box(3, 3)
    bad = gg.Grid(**{**g.__dict__, "face_p1": g.face_p1.copy()})
    bad.face_p1[4] = g.num_points + 5
    assert _mesh_err(bad, case).code == -2
    bad = gg.Grid(**{**g.__dict__, "face_c1": g.face_c1.copy()})
    bad.face_c1[0] = -7
    assert _mesh_err(bad, case).code == -2
    # a cell left with two faces
    keep = np.ones(g.num_faces, bool)
    c0 = np.nonzero((g.face_c1 == 0) | (g.face_c2 == 0))[0]
    keep[c0[:2]] = False
    bad = gg.Grid(x=g.x, y=g.y, face_p1=g.face_p1[keep], face_p2=g.face_p2[keep],
                  face_c1=g.face_c1[keep], face_c2=g.face_c2[keep], bc_marker=g.bc_marker[keep],
                  num_cells=g.num_cells)
    assert "faces" in str(_mesh_err(bad, case))


def test_noslip_second_order_needs_face_index_cell():
    # a no-slip face whose index exceeds the cell count: the reference would
    # read past its state arrays (cell_gradient.cpp:63-66); refused here
    g = gg.box(2, 2, markers=(-1, -1, -1, -3))
    idx = np.nonzero(g.bc_marker == -3)[0]
    assert idx.max() >= g.num_cells
    case = inputs.Case(poly=1, BC1="SLIP_WALL", BC2="FREESTREAM")
    assert _mesh_err(g, case).code == -2
    F.host_geometry(g, inputs.Case(poly=0), "cell_vol")   # first order never reads it


def test_polygons_with_more_than_six_sides():
    # a regular nonagon surrounded by triangles: closed-form area and centroid
    n = 9
    th = 2 * np.pi * np.arange(n) / n
    x = np.concatenate([np.cos(th), 2 * np.cos(th + np.pi / n)])
    y = np.concatenate([np.sin(th), 2 * np.sin(th + np.pi / n)])
    cells = [list(range(n))]
    for k in range(n):
        cells.append([k, (k + 1) % n, n + k])
    padded = np.full((len(cells), n), -1)
    nv = np.array([len(c) for c in cells])
    for i, c in enumerate(cells):
        padded[i, :len(c)] = c
    g = gg.build_grid(x, y, padded, nv, lambda a, b: np.full(a.shape[0], -1), "nonagon")
    vol = F.host_geometry(g, inputs.Case(BC1="FREESTREAM"), "cell_vol")
    assert abs(vol[0] - 0.5 * n * np.sin(2 * np.pi / n)) < 1e-14
    cx = F.host_geometry(g, inputs.Case(BC1="FREESTREAM"), "cell_centerx")
    assert abs(cx[0]) < 1e-15


def test_generators_are_canonical_and_sized():
    b = gg.bump()
    assert (b.num_cells, b.num_faces, int((b.face_c2 < 0).sum())) == (2048, 4192, 192)
    r = gg.ramp()
    assert (r.num_cells, r.num_faces, int((r.face_c2 < 0).sum())) == (196608, 328448, 1536)
    o = gg.ogrid(64, 32)
    for g in (b, r, o):
        assert gg.check_canonical(g) == []
    # wall face k is owned by cell k on the O-grid (no-slip gradient ghost)
    wall = np.nonzero(o.bc_marker == 0)[0]
    assert np.array_equal(o.face_c1[wall], wall)
    # same seed, same grid
    assert np.array_equal(gg.ramp(32, 16).face_p1, gg.ramp(32, 16).face_p1)


def test_grid_file_round_trip(tmp_path):
    g = gg.box(5, 4, jitter=0.3, split_fraction=0.5)
    p = tmp_path / "g.in"
    g.write(str(p))
    h = gg.read_grid(str(p))
    for k in ("x", "y"):
        assert np.array_equal(getattr(g, k), getattr(h, k))
    for k in ("face_p1", "face_p2", "face_c1", "face_c2", "bc_marker"):
        assert np.array_equal(getattr(g, k), getattr(h, k))


def _read_grid_cache(path):
    """The CLI's binary grid cache: 40-byte header, then x, y, p1, p2, c1, c2."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"FVKGRID1"
    npt, nf, nc = np.frombuffer(raw, np.int32, 3, 24)
    x = np.frombuffer(raw, np.float64, npt, 40)
    y = np.frombuffer(raw, np.float64, npt, 40 + 8 * npt)
    ids = np.frombuffer(raw, np.int32, 4 * nf, 40 + 16 * npt).reshape(4, nf)
    return int(nc), x, y, ids


def test_cli_parallel_grid_reader_and_cache(tmp_path):
    """`fv2d --cache-grid` parses a grid file with several threads (the ramp's
    14 MB text is split) into exactly the generator's arrays, and a run with
    the cache reads the binary copy only while the text keeps its size and
    modification time."""
    cli = os.path.join(os.path.dirname(F.LIB_PATH), "fv2d")
    g = gg.ramp()
    p = str(tmp_path / "ramp.in")
    g.write(p)
    assert os.path.getsize(p) > (4 << 20)
    r = subprocess.run([cli, "--cache-grid", p], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert f"{g.num_points} Points" in r.stdout and f"{g.num_cells} Cells" in r.stdout
    nc, x, y, ids = _read_grid_cache(p + ".fvkgrid")
    assert nc == g.num_cells
    assert np.array_equal(x.view(np.uint64), g.x.view(np.uint64))
    assert np.array_equal(y.view(np.uint64), g.y.view(np.uint64))
    assert np.array_equal(ids[0], g.face_p1) and np.array_equal(ids[1], g.face_p2)
    assert np.array_equal(ids[2], g.face_c1) and np.array_equal(ids[3], g.reader_c2())
    # same size and time stamp: the cache is used, even over a text that no
    # longer parses; a new time stamp sends the reader back to the text
    st = os.stat(p)
    with open(p, "r+b") as f:
        f.seek(st.st_size // 2)
        f.write(b"#")
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns))
    r2 = subprocess.run([cli, "--cache-grid", p], capture_output=True, text=True)
    assert r2.returncode == 0 and r2.stdout == r.stdout
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns + 1000))
    r3 = subprocess.run([cli, "--cache-grid", p], capture_output=True, text=True)
    assert r3.returncode != 0 and "not a number" in r3.stdout


def test_reference_grid_file_reads():
    """grids/grid_file.in ships with the reference; its header and counts are
    recorded here (7026 points, 15222 faces, 8196 cells, 219 + 40 BC faces)."""
    path = "/root/reference/grids/grid_file.in"
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    g = gg.read_grid(path)
    assert (g.num_points, g.num_faces, g.num_cells) == (7026, 15222, 8196)
    assert int((g.bc_marker == 0).sum()) == 219 and int((g.bc_marker == -1).sum()) == 40
    assert gg.check_canonical(g) == []
    vol = F.host_geometry(g, inputs.Case(BC1="NO_SLIP_WALL"), "cell_vol")
    assert np.all(vol > 0)


def test_airfoil_dual_generator(tmp_path):
    """The polygonal Delaunay-dual mesh: 4- to 9-sided cells, a slip wall
    around the section and freestream outside, closed cells."""
    g = gg.airfoil_dual(20000)
    deg = np.bincount(g.face_c1, minlength=g.num_cells)
    deg += np.bincount(g.face_c2[g.face_c2 >= 0], minlength=g.num_cells)
    assert deg.min() >= 3 and deg.max() <= 9 and (deg > 6).any()
    assert set(np.unique(g.bc_marker[g.face_c2 < 0]).tolist()) == {0, -1}
    assert 2.9 < g.num_faces / g.num_cells < 3.1
    # wall faces lie on the section's band, freestream faces near the outer rectangle
    mx = 0.5 * (g.x[g.face_p1] + g.x[g.face_p2])
    wall = g.bc_marker == 0
    assert (mx[wall] > 0.99).all() and (mx[wall] < 2.01).all()
    # sum over each cell's faces of the outward n*A vanishes (closed polygons)
    case = inputs.Case(poly=1, limiter="SQUEEZE")
    nx, ny, area = (F.host_geometry(g, case, k) for k in ("face_nx", "face_ny", "face_area"))
    sx = np.bincount(g.face_c1, nx * area, g.num_cells)
    sy = np.bincount(g.face_c1, ny * area, g.num_cells)
    inner = g.face_c2 >= 0
    sx -= np.bincount(g.face_c2[inner], (nx * area)[inner], g.num_cells)
    sy -= np.bincount(g.face_c2[inner], (ny * area)[inner], g.num_cells)
    assert np.abs(sx).max() < 1e-15 and np.abs(sy).max() < 1e-15
    # the binary cache gives back the same mesh
    path = os.path.join(str(tmp_path), "g.npz")
    g.save_npz(path)
    h = gg.Grid.load_npz(path)
    for k in ("x", "y", "face_p1", "face_p2", "face_c1", "face_c2", "bc_marker"):
        a, b = getattr(g, k), getattr(h, k)
        assert a.dtype == b.dtype and np.array_equal(a, b), k
    assert (h.num_cells, h.name) == (g.num_cells, g.name)


def test_wall_coefficients_of_a_uniform_flow():
    """A uniform freestream exerts no net force on a closed body and has zero
    pressure coefficient everywhere (fv2d.analysis)."""
    from fv2d import analysis
    g = gg.ogrid(64, 32)
    u, v = inputs.mach_velocity(0.2, angle_deg=3.0)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS", mu=1.225 * u / 100.0,
                       BC1="NO_SLIP_WALL")
    qinf = np.array([case.rho, case.rho * u, case.rho * v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (u * u + v * v)])
    q = np.repeat(qinf[:, None], g.num_cells, axis=1)
    cp, xw, cl, cd = analysis.wall_coefficients(g, inputs.Case(**{**case.__dict__, "mu": 0.0}), q)
    assert cp.size == 64 and np.abs(cp).max() < 1e-9
    assert abs(cl) < 1e-9 and abs(cd) < 1e-9
    # with viscosity the no-slip shear of a uniform stream is pure drag, along the flow
    cp, xw, cl, cd = analysis.wall_coefficients(g, case, q)
    assert cd > 0 and abs(cl) < 1e-9 * cd + 1e-12


def test_strouhal_estimator_on_a_known_signal():
    """An evenly sampled lift history with a known frequency (a start-up
    ramp, then a sinusoid with a harmonic): the estimator returns
    St = f D / U."""
    from fv2d import analysis
    f, D, U = 11.3, 1.0, 67.6
    t = np.arange(4000) * 1.7e-3
    cl = 0.3 * np.sin(2 * np.pi * f * t + 0.4) + 0.05 * np.sin(4 * np.pi * f * t) + 0.02
    cl[:1000] *= np.linspace(0, 1, 1000)
    r = analysis.strouhal(t, cl, D, U)
    assert abs(r["st"] - f * D / U) < 2e-3 * f * D / U
    assert abs(r["st_fft"] - f * D / U) < 1e-2 * f * D / U
    assert r["periods"] >= 30 and abs(r["cl_amplitude"] - 0.3) < 0.06
    assert analysis.strouhal(t[:10], np.ones(10), D, U)["st"] is None
