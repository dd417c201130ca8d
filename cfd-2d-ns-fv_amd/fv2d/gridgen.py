"""Canonical native-grid generators for the FV solver's benchmark and test cases.

Every generator returns a :class:`Grid` in the reference's native grid-file
layout (``src/base/readmesh.cpp:31-85``):

* ``x, y``            point coordinates (float64)
* ``face_p1, face_p2`` 0-based point ids per face
* ``face_c1``         0-based owner cell
* ``face_c2``         0-based neighbour cell, or a file boundary marker:
                      ``-1`` (file 0) -> BC1 kind, ``-2`` (file -1) -> BC2 kind,
                      ``-3`` (file -2) -> EXTRAPOLATION, ``-4`` (file -3) ->
                      NO_SLIP_WALL.  Internally we keep the *file* marker
                      (0, -1, -2, -3) in ``bc_marker`` and ``face_c2 = -1`` for
                      every boundary face.

and obeys the ordering the reference needs to stay inside its arrays
(``readmesh.cpp:373,394``, ``save_sol.cpp:23-24``, ``cell_gradient.cpp:63-66``):
boundary faces first (BC1 first, then BC2, then raw-marker faces), the owner
is on the left of p1->p2 so ``(dy, -dx)`` points out of the owner, and the
owner of the k-th BC1 face is cell k whenever the generator numbers the wall
row first (all generators here do).

Randomness comes from a counter-based splitmix64 stream so the same seed gives
the same grid on every machine and numpy version.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable

import numpy as np

DEFAULT_SEED = 0x5EED

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix_u01(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniform doubles in [0,1) from splitmix64 keyed by (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = base + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Grid:
    x: np.ndarray
    y: np.ndarray
    face_p1: np.ndarray
    face_p2: np.ndarray
    face_c1: np.ndarray
    face_c2: np.ndarray          # neighbour cell or -1 on boundary faces
    bc_marker: np.ndarray        # file marker on boundary faces (0,-1,-2,-3); 1 on interior
    num_cells: int
    name: str = "grid"

    @property
    def num_points(self) -> int:
        return int(self.x.shape[0])

    @property
    def num_faces(self) -> int:
        return int(self.face_p1.shape[0])

    def file_c2(self) -> np.ndarray:
        """Neighbour column exactly as written in the file (1-based / marker)."""
        return np.where(self.face_c2 >= 0, self.face_c2 + 1, self.bc_marker).astype(np.int64)

    def reader_c2(self) -> np.ndarray:
        """Neighbour column as readmesh.cpp:59-69 stores it, before BC kinds
        are substituted: >=0 neighbour, -1 BC1 group, -2 BC2 group, -3/-4 raw."""
        return np.where(self.face_c2 >= 0, self.face_c2, self.bc_marker - 1).astype(np.int32)

    def save_npz(self, path: str) -> None:
        """Binary copy of the arrays (a cache for large generated meshes)."""
        np.savez(path, x=self.x, y=self.y, face_p1=self.face_p1, face_p2=self.face_p2,
                 face_c1=self.face_c1, face_c2=self.face_c2, bc_marker=self.bc_marker,
                 num_cells=np.int64(self.num_cells), name=np.array(self.name))

    @classmethod
    def load_npz(cls, path: str) -> "Grid":
        with np.load(path, allow_pickle=False) as z:
            return cls(x=z["x"], y=z["y"], face_p1=z["face_p1"], face_p2=z["face_p2"],
                       face_c1=z["face_c1"], face_c2=z["face_c2"], bc_marker=z["bc_marker"],
                       num_cells=int(z["num_cells"]), name=str(z["name"]))

    def write(self, path: str) -> None:
        """Write the native text format (readmesh.cpp:31-85), round-trip exact."""
        with open(path, "w") as f:
            f.write(f"{self.num_points} {self.num_faces} {self.num_cells}\n")
            xy = np.stack([self.x, self.y], axis=1)
            f.write("\n".join(f"{a!r} {b!r}" for a, b in xy.tolist()))
            f.write("\n")
            pp = np.stack([self.face_p1 + 1, self.face_p2 + 1], axis=1)
            np.savetxt(f, pp, fmt="%d")
            cc = np.stack([self.face_c1 + 1, self.file_c2()], axis=1)
            np.savetxt(f, cc, fmt="%d")


def read_grid(path: str) -> Grid:
    """Read a native grid file (same token stream as readmesh.cpp:31-69)."""
    with open(path) as f:
        tok = f.read().split()
    n_p, n_f, n_c = int(tok[0]), int(tok[1]), int(tok[2])
    pos = 3
    xy = np.array([float(t) for t in tok[pos:pos + 2 * n_p]], dtype=np.float64).reshape(n_p, 2)
    pos += 2 * n_p
    fp = np.array(tok[pos:pos + 2 * n_f], dtype=np.int64).reshape(n_f, 2)
    pos += 2 * n_f
    fc = np.array(tok[pos:pos + 2 * n_f], dtype=np.int64).reshape(n_f, 2)
    c2 = fc[:, 1]
    return Grid(
        x=xy[:, 0].copy(), y=xy[:, 1].copy(),
        face_p1=(fp[:, 0] - 1).astype(np.int32), face_p2=(fp[:, 1] - 1).astype(np.int32),
        face_c1=(fc[:, 0] - 1).astype(np.int32),
        face_c2=np.where(c2 > 0, c2 - 1, -1).astype(np.int32),
        bc_marker=np.where(c2 > 0, 1, c2).astype(np.int32),
        num_cells=n_c, name=path)


def _polygon_area2(x, y, cells, nv):
    """Twice the signed area of each padded polygon row."""
    n, m = cells.shape
    a2 = np.zeros(n)
    for k in range(m):
        kn = np.where(k + 1 < nv, k + 1, 0)
        valid = k < nv
        i0 = np.where(valid, cells[:, k], 0)
        i1 = cells[np.arange(n), np.where(valid, kn, 0)]
        a2 += np.where(valid, x[i0] * y[i1] - x[i1] * y[i0], 0.0)
    return a2


def build_grid(x: np.ndarray, y: np.ndarray, cells: np.ndarray, nv: np.ndarray,
               marker_fn: Callable[[np.ndarray, np.ndarray], np.ndarray],
               name: str) -> Grid:
    """Faces from polygon cells, in canonical order.

    cells: (N_c, m) padded point ids; nv: vertex count per cell.
    marker_fn(a, b) -> file marker (0,-1,-2,-3) for boundary edges a-b.
    """
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    cells = cells.astype(np.int64).copy()
    nv = nv.astype(np.int64)
    n_c, m = cells.shape
    # make every polygon counter-clockwise so that the interior is on the left
    neg = _polygon_area2(x, y, cells, nv) < 0
    for c in np.nonzero(neg)[0]:
        k = nv[c]
        cells[c, :k] = cells[c, :k][::-1]
    # directed edges a->b with the cell on their left
    rows, ks = np.nonzero(np.arange(m)[None, :] < nv[:, None])
    a = cells[rows, ks]
    b = cells[rows, np.where(ks + 1 < nv[rows], ks + 1, 0)]
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    key = lo * (np.int64(x.shape[0]) + 1) + hi
    order = np.lexsort((rows, key))
    key_s = key[order]
    first = np.ones(key_s.shape[0], dtype=bool)
    first[1:] = key_s[1:] != key_s[:-1]
    starts = np.nonzero(first)[0]
    counts = np.diff(np.append(starts, key_s.shape[0]))
    if np.any(counts > 2):
        raise ValueError("non-manifold edge")
    own = order[starts]                       # lowest cell index owns the face
    c1 = rows[own]
    p1 = a[own]
    p2 = b[own]
    c2 = np.full(c1.shape[0], -1, dtype=np.int64)
    two = counts == 2
    c2[two] = rows[order[starts[two] + 1]]
    marker = np.ones(c1.shape[0], dtype=np.int64)
    bnd = ~two
    marker[bnd] = marker_fn(p1[bnd], p2[bnd])
    # canonical order: BC1 (0), BC2 (-1), raw -2, raw -3, then interior;
    # boundary faces by owner, interior faces by (owner, neighbour)
    rank = np.where(bnd, -marker, 4)
    fo = np.lexsort((c2, c1, rank))
    g = Grid(x=x, y=y,
             face_p1=p1[fo].astype(np.int32), face_p2=p2[fo].astype(np.int32),
             face_c1=c1[fo].astype(np.int32), face_c2=c2[fo].astype(np.int32),
             bc_marker=marker[fo].astype(np.int32), num_cells=int(n_c), name=name)
    return g


# --------------------------------------------------------------------------
# bump channel: [0,3]x[0,1] with a circular-arc bump (thickness 0.1) on x in [1,2]
# --------------------------------------------------------------------------
def _bump_y(xv: np.ndarray, thickness: float = 0.1) -> np.ndarray:
    r = (0.25 + thickness * thickness) / (2.0 * thickness)
    yc = thickness - r
    inside = (xv > 1.0) & (xv < 2.0)
    return np.where(inside, np.sqrt(np.maximum(r * r - (xv - 1.5) ** 2, 0.0)) + yc, 0.0)


def bump(nx: int = 64, ny: int = 32, jitter: float = 0.1, seed: int = DEFAULT_SEED,
         lower_marker: int = 0, upper_marker: int = 0, inout_marker: int = -1) -> Grid:
    """Bump channel: SLIP walls bottom/top (BC1), FREESTREAM inlet/outlet (BC2)."""
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    xs = 3.0 * i / nx
    yb = _bump_y(xs)
    X = np.repeat(xs[None, :], ny + 1, axis=0)
    Y = yb[None, :] + (1.0 - yb[None, :]) * (j[:, None] / ny)
    if jitter > 0:
        n = (nx + 1) * (ny + 1)
        ux = splitmix_u01(seed, 1, n).reshape(ny + 1, nx + 1) * 2.0 - 1.0
        uy = splitmix_u01(seed, 2, n).reshape(ny + 1, nx + 1) * 2.0 - 1.0
        hx = 3.0 / nx
        hy = (1.0 - yb[None, :]) / ny
        interior = np.zeros((ny + 1, nx + 1), dtype=bool)
        interior[1:-1, 1:-1] = True
        X = np.where(interior, X + jitter * hx * ux, X)
        Y = np.where(interior, Y + jitter * hy * uy, Y)
    x = X.ravel()
    y = Y.ravel()
    pid = lambda ii, jj: jj * (nx + 1) + ii
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny))
    ii = ii.ravel()
    jj = jj.ravel()
    cells = np.stack([pid(ii, jj), pid(ii + 1, jj), pid(ii + 1, jj + 1), pid(ii, jj + 1)], axis=1)
    nv = np.full(cells.shape[0], 4)

    def marker(a, b):
        ja, jb = a // (nx + 1), b // (nx + 1)
        ia, ib = a % (nx + 1), b % (nx + 1)
        out = np.full(a.shape[0], inout_marker)
        out[(ja == 0) & (jb == 0)] = lower_marker
        out[(ja == ny) & (jb == ny)] = upper_marker
        out[((ia == 0) & (ib == 0)) | ((ia == nx) & (ib == nx))] = inout_marker
        return out

    return build_grid(x, y, cells, nv, marker, f"bump{nx}x{ny}")


# --------------------------------------------------------------------------
# 15-degree compression ramp, mixed tri/quad
# --------------------------------------------------------------------------
def ramp(nx: int = 512, ny: int = 256, split_fraction: float = 0.5,
         seed: int = DEFAULT_SEED, angle_deg: float = 15.0) -> Grid:
    """Ramp: SLIP bottom (BC1, file 0), FREESTREAM inlet (BC2, file -1),
    EXTRAPOLATION outlet and top (raw file marker -2)."""
    xs = 2.0 * np.arange(nx + 1) / nx
    yb = np.where(xs > 0.5, (xs - 0.5) * math.tan(math.radians(angle_deg)), 0.0)
    j = np.arange(ny + 1)
    X = np.repeat(xs[None, :], ny + 1, axis=0)
    Y = yb[None, :] + (1.0 - yb[None, :]) * (j[:, None] / ny)
    x = X.ravel()
    y = Y.ravel()
    pid = lambda ii, jj: jj * (nx + 1) + ii
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny))
    ii = ii.ravel()
    jj = jj.ravel()
    nq = nx * ny
    nsplit = int(round(split_fraction * nq))
    keys = splitmix_u01(seed, 3, nq)
    split = np.zeros(nq, dtype=bool)
    split[np.argsort(keys, kind="stable")[:nsplit]] = True
    diag = splitmix_u01(seed, 4, nq) < 0.5
    q0, q1, q2, q3 = pid(ii, jj), pid(ii + 1, jj), pid(ii + 1, jj + 1), pid(ii, jj + 1)
    rows = []
    nvs = []
    # cells are emitted row by row (wall row first), a split quad as two triangles
    cells = np.full((nq + nsplit, 4), -1, dtype=np.int64)
    nv = np.zeros(nq + nsplit, dtype=np.int64)
    extra = np.cumsum(split) - split          # splits before this quad
    base = np.arange(nq) + extra
    s = split
    # unsplit quads
    u = ~s
    cells[base[u]] = np.stack([q0[u], q1[u], q2[u], q3[u]], axis=1)
    nv[base[u]] = 4
    # split along q0-q2 (diag) or q1-q3
    d = s & diag
    cells[base[d], :3] = np.stack([q0[d], q1[d], q2[d]], axis=1)
    cells[base[d] + 1, :3] = np.stack([q0[d], q2[d], q3[d]], axis=1)
    e = s & ~diag
    cells[base[e], :3] = np.stack([q0[e], q1[e], q3[e]], axis=1)
    cells[base[e] + 1, :3] = np.stack([q1[e], q2[e], q3[e]], axis=1)
    nv[base[s]] = 3
    nv[base[s] + 1] = 3
    del rows, nvs

    def marker(a, b):
        ja, jb = a // (nx + 1), b // (nx + 1)
        ia, ib = a % (nx + 1), b % (nx + 1)
        out = np.full(a.shape[0], -2)
        out[(ja == 0) & (jb == 0)] = 0
        out[(ia == 0) & (ib == 0)] = -1
        return out

    return build_grid(x, y, cells, nv, marker, f"ramp{nx}x{ny}")


# --------------------------------------------------------------------------
# O-grid around a circular cylinder of diameter 1
# --------------------------------------------------------------------------
def ogrid_ratio(first: float, total: float, n: int) -> float:
    """Geometric ratio q with first*(q^n-1)/(q-1) == total (bisection)."""
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first * (mid ** n - 1.0) / (mid - 1.0) > total:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def ogrid(ni: int = 1024, nj: int = 512, r0: float = 0.5, r1: float = 30.0,
          first: float = 1e-3) -> Grid:
    """O-grid: NO_SLIP cylinder (BC1, file 0; wall face k owned by cell k),
    FREESTREAM far field (BC2, file -1)."""
    q = ogrid_ratio(first, r1 - r0, nj)
    h = first * q ** np.arange(nj)
    r = r0 + np.concatenate([[0.0], np.cumsum(h)])
    r[-1] = r1
    th = 2.0 * math.pi * np.arange(ni) / ni
    X = r[:, None] * np.cos(th)[None, :]
    Y = r[:, None] * np.sin(th)[None, :]
    x = X.ravel()
    y = Y.ravel()
    pid = lambda ii, jj: jj * ni + (ii % ni)
    ii, jj = np.meshgrid(np.arange(ni), np.arange(nj))
    ii = ii.ravel()
    jj = jj.ravel()
    cells = np.stack([pid(ii, jj), pid(ii + 1, jj), pid(ii + 1, jj + 1), pid(ii, jj + 1)], axis=1)
    nv = np.full(cells.shape[0], 4)

    def marker(a, b):
        ja, jb = a // ni, b // ni
        out = np.full(a.shape[0], -1)
        out[(ja == 0) & (jb == 0)] = 0
        return out

    return build_grid(x, y, cells, nv, marker, f"ogrid{ni}x{nj}")


# --------------------------------------------------------------------------
# centroidal dual of a Delaunay triangulation around a NACA0012 section
# --------------------------------------------------------------------------
def naca0012_half_thickness(xr: np.ndarray) -> np.ndarray:
    """Half thickness of the NACA 0012 section at chord fraction xr in [0, 1]
    (closed trailing edge form of the 4-digit thickness polynomial)."""
    xr = np.clip(xr, 0.0, 1.0)
    return 5.0 * 0.12 * (0.2969 * np.sqrt(xr) - 0.1260 * xr - 0.3516 * xr ** 2
                         + 0.2843 * xr ** 3 - 0.1036 * xr ** 4)


def airfoil_dual(n_points: int = 8_000_000, seed: int = DEFAULT_SEED, length: float = 4.0,
                 chord: float = 1.0, x_le: float = 1.0, y_c: float = 0.5,
                 jitter: float = 0.25, log: Callable[[str], None] = None) -> Grid:
    """Polygonal mesh: the centroidal dual of the Delaunay triangulation of
    about n_points seed-jittered lattice points in [0, length] x [0, 1] with a
    NACA0012 section (chord `chord`, leading edge at (x_le, y_c), zero angle)
    cut out.  Each dual cell joins the centroids of the triangles around one
    Delaunay vertex, so it has as many sides as the vertex has neighbours
    (3 to about 9).  Vertices on the convex hull, on the section and of
    triangles removed inside it get no cell; the dual faces towards them are
    boundary faces: SLIP wall (BC1, file 0) towards the section, FREESTREAM
    (BC2, file -1) towards the outer hull."""
    from scipy.spatial import Delaunay
    ny = max(8, int(round(math.sqrt(n_points / length))))
    nx = int(round(length * ny))
    h = 1.0 / ny
    X, Y = np.meshgrid((np.arange(nx) + 0.5) * (length / nx), (np.arange(ny) + 0.5) * h)
    n = X.size
    px = X.ravel() + (splitmix_u01(seed, 21, n) - 0.5) * 2.0 * jitter * h
    py = Y.ravel() + (splitmix_u01(seed, 22, n) - 0.5) * 2.0 * jitter * h
    # drop lattice points inside the section or within half a spacing of it
    xr = (px - x_le) / chord
    near = (xr > -0.5 * h / chord) & (xr < 1.0 + 0.5 * h / chord) & \
        (np.abs(py - y_c) < chord * naca0012_half_thickness(xr) + 0.5 * h)
    px, py = px[~near], py[~near]
    # surface points about one lattice spacing apart in arc length
    tf = 0.5 * (1.0 - np.cos(np.linspace(0.0, math.pi, 20001)))
    arc = np.r_[0.0, np.cumsum(np.hypot(np.diff(tf) * chord,
                                        np.diff(chord * naca0012_half_thickness(tf))))]
    ns = max(16, int(math.ceil(arc[-1] / h)) + 1)
    t = np.interp(np.linspace(0.0, arc[-1], ns), arc, tf)
    yt = chord * naca0012_half_thickness(t)
    sx = np.concatenate([x_le + chord * t, x_le + chord * t[1:-1]])
    sy = np.concatenate([y_c + yt, y_c - yt[1:-1]])
    # points on the rectangle's sides, half a spacing out of phase with the
    # lattice rows and columns, so that the hull is the rectangle itself
    bx = np.arange(nx + 1) * (length / nx)
    by = np.arange(1, ny) * h
    rx = np.concatenate([bx, bx, np.zeros(ny - 1), np.full(ny - 1, length)])
    ry = np.concatenate([np.zeros(nx + 1), np.ones(nx + 1), by, by])
    n_lat = px.size
    px = np.concatenate([px, rx, sx])
    py = np.concatenate([py, ry, sy])
    n_lat += rx.size
    if log:
        log(f"airfoil_dual: Delaunay of {px.size} points")
    tri = Delaunay(np.stack([px, py], axis=1)).simplices.astype(np.int64)
    if log:
        log(f"airfoil_dual: {tri.shape[0]} triangles; building the dual")
    cx = px[tri].mean(axis=1)
    cy = py[tri].mean(axis=1)
    xr = (cx - x_le) / chord
    inside = (xr >= 0.0) & (xr <= 1.0) & (np.abs(cy - y_c) < chord * naca0012_half_thickness(xr))
    # boundary vertices: hull, section surface, vertices of removed triangles
    npts = px.size
    bnd = np.zeros(npts, dtype=bool)
    wall = np.zeros(npts, dtype=bool)
    wall[n_lat:] = True
    wall[tri[inside].ravel()] = True
    deg_all = np.bincount(tri.ravel(), minlength=npts)
    tri = tri[~inside]
    cx, cy = cx[~inside], cy[~inside]
    # a vertex is interior when its triangle fan closes: every edge at it is
    # shared by two kept triangles (hull vertices and vertices next to the
    # cut-out have an edge with one triangle only)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * npts + e[:, 1]
    uk, cnt = np.unique(key, return_counts=True)
    open_e = uk[cnt == 1]
    bnd[open_e // npts] = True
    bnd[open_e % npts] = True
    bnd |= wall
    del deg_all
    # fans: the triangles around each interior vertex, counter-clockwise
    vert = tri.ravel()
    tid = np.repeat(np.arange(tri.shape[0]), 3)
    keep = ~bnd[vert]
    vert, tid = vert[keep], tid[keep]
    ang = np.arctan2(cy[tid] - py[vert], cx[tid] - px[vert])
    order = np.lexsort((ang, vert))
    vert, tid = vert[order], tid[order]
    starts = np.flatnonzero(np.r_[True, vert[1:] != vert[:-1]])
    counts = np.diff(np.r_[starts, vert.size])
    ncell = starts.size
    m = int(counts.max())
    pos = np.arange(vert.size) - np.repeat(starts, counts)
    cells = np.full((ncell, m), -1, dtype=np.int64)
    cells[np.repeat(np.arange(ncell), counts), pos] = tid
    # boundary dual edges: between consecutive fan triangles t_k, t_k+1 of
    # vertex v lies the Delaunay edge (v, w); w is a boundary vertex there
    nxt = np.where(pos + 1 < np.repeat(counts, counts), np.arange(vert.size) + 1,
                   np.repeat(starts, counts))
    ta, tb = tid, tid[nxt]
    va = tri[ta]
    vb = tri[tb]
    common = (va[:, :, None] == vb[:, None, :]).any(axis=2)
    other = np.where(common & (va != vert[:, None]), va, -1).max(axis=1)
    is_b = bnd[other]
    lo = np.minimum(ta, tb)[is_b]
    hi = np.maximum(ta, tb)[is_b]
    mk = np.where(wall[other[is_b]], 0, -1)
    bkey = lo * (tri.shape[0] + 1) + hi
    srt = np.argsort(bkey)
    bkey, mk = bkey[srt], mk[srt]

    def marker(a, b):
        k = np.minimum(a, b) * (tri.shape[0] + 1) + np.maximum(a, b)
        i = np.searchsorted(bkey, k)
        if np.any(i >= bkey.size) or np.any(bkey[np.minimum(i, bkey.size - 1)] != k):
            raise ValueError("boundary dual edge without a Delaunay boundary vertex")
        return mk[i]

    if log:
        log(f"airfoil_dual: {ncell} cells; faces and canonical order")
    return build_grid(cx, cy, cells, counts, marker, f"airfoil_dual{n_points}")


def box(nx: int = 4, ny: int = 3, jitter: float = 0.0, seed: int = DEFAULT_SEED,
        split_fraction: float = 0.0, markers=(0, -1, -2, -3)) -> Grid:
    """Unit-square test grid; boundary sides bottom/right/top/left carry the
    given file markers; optional jitter and triangle splits (mixed cells)."""
    xs = np.arange(nx + 1) / nx
    ys = np.arange(ny + 1) / ny
    X, Y = np.meshgrid(xs, ys)
    X = X.copy()
    Y = Y.copy()
    if jitter > 0:
        n = (nx + 1) * (ny + 1)
        ux = splitmix_u01(seed, 11, n).reshape(ny + 1, nx + 1) * 2.0 - 1.0
        uy = splitmix_u01(seed, 12, n).reshape(ny + 1, nx + 1) * 2.0 - 1.0
        X[1:-1, 1:-1] += jitter / nx * ux[1:-1, 1:-1]
        Y[1:-1, 1:-1] += jitter / ny * uy[1:-1, 1:-1]
    x = X.ravel()
    y = Y.ravel()
    pid = lambda ii, jj: jj * (nx + 1) + ii
    cl = []
    nvl = []
    keys = splitmix_u01(seed, 13, nx * ny)
    for jj in range(ny):
        for ii in range(nx):
            q = [pid(ii, jj), pid(ii + 1, jj), pid(ii + 1, jj + 1), pid(ii, jj + 1)]
            if keys[jj * nx + ii] < split_fraction:
                cl.append([q[0], q[1], q[2], -1]); nvl.append(3)
                cl.append([q[0], q[2], q[3], -1]); nvl.append(3)
            else:
                cl.append(q); nvl.append(4)
    cells = np.array(cl, dtype=np.int64)
    nv = np.array(nvl, dtype=np.int64)

    def marker(a, b):
        ja, jb = a // (nx + 1), b // (nx + 1)
        ia, ib = a % (nx + 1), b % (nx + 1)
        out = np.zeros(a.shape[0], dtype=np.int64)
        out[(ja == 0) & (jb == 0)] = markers[0]
        out[(ia == nx) & (ib == nx)] = markers[1]
        out[(ja == ny) & (jb == ny)] = markers[2]
        out[(ia == 0) & (ib == 0)] = markers[3]
        return out

    return build_grid(x, y, cells, nv, marker, f"box{nx}x{ny}")


def check_canonical(g: Grid) -> list:
    """Return the list of canonical-ordering violations (empty when canonical)."""
    errs = []
    bnd = g.face_c2 < 0
    nb = int(bnd.sum())
    if nb and not np.all(bnd[:nb]):
        errs.append("boundary faces are not first")
    ranks = -g.bc_marker[:nb]
    if nb and np.any(np.diff(ranks) < 0):
        errs.append("boundary groups not in marker order 0,-1,-2,-3")
    counts = np.bincount(g.face_c1, minlength=g.num_cells) + np.bincount(
        g.face_c2[~bnd], minlength=g.num_cells)
    if counts.max() > 6:
        errs.append("cell with more than 6 faces")
    return errs
