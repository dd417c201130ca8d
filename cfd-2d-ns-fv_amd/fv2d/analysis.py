"""Wall quantities of a solution: pressure coefficient along the wall and the
force coefficients, for comparing the FP32-mixed and FP64 solutions.

The reference's own CL/CD (save_sol.cpp:20-75) truncates centroid and face
coordinates to int before reconstructing the wall state, which makes its
value meaningless on meshes whose cells are smaller than one length unit
(the drop-in CLI reproduces it as it is).  This is the integrator the
comparison uses instead: the wall state is the owner cell's state
reconstructed to the face centre with its limited gradient (or the cell
state at first order), pressure acts along the outward face normal, and at
no-slip walls the shear is mu * |u_c| / d_n with d_n the centroid's distance
to the face centre, as reconstruction.cpp:110-123 forms the wall gradient.
"""
from __future__ import annotations

import numpy as np

from .lib import host_geometry


GEOMETRY = ("face_nx", "face_ny", "face_area", "face_centerx", "face_centery", "cell_centerx",
            "cell_centery")


def wall_geometry(grid, case, solver=None):
    """The host geometry arrays the wall quantities use (compute once per mesh
    and pass as `geo` when comparing several solutions); from an existing
    solver context of the mesh when given, which holds them already."""
    if solver is not None:
        return {k: solver.geometry(k) for k in GEOMETRY}
    return {k: host_geometry(grid, case, k) for k in GEOMETRY}


def wall_state(grid, case, q, grad=None, geo=None):
    """Face index, conserved state at the face centre, and geometry of every
    face of the BC1 group (file marker 0)."""
    walls = np.flatnonzero((grid.face_c2 < 0) & (grid.bc_marker == 0))
    geo = geo or wall_geometry(grid, case)
    c = grid.face_c1[walls]
    qf = q[:, c].copy()
    dx = geo["face_centerx"][walls] - geo["cell_centerx"][c]
    dy = geo["face_centery"][walls] - geo["cell_centery"][c]
    if grad is not None:
        qf += grad[0:4, c] * dx + grad[4:8, c] * dy
    return walls, qf, q[:, c], geo, dx, dy


def wall_coefficients(grid, case, q, grad=None, length: float = 1.0, geo=None):
    """(cp, x_wall, cl, cd): pressure coefficient at each wall face (ordered as
    the faces), the wall faces' x, and the lift and drag coefficients per
    unit span referred to `length`."""
    walls, qf, qc, geo, dx, dy = wall_state(grid, case, q, grad, geo)
    g = case.gamma
    rho, mx, my, E = qf
    P = (E - 0.5 * (mx * mx + my * my) / rho) * (g - 1)
    vinf2 = case.u ** 2 + case.v ** 2
    qdyn = 0.5 * case.rho * vinf2
    nx, ny, A = geo["face_nx"][walls], geo["face_ny"][walls], geo["face_area"][walls]
    fx = np.sum(P * A * nx)
    fy = np.sum(P * A * ny)
    if case.eqn == "NS" and case.BC1 == "NO_SLIP_WALL":
        dn = np.hypot(dx, dy)
        u = qc[1] / qc[0]
        v = qc[2] / qc[0]
        # shear on the body: mu * (velocity at the centroid) / d_n, along it
        fx += np.sum(case.mu * u / dn * A)
        fy += np.sum(case.mu * v / dn * A)
    vinf = np.sqrt(vinf2)
    ex, ey = case.u / vinf, case.v / vinf
    drag = fx * ex + fy * ey
    lift = -fx * ey + fy * ex
    cp = (P - case.P) / qdyn
    return cp, geo["face_centerx"][walls], lift / (qdyn * length), drag / (qdyn * length)


def compare_walls(grid, case, q32, g32, q64, g64, geo=None):
    """FP32-mixed against FP64 wall quantities."""
    geo = geo or wall_geometry(grid, case)
    cp32, _, cl32, cd32 = wall_coefficients(grid, case, q32, g32, geo=geo)
    cp64, _, cl64, cd64 = wall_coefficients(grid, case, q64, g64, geo=geo)
    return {"cd_fp64": cd64, "cd_fp32": cd32, "cl_fp64": cl64, "cl_fp32": cl32,
            "cd_rel_err": abs(cd32 - cd64) / abs(cd64),
            "cp_l2_err": float(np.linalg.norm(cp32 - cp64) / np.sqrt(cp64.size)),
            "cp_linf_err": float(np.max(np.abs(cp32 - cp64))),
            "wall_faces": int(cp64.size)}


def strouhal(t, cl, length: float, speed: float, tail: float = 0.5):
    """Shedding frequency of a lift history: St = f * length / speed.

    Over the last `tail` fraction of the record (the periodic part), f is
    the number of whole periods between the first and last upward zero
    crossings of cl - mean(cl) (crossing times interpolated linearly) over
    the time between them; the peak of the record's spectrum (parabolic
    interpolation between bins) is given beside it as a cross-check.
    Returns {"st", "st_fft", "periods", "cl_amplitude", "cl_mean"}; st is None
    when fewer than two crossings are found."""
    t = np.asarray(t, dtype=np.float64)
    cl = np.asarray(cl, dtype=np.float64)
    k0 = int(len(t) * (1.0 - tail))
    t, cl = t[k0:], cl[k0:]
    x = cl - cl.mean()
    up = np.flatnonzero((x[:-1] < 0) & (x[1:] >= 0))
    tc = t[up] - x[up] * (t[up + 1] - t[up]) / (x[up + 1] - x[up])
    out = {"st": None, "st_fft": None, "periods": max(0, len(tc) - 1),
           "cl_amplitude": float(0.5 * (cl.max() - cl.min())) if len(cl) else None,
           "cl_mean": float(cl.mean()) if len(cl) else None}
    if len(tc) >= 2:
        f = (len(tc) - 1) / (tc[-1] - tc[0])
        out["st"] = float(f * length / speed)
    if len(t) >= 8:
        dt = (t[-1] - t[0]) / (len(t) - 1)
        spec = np.abs(np.fft.rfft(x * np.hanning(len(x))))
        j = int(np.argmax(spec[1:])) + 1
        if 1 <= j < len(spec) - 1:
            a, b, c = spec[j - 1], spec[j], spec[j + 1]
            j = j + 0.5 * (a - c) / (a - 2 * b + c)
        out["st_fft"] = float(j / (len(x) * dt) * length / speed)
    return out
