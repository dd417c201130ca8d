"""ctypes binding of liboracle.so and the FVDUMP reader (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

BC_CODES = {"FREESTREAM": -1, "SLIP_WALL": -2, "EXTRAPOLATION": -3, "NO_SLIP_WALL": -4}


class OrOptions(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("limiter", ctypes.c_int32), ("eqn", ctypes.c_int32),
                ("flux", ctypes.c_int32), ("dt_type", ctypes.c_int32), ("bc1", ctypes.c_int32),
                ("bc2", ctypes.c_int32), ("scheme", ctypes.c_int32), ("cfl", ctypes.c_double),
                ("gamma", ctypes.c_double), ("R", ctypes.c_double), ("mu", ctypes.c_double),
                ("Pr", ctypes.c_double), ("P", ctypes.c_double), ("rho", ctypes.c_double),
                ("u", ctypes.c_double), ("v", ctypes.c_double)]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (always) and the reference harness when the
    reference tree is present."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.check_call(["make", "-s", "-C", HERE, "-B", "liboracle.so"], stdout=out)
    if os.path.isdir("/root/reference/src/solver"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"], stdout=out)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int32] * 3 + [P(ctypes.c_double)] * 2 + \
            [P(ctypes.c_int32)] * 4 + [P(OrOptions)]
        L.or_error.restype = ctypes.c_char_p
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_set_state.argtypes = [ctypes.c_void_p] + [P(ctypes.c_double)] * 4
        L.or_step.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.or_res.argtypes = [ctypes.c_void_p]
        L.or_residual_l2.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        L.or_wall_forces.argtypes = [ctypes.c_void_p, ctypes.c_double, P(ctypes.c_double)]
        L.or_set_fixed_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_roe_flux.argtypes = [P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, P(ctypes.c_double)]
        L.or_array.restype = ctypes.c_void_p
        L.or_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(ctypes.c_int64), P(ctypes.c_char)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def options_from_case(case) -> OrOptions:
    """OrOptions from a fv2d.inputs.Case (same enum codes as read_inputs.cpp)."""
    lim = {"NONE": 0, "MINMOD": 1, "SQUEEZE": 2, "BARTH_JESPERSEN": 3}[case.limiter]
    return OrOptions(order=case.poly + 1, limiter=lim, eqn={"EULER": 1, "NS": 2}[case.eqn],
                     flux={"RUSANOV": 1, "ROE": 2}[case.flux],
                     dt_type={"GLOBAL": 1, "LOCAL": 2, "UNIFORM": 3}[case.dt_type],
                     bc1=BC_CODES[case.BC1], bc2=BC_CODES[case.BC2],
                     scheme={"SSP_RK2": 1, "SSP_RK3": 2, "SSP_RK2_TWO_STAGE": 3}[case.scheme],
                     cfl=case.CFL,
                     gamma=case.gamma, R=case.R, mu=case.mu, Pr=case.Pr, P=case.P, rho=case.rho,
                     u=case.u, v=case.v)


class Oracle:
    """One CPU solver instance (the reference's algorithm restated in C)."""

    def __init__(self, grid, case):
        L = lib()
        self._keep = [np.ascontiguousarray(grid.x, dtype=np.float64),
                      np.ascontiguousarray(grid.y, dtype=np.float64),
                      np.ascontiguousarray(grid.face_p1, dtype=np.int32),
                      np.ascontiguousarray(grid.face_p2, dtype=np.int32),
                      np.ascontiguousarray(grid.face_c1, dtype=np.int32),
                      np.ascontiguousarray(grid.reader_c2(), dtype=np.int32)]
        x, y, p1, p2, c1, c2 = self._keep
        self.opt = options_from_case(case)
        self.h = L.or_create(grid.num_points, grid.num_faces, grid.num_cells, _dp(x), _dp(y),
                             _ip(p1), _ip(p2), _ip(c1), _ip(c2), ctypes.byref(self.opt))
        if not self.h:
            raise ValueError(L.or_error().decode())
        self.nc = grid.num_cells
        self.nf = grid.num_faces

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def step(self, n: int = 1) -> None:
        lib().or_step(self.h, n)

    def res(self) -> None:
        lib().or_res(self.h)

    def set_state(self, Q: np.ndarray) -> None:
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(4, self.nc)
        lib().or_set_state(self.h, _dp(Q[0]), _dp(Q[1]), _dp(Q[2]), _dp(Q[3]))

    def get(self, name: str) -> np.ndarray:
        n = ctypes.c_int64()
        k = ctypes.c_char()
        p = lib().or_array(self.h, name.encode(), ctypes.byref(n), ctypes.byref(k))
        if not p:
            raise KeyError(name)
        ct = ctypes.c_double if k.value == b"d" else ctypes.c_int32
        buf = (ct * n.value).from_address(p)
        return np.ctypeslib.as_array(buf).copy()

    def state(self) -> np.ndarray:
        return np.stack([self.get(f"p{k}") for k in range(1, 5)])

    def grad(self) -> np.ndarray:
        return np.stack([self.get(f"Qxp{k}") for k in range(1, 5)] +
                        [self.get(f"Qyp{k}") for k in range(1, 5)])

    def residual(self) -> np.ndarray:
        return np.stack([self.get(f"R{k}") for k in range(1, 5)])

    def set_fixed_dt(self, dt: float) -> None:
        """dt > 0: every following step takes this dt in every cell (tests)."""
        lib().or_set_fixed_dt(self.h, dt)

    def residual_l2(self) -> np.ndarray:
        out = (ctypes.c_double * 4)()
        lib().or_residual_l2(self.h, out)
        return np.array(out[:])

    def wall_forces(self, length: float = 1.0) -> np.ndarray:
        out = (ctypes.c_double * 2)()
        lib().or_wall_forces(self.h, length, out)
        return np.array(out[:])


def roe_flux(qL, qR, nx: float, ny: float, gamma: float = 1.4) -> np.ndarray:
    """The oracle's Roe flux F.n (the flux = 2 extension) between two conserved
    states across the unit normal (nx, ny)."""
    a = np.ascontiguousarray(qL, dtype=np.float64)
    b = np.ascontiguousarray(qR, dtype=np.float64)
    out = np.zeros(4)
    lib().or_roe_flux(_dp(a), _dp(b), nx, ny, gamma, _dp(out))
    return out


def read_dump(path: str) -> Dict[str, np.ndarray]:
    """Parse an FVDUMP01 file written by oracle/_ref/ref_harness."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"FVDUMP01":
        raise ValueError("not an FVDUMP01 file")
    out = {}
    pos = 8
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "little"); pos += 4
        name = data[pos:pos + n].decode(); pos += n
        kind = chr(data[pos]); pos += 1
        cnt = int.from_bytes(data[pos:pos + 8], "little"); pos += 8
        dt = np.float64 if kind == "d" else np.int32
        nb = cnt * np.dtype(dt).itemsize
        out[name] = np.frombuffer(data[pos:pos + nb], dtype=dt).copy()
        pos += nb
    return out


def run_reference(input_path: str, out_path: str, nsteps: int, dumps=None, state_path=None,
                  res_only: bool = False, timeit: bool = False) -> str:
    """Run the compiled reference harness (cwd = the input file's folder)."""
    cmd = [REF_HARNESS, os.path.basename(input_path), os.path.abspath(out_path), str(nsteps)]
    if dumps:
        cmd += ["--dump", ",".join(str(d) for d in dumps)]
    if state_path:
        cmd += ["--state", os.path.abspath(state_path)]
    if res_only:
        cmd += ["--res-only"]
    if timeit:
        cmd += ["--time"]
    r = subprocess.run(cmd, cwd=os.path.dirname(os.path.abspath(input_path)), capture_output=True,
                       text=True, check=True)
    return r.stdout
