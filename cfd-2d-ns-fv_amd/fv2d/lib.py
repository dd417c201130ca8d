"""ctypes binding of libfvk.so (include/fvk.h).

The binding is deliberately thin: it only converts numpy arrays to the C ABI's
plain pointers.  It never falls back to any CPU path: if libfvk.so is missing
it raises, and every solver call runs the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfvk.so")
CLI_PATH = os.path.join(PKG_DIR, "lib", "fv2d")

FVK_OK = 0
ERRORS = {-1: "FVK_ERR_ARG", -2: "FVK_ERR_MESH", -3: "FVK_ERR_OPTION", -4: "FVK_ERR_HIP",
          -5: "FVK_ERR_NONFINITE", -6: "FVK_ERR_UNSUPPORTED"}
BC_CODES = {"FREESTREAM": -1, "SLIP_WALL": -2, "EXTRAPOLATION": -3, "NO_SLIP_WALL": -4}
# precision forms (FVK_FP64, FVK_FP32_MIXED)
FP64 = 0
FP32_MIXED = 1
# why a batch instance is not active (FVK_RETIRED_*)
RETIRED_NONE, RETIRED_DROP, RETIRED_NONFINITE, RETIRED_STALL, RETIRED_CALLER = range(5)

# every symbol include/fvk.h declares
EXPORTS = ["fvk_default_options", "fvk_create", "fvk_destroy", "fvk_last_error",
           "fvk_init_freestream", "fvk_set_state", "fvk_step", "fvk_get_state",
           "fvk_use_graph", "fvk_get_residual", "fvk_residual_l2", "fvk_residual_l2_noalloc",
           "fvk_conservation_sums",
           "fvk_get_geometry", "fvk_geometry",
           "fvk_get_info", "fvk_get_time", "fvk_get_options", "fvk_time_steps", "fvk_time_kernels",
           "fvk_batch_create", "fvk_batch_destroy", "fvk_batch_last_error", "fvk_batch_init",
           "fvk_batch_step", "fvk_batch_set_active", "fvk_batch_residual_l2",
           "fvk_batch_conservation_sums", "fvk_batch_set_stall", "fvk_batch_retired",
           "fvk_batch_set_reference",
           "fvk_batch_run_to_drop", "fvk_batch_get_state", "fvk_batch_get_info",
           "fvk_batch_get_time", "fvk_batch_get_options", "fvk_batch_use_graph",
           "fvk_batch_time_kernels",
           "fvk_checkpoint_size", "fvk_checkpoint_save", "fvk_checkpoint_load",
           "fvk_batch_checkpoint_size", "fvk_batch_checkpoint_save", "fvk_batch_checkpoint_load"]


class FvkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class FvkMesh(ctypes.Structure):
    _fields_ = [("num_points", ctypes.c_int32), ("num_faces", ctypes.c_int32),
                ("num_cells", ctypes.c_int32), ("reserved_", ctypes.c_int32),
                ("x", ctypes.POINTER(ctypes.c_double)), ("y", ctypes.POINTER(ctypes.c_double)),
                ("face_point1", ctypes.POINTER(ctypes.c_int32)),
                ("face_point2", ctypes.POINTER(ctypes.c_int32)),
                ("face_cell1", ctypes.POINTER(ctypes.c_int32)),
                ("face_cell2", ctypes.POINTER(ctypes.c_int32))]


class FvkOptions(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("limiter", ctypes.c_int32), ("eqn", ctypes.c_int32),
                ("flux_solver", ctypes.c_int32), ("time_scheme", ctypes.c_int32),
                ("time_step_type", ctypes.c_int32), ("bc1", ctypes.c_int32), ("bc2", ctypes.c_int32),
                ("cfl", ctypes.c_double), ("gamma", ctypes.c_double), ("R", ctypes.c_double),
                ("mu", ctypes.c_double), ("Pr", ctypes.c_double), ("P", ctypes.c_double),
                ("rho", ctypes.c_double), ("u", ctypes.c_double), ("v", ctypes.c_double),
                ("precision", ctypes.c_int32), ("renumber", ctypes.c_int32),
                ("block_size", ctypes.c_int32), ("xcd_remap", ctypes.c_int32),
                ("kernel_variant", ctypes.c_int32), ("accumulate", ctypes.c_int32)]


_lib = None


def lib():
    """Load libfvk.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        dp = P(ctypes.c_double)
        L.fvk_default_options.argtypes = [P(FvkOptions)]
        L.fvk_default_options.restype = None
        L.fvk_create.argtypes = [P(FvkMesh), P(FvkOptions), P(vp)]
        L.fvk_destroy.argtypes = [vp]
        L.fvk_last_error.argtypes = [vp]
        L.fvk_last_error.restype = ctypes.c_char_p
        L.fvk_init_freestream.argtypes = [vp]
        L.fvk_set_state.argtypes = [vp, dp]
        L.fvk_step.argtypes = [vp, ctypes.c_int32]
        L.fvk_get_state.argtypes = [vp, dp, dp]
        L.fvk_get_residual.argtypes = [vp, dp]
        L.fvk_residual_l2.argtypes = [vp, dp]
        L.fvk_residual_l2_noalloc.argtypes = [vp, dp]
        L.fvk_use_graph.argtypes = [vp, ctypes.c_int32]
        L.fvk_conservation_sums.argtypes = [vp, dp]
        L.fvk_get_geometry.argtypes = [vp, ctypes.c_char_p, dp]
        L.fvk_geometry.argtypes = [P(FvkMesh), P(FvkOptions), ctypes.c_char_p, dp]
        L.fvk_get_info.argtypes = [vp, P(ctypes.c_int64)]
        L.fvk_get_options.argtypes = [vp, P(FvkOptions)]
        L.fvk_get_time.argtypes = [vp, dp]
        L.fvk_time_steps.argtypes = [vp, ctypes.c_int32, dp, P(ctypes.c_int64)]
        L.fvk_time_kernels.argtypes = [vp, ctypes.c_int32, dp]
        ip = P(ctypes.c_int32)
        L.fvk_batch_create.argtypes = [ctypes.c_int32, P(FvkMesh), P(FvkOptions), P(vp)]
        L.fvk_batch_destroy.argtypes = [vp]
        L.fvk_batch_last_error.argtypes = [vp]
        L.fvk_batch_last_error.restype = ctypes.c_char_p
        L.fvk_batch_init.argtypes = [vp, P(dp)]
        L.fvk_batch_step.argtypes = [vp, ctypes.c_int32]
        L.fvk_batch_set_active.argtypes = [vp, P(ctypes.c_uint8)]
        L.fvk_batch_residual_l2.argtypes = [vp, dp]
        L.fvk_batch_conservation_sums.argtypes = [vp, dp]
        L.fvk_batch_set_stall.argtypes = [vp, ctypes.c_int32, ctypes.c_double]
        L.fvk_batch_retired.argtypes = [vp, ip]
        L.fvk_batch_set_reference.argtypes = [vp, dp]
        L.fvk_batch_run_to_drop.argtypes = [vp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                            ip, dp, dp]
        L.fvk_batch_get_state.argtypes = [vp, ctypes.c_int32, dp, dp]
        L.fvk_batch_get_info.argtypes = [vp, P(ctypes.c_int64)]
        L.fvk_batch_get_options.argtypes = [vp, ctypes.c_int32, P(FvkOptions)]
        L.fvk_batch_get_time.argtypes = [vp, dp]
        L.fvk_batch_use_graph.argtypes = [vp, ctypes.c_int32]
        L.fvk_batch_time_kernels.argtypes = [vp, ctypes.c_int32, dp]
        for pre in ("fvk_", "fvk_batch_"):
            getattr(L, pre + "checkpoint_size").argtypes = [vp, P(ctypes.c_int64)]
            getattr(L, pre + "checkpoint_save").argtypes = [vp, vp, ctypes.c_int64]
            getattr(L, pre + "checkpoint_load").argtypes = [vp, vp, ctypes.c_int64]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def make_options(case, **overrides) -> FvkOptions:
    """FvkOptions from a fv2d.inputs.Case (read_inputs.cpp enum codes)."""
    o = FvkOptions()
    lib().fvk_default_options(ctypes.byref(o))
    o.order = case.poly + 1
    o.limiter = {"NONE": 0, "MINMOD": 1, "SQUEEZE": 2, "BARTH_JESPERSEN": 3}[case.limiter]
    o.eqn = {"EULER": 1, "NS": 2}[case.eqn]
    o.flux_solver = {"RUSANOV": 1, "ROE": 2}[case.flux]
    o.time_scheme = {"SSP_RK2": 1, "SSP_RK3": 2, "SSP_RK2_TWO_STAGE": 3}[case.scheme]
    o.time_step_type = {"GLOBAL": 1, "LOCAL": 2, "UNIFORM": 3}[case.dt_type]
    o.bc1 = BC_CODES[case.BC1]
    o.bc2 = BC_CODES[case.BC2]
    o.cfl = case.CFL
    o.gamma, o.R, o.mu, o.Pr = case.gamma, case.R, case.mu, case.Pr
    o.P, o.rho, o.u, o.v = case.P, case.rho, case.u, case.v
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


class _MeshArrays:
    def __init__(self, grid):
        self.x = np.ascontiguousarray(grid.x, dtype=np.float64)
        self.y = np.ascontiguousarray(grid.y, dtype=np.float64)
        self.p1 = np.ascontiguousarray(grid.face_p1, dtype=np.int32)
        self.p2 = np.ascontiguousarray(grid.face_p2, dtype=np.int32)
        self.c1 = np.ascontiguousarray(grid.face_c1, dtype=np.int32)
        self.c2 = np.ascontiguousarray(grid.reader_c2(), dtype=np.int32)
        self.m = FvkMesh(grid.num_points, grid.num_faces, grid.num_cells, 0, _dp(self.x),
                         _dp(self.y), _ip(self.p1), _ip(self.p2), _ip(self.c1), _ip(self.c2))


def host_geometry(grid, case, name: str) -> np.ndarray:
    """Geometry array from the host-only fvk_geometry (no device needed)."""
    L = lib()
    ma = _MeshArrays(grid)
    opt = make_options(case)
    n = grid.num_cells if name in ("cell_vol", "cell_centerx", "cell_centery", "Ixx", "Iyy", "Ixy") \
        else grid.num_faces
    out = np.empty(n, dtype=np.float64)
    rc = L.fvk_geometry(ctypes.byref(ma.m), ctypes.byref(opt), name.encode(), _dp(out))
    if rc != FVK_OK:
        raise FvkError(rc, L.fvk_last_error(None).decode())
    return out


class Solver:
    """One device solver context (fvk_ctx)."""

    def __init__(self, grid, case, **overrides):
        L = lib()
        self._ma = _MeshArrays(grid)
        self.opt = make_options(case, **overrides)
        h = ctypes.c_void_p()
        rc = L.fvk_create(ctypes.byref(self._ma.m), ctypes.byref(self.opt), ctypes.byref(h))
        if rc != FVK_OK:
            raise FvkError(rc, L.fvk_last_error(None).decode())
        self.h = h
        self.nc = grid.num_cells
        self.nf = grid.num_faces

    def _check(self, rc):
        if rc != FVK_OK:
            raise FvkError(rc, lib().fvk_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().fvk_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init_freestream(self):
        self._check(lib().fvk_init_freestream(self.h))

    def set_state(self, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(4 * self.nc)
        self._check(lib().fvk_set_state(self.h, _dp(q)))

    def step(self, n: int = 1):
        self._check(lib().fvk_step(self.h, n))

    def state(self) -> np.ndarray:
        q = np.empty(4 * self.nc)
        self._check(lib().fvk_get_state(self.h, _dp(q), None))
        return q.reshape(4, self.nc)

    def state_and_grad(self):
        q = np.empty(4 * self.nc)
        g = np.empty(8 * self.nc)
        self._check(lib().fvk_get_state(self.h, _dp(q), _dp(g)))
        return q.reshape(4, self.nc), g.reshape(8, self.nc)

    def residual(self) -> np.ndarray:
        r = np.empty(4 * self.nc)
        self._check(lib().fvk_get_residual(self.h, _dp(r)))
        return r.reshape(4, self.nc)

    def residual_l2(self) -> np.ndarray:
        out = np.empty(4)
        rc = lib().fvk_residual_l2(self.h, _dp(out))
        if rc not in (FVK_OK, -5):
            self._check(rc)
        return out

    def residual_l2_noalloc(self) -> np.ndarray:
        """residual_l2 through the allocation-free entry point."""
        out = np.empty(4)
        rc = lib().fvk_residual_l2_noalloc(self.h, _dp(out))
        if rc not in (FVK_OK, -5):
            self._check(rc)
        return out

    def conservation_sums(self) -> np.ndarray:
        """sum_i V_i Q_k,i of the current state (compensated), k = 1..4."""
        out = np.empty(4)
        self._check(lib().fvk_conservation_sums(self.h, _dp(out)))
        return out

    def use_graph(self, on: bool):
        """Replay the captured step graph (default) or launch step by step."""
        self._check(lib().fvk_use_graph(self.h, 1 if on else 0))

    def geometry(self, name: str) -> np.ndarray:
        n = self.nc if name in ("cell_vol", "cell_centerx", "cell_centery", "Ixx", "Iyy", "Ixy") \
            else self.nf
        out = np.empty(n)
        self._check(lib().fvk_get_geometry(self.h, name.encode(), _dp(out)))
        return out

    def info(self) -> dict:
        a = (ctypes.c_int64 * 6)()
        self._check(lib().fvk_get_info(self.h, a))
        keys = ["num_cells", "num_faces", "num_boundary_faces", "num_slots", "max_faces",
                "num_wall"]
        return dict(zip(keys, list(a)))

    def time(self) -> float:
        """Physical time reached (uniform time step only)."""
        t = np.zeros(1)
        self._check(lib().fvk_get_time(self.h, _dp(t)))
        return float(t[0])

    def launch_options(self) -> dict:
        """Cells per block, XCD order and kernel shape in force."""
        o = FvkOptions()
        self._check(lib().fvk_get_options(self.h, ctypes.byref(o)))
        return {"block_size": o.block_size, "xcd_remap": o.xcd_remap,
                "kernel_variant": o.kernel_variant, "accumulate": o.accumulate}

    def time_kernels(self, reps: int):
        """Average microseconds per launch of pass A and pass B (stage 2)."""
        out = np.zeros(2)
        self._check(lib().fvk_time_kernels(self.h, reps, _dp(out)))
        return {"pass_a": float(out[0]), "pass_b": float(out[1])}

    def checkpoint(self) -> bytes:
        """The complete device state (fvk_checkpoint_save)."""
        n = ctypes.c_int64()
        self._check(lib().fvk_checkpoint_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(lib().fvk_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, data: bytes):
        """Resume from a checkpoint of a solver with the same mesh and options."""
        buf = ctypes.create_string_buffer(data, len(data))
        self._check(lib().fvk_checkpoint_load(self.h, buf, len(data)))

    def time_steps(self, n: int):
        ms = np.zeros(4)
        cnt = (ctypes.c_int64 * 4)()
        self._check(lib().fvk_time_steps(self.h, n, _dp(ms), cnt))
        return ms, list(cnt)


class Batch:
    """Many independent instances advanced together (fvk_batch).

    `cases` is a list of (grid, case) pairs; `overrides` apply to every
    instance's options (the kernel-shaping ones must agree)."""

    def __init__(self, cases, **overrides):
        L = lib()
        self._ma = [_MeshArrays(g) for g, _ in cases]
        self.nc = [g.num_cells for g, _ in cases]
        n = len(cases)
        meshes = (FvkMesh * n)(*[m.m for m in self._ma])
        opts = (FvkOptions * n)(*[make_options(c, **overrides) for _, c in cases])
        self.opts = opts
        h = ctypes.c_void_p()
        rc = L.fvk_batch_create(n, meshes, opts, ctypes.byref(h))
        if rc != FVK_OK:
            raise FvkError(rc, L.fvk_batch_last_error(None).decode())
        self.h = h
        self.n = n

    def _check(self, rc):
        if rc != FVK_OK:
            raise FvkError(rc, lib().fvk_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().fvk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, states=None):
        """Freestream start (states None) or one conserved state per instance
        (an entry None: that instance's freestream start)."""
        if states is None:
            self._check(lib().fvk_batch_init(self.h, None))
            return
        if len(states) != self.n:
            raise ValueError(f"init: {len(states)} states for {self.n} instances")
        keep = [None if q is None else np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
                for q in states]
        for i, q in enumerate(keep):
            if q is not None and q.size != 4 * self.nc[i]:
                raise ValueError(f"init: state {i} has {q.size} values, not 4 x {self.nc[i]}")
        arr = (ctypes.POINTER(ctypes.c_double) * self.n)(
            *[ctypes.POINTER(ctypes.c_double)() if q is None else _dp(q) for q in keep])
        self._check(lib().fvk_batch_init(self.h, arr))

    def step(self, n: int = 1):
        self._check(lib().fvk_batch_step(self.h, n))

    def set_active(self, flags):
        a = np.ascontiguousarray(np.asarray(flags, dtype=np.uint8)).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"set_active: {a.size} flags for {self.n} instances")
        self._check(lib().fvk_batch_set_active(
            self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def residual_l2(self) -> np.ndarray:
        out = np.empty(4 * self.n)
        rc = lib().fvk_batch_residual_l2(self.h, _dp(out))
        if rc not in (FVK_OK, -5):
            self._check(rc)
        return out.reshape(self.n, 4)

    def set_stall(self, checks: int, factor: float = 0.5):
        """Stall retirement in run_to_drop (checks = 0: off)."""
        self._check(lib().fvk_batch_set_stall(self.h, checks, factor))

    def set_reference(self, res0=None):
        """Reference density residual norms for the following run_to_drop
        calls (None, or an entry <= 0: the first step's norm)."""
        if res0 is None:
            self._check(lib().fvk_batch_set_reference(self.h, None))
            return
        r = np.ascontiguousarray(np.asarray(res0, dtype=np.float64)).reshape(-1)
        if r.size != self.n:
            raise ValueError(f"set_reference: {r.size} norms for {self.n} instances")
        self._check(lib().fvk_batch_set_reference(self.h, _dp(r)))

    def retired(self) -> np.ndarray:
        """Per instance: RETIRED_* code (0 active, 1 drop, 2 non-finite,
        3 stall, 4 by the caller)."""
        why = np.zeros(self.n, dtype=np.int32)
        self._check(lib().fvk_batch_retired(self.h, _ip(why)))
        return why

    def conservation_sums(self) -> np.ndarray:
        """Per instance: sum_i V_i Q_k,i of the current state, shape (n, 4)."""
        out = np.empty(4 * self.n)
        self._check(lib().fvk_batch_conservation_sums(self.h, _dp(out)))
        return out.reshape(self.n, 4)

    def run_to_drop(self, drop: float, max_steps: int, check_every: int):
        steps = np.zeros(self.n, dtype=np.int32)
        r0 = np.zeros(self.n)
        r1 = np.zeros(self.n)
        self._check(lib().fvk_batch_run_to_drop(self.h, drop, max_steps, check_every, _ip(steps),
                                                _dp(r0), _dp(r1)))
        return steps, r0, r1

    def state(self, i: int) -> np.ndarray:
        q = np.empty(4 * self.nc[i])
        self._check(lib().fvk_batch_get_state(self.h, i, _dp(q), None))
        return q.reshape(4, self.nc[i])

    def state_and_grad(self, i: int):
        q = np.empty(4 * self.nc[i])
        g = np.empty(8 * self.nc[i])
        self._check(lib().fvk_batch_get_state(self.h, i, _dp(q), _dp(g)))
        return q.reshape(4, self.nc[i]), g.reshape(8, self.nc[i])

    def info(self) -> dict:
        a = (ctypes.c_int64 * 6)()
        self._check(lib().fvk_batch_get_info(self.h, a))
        keys = ["instances", "cells", "faces", "blocks", "active_blocks", "active_instances"]
        return dict(zip(keys, list(a)))

    def time(self) -> np.ndarray:
        """Physical time reached by each instance (uniform time step only)."""
        t = np.zeros(self.n)
        self._check(lib().fvk_batch_get_time(self.h, _dp(t)))
        return t

    def launch_options(self) -> dict:
        """Cells per block, XCD order and kernel shape in force (common to all
        instances)."""
        o = FvkOptions()
        self._check(lib().fvk_batch_get_options(self.h, 0, ctypes.byref(o)))
        return {"block_size": o.block_size, "xcd_remap": o.xcd_remap,
                "kernel_variant": o.kernel_variant, "accumulate": o.accumulate}

    def checkpoint(self) -> bytes:
        n = ctypes.c_int64()
        self._check(lib().fvk_batch_checkpoint_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(lib().fvk_batch_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, data: bytes):
        buf = ctypes.create_string_buffer(data, len(data))
        self._check(lib().fvk_batch_checkpoint_load(self.h, buf, len(data)))

    def use_graph(self, on: bool):
        self._check(lib().fvk_batch_use_graph(self.h, 1 if on else 0))

    def time_kernels(self, reps: int):
        out = np.zeros(2)
        self._check(lib().fvk_batch_time_kernels(self.h, reps, _dp(out)))
        return {"pass_a": float(out[0]), "pass_b": float(out[1])}


def finish_stalled(src: Batch, dst: Batch, res0, drop: float, max_steps: int,
                   check_every: int, which=None):
    """Continue, in `dst` (a batch of the same instances, typically FP64),
    the instances that `src` (typically FP32-mixed) retired by a stall
    (or those `which` selects), from their states in `src`, to the same
    density-residual drop measured from `src`'s reference norms `res0`.
    The other instances of `dst` start from their freestream and stay
    inactive.  `which` is needed once `src.set_active` has reset the
    retirement codes.

    Returns (finished, steps, res_end, why): the boolean mask of the
    instances continued, and dst.run_to_drop's steps and last norms and
    dst's retirement codes (RETIRED_CALLER for the instances not continued).
    """
    if dst.n != src.n or list(dst.nc) != list(src.nc):
        raise ValueError("finish_stalled: the two batches hold different instances")
    finished = (src.retired() == RETIRED_STALL) if which is None else \
        np.asarray(which, dtype=bool).reshape(-1)
    if finished.size != src.n:
        raise ValueError(f"finish_stalled: {finished.size} flags for {src.n} instances")
    dst.init([src.state(i) if finished[i] else None for i in range(src.n)])
    dst.set_active(finished)
    dst.set_reference(np.where(finished, np.asarray(res0, dtype=np.float64), 0.0))
    steps, _, r1 = dst.run_to_drop(drop, max_steps, check_every)
    return finished, steps, r1, dst.retired()
