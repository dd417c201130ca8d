"""Shared helpers of the test suite: golden fixtures and exact comparisons."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from fv2d import gridgen as gg
from fv2d import inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

STEP_KEYS = ["p1", "p2", "p3", "p4", "Qxp1", "Qxp2", "Qxp3", "Qxp4", "Qyp1", "Qyp2", "Qyp3", "Qyp4",
             "R1", "R2", "R3", "R4"]


def golden_names():
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


class Golden:
    """One fixture written by tests/golden/make_golden.py."""

    def __init__(self, name: str):
        z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.a = {k: z[k] for k in z.files}
        self.meta = json.loads(bytes(self.a["meta"]).decode())
        self.grid = gg.Grid(x=self.a["grid/x"], y=self.a["grid/y"], face_p1=self.a["grid/face_p1"],
                            face_p2=self.a["grid/face_p2"], face_c1=self.a["grid/face_c1"],
                            face_c2=self.a["grid/face_c2"], bc_marker=self.a["grid/bc_marker"],
                            num_cells=int(self.a["grid/num_cells"][0]), name=name)
        self.case = inputs.Case(grid_file="grid.in", **self.meta["case"])
        self.name = name

    def step(self, n: int, key: str) -> np.ndarray:
        return self.a[f"step{n}/{key}"]

    def state(self, n: int) -> np.ndarray:
        return np.stack([self.step(n, f"p{k}") for k in range(1, 5)])

    def grad(self, n: int) -> np.ndarray:
        return np.stack([self.step(n, f"Qxp{k}") for k in range(1, 5)] +
                        [self.step(n, f"Qyp{k}") for k in range(1, 5)])

    def residual(self, n: int) -> np.ndarray:
        return np.stack([self.step(n, f"R{k}") for k in range(1, 5)])


def same(a, b) -> bool:
    """Equal as IEEE values (+0 == -0), NaN equal to NaN."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def max_rel(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(s > 0, d / s, 0.0)
    return float(np.nanmax(r)) if r.size else 0.0


def mismatch(a, b) -> str:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    idx = np.argwhere(bad)
    return f"{int(bad.sum())} of {a.size} differ, max rel {max_rel(a, b):.3e}, first at {idx[:3].tolist()}"
