"""The benchmark workloads: meshes and input options, all deterministic from
one seed (gridgen.DEFAULT_SEED unless given).

* channel  -- 64x32 bump channel, Euler, first order, GLOBAL dt, CFL 0.5, M 0.5
* ramp     -- 15-degree ramp 512x256 with a 50% triangle split, Euler, second
              order, SQUEEZE, LOCAL dt, CFL 0.8, M 2.5
* cylinder -- O-grid 1024x512 around a unit cylinder, laminar NS at Re 100,
              M 0.2, second order, SQUEEZE, GLOBAL dt, CFL 0.5
* airfoil  -- centroidal dual of the Delaunay triangulation of 8M jittered
              points in [0,4]x[0,1] around a NACA0012 section (chord 1): about
              7.8M polygonal cells of 4 to 9 sides, Euler, second order,
              SQUEEZE, LOCAL dt, CFL 0.5, M 0.7 at 1.25 degrees
* batch    -- 256 independent bump-channel or ramp instances with cell counts
              log-uniform in [2k, 200k] (about 10M cells in all), Mach uniform
              in [0.3, 0.8] (channel) or [1.5, 3.0] (ramp), flow angle uniform
              in [0, 5] degrees; Euler, second order, SQUEEZE, LOCAL dt, CFL 0.5
              (at CFL 0.8 the FP64 scheme diverges on two of the 256 ramp
              instances: M 2.74 at 3.6 degrees and M 2.41 at 4.9 degrees).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import gridgen as gg
from . import inputs


def channel():
    u, v = inputs.mach_velocity(0.5)
    return ("bump64x32 Euler o1 GLOBAL CFL0.5 M0.5", gg.bump(64, 32),
            inputs.Case(u=u, v=v, poly=0, dt_type="GLOBAL", CFL=0.5))


def ramp():
    u, v = inputs.mach_velocity(2.5)
    return ("ramp512x256-split50 Euler o2 SQUEEZE LOCAL CFL0.8 M2.5", gg.ramp(512, 256),
            inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8))


def cylinder():
    u, v = inputs.mach_velocity(0.2)
    return ("ogrid1024x512 NS Re100 o2 SQUEEZE GLOBAL CFL0.5 M0.2", gg.ogrid(1024, 512),
            inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", eqn="NS", mu=1.225 * u / 100.0,
                        BC1="NO_SLIP_WALL", dt_type="GLOBAL", CFL=0.5))


def airfoil(n_points: int = 8_000_000, log=None, cache: str = None):
    """The airfoil workload; `cache` names a Grid.save_npz file of the same
    mesh to load instead of generating it."""
    u, v = inputs.mach_velocity(0.7, angle_deg=1.25)
    grid = gg.Grid.load_npz(cache) if cache else gg.airfoil_dual(n_points, log=log)
    return (f"airfoil dual of Delaunay({n_points} points) NACA0012 Euler o2 SQUEEZE LOCAL "
            "CFL0.5 M0.7 AoA1.25", grid,
            inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.5,
                        BC1="SLIP_WALL", BC2="FREESTREAM"))


@dataclasses.dataclass
class Instance:
    kind: str            # "channel" or "ramp"
    target_cells: int
    mach: float
    angle_deg: float
    seed: int
    grid: gg.Grid = None
    case: inputs.Case = None


def batch_plan(n: int = 256, lo: int = 2000, hi: int = 200000, seed: int = gg.DEFAULT_SEED):
    """The instance list of the batch workload (meshes not yet built)."""
    kind_u = gg.splitmix_u01(seed, 101, n)
    size_u = gg.splitmix_u01(seed, 102, n)
    mach_u = gg.splitmix_u01(seed, 103, n)
    ang_u = gg.splitmix_u01(seed, 104, n)
    out = []
    for i in range(n):
        kind = "channel" if kind_u[i] < 0.5 else "ramp"
        cells = int(round(math.exp(math.log(lo) + size_u[i] * (math.log(hi) - math.log(lo)))))
        mach = 0.3 + 0.5 * mach_u[i] if kind == "channel" else 1.5 + 1.5 * mach_u[i]
        out.append(Instance(kind, cells, float(mach), float(5.0 * ang_u[i]), seed + 7919 * (i + 1)))
    return out


def build_instance(inst: Instance, cfl: float = 0.5) -> Instance:
    """Mesh and options of one batch instance: nx = 2 ny, with ny chosen so
    the cell count is nearest the target (channel 2 ny^2 quads; ramp 3 ny^2
    cells after splitting half of its 2 ny^2 quads)."""
    u, v = inputs.mach_velocity(inst.mach, angle_deg=inst.angle_deg)
    if inst.kind == "channel":
        ny = max(4, int(round(math.sqrt(inst.target_cells / 2.0))))
        inst.grid = gg.bump(2 * ny, ny, seed=inst.seed)
    else:
        ny = max(4, int(round(math.sqrt(inst.target_cells / 3.0))))
        inst.grid = gg.ramp(2 * ny, ny, seed=inst.seed)
    inst.case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=cfl)
    return inst


def batch(n: int = 256, seed: int = gg.DEFAULT_SEED):
    plan = [build_instance(i) for i in batch_plan(n, seed=seed)]
    cells = sum(i.grid.num_cells for i in plan)
    name = (f"batch{n}: channel/ramp instances, cells log-uniform [2k,200k] (total {cells}), "
            "Euler o2 SQUEEZE LOCAL CFL0.5")
    return name, plan


SINGLE = {"channel": channel, "ramp": ramp, "cylinder": cylinder, "airfoil": airfoil}


def cells_of(plan) -> np.ndarray:
    return np.array([i.grid.num_cells for i in plan], dtype=np.int64)
