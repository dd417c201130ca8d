"""input.in options: the reference's nine sections (src/base/read_inputs.cpp:20-209).

`Case` holds one run's options with the reference's spelling; `write()` emits a
file both the reference and the drop-in `fv2d` CLI read.  Numbers are written
with repr() so they round-trip exactly through stod().
"""
from __future__ import annotations

import dataclasses
import math

# enum codes, read_inputs.cpp:45-56,75-92,133-151,171-205
LIMITERS = {"NONE": 0, "MINMOD": 1, "SQUEEZE": 2}
EQNS = {"EULER": 1, "NS": 2}
FLUXES = {"RUSANOV": 1, "ROE": 2}
DT_TYPES = {"GLOBAL": 1, "LOCAL": 2}
SCHEMES = {"SSP_RK2": 1, "SSP_RK3": 2}
BCS = {"FREESTREAM": -1, "SLIP_WALL": -2, "EXTRAPOLATION": -3, "NO_SLIP_WALL": -4}


@dataclasses.dataclass
class Case:
    grid_file: str = "grid.in"
    nmax: int = 100
    monitor_step: int = 10
    output_step: int = 1000000
    poly: int = 0
    restart: int = 0
    eqn: str = "EULER"
    flux: str = "RUSANOV"
    gamma: float = 1.4
    R: float = 287.0
    mu: float = 0.0
    Pr: float = 0.72
    P: float = 1e5
    rho: float = 1.225
    u: float = 0.0
    v: float = 0.0
    scheme: str = "SSP_RK3"
    dt_type: str = "GLOBAL"
    CFL: float = 0.5
    limiter: str = "NONE"
    length: float = 1.0
    BC1: str = "SLIP_WALL"
    BC2: str = "FREESTREAM"

    def write(self, path: str) -> None:
        r = repr
        txt = f"""[case]
    nmax = {self.nmax}
    monitor_step = {self.monitor_step}
    output_step = {self.output_step}
    solution_poly_order = {self.poly}
    restart = {self.restart}

[grid]
    grid_file = {self.grid_file}

[equations]
    eqn = {self.eqn}
    flux_solver = {self.flux}

[fluid]
    gamma = {r(float(self.gamma))}
    R = {r(float(self.R))}
    mu = {r(float(self.mu))}
    Pr = {r(float(self.Pr))}

[init]
    P = {r(float(self.P))}
    rho = {r(float(self.rho))}
    u = {r(float(self.u))}
    v = {r(float(self.v))}

[time]
    time_scheme = {self.scheme}
    time_step_type = {self.dt_type}
    CFL = {r(float(self.CFL))}

[limiter]
    limiter = {self.limiter}

[report]
    length = {r(float(self.length))}

[BC]
    BC1 = {self.BC1}
    BC2 = {self.BC2}
"""
        with open(path, "w") as f:
            f.write(txt)


def mach_velocity(mach: float, P: float = 1e5, rho: float = 1.225, gamma: float = 1.4,
                  angle_deg: float = 0.0):
    """(u, v) of a freestream at the given Mach number and flow angle."""
    c = math.sqrt(gamma * (P / rho))
    a = math.radians(angle_deg)
    return mach * c * math.cos(a), mach * c * math.sin(a)
