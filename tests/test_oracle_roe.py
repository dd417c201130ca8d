"""The Roe flux (flux_solver ROE, an extension) of the oracle, pinned by what
the mathematics fixes.

The reference reads ROE (read_inputs.cpp:89-91) but its roe_flux.cpp is an
unfinished draft outside its solver library, and res.cpp:66-68 leaves the flux
at -101, so there is no reference output to compare with.  The oracle's Roe
flux (oracle/fv_oracle.c roe_dissipation) is pinned instead by:
  * consistency: F(q, q) = F(q) exactly;
  * Roe's property A~ (q_R - q_L) = F(q_R) - F(q_L): with every eigenvalue of
    one sign the flux is the upwind side's flux, to rounding;
  * the stationary contact, which Roe's flux keeps exactly (zero mass flux)
    and the Rusanov flux does not;
  * symmetry F(q_L, q_R, n) = -F(q_R, q_L, -n) and rotation invariance;
  * in full runs: conservation in a closed box under the uniform time step,
    and freestream preservation.
The device path evaluates the same expressions; tests/test_gpu_roe.py holds it
to the oracle bit for bit.
"""
import numpy as np
import pytest

import oracle as O

GAMMA = 1.4


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def conserved(rho, u, v, P):
    return np.array([rho, rho * u, rho * v, P / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])


def flux(q, nx, ny):
    rho, u, v, E = q[0], q[1] / q[0], q[2] / q[0], q[3]
    P = (E - 0.5 * rho * (u * u + v * v)) * (GAMMA - 1)
    vn = u * nx + v * ny
    return np.array([rho * vn, rho * u * vn + P * nx, rho * v * vn + P * ny, vn * (E + P)])


def random_state(rng, vn, nx, ny):
    """rho, P of order one, normal velocity vn, a random tangential one"""
    rho = rng.uniform(0.3, 3.0)
    P = rng.uniform(0.3, 3.0)
    vt = rng.uniform(-2.0, 2.0)
    u = vn * nx - vt * ny
    v = vn * ny + vt * nx
    return conserved(rho, u, v, P), np.sqrt(GAMMA * P / rho)


def unit(rng):
    a = rng.uniform(0, 2 * np.pi)
    return np.cos(a), np.sin(a)


def test_consistency_is_exact():
    rng = np.random.default_rng(1)
    for _ in range(50):
        nx, ny = unit(rng)
        q, _ = random_state(rng, rng.uniform(-3, 3), nx, ny)
        assert np.array_equal(O.roe_flux(q, q, nx, ny, GAMMA), flux(q, nx, ny))


@pytest.mark.parametrize("sign", [1, -1])
def test_supersonic_states_take_the_upwind_flux(sign):
    """Both sides supersonic along sign*n: every eigenvalue of the Roe matrix
    has that sign (far from the entropy fix), so |A~| = sign*A~ and Roe's
    property A~ dq = dF gives F = F(upwind side)."""
    rng = np.random.default_rng(2 + sign)
    checked = 0
    for _ in range(200):
        nx, ny = unit(rng)
        qL, cL = random_state(rng, 0.0, nx, ny)
        qR, cR = random_state(rng, 0.0, nx, ny)
        # normal velocities of 2-4 sound speeds of the faster side, same sign
        c = max(cL, cR)
        for q, m in ((qL, rng.uniform(2.0, 4.0)), (qR, rng.uniform(2.0, 4.0))):
            rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
            P = (q[3] - 0.5 * rho * (u * u + v * v)) * (GAMMA - 1)
            vt = -u * ny + v * nx
            vn = sign * m * c
            q[:] = conserved(rho, vn * nx - vt * ny, vn * ny + vt * nx, P)
        F = O.roe_flux(qL, qR, nx, ny, GAMMA)
        up = flux(qL if sign > 0 else qR, nx, ny)
        scale = np.abs(flux(qL, nx, ny)).max() + np.abs(flux(qR, nx, ny)).max()
        assert np.abs(F - up).max() <= 1e-13 * scale, (F, up)
        checked += 1
    assert checked == 200


def test_stationary_contact_is_kept_exactly():
    """Zero normal velocity, equal pressure and tangential velocity, different
    densities: the Roe flux has no dissipation (zero mass flux), the Rusanov
    flux diffuses the density jump."""
    rng = np.random.default_rng(4)
    for _ in range(50):
        nx, ny = unit(rng)
        P, vt = rng.uniform(0.5, 2.0), rng.uniform(-1.0, 1.0)
        u, v = -vt * ny, vt * nx
        qL = conserved(rng.uniform(0.5, 1.0), u, v, P)
        qR = conserved(rng.uniform(1.5, 3.0), u, v, P)
        F = O.roe_flux(qL, qR, nx, ny, GAMMA)
        assert abs(F[0]) < 1e-15 and abs(F[3]) < 1e-14
        assert np.allclose(F[1:3], [P * nx, P * ny], rtol=1e-14, atol=1e-15)
        # the Rusanov flux of the same states (rusanov.cpp:32-35) moves mass
        cL = np.sqrt(GAMMA * P / qL[0])
        cR = np.sqrt(GAMMA * P / qR[0])
        rus = -(0.5 * (cL + cR)) / 2 * (qR[0] - qL[0])
        assert abs(rus) > 0.1


def test_symmetry_and_rotation():
    rng = np.random.default_rng(5)
    for _ in range(100):
        nx, ny = unit(rng)
        qL, _ = random_state(rng, rng.uniform(-1.5, 1.5), nx, ny)
        qR, _ = random_state(rng, rng.uniform(-1.5, 1.5), nx, ny)
        F = O.roe_flux(qL, qR, nx, ny, GAMMA)
        scale = np.abs(flux(qL, nx, ny)).max() + np.abs(flux(qR, nx, ny)).max()
        # swapping the sides and the normal reverses the flux
        G = O.roe_flux(qR, qL, -nx, -ny, GAMMA)
        assert np.abs(F + G).max() <= 1e-13 * scale
        # rotating velocities and normal rotates the momentum flux only
        a = rng.uniform(0, 2 * np.pi)
        ca, sa = np.cos(a), np.sin(a)

        def rot(q):
            return np.array([q[0], ca * q[1] - sa * q[2], sa * q[1] + ca * q[2], q[3]])

        H = O.roe_flux(rot(qL), rot(qR), ca * nx - sa * ny, sa * nx + ca * ny, GAMMA)
        assert np.abs(H - rot(F)).max() <= 1e-13 * scale


def _box(dt_type, poly):
    from fv2d import gridgen as gg
    from fv2d import inputs
    g = gg.box(24, 18, jitter=0.25, split_fraction=0.5, markers=(0, 0, 0, 0))
    u, v = inputs.mach_velocity(0.4, angle_deg=30)
    return g, inputs.Case(u=u, v=v, poly=poly, limiter="SQUEEZE" if poly else "NONE",
                          dt_type=dt_type, CFL=0.5, BC1="SLIP_WALL", BC2="FREESTREAM",
                          flux="ROE")


@pytest.mark.parametrize("poly", [0, 1])
def test_roe_runs_conserve_in_a_closed_box(poly):
    """Slip walls all round and one dt for every cell: the Roe face fluxes
    cancel pairwise, so mass and energy stay constant to rounding while the
    flow inside (started at M 0.4 against the walls) changes."""
    g, case = _box("UNIFORM", poly)
    o = O.Oracle(g, case)
    vol = o.get("geom/cell_vol")
    q0 = o.state()
    m0, e0 = (vol * q0[0]).sum(), (vol * q0[3]).sum()
    o.step(100)
    q = o.state()
    assert np.isfinite(q).all()
    assert abs((vol * q[0]).sum() / m0 - 1) < 1e-13
    assert abs((vol * q[3]).sum() / e0 - 1) < 1e-13
    assert np.abs(q[0] - q0[0]).max() > 1e-3


def test_roe_preserves_the_freestream():
    from fv2d import gridgen as gg
    from fv2d import inputs
    g = gg.box(20, 14, jitter=0.3, split_fraction=0.5, markers=(0, -1, 0, -1))
    u, v = inputs.mach_velocity(0.7, angle_deg=20)
    case = inputs.Case(u=u, v=v, poly=1, limiter="SQUEEZE", dt_type="LOCAL", CFL=0.8,
                       BC1="FREESTREAM", BC2="FREESTREAM", flux="ROE")
    o = O.Oracle(g, case)
    q0 = o.state()
    o.step(50)
    q = o.state()
    assert np.abs(q - q0).max() <= 1e-12 * np.abs(q0).max()


def test_entropy_fix_is_continuous_through_the_sonic_point():
    """Harten's fix replaces |l| by (l^2 + d^2)/(2d) below d = c~/10, which
    meets |l| with the same value and slope at l = d: sweeping both sides'
    normal Mach number through one, the flux changes by steps of one size
    (a wrong fix would jump where |l| crosses d)."""
    nx, ny = 0.6, 0.8
    F = []
    for m in np.linspace(0.7, 1.3, 3001):
        qL = conserved(1.0, 0, 0, 1.0)
        qR = conserved(1.3, 0, 0, 1.1)
        for q in (qL, qR):
            rho, P = q[0], (GAMMA - 1) * q[3]
            vn = m * np.sqrt(GAMMA * P / rho)
            q[:] = conserved(rho, vn * nx + 0.2 * ny, vn * ny - 0.2 * nx, P)
        F.append(O.roe_flux(qL, qR, nx, ny, GAMMA))
    d = np.abs(np.diff(np.array(F), axis=0)).max(axis=1)
    assert d.max() <= 3 * np.median(d), (d.max(), np.median(d))
