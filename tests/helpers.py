"""Scenario builders shared by the oracle tests and the GPU parity tests.

The flag layouts follow the reference drivers' setup_initial_conditions
(examples/poiseulle_2d_gpu/poiseulle_2d_gpu.cpp:55-100): everything wall
(1), interior fluid (2), inflow column x=0 (3) and outflow column x=nx-1 (4)
without the corner rows.
"""
import numpy as np

from lbm_mirror_amd import abi


def channel_flags_2d(nx, ny):
    f = np.ones((ny, nx), dtype=np.uint8)
    f[1:-1, 1:-1] = 2
    f[1:-1, 0] = 3
    f[1:-1, nx - 1] = 4
    return f.reshape(-1)


def duct_flags_3d(nx, ny, nz, periodic_x=True):
    """Walls on y and z faces, fluid elsewhere (the 3D channel with periodic x)."""
    f = np.full((nz, ny, nx), 2, dtype=np.uint8)
    f[:, 0, :] = 1
    f[:, ny - 1, :] = 1
    f[0, :, :] = 1
    f[nz - 1, :, :] = 1
    if not periodic_x:
        f[:, :, 0] = 1
        f[:, :, nx - 1] = 1
    return f.reshape(-1)


def parabolic_inlet(nx, ny, u_max, D=2):
    """Per-cell inlet velocity for the x=0 column: u(y) = u_max·4(y−½)(ny−1.5−y)/(ny−2)²,
    the half-way wall positions of full-way bounce-back at y=0 and y=ny−1."""
    u = np.zeros((D, ny, nx))
    y = np.arange(ny, dtype=np.float64)
    prof = u_max * 4.0 * (y - 0.5) * (ny - 1.5 - y) / (ny - 2.0) ** 2
    prof[0] = 0.0
    prof[-1] = 0.0
    u[0, :, 0] = prof
    return u.reshape(D, -1)


def c1_ops(tau=0.8, u_max=0.05):
    """Config-1 operators: BB walls, BGK fluid, Zou/He velocity inlet + BGK,
    Zou/He pressure outlet (rho=1) + BGK (stdkr_2d_bgk/step.hpp:21-67)."""
    return {
        1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
        2: abi.op_desc(collision=abi.COLL_BGK, tau=tau),
        3: abi.op_desc(boundary=abi.BC_ZOUHE_VELOCITY_X, collision=abi.COLL_BGK, east=1, tau=tau),
        4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0,
                       tau=tau, rho=1.0),
    }


def c2_ops(tau=0.6, g=1e-6):
    return {
        1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
        2: abi.op_desc(collision=abi.COLL_BGK_GUO, tau=tau, vec=(g, 0.0, 0.0)),
    }


def make_particle(pos, radius, density, D, storage, vel=(0.0, 0.0, 0.0), mass=None,
                  inertia=None, orc=None):
    p = abi.Particle()
    for d in range(3):
        x = pos[d] if d < len(pos) else 0.0
        v = vel[d] if d < len(vel) else 0.0
        p.position[d] = x
        p.position_old[d] = x - v
    p.radius = radius
    p.density = density
    if mass is None:
        mass, inertia = orc.spheroid_mass(D, storage, radius, density)
    p.mass = mass
    p.inertia = inertia
    return p


def seeded_discs(n, extent, r_lo, r_hi, seed, margin=2.0, gap=1.0, max_tries=200000):
    """Non-overlapping discs by seeded rejection sampling (config 3 placement)."""
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n and tries < max_tries:
        tries += 1
        r = rng.uniform(r_lo, r_hi)
        x = rng.uniform(r + margin, extent[0] - r - margin)
        y = rng.uniform(r + margin, extent[1] - r - margin)
        ok = True
        for (xx, yy, rr) in out:
            if (x - xx) ** 2 + (y - yy) ** 2 < (r + rr + gap) ** 2:
                ok = False
                break
        if ok:
            out.append((x, y, r))
    return out


def restitution_to_zeta(e):
    """Damping ratio of a linear spring-dashpot with restitution e:
    ζ = −ln e / sqrt(π² + ln² e)."""
    le = np.log(e)
    return -le / np.sqrt(np.pi ** 2 + le ** 2)
