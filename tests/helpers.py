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


def c4_member(rng, side_lo=64, side_hi=512, frac=(0.05, 0.20), r_range=(4.0, 24.0),
              u_range=(0.01, 0.08), tau_range=(0.55, 1.2)):
    """One ensemble member of the batched config: D2Q9 channel of uneven
    sides, seeded random circular bounce-back obstacles covering a target
    solid fraction, parabolic Zou/He velocity inlet (u_max drawn), Zou/He
    pressure outlet at rho = 1, BGK with tau drawn.  Returns
    (extent, flags, ops, inlet_velocity, params)."""
    nx = int(rng.integers(side_lo, side_hi + 1))
    ny = int(rng.integers(side_lo, side_hi + 1))
    u_max = float(rng.uniform(*u_range))
    tau = float(rng.uniform(*tau_range))
    target = float(rng.uniform(*frac))
    flags = channel_flags_2d(nx, ny).reshape(ny, nx)
    yy, xx = np.mgrid[0:ny, 0:nx]
    solid = np.zeros((ny, nx), bool)
    interior = (xx >= 3) & (xx <= nx - 4) & (yy >= 1) & (yy <= ny - 2)
    tries = 0
    while solid[interior].mean() < target and tries < 500:
        tries += 1
        r = float(rng.uniform(*r_range))
        r = min(r, 0.3 * min(nx, ny))
        cx = float(rng.uniform(3 + r, nx - 4 - r)) if nx - 7 > 2 * r else nx / 2
        cy = float(rng.uniform(1 + r, ny - 2 - r)) if ny - 3 > 2 * r else ny / 2
        solid |= ((xx - cx) ** 2 + (yy - cy) ** 2 < r * r) & interior
    flags[solid] = 1
    ops = c1_ops(tau, u_max)
    bcu = parabolic_inlet(nx, ny, u_max)
    return (nx, ny), flags.reshape(-1), ops, bcu, {"u_max": u_max, "tau": tau,
                                                    "solid": float(solid[interior].mean())}
