"""3D particle paths on the GPU against the oracle:

* ParticleAxisIntegrate (integrate.hpp:143-190, 298-327), the pinned body
  spinning about a fixed axis, bitwise in 2D and 3D;
* the two 3D HLBM drivers at their own shape, D3Q27 32^3:
  examples/poiseulle_particle_3d_hlbm_gpu/sim.cpp (a free sphere in a
  pressure-driven channel with DEM wall contact) and
  examples/vawt_3d_hlbm_gpu/sim.cpp (the same sphere pinned, spinning about
  +y: ParticleAxisIntegrate, no contact)."""
import numpy as np
import pytest

from helpers import make_particle
from lbm_mirror_amd import abi

pytestmark = pytest.mark.gpu

DIM = 32
RADIUS = 0.15 * DIM          # sim.cpp:32 of both drivers
TAU = 3 * 0.1 + 0.5          # kinematic_viscosity_si_to_tau(0.1) with dx = dt = 1


def _pair(orc, st, ext, lat):
    from lbm_mirror_amd.lattice import Lattice
    return Lattice(lat, st, ext), orc.Sim(lat, st, ext)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
@pytest.mark.parametrize("D", [2, 3])
def test_fixed_axis_integrate_bitwise(hip, orc, st, D):
    lat = abi.D2Q9 if D == 2 else abi.D3Q27
    ext = (40, 30) if D == 2 else (20, 18, 16)
    g, o = _pair(orc, st, ext, lat)
    rng = np.random.default_rng(3 + D)
    for s in (g, o):
        s.set_flags(np.full(int(np.prod(ext)), 2, np.uint8))
        s.init_equilibrium()
    parts = [make_particle([8.0 + 3 * k] * D, 2.0, 1.2, D, st, orc=orc) for k in range(3)]
    for axis in ((0.3, -2.0, 0.7), (-0.5,)) if D == 3 else ((-0.5,), (2.0,)):
        prm = abi.particle_params(k_n=0.0, sub_steps=8, max_radius=2.0, spin_axis=axis)
        for s in (g, o):
            s.particles_set(parts, prm)
        for _ in range(20):
            ps = o.particles_get()
            for p in ps:
                for d in range(3 if D == 3 else 1):
                    p.torque[d] = float(np.float32(rng.uniform(-1, 1)))
                for d in range(D):
                    p.force[d] = float(np.float32(rng.uniform(-1, 1)))
            for s in (g, o):
                s.particles_set(ps, prm)
                s.integrate()
            for a, b in zip(g.particles_get(), o.particles_get()):
                for fld in ("position", "position_old", "acceleration", "rotation", "rotation_old",
                            "angular_acceleration"):
                    assert list(getattr(a, fld)) == list(getattr(b, fld)), fld
        parts = o.particles_get()


def _driver_flags():
    """setup_initial_conditions of both drivers (sim.cpp:100-140): walls, the
    interior fluid, the x = 0 and x = dim-1 faces inflow (3) and outflow (4)
    without their edges."""
    f = np.ones((DIM, DIM, DIM), np.uint8)
    f[1:-1, 1:-1, 1:-1] = 2
    f[1:-1, 1:-1, 0] = 3
    f[1:-1, 1:-1, DIM - 1] = 4
    return f.reshape(-1)


def _driver(orc, st, spin, zflags=0):
    """Both drivers' set-up: rho 1 everywhere, u = (0.01, 0, 0) in the
    interior (sim.cpp:142-180), one sphere r = 4.8, density 1, at
    (0.3, 0.5, 0.5) * dim, at rest; HLBM collision tau 0.8; Zou/He pressure
    inflow ramping 1 -> 1.001 over 256 steps, outflow 1.  The channel driver
    adds DEM wall contact {1e-2, 0.3, 0, 0}; the rotor pins the sphere and
    spins it about +y (ParticleAxisIntegrate) with no contact pass."""
    ext = (DIM, DIM, DIM)
    g, o = _pair(orc, st, ext, abi.D3Q27)
    N = DIM ** 3
    flags = _driver_flags()
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
           2: abi.op_desc(collision=abi.COLL_HLBM, tau=TAU),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=1, tau=TAU,
                          ramp=(1.0, 0.001 / 256, 1.001, 256), flags=zflags),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=0, tau=TAU, rho=1.0,
                          flags=zflags)}
    interior = np.zeros((DIM, DIM, DIM), bool)
    interior[1:-1, 1:-1, 1:-1] = True
    interior = interior.reshape(-1)
    eq0 = orc.equilibrium(abi.D3Q27, st, 1.0, (0.0, 0.0, 0.0))
    eq1 = orc.equilibrium(abi.D3Q27, st, 1.0, (0.01, 0.0, 0.0))
    f0 = np.where(interior[None, :], eq1[:, None], eq0[:, None])
    u0 = np.zeros((3, N))
    u0[0, interior] = 0.01
    for s in (g, o):
        s.set_flags(flags)
        for fl, op in ops.items():
            s.set_operator(fl, op)
        s.set_populations(f0)
        s.set_macros(np.ones(N), u0)
    p = make_particle([0.3 * DIM, 0.5 * DIM, 0.5 * DIM], RADIUS, 1.0, 3, st, orc=orc)
    if spin:
        prm = abi.particle_params(k_n=0.0, zeta_n=0.0, walls=False, sub_steps=16, max_radius=RADIUS,
                                  coupling=abi.COUPLING_HLBM, spin_axis=(0.0, 1.0, 0.0))
    else:
        prm = abi.particle_params(k_n=1.0e-2, zeta_n=0.3, walls=True, sub_steps=16, max_radius=RADIUS,
                                  coupling=abi.COUPLING_HLBM)
    for s in (g, o):
        s.particles_set([p], prm)
    o.match_gpu_arithmetic()  # FP64 force/torque sums, the engine's ramp sine
    return g, o, p


def _compare(g, o, st):
    # FP64 force/torque sums on both sides (tree on the GPU, serial in the oracle)
    tol = 1e-11 if st == abi.F64 else 1e-5
    a, b = g.particles_get()[0], o.particles_get()[0]
    for d in range(3):
        assert abs(a.position[d] - b.position[d]) <= tol * max(abs(b.position[d]), 1.0)
        assert abs(a.rotation[d] - b.rotation[d]) <= tol * max(abs(b.rotation[d]), 1e-6) + 1e-15
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    ftol = 1e-12 if st == abi.F64 else 1e-5
    assert np.max(np.abs(rg - ro)) <= ftol
    assert np.max(np.abs(ug - uo)) <= ftol * np.max(np.abs(uo))
    assert np.max(np.abs(g.get_porosity() - o.get_porosity())) <= (1e-13 if st == abi.F64 else 1e-5)
    return a, b


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_sphere_in_channel_3d_hlbm_driver(hip, orc, st):
    g, o, p0 = _driver(orc, st, spin=False)
    g.step(60)
    o.step(60)
    a, b = _compare(g, o, st)
    # the flow pushes the sphere downstream
    assert b.position[0] > p0.position[0]
    g.close()


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_vawt_rotor_3d_hlbm_driver(hip, orc, st):
    g, o, _ = _driver(orc, st, spin=True)
    start = list(o.particles_get()[0].position)
    assert list(g.particles_get()[0].position) == start
    g.step(60)
    o.step(60)
    a, b = _compare(g, o, st)
    # pinned exactly, spinning about +y only
    assert list(a.position) == start and list(b.position) == start
    assert a.rotation[0] == 0.0 and a.rotation[2] == 0.0
    g.close()


@pytest.mark.parametrize("spin", [False, True])
def test_3d_drivers_full_length_with_normal_only_zouhe(hip, orc, spin):
    """The drivers' full 256 steps.  With the reference's D3Q27 Zou/He the
    channel leaves the finite numbers after ~100 steps (its reconstruction
    returns 4/3 of the tangential momentum each step); with
    LBM_OPF_ZOUHE_NORMAL_ONLY the run stays finite and the GPU follows the
    oracle: the free sphere drifts downstream, the rotor spins about +y."""
    st = abi.F32
    g, o, p0 = _driver(orc, st, spin, zflags=abi.OPF_ZOUHE_NORMAL_ONLY)
    g.step(256)
    o.step(256)
    a, b = _compare(g, o, st)
    assert np.all(np.isfinite(g.get_macros()[1]))
    if spin:
        assert abs(b.rotation[1]) > 1e-3 and a.rotation[0] == 0.0 and a.rotation[2] == 0.0
    else:
        assert b.position[0] > p0.position[0] + 0.1
    g.close()
