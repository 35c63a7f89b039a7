"""The oracle against the reference's own code (oracle/_ref, built in place
from the reference tree by oracle/ref/Makefile): equilibrium, moments,
porosity, AABB, spheroid mass, DEM contact + integration and the cell-list
product porosity, bit for bit in FP64 and FP32 on seeded random inputs.

Skipped where oracle/_ref was not built (no reference tree at build time).
"""
import ctypes as C
import os

import numpy as np
import pytest

from lbm_mirror_amd import abi
from helpers import make_particle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")

STS = [abi.F64, abi.F32]


@pytest.fixture(scope="module")
def ref():
    L = C.CDLL(REF)
    dp = C.POINTER(C.c_double)
    u64p = C.POINTER(C.c_uint64)
    L.ref_equilibrium.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
    L.ref_rho_u.argtypes = [C.c_int, C.c_int, dp, dp, dp]
    L.ref_porosity.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double]
    L.ref_porosity.restype = C.c_double
    L.ref_aabb.argtypes = [C.c_int, C.POINTER(abi.Particle), u64p, u64p, u64p]
    L.ref_spheroid_mass.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp, dp]
    L.ref_dem.argtypes = [C.c_int, C.c_int, C.c_uint32, C.POINTER(abi.Particle),
                          C.POINTER(abi.ParticleParams), u64p, C.c_int, C.c_int, C.c_int]
    L.ref_couple.argtypes = [C.c_int, C.c_uint32, C.POINTER(abi.Particle), u64p, C.c_double,
                             C.c_double, dp]
    L.ref_couple.restype = C.c_int64
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _q(x, st):
    return float(np.float32(x)) if st == abi.F32 else float(x)


@pytest.mark.parametrize("lat", [abi.D2Q9, abi.D3Q19, abi.D3Q27])
@pytest.mark.parametrize("st", STS)
def test_equilibrium_and_moments_bitwise(ref, orc, lat, st):
    D, Q = abi.LATTICE_DQ[lat]
    rng = np.random.default_rng(100 + lat)
    for _ in range(300):
        rho = _q(rng.uniform(0.8, 1.2), st)
        u = np.array([_q(v, st) for v in rng.uniform(-0.1, 0.1, size=3)])
        u[D:] = 0
        want = np.zeros(27)
        ref.ref_equilibrium(lat, st, rho, _dp(u), _dp(want))
        got = orc.equilibrium(lat, st, rho, u)
        assert np.array_equal(got, want[:Q])
        f = np.array([_q(v, st) for v in want[:Q] * (1 + 0.1 * rng.uniform(-1, 1, size=Q))])
        r_want, u_want = np.zeros(1), np.zeros(3)
        ff = np.zeros(27)
        ff[:Q] = f
        ref.ref_rho_u(lat, st, _dp(ff), _dp(r_want), _dp(u_want))
        r_got, u_got = orc.rho_u(lat, st, f)
        assert r_got == r_want[0] and np.array_equal(u_got, u_want[:D])


@pytest.mark.parametrize("st", STS)
@pytest.mark.parametrize("D", [2, 3])
def test_porosity_bitwise(ref, orc, st, D):
    rng = np.random.default_rng(7 + D)
    for _ in range(2000):
        R = _q(rng.uniform(1, 16), st)
        rel = np.array([_q(v, st) for v in rng.uniform(-R - 2, R + 2, size=3)])
        rel[D:] = 0
        assert orc.porosity(D, st, rel, R, 1.5) == ref.ref_porosity(D, st, _dp(rel), R, 1.5)


@pytest.mark.parametrize("st", STS)
def test_aabb_bitwise(ref, orc, st):
    rng = np.random.default_rng(3)
    ext = (C.c_uint64 * 3)(64, 32, 1)
    for _ in range(500):
        pos = [_q(rng.uniform(-2, 70), st), _q(rng.uniform(-2, 40), st)]
        r = _q(rng.uniform(0.5, 12), st)
        p = abi.Particle()
        p.position[0], p.position[1] = pos
        p.radius, p.density, p.mass, p.inertia = r, 1.0, 1.0, 1.0
        a, b = (C.c_uint64 * 3)(), (C.c_uint64 * 3)()
        if pos[0] + r + 1 < 0 or pos[1] + r + 1 < 0:
            continue  # the reference casts a negative float to uint64 there
        assert ref.ref_aabb(st, C.byref(p), ext, a, b) == 0
        oa, ob = orc.aabb(2, st, pos, r, (64, 32))
        assert oa == list(a[:2]) and ob == list(b[:2])


@pytest.mark.parametrize("st", STS)
@pytest.mark.parametrize("D", [2, 3])
def test_spheroid_mass_bitwise(ref, orc, st, D):
    for r in (1.0, 2.2, 7.3, 16.0):
        for rho in (1.0, 1.5, 1 / np.pi):
            m, i = C.c_double(), C.c_double()
            ref.ref_spheroid_mass(D, st, r, rho, C.byref(m), C.byref(i))
            assert orc.spheroid_mass(D, st, r, rho) == (m.value, i.value)


def _rand_group(rng, n, D, st, orc, spread=6.0):
    parts = []
    for k in range(n):
        pos = [_q(20 + spread * rng.uniform(-1, 1), st) for _ in range(D)]
        vel = [0.05 * rng.uniform(-1, 1) for _ in range(D)]
        p = make_particle(pos, 2.0, 1.0, D, st, vel=vel, orc=orc)
        for d in range(D):
            p.position_old[d] = _q(p.position_old[d], st)
        if D == 2:
            p.rotation[0] = _q(0.1 * rng.uniform(-1, 1), st)
            p.rotation_old[0] = _q(p.rotation[0] - 0.01 * rng.uniform(-1, 1), st)
        parts.append(p)
    return parts


@pytest.mark.parametrize("st", STS)
@pytest.mark.parametrize("D", [2, 3])
def test_dem_contact_and_integrate_bitwise(ref, orc, st, D):
    """Reset, pair + wall contacts, Verlet sub-steps for 2-4 bodies with
    damping and Coulomb-capped friction, over 40 steps: the oracle's state
    equals the reference's bit for bit."""
    rng = np.random.default_rng(11 * D)
    ext = [64, 40, 40][:D]
    for n in (2, 3, 4):
        parts = _rand_group(rng, n, D, st, orc)
        prm = abi.particle_params(k_n=5.0, zeta_n=0.3, gamma_t=2.0, mu=0.4, walls=True,
                                  sub_steps=4, body_acceleration=(0.0, -1e-3, 0.0),
                                  max_radius=2.0)
        s = orc.Sim(abi.D2Q9 if D == 2 else abi.D3Q19, st, ext)
        s.particles_set(parts, prm)
        arr = (abi.Particle * n)(*parts)
        e3 = (C.c_uint64 * 3)(*(ext + [1] * (3 - D)))
        for _ in range(40):
            ps = s.particles_get()
            for p in ps:
                for d in range(3):
                    p.force[d] = 0.0
                    p.torque[d] = 0.0
            s.particles_set(ps, prm)
            s.dem_contact()
            s.integrate()
            assert ref.ref_dem(D, st, n, arr, C.byref(prm), e3, 1, 1, 1) == 0
            for a, b in zip(s.particles_get(), arr):
                for fld in ("position", "position_old", "rotation", "rotation_old", "force",
                            "torque", "acceleration", "angular_acceleration"):
                    assert list(getattr(a, fld))[:D] == list(getattr(b, fld))[:D], fld


@pytest.mark.parametrize("st", STS)
def test_cell_list_product_porosity_bitwise(ref, orc, st):
    rng = np.random.default_rng(5)
    ext = (40, 32)
    for n in (1, 2, 3, 4):
        parts = []
        for k in range(n):
            pos = [_q(rng.uniform(4, 36), st), _q(rng.uniform(4, 28), st)]
            parts.append(make_particle(pos, 3.0, 1.0, 2, st, orc=orc))
        out = np.zeros(ext[0] * ext[1])
        arr = (abi.Particle * n)(*parts)
        ovf = ref.ref_couple(st, n, arr, (C.c_uint64 * 3)(ext[0], ext[1], 1), 3.0, 1.5, _dp(out))
        s = orc.Sim(abi.D2Q9, st, ext)
        s.set_flags(np.full(ext[0] * ext[1], 2, np.uint8))
        s.init_equilibrium()
        s.particles_set(parts, abi.particle_params(porosity_mode=abi.POROSITY_PRODUCT,
                                                   max_radius=3.0, cell_capacity=4, eps=1.5))
        s.psm_couple()
        assert np.array_equal(s.get_porosity(), out)
        assert s.overflow() == ovf


@pytest.mark.parametrize("st", STS)
@pytest.mark.parametrize("D", [2, 3])
def test_fixed_axis_integrate_bitwise(ref, orc, st, D):
    """ParticleAxisIntegrate (integrate.hpp:143-190, 298-327): pinned bodies
    spinning about a fixed, unnormalised axis under a torque that changes
    every step; the oracle equals the reference's integrate_particle_fixed_axis
    (with its normalise_spin_axis) bit for bit, over 30 steps."""
    rng = np.random.default_rng(7 + D)
    ext = [64, 40, 40][:D]
    axes = [(0.0, 1.0, 0.0), (0.3, -2.0, 0.7), (0.0, 0.0, 0.0)] if D == 3 else [(1.0,), (-0.2,), (3.0,)]
    for axis in axes:
        n = 2
        parts = _rand_group(rng, n, D, st, orc)
        for p in parts:  # at rest: x_old == x, which is what keeps a pinned body in place
            for d in range(D):
                p.position_old[d] = p.position[d]
        prm = abi.particle_params(k_n=0.0, sub_steps=8, max_radius=2.0, spin_axis=axis)
        s = orc.Sim(abi.D2Q9 if D == 2 else abi.D3Q19, st, ext)
        s.particles_set(parts, prm)
        arr = (abi.Particle * n)(*parts)
        e3 = (C.c_uint64 * 3)(*(ext + [1] * (3 - D)))
        start = [list(p.position) for p in parts]
        for _ in range(30):
            ps = s.particles_get()
            for k, p in enumerate(ps):
                for d in range(3 if D == 3 else 1):
                    p.torque[d] = arr[k].torque[d] = _q(rng.uniform(-1, 1), st)
                for d in range(D):
                    p.force[d] = arr[k].force[d] = _q(rng.uniform(-1, 1), st)
            s.particles_set(ps, prm)
            s.integrate()
            assert ref.ref_dem(D, st, n, arr, C.byref(prm), e3, 0, 0, 1) == 0
            R = 1 if D == 2 else 3
            for a, b in zip(s.particles_get(), arr):
                for fld, m in (("position", D), ("position_old", D), ("acceleration", D),
                               ("rotation", R), ("rotation_old", R), ("angular_acceleration", R)):
                    assert list(getattr(a, fld))[:m] == list(getattr(b, fld))[:m], (axis, fld)
        # pinned: the force is ignored and the body never moves
        for a, p0 in zip(s.particles_get(), start):
            assert list(a.position)[:D] == p0[:D]


@pytest.mark.parametrize("D", [2, 3])
def test_fixed_axis_closed_form(orc, D):
    """Constant torque tau about the axis: position Verlet from rest
    (theta_old == theta) under constant angular acceleration alpha gives
    theta_k = alpha dt^2 k (k + 1) / 2 after k sub-steps of dt, so after n
    steps of s sub-steps theta = axis * alpha (n s)(n s + 1) / (2 s^2); an
    off-axis torque component does not spin the body; a zero axis never
    spins it."""
    st = abi.F64
    p = make_particle([20.0] * D, 2.0, 1.0, D, st, orc=orc)
    axis = (0.0, 0.0, 2.0) if D == 3 else (-1.0,)
    prm = abi.particle_params(k_n=0.0, sub_steps=4, max_radius=2.0, spin_axis=axis)
    s = orc.Sim(abi.D2Q9 if D == 2 else abi.D3Q19, st, [40] * D)
    s.particles_set([p], prm)
    steps, tau = 25, 0.125
    for _ in range(steps):
        ps = s.particles_get()
        if D == 3:
            ps[0].torque[0], ps[0].torque[1], ps[0].torque[2] = 0.5, -0.25, tau
        else:
            ps[0].torque[0] = tau
        s.particles_set(ps, prm)
        s.integrate()
    q = s.particles_get()[0]
    alpha = tau / p.inertia
    k = steps * 4
    theta = alpha * k * (k + 1) / (2 * 4 ** 2)
    if D == 3:
        assert q.rotation[0] == 0.0 and q.rotation[1] == 0.0
        assert np.isclose(q.rotation[2], theta, rtol=1e-12)
    else:
        assert np.isclose(q.rotation[0], -theta, rtol=1e-12)
    zero = abi.particle_params(k_n=0.0, sub_steps=4, max_radius=2.0, spin_axis=(0.0, 0.0, 0.0))
    if D == 3:
        s2 = orc.Sim(abi.D3Q19, st, [40] * 3)
        s2.particles_set([p], zero)
        ps = s2.particles_get()
        ps[0].torque[0] = ps[0].torque[1] = ps[0].torque[2] = 1.0
        s2.particles_set(ps, zero)
        s2.integrate()
        assert list(s2.particles_get()[0].rotation) == [0.0, 0.0, 0.0]
