"""GPU parity: the HIP path through the C ABI against the oracle on the same
seeded inputs.

Bar (per the tolerance stated in each test): FP64 runs equal the oracle bit
for bit (the kernels keep the reference's operation order and are built
without FMA contraction; 1e-12 relative is the stated limit), FP32 runs bit
for bit or within 1e-5 relative on density and velocity.  Particle force sums
are FP64 tree reductions against the oracle's serial sum: |ΔF| ≤ 1e-12·Σ|F_cell|
(FP64) and 1e-5·Σ|F_cell| (FP32) — the summation-order bound.
"""
import numpy as np
import pytest

from lbm_mirror_amd import abi
from helpers import (channel_flags_2d, duct_flags_3d, parabolic_inlet, c1_ops, c2_ops,
                     make_particle, seeded_discs, restitution_to_zeta)

pytestmark = pytest.mark.gpu


def _pair(hip, orc, lat, st, ext):
    from lbm_mirror_amd.lattice import Lattice
    return Lattice(lat, st, ext), orc.Sim(lat, st, ext)


def _setup(sims, flags, ops, periodic=0, bcu=None, init=(1.0, (0.0, 0.0, 0.0)), f0=None,
           force=None, por=None):
    for s in sims:
        s.set_flags(flags)
        if periodic:
            s.set_periodic(periodic)
        for fl, op in ops.items():
            s.set_operator(fl, op)
        if bcu is not None:
            s.set_boundary_velocity(bcu)
        if f0 is not None:
            s.set_populations(f0)
        else:
            s.init_equilibrium(*init)
        if force is not None:
            s.set_force_field(force)
        if por is not None:
            s.set_porosity(por)


def _assert_same(g, o, exact=True, rtol=1e-5):
    fg, fo = g.get_populations(), o.get_populations()
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    if exact:
        assert np.array_equal(fg, fo), f"populations differ, max {np.max(np.abs(fg - fo)):.3e}"
        assert np.array_equal(rg, ro), f"density differs, max {np.max(np.abs(rg - ro)):.3e}"
        assert np.array_equal(ug, uo), f"velocity differs, max {np.max(np.abs(ug - uo)):.3e}"
    else:
        scale_u = max(np.max(np.abs(uo)), 1e-30)
        assert np.max(np.abs(rg - ro) / np.abs(ro)) <= rtol
        assert np.max(np.abs(ug - uo)) <= rtol * scale_u


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
@pytest.mark.parametrize("steps", [1, 2, 37, 400])
def test_c1_channel_small(hip, orc, st, steps):
    nx, ny = 64, 18
    g, o = _pair(hip, orc, abi.D2Q9, st, (nx, ny))
    _setup((g, o), channel_flags_2d(nx, ny), c1_ops(0.8, 0.05), bcu=parabolic_inlet(nx, ny, 0.05))
    g.step(steps)
    o.step(steps)
    _assert_same(g, o, exact=True)


def test_c1_full_config_fp64(hip, orc):
    """Config 1 at its full lattice (256x64, tau 0.8, u_max 0.05), 3000 steps:
    bitwise equal populations, density and velocity in FP64."""
    nx, ny = 256, 64
    g, o = _pair(hip, orc, abi.D2Q9, abi.F64, (nx, ny))
    _setup((g, o), channel_flags_2d(nx, ny), c1_ops(0.8, 0.05), bcu=parabolic_inlet(nx, ny, 0.05))
    g.step(3000)
    o.step(3000, threads=8)
    _assert_same(g, o, exact=True)


@pytest.mark.parametrize("lat,ext", [(abi.D3Q19, (24, 10, 9)), (abi.D3Q27, (16, 9, 8))])
@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_c2_periodic_duct_guo(hip, orc, lat, ext, st):
    nx, ny, nz = ext
    g, o = _pair(hip, orc, lat, st, ext)
    rng = np.random.default_rng(20260101)
    rho0 = 1 + 1e-3 * rng.uniform(-1, 1, size=nx * ny * nz)
    D, Q, c, w, _ = orc.lattice(lat)
    f0 = np.outer(w, rho0)
    _setup((g, o), duct_flags_3d(nx, ny, nz), c2_ops(0.6, 1e-4), periodic=1, f0=f0)
    g.step(61)
    o.step(61)
    _assert_same(g, o, exact=True)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
@pytest.mark.parametrize("zflags", [0, abi.OPF_ZOUHE_NORMAL_ONLY])
def test_d3q27_zouhe_pressure_channel(hip, orc, st, zflags):
    nx, ny, nz = 20, 8, 8
    fl = np.ones((nz, ny, nx), np.uint8)
    fl[1:-1, 1:-1, 1:-1] = 2
    fl[1:-1, 1:-1, 0] = 3
    fl[1:-1, 1:-1, nx - 1] = 4
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
           2: abi.op_desc(collision=abi.COLL_BGK, tau=0.7),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.7,
                          ramp=(1.0, 0.001 / 16, 1.001, 16), flags=zflags),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.7,
                          rho=1.0, flags=zflags)}
    g, o = _pair(hip, orc, abi.D3Q27, st, (nx, ny, nz))
    _setup((g, o), fl.reshape(-1), ops)
    g.step(45)
    o.step(45)
    _assert_same(g, o, exact=True)


@pytest.mark.parametrize("lat,ext", [(abi.D2Q9, (21, 13)), (abi.D3Q19, (9, 7, 6)),
                                     (abi.D3Q27, (8, 7, 6)),
                                     # vectorised kernel: 5 threads per row (partial
                                     # wavefront), and 65 threads per row (a row over
                                     # two wavefronts, one active lane in the second)
                                     (abi.D2Q9, (20, 7)), (abi.D2Q9, (260, 5)),
                                     (abi.D3Q19, (12, 5, 4)), (abi.D3Q27, (260, 3, 3)),
                                     # one-row blocks (row-uniform addressing)
                                     (abi.D3Q19, (260, 3, 3))])
@pytest.mark.parametrize("periodic", [0, 3])
def test_random_flags_every_operator(hip, orc, lat, ext, periodic):
    """Random flag mixture over every operator the lattice supports, random
    populations, odd and even step counts: exercises both AA phases, ghost
    reads/writes, the unpacking of the swapped layout, and both the scalar
    kernel (row length not a multiple of the vector width) and the vectorised
    kernel's wavefront-edge handling."""
    rng = np.random.default_rng(99 + lat)
    N = int(np.prod(ext))
    D, Q, c, w, _ = orc.lattice(lat)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
           2: abi.op_desc(collision=abi.COLL_BGK, tau=0.8),
           3: abi.op_desc(collision=abi.COLL_BGK_GUO, tau=0.65, vec=(2e-4, -1e-4, 5e-5),
                          flags=abi.OPF_GUO_CELL_FORCE),
           4: abi.op_desc(collision=abi.COLL_PSM, tau=0.9, flags=abi.OPF_PSM_FORCE_GATHER),
           5: abi.op_desc(boundary=abi.BC_EQUILIBRIUM, rho_eq=1.01, vec=(0.01, 0.0, 0.0)),
           6: abi.op_desc(),  # stream only (the drivers' `default: break`)
           7: abi.op_desc(collision=abi.COLL_PSM, tau=0.7, flags=abi.OPF_PSM_CELL_FORCE)}
    if lat == abi.D2Q9:
        ops[8] = abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1,
                             tau=0.8, rho=1.002)
        ops[9] = abi.op_desc(boundary=abi.BC_ZOUHE_VELOCITY_X, collision=abi.COLL_BGK, east=0,
                             tau=0.8)
        ops[10] = abi.op_desc(boundary=abi.BC_ANTI_BOUNCE_BACK)
    if lat == abi.D3Q27:
        ops[8] = abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0,
                             tau=0.8, rho=0.999)
    flags = rng.choice(sorted(ops), size=N).astype(np.uint8)
    f0 = np.outer(w, np.ones(N)) * (1 + 0.05 * rng.uniform(-1, 1, size=(Q, N)))
    force = 1e-4 * rng.uniform(-1, 1, size=(D, N))
    por = rng.uniform(0, 1, size=N)
    bcu = 0.03 * rng.uniform(-1, 1, size=(D, N))
    for st in (abi.F64, abi.F32):
        g, o = _pair(hip, orc, lat, st, ext)
        _setup((g, o), flags, ops, periodic=periodic, bcu=bcu if lat == abi.D2Q9 else None,
               f0=f0, force=force, por=por)
        for n in (1, 2, 3):
            g.step(n)
            o.step(n)
            _assert_same(g, o, exact=True)
        assert np.array_equal(g.get_force_field(), o.get_force_field())
        assert np.array_equal(g.get_porosity(), o.get_porosity())


def test_rejects_interior_ghost_flag_and_unsupported(hip):
    from lbm_mirror_amd import LbmError
    from lbm_mirror_amd.lattice import Lattice
    g = Lattice(abi.D2Q9, abi.F64, (8, 8))
    fl = np.full(64, 2, np.uint8)
    fl[20] = 0
    with pytest.raises(LbmError) as e:
        g.set_flags(fl)
    assert e.value.status == abi.ERR_NOT_SUPPORTED
    g3 = Lattice(abi.D3Q19, abi.F32, (8, 8, 8))
    with pytest.raises(LbmError) as e:
        g3.set_operator(3, abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK,
                                       tau=0.8))
    assert e.value.status == abi.ERR_NOT_SUPPORTED
    with pytest.raises(LbmError):
        g.step(1)  # flags never set


# ------------------------------------------------------------------ particles
def _particle_case(orc, st, n=6, ext=(96, 48), seed=7, mode=abi.POROSITY_SCATTER, torque=False):
    discs = seeded_discs(n, ext, 4.0, 7.0, seed)
    parts = []
    rng = np.random.default_rng(seed)
    for (x, y, r) in discs:
        v = rng.uniform(-0.01, 0.01, size=2)
        parts.append(make_particle([x, y], r, 1.5, 2, st, vel=list(v / 16), orc=orc))
    prm = abi.particle_params(k_n=1e3, zeta_n=restitution_to_zeta(0.9), walls=True,
                              porosity_mode=mode, psm_torque=torque, sub_steps=16,
                              body_acceleration=(0.0, -1e-4), eps=1.5,
                              max_radius=max(r for *_, r in discs), cell_capacity=4)
    return parts, prm


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
@pytest.mark.parametrize("mode", [abi.POROSITY_SCATTER, abi.POROSITY_PRODUCT])
def test_psm_couple_single_pass(hip, orc, st, mode):
    ext = (96, 48)
    g, o = _pair(hip, orc, abi.D2Q9, st, ext)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_PSM, tau=0.7),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.7,
                          rho=1.01),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.7,
                          rho=1.0)}
    rng = np.random.default_rng(3)
    D, Q, c, w, _ = orc.lattice(abi.D2Q9)
    N = ext[0] * ext[1]
    f0 = np.outer(w, np.ones(N)) * (1 + 0.05 * rng.uniform(-1, 1, size=(Q, N)))
    _setup((g, o), flags, ops, f0=f0)
    parts, prm = _particle_case(orc, st, mode=mode, torque=True)
    for s in (g, o):
        s.particles_set(parts, prm)
    o.match_gpu_arithmetic()
    # both AA phases: couple on the natural layout, then after one step
    for k in range(2):
        Fg, Tg = g.psm_couple()
        Fo, To, A = o.psm_couple()
        tol = (1e-12 if st == abi.F64 else 1e-5)
        assert np.all(np.abs(Fg - Fo) <= tol * A[:, None] + 1e-300)
        assert np.all(np.abs(Tg - To) <= tol * A[:, None] * 64 + 1e-300)
        # the porosity ramp's sine (porosity.hpp:82) is the engine's series on
        # both sides: porosity bit for bit
        pg, po = g.get_porosity(), o.get_porosity()
        assert np.array_equal(pg, po)
        assert np.sum(pg < 1) == np.sum(po < 1) and np.sum(pg < 1) > 0
        ffg, ffo = g.get_force_field(), o.get_force_field()
        fscale = max(np.max(np.abs(ffo)), 1e-300)
        assert np.max(np.abs(ffg - ffo)) <= (1e-13 if st == abi.F64 else 1e-5) * fscale
        for s in (g, o):
            s.step(1) if k == 0 else None
        if k == 0:
            # the particles moved during step(); put both back in sync
            pg = g.particles_get()
            o.particles_set(pg, prm)
            g.particles_set(pg, prm)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_particle_laden_run(hip, orc, st):
    """Config 3 structure at a small size: PSM fluid, Zou/He in/out, BB walls,
    discs with DEM contact and gravity, 60 full steps through lbm_step."""
    ext = (96, 48)
    g, o = _pair(hip, orc, abi.D2Q9, st, ext)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_PSM, tau=0.7),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.7,
                          ramp=(1.0, 0.01 / 8, 1.01, 8)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.7,
                          rho=1.0)}
    _setup((g, o), flags, ops)
    parts, prm = _particle_case(orc, st, n=8)
    for s in (g, o):
        s.particles_set(parts, prm)
    # the oracle with the GPU's arithmetic: FP64 particle force sums (in FP64
    # identical to the reference's serial sums) and the engine's ramp sine
    o.match_gpu_arithmetic()
    g.step(60)
    o.step(60)
    pg, po = g.particles_get(), o.particles_get()
    tol = 1e-11 if st == abi.F64 else 1e-5
    for a, b in zip(pg, po):
        for d in range(2):
            scale = max(abs(b.position[d]), 1.0)
            assert abs(a.position[d] - b.position[d]) <= tol * scale
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    if st == abi.F64:
        assert np.max(np.abs(rg - ro)) <= 1e-12
        assert np.max(np.abs(ug - uo)) <= 1e-12 * np.max(np.abs(uo))
        return
    # FP32 against the oracle with the same FP64 force sums: 1e-5 relative on
    # density and velocity; and against the FP64 oracle the GPU's FP32 error
    # is no larger than that of the reference's own FP32 arithmetic (serial
    # FP32 force sums)
    assert np.max(np.abs(ug - uo)) <= 1e-5 * np.max(np.abs(uo))
    assert np.max(np.abs(rg - ro) / ro) <= 1e-5
    o32 = orc.Sim(abi.D2Q9, abi.F32, ext)
    _setup((o32,), flags, ops)
    o32.particles_set(parts, prm)
    o32.step(60)
    uo = o32.get_macros()[1]
    o64 = orc.Sim(abi.D2Q9, abi.F64, ext)
    _setup((o64,), flags, ops)
    o64.particles_set(parts, prm)
    o64.step(60)
    r64, u64 = o64.get_macros()
    us = np.max(np.abs(u64))
    err_gpu = np.max(np.abs(ug - u64)) / us
    err_ref32 = np.max(np.abs(uo - u64)) / us
    print(f"FP32 particle-laden: GPU err vs FP64 {err_gpu:.2e}, reference-arithmetic FP32 err {err_ref32:.2e}")
    assert err_gpu <= 2.0 * err_ref32 + 1e-6


def test_dem_contact_and_integrate_bitwise(hip, orc):
    """DEM and Verlet in isolation (no fluid force): bitwise equal in FP64."""
    ext = (64, 32)
    for st in (abi.F64, abi.F32):
        g, o = _pair(hip, orc, abi.D2Q9, st, ext)
        states = [[0.0 + 20, 0.3 + 16, 0.05, 0.0], [3.0 + 20, 16.0, -0.05, 0.0],
                  [1.5, 10.0, -0.02, 0.01], [40.0, 30.5, 0.0, 0.03]]
        parts = [make_particle(s[:2], 2.0, 1.0, 2, st, vel=s[2:], orc=orc) for s in states]
        prm = abi.particle_params(k_n=10.0, zeta_n=0.4, gamma_t=2.0, mu=0.5, walls=True,
                                  sub_steps=4, body_acceleration=(0.0, -1e-3), max_radius=2.0)
        for s in (g, o):
            s.set_flags(np.full(ext[0] * ext[1], 2, np.uint8))
            s.particles_set(parts, prm)
        for _ in range(30):
            for s in (g, o):
                s.dem_contact()
                s.integrate()
        for a, b in zip(g.particles_get(), o.particles_get()):
            for fld in ("position", "position_old", "rotation", "force", "torque", "acceleration"):
                assert list(getattr(a, fld)) == list(getattr(b, fld)), fld


@pytest.mark.parametrize("st", [abi.F16S, abi.BF16S])
def test_shifted_16bit_storage_error_vs_fp64(hip, orc, st):
    """FP16/BF16 storage of f − w with FP32 arithmetic: reported error against
    the FP64 oracle on the config-1 channel (no reference counterpart; the
    bound is the storage precision times the deviation scale)."""
    from lbm_mirror_amd.lattice import Lattice
    nx, ny = 64, 18
    g = Lattice(abi.D2Q9, st, (nx, ny))
    o = orc.Sim(abi.D2Q9, abi.F64, (nx, ny))
    _setup((g, o), channel_flags_2d(nx, ny), c1_ops(0.8, 0.05), bcu=parabolic_inlet(nx, ny, 0.05))
    g.step(800)
    o.step(800)
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    err_u = np.max(np.abs(ug - uo)) / np.max(np.abs(uo))
    err_r = np.max(np.abs(rg - ro))
    print(f"storage {st}: max rel velocity error {err_u:.2e}, max density error {err_r:.2e}")
    bound = 2e-2 if st == abi.F16S else 8e-2
    assert err_u < bound and err_r < bound / 10


def test_particle_laden_precision_vs_fp64(hip, orc):
    """The particle-laden channel in FP32 and with FP16/BF16 shifted storage
    (FP32 arithmetic), against the FP64 run (itself bitwise equal to the
    oracle): velocity-field and particle-trajectory errors, reported and
    bounded over 80 steps; particle trajectories reported.  The discs start
    well apart, so no contact switches on.  Particle state is kept in the
    compute type T as in the reference, and in FP32 the Verlet increment
    a*ts^2 (here 1e-5/16^2) is far below half an ulp of a position near 50:
    FP32 discs do not move where FP64 discs drift, so the trajectory error of
    every FP32-arithmetic run is the whole FP64 displacement; the bar is on
    the velocity field.  No reference counterpart for the 16-bit storage."""
    from lbm_mirror_amd.lattice import Lattice
    ext = (192, 96)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_PSM, tau=0.7),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.7,
                          ramp=(1.0, 0.005 / 50, 1.005, 50)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.7,
                          rho=1.0)}
    discs = seeded_discs(8, ext, 4.0, 7.0, 11, margin=6.0, gap=8.0)
    runs = {}
    for st, s64 in ((abi.F64, False), (abi.F32, False), (abi.F16S, False), (abi.BF16S, False),
                    (abi.F32, True), (abi.F16S, True), (abi.BF16S, True)):
        lat = Lattice(abi.D2Q9, st, ext)
        _setup((lat,), flags, ops)
        ms = abi.F64 if (st == abi.F64 or s64) else abi.F32
        parts = [make_particle([x, y], r, 1.5, 2, ms, orc=orc) for (x, y, r) in discs]
        prm = abi.particle_params(k_n=1e3, zeta_n=restitution_to_zeta(0.9), walls=True, sub_steps=16,
                                  body_acceleration=(0.0, -1e-5), eps=1.5, max_radius=7.0, state_fp64=s64)
        lat.particles_set(parts, prm)
        lat.step(80)
        ps = lat.particles_get()
        runs[(st, s64)] = (lat.get_macros()[1], np.array([[p.position[0], p.position[1]] for p in ps]))
    u64, x64 = runs[(abi.F64, False)]
    x0 = np.array([[x, y] for (x, y, _) in discs])
    disp = np.max(np.linalg.norm(x64 - x0, axis=1))
    assert 1e-4 < disp < 4.0
    bounds = {abi.F32: 1e-3, abi.F16S: 5e-2, abi.BF16S: 0.2}
    # with FP64 particle state the trajectories follow the FP64 run to the
    # storage's effect on the fluid force
    traj = {abi.F32: 1e-3, abi.F16S: 5e-2, abi.BF16S: 0.2}
    for (st, s64), (u, x) in runs.items():
        if st == abi.F64:
            continue
        eu = np.linalg.norm(u - u64) / np.linalg.norm(u64)
        ex = np.max(np.linalg.norm(x - x64, axis=1))
        moved = np.max(np.linalg.norm(x - x0, axis=1))
        print(f"storage {st}, FP64 particle state {s64}: velocity rel L2 error {eu:.2e}; trajectory error "
              f"{ex:.2e} cells (FP64 displacement {disp:.4f}, this run's {moved:.2e})")
        assert eu < bounds[st]
        assert np.all(np.isfinite(x)) and ex <= disp * 1.0001
        if s64:
            assert ex < traj[st] * disp


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_static_disc_psm_driver(hip, orc, st):
    """The static-obstacle PSM driver (poiseulle_particles_2d_psm_gpu/sim.cpp:
    207-295): OneParticleAt stamp -> Psm -> ForceGather on fluid cells, BGK
    after Zou/He pressure at inlet/outlet, bounce-back walls and corner
    flag 5.  The stamp evaluates sin (porosity.hpp:82): the oracle takes the
    engine's series sine, so porosity is bitwise; populations and macros
    within 1e-12 (FP64) / 1e-5 (FP32) relative."""
    nx, ny = 128, 40
    fl = channel_flags_2d(nx, ny).reshape(ny, nx)
    fl[0, 0] = fl[ny - 1, 0] = fl[0, nx - 1] = fl[ny - 1, nx - 1] = 5
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
           2: abi.op_desc(collision=abi.COLL_PSM, tau=0.8,
                          flags=abi.OPF_PSM_STAMP_DISC | abi.OPF_PSM_FORCE_GATHER,
                          disc=(nx * 0.25, ny * 0.5, 0.0, ny * 0.1, 1.5)),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1,
                          tau=0.8, ramp=(1.0, 0.0002 / 64, 1.0002, 64)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0,
                          tau=0.8, rho=1.0),
           5: abi.op_desc(boundary=abi.BC_BOUNCE_BACK)}
    g, o = _pair(hip, orc, abi.D2Q9, st, (nx, ny))
    o.set_series_sin(True)
    _setup((g, o), fl.reshape(-1), ops)
    g.step(301)
    o.step(301)
    pg, po = g.get_porosity(), o.get_porosity()
    assert np.array_equal(pg, po)
    assert np.sum(pg < 1) > 0
    tol = 1e-12 if st == abi.F64 else 1e-5
    fg, fo = g.get_populations(), o.get_populations()
    assert np.max(np.abs(fg - fo)) <= tol * np.max(np.abs(fo))
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    assert np.max(np.abs(rg - ro)) <= tol
    assert np.max(np.abs(ug - uo)) <= tol * np.max(np.abs(uo))
    ffg, ffo = g.get_force_field(), o.get_force_field()
    assert np.max(np.abs(ffg - ffo)) <= tol * max(np.max(np.abs(ffo)), 1e-300)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_hlbm_collision_on_porosity_field_bitwise(hip, orc, st):
    """The HLBM collision (hlbm.hpp:62-105) with a given porosity field and no
    particles (N = D = 0, so u -> u*B): bitwise against the oracle, both AA
    phases, inlet/outlet Zou/He + HLBM as in the HLBM drivers."""
    ext = (40, 18)
    g, o = _pair(hip, orc, abi.D2Q9, st, ext)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_HLBM, tau=0.8),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=1, tau=0.8,
                          ramp=(1.0, 0.01 / 16, 1.01, 16)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=0, tau=0.8, rho=1.0)}
    rng = np.random.default_rng(41)
    por = np.clip(rng.uniform(0.2, 1.4, ext[0] * ext[1]), 0, 1)
    _setup((g, o), flags, ops, por=por)
    for n in (1, 2, 37):
        g.step(n)
        o.step(n)
        _assert_same(g, o, exact=True)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_hlbm_particle_laden_run(hip, orc, st):
    """HLBM particle coupling (HlbmReset, HlbmOneParticleGather, integrate,
    HlbmOneParticleStamp, HLBM collision; examples/poiseulle_particles_2d_gpu/
    step.hpp:21-93) with 6 discs for 60 steps against the oracle."""
    ext = (96, 48)
    g, o = _pair(hip, orc, abi.D2Q9, st, ext)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_HLBM, tau=0.8),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=1, tau=0.8,
                          ramp=(1.0, 0.01 / 16, 1.01, 16)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=0, tau=0.8, rho=1.0)}
    _setup((g, o), flags, ops)
    discs = seeded_discs(6, ext, 4.0, 7.0, 17, margin=4.0, gap=4.0)
    parts = [make_particle([x, y], r, 1.5, 2, st, orc=orc) for (x, y, r) in discs]
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, walls=False, sub_steps=16,
                              body_acceleration=(0.0, -1e-4), max_radius=7.0, coupling=abi.COUPLING_HLBM)
    for s in (g, o):
        s.particles_set(parts, prm)
    o.match_gpu_arithmetic()  # FP64 force/torque sums, the engine's ramp sine
    g.step(60)
    o.step(60)
    pg, po = g.particles_get(), o.particles_get()
    moved = max(abs(a.position[d] - b_[d]) for a, b_ in zip(po, [(x, y) for (x, y, _) in discs])
                for d in range(2))
    assert moved > 1e-4
    # FP64 force/torque sums on both sides (tree on the GPU, serial in the
    # oracle): FP64 runs agree to the summation order, FP32 runs to 1e-5
    tol = 1e-11 if st == abi.F64 else 1e-5
    for a, b in zip(pg, po):
        for d in range(2):
            assert abs(a.position[d] - b.position[d]) <= tol * max(abs(b.position[d]), 1.0)
        assert abs(a.torque[0] - b.torque[0]) <= (1e-9 if st == abi.F64 else 1e-5) * max(abs(b.torque[0]), 1e-3)
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    ftol = 1e-12 if st == abi.F64 else 1e-5
    assert np.max(np.abs(rg - ro)) <= ftol
    assert np.max(np.abs(ug - uo)) <= ftol * np.max(np.abs(uo))
    assert np.max(np.abs(g.get_porosity() - o.get_porosity())) <= (1e-13 if st == abi.F64 else 1e-5)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_fplbm_collision_on_force_field_bitwise(hip, orc, st):
    """The FPLBM collision (fplbm.hpp:60-133) with a given cell force field and
    no particles: bitwise against the oracle, both AA phases; the macros hold
    rho and the momentum."""
    ext = (40, 18)
    g, o = _pair(hip, orc, abi.D2Q9, st, ext)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_FPLBM, tau=0.8),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.8,
                          ramp=(1.0, 0.01 / 8, 1.01, 8)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.8, rho=1.0)}
    rng = np.random.default_rng(43)
    force = rng.uniform(-1e-4, 1e-4, 2 * ext[0] * ext[1])
    _setup((g, o), flags, ops, force=force)
    for n in (1, 2, 37):
        g.step(n)
        o.step(n)
        _assert_same(g, o, exact=True)


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_fplbm_particle_laden_run(hip, orc, st):
    """FPLBM coupling (FpLbmReset, FpLbmOneParticle, FPLBM collision; driver
    examples/moving_poiseulle_particles_2d_fplbm_gpu/sim.cpp:185-258) with 6
    discs for 60 steps against the oracle."""
    ext = (96, 48)
    g, o = _pair(hip, orc, abi.D2Q9, st, ext)
    flags = channel_flags_2d(*ext)
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=abi.COLL_FPLBM, tau=0.8),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.8,
                          ramp=(1.0, 0.01 / 8, 1.01, 8)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.8, rho=1.0)}
    _setup((g, o), flags, ops)
    discs = seeded_discs(6, ext, 4.0, 7.0, 19, margin=4.0, gap=4.0)
    parts = [make_particle([x, y], r, 1.5, 2, st, orc=orc) for (x, y, r) in discs]
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, walls=False, sub_steps=16,
                              body_acceleration=(0.0, -1e-4), max_radius=7.0, coupling=abi.COUPLING_FPLBM)
    for s in (g, o):
        s.particles_set(parts, prm)
    o.match_gpu_arithmetic()  # FP64 force sums, the engine's ramp sine
    g.step(60)
    o.step(60)
    pg, po = g.particles_get(), o.particles_get()
    moved = max(abs(a.position[d] - (x, y)[d]) for a, (x, y, _) in zip(po, discs) for d in range(2))
    assert moved > 1e-4
    tol = 1e-11 if st == abi.F64 else 1e-5
    for a, b in zip(pg, po):
        for d in range(2):
            assert abs(a.position[d] - b.position[d]) <= tol * max(abs(b.position[d]), 1.0)
    rg, ug = g.get_macros()
    ro, uo = o.get_macros()
    ftol = 1e-12 if st == abi.F64 else 1e-5
    assert np.max(np.abs(rg - ro)) <= ftol
    assert np.max(np.abs(ug - uo)) <= ftol * np.max(np.abs(uo))
    assert np.max(np.abs(g.get_porosity() - o.get_porosity())) <= (1e-13 if st == abi.F64 else 1e-5)
    ff_tol = 1e-12 if st == abi.F64 else 1e-5
    fg, fo = g.get_force_field(), o.get_force_field()
    assert np.max(np.abs(fg - fo)) <= ff_tol * max(np.max(np.abs(fo)), 1e-12)
