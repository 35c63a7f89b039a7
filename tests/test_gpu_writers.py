"""Result writers on a GPU lattice (lbm_write_vtk / _csv / _forst,
lbm_read_forst): files byte-identical to the reference's writers
(write_vtk.hpp, write_csv.hpp, abstract/forst.hpp, built from the reference
sources into oracle/_ref) for the fields the lattice holds, and a Forst
checkpoint that resumes bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import c1_ops, c2_ops, channel_flags_2d, duct_flags_3d, make_particle, parabolic_inlet
from lbm_mirror_amd import LbmError, abi, output

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "libref_harness.so")
FULL = "velocity,density,force,porosity,info,dfs"


def _tree(d):
    out = {}
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith((".vtk", ".csv", ".forst")):
                out[os.path.relpath(os.path.join(root, f), d)] = open(os.path.join(root, f), "rb").read()
    return out


def _seeded_lattice(kind, st, ext, seed):
    from lbm_mirror_amd.lattice import Lattice
    lat = Lattice(kind, st, ext)
    D, Q, N = lat.D, lat.Q, lat.N
    rng = np.random.default_rng(seed)
    lat.set_flags(rng.integers(1, 6, N).astype(np.uint8))
    lat.set_operator(1, abi.op_desc(boundary=abi.BC_BOUNCE_BACK))
    for fl in (2, 3, 4, 5):
        lat.set_operator(fl, abi.op_desc(collision=abi.COLL_BGK, tau=0.8))
    f = np.tile(np.array([4 / 9] + [1 / 9] * (Q - 1))[:, None], (1, N)) * (1 + 0.05 * rng.standard_normal((Q, N)))
    lat.set_populations(f)
    lat.step(3)
    lat.set_force_field(1e-5 * rng.standard_normal((D, N)))
    lat.set_porosity(rng.uniform(0, 1, N))
    return lat


@pytest.mark.parametrize("kind,ext,shape", [(abi.D2Q9, (5, 3), 0), (abi.D3Q19, (3, 4, 2), 1)])
@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_lattice_writers_match_reference(hip, tmp_path, kind, ext, shape, st):
    lat = _seeded_lattice(kind, st, ext, 40 + shape)
    mine = tmp_path / "mine"
    output.write_vtk(lat, str(mine), "m", 7, FULL)
    output.write_csv(lat, str(mine), "m", 7, FULL)
    output.write_forst(lat, str(mine / "m.forst"), FULL)
    # the same fields, read back, through the host field-set path
    rho, u = lat.get_macros()
    x = {"velocity": u, "density": rho, "force": lat.get_force_field(), "porosity": lat.get_porosity(),
         "info": lat.get_flags(), "dfs": lat.get_populations()}
    host = tmp_path / "host"
    fs = output.Fields(kind, st, ext, **x)
    output.write_vtk_fields(fs, str(host), "m", 7, FULL)
    output.write_csv_fields(fs, str(host), "m", 7, FULL)
    output.write_forst_fields(fs, str(host / "m.forst"), FULL)
    assert _tree(str(mine)) == _tree(str(host))
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built")
    L = C.CDLL(REF)
    L.ref_write_fields.restype = C.c_int64
    dp = C.POINTER(C.c_double)
    ref = tmp_path / "ref"
    a = {k: np.ascontiguousarray(v).reshape(-1) for k, v in x.items()}
    buf = np.zeros(1 << 20, np.uint8)
    for what in (0, 1, 2):
        n = L.ref_write_fields(shape, st, 1, *[a[k].ctypes.data_as(dp) for k in ("velocity", "density", "force", "porosity")],
                               a["info"].ctypes.data_as(C.POINTER(C.c_uint8)), a["dfs"].ctypes.data_as(dp),
                               str(ref).encode(), b"m", C.c_uint64(7), what,
                               buf.ctypes.data_as(C.POINTER(C.c_uint8)), C.c_uint64(buf.size), None)
        assert n >= 0
    with open(ref / "m.forst", "wb") as fh:
        fh.write(bytes(buf[:n]))
    assert _tree(str(mine)) == _tree(str(ref))
    lat.close()


def _c1(st):
    from lbm_mirror_amd.lattice import Lattice
    nx, ny = 64, 18
    lat = Lattice(abi.D2Q9, st, (nx, ny))
    lat.set_flags(channel_flags_2d(nx, ny))
    for fl, op in c1_ops().items():
        lat.set_operator(fl, op)
    lat.set_boundary_velocity(parabolic_inlet(nx, ny, 0.05))
    return lat


def _duct(st):
    from lbm_mirror_amd.lattice import Lattice
    ext = (12, 8, 6)
    lat = Lattice(abi.D3Q19, st, ext)
    lat.set_flags(duct_flags_3d(*ext))
    lat.set_periodic(1)
    for fl, op in c2_ops(g=1e-5).items():
        lat.set_operator(fl, op)
    return lat


@pytest.mark.parametrize("make,members,st", [
    (_c1, "info,dfs,velocity,density", abi.F32),
    (_c1, "info,dfs,velocity,density", abi.F64),
    (_duct, "info,dfs,velocity,density,force,porosity", abi.F32),
])
def test_forst_checkpoint_resumes_bit_for_bit(hip, tmp_path, make, members, st):
    t1, t2 = 37, 50
    straight = make(st)
    straight.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    straight.step(t1 + t2, t0=0)
    first = make(st)
    first.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    first.step(t1, t0=0)
    p = str(tmp_path / "ckpt.forst")
    output.write_forst(first, p, members)
    first.close()
    resumed = make(st)
    output.read_forst(resumed, p, members)
    resumed.step(t2, t0=t1)
    assert np.array_equal(resumed.get_populations(), straight.get_populations())
    r0, u0 = straight.get_macros()
    r1, u1 = resumed.get_macros()
    assert np.array_equal(r0, r1) and np.array_equal(u0, u1)
    resumed.close()
    straight.close()


def test_sixteen_bit_populations_have_no_forst_layout(hip, tmp_path):
    lat = _c1(abi.F16S)
    lat.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    with pytest.raises(LbmError, match="16-bit"):
        output.write_forst(lat, str(tmp_path / "x.forst"), "dfs")
    output.write_forst(lat, str(tmp_path / "m.forst"), "velocity,density")
    assert os.path.getsize(tmp_path / "m.forst") == lat.N * 3 * 4
    lat.close()


def _particle_channel(st, coupling):
    from lbm_mirror_amd.lattice import Lattice, spheroid_mass
    from helpers import seeded_discs
    nx, ny = 160, 64
    lat = Lattice(abi.D2Q9, st, (nx, ny))
    lat.set_flags(channel_flags_2d(nx, ny))
    coll = {abi.COUPLING_PSM: abi.COLL_PSM, abi.COUPLING_HLBM: abi.COLL_HLBM,
            abi.COUPLING_FPLBM: abi.COLL_FPLBM}[coupling]
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=coll, tau=0.7),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.7,
                          ramp=(1.0, 0.002 / 30, 1.002, 30)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.7, rho=1.0)}
    for fl, op in ops.items():
        lat.set_operator(fl, op)
    parts = []
    for k, (x, y, r) in enumerate(seeded_discs(6, (nx, ny), 4.0, 7.0, 9, margin=4.0, gap=2.0)):
        m, i = spheroid_mass(2, st, r, 1.2)
        parts.append(make_particle([x, y], r, 1.2, 2, st, vel=[0.2 / 8 * (-1) ** k, 0.05 / 8], mass=m, inertia=i))
    prm = abi.particle_params(k_n=1e2, zeta_n=0.3, walls=True, sub_steps=8, body_acceleration=(0.0, -1e-4),
                              max_radius=7.0, coupling=coupling)
    return lat, parts, prm


@pytest.mark.parametrize("coupling", [abi.COUPLING_PSM, abi.COUPLING_HLBM, abi.COUPLING_FPLBM])
def test_particle_run_resumes_from_checkpoint_bit_for_bit(hip, tmp_path, coupling):
    """A particle-laden run checkpointed mid-way (Forst: info, dfs, velocity,
    density, force, porosity; the particle records through lbm_particles_get)
    and resumed in a new context (lbm_read_forst, lbm_particles_set) continues
    bit for bit like the run that never stopped, across the inlet ramp."""
    st = abi.F32
    t1, t2 = 23, 41
    members = "info,dfs,velocity,density,force,porosity"
    straight, parts, prm = _particle_channel(st, coupling)
    straight.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    straight.particles_set(parts, prm)
    straight.step(t1 + t2, t0=0)
    first, parts, prm = _particle_channel(st, coupling)
    first.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    first.particles_set(parts, prm)
    first.step(t1, t0=0)
    path = str(tmp_path / "ckpt.forst")
    output.write_forst(first, path, members)
    records = first.particles_get()
    first.close()
    resumed, _, prm = _particle_channel(st, coupling)
    output.read_forst(resumed, path, members)
    resumed.particles_set(records, prm)
    resumed.step(t2, t0=t1)
    assert np.array_equal(resumed.get_populations(), straight.get_populations())
    for a, b in zip(resumed.get_macros(), straight.get_macros()):
        assert np.array_equal(a, b)
    assert np.array_equal(resumed.get_porosity(), straight.get_porosity())
    assert np.array_equal(resumed.get_force_field(), straight.get_force_field())
    for a, b in zip(resumed.particles_get(), straight.particles_get()):
        for fld in ("position", "position_old", "rotation", "rotation_old", "acceleration", "force", "torque"):
            assert list(getattr(a, fld)) == list(getattr(b, fld)), fld
    resumed.close()
    straight.close()
