"""lbm_step's graph replays (lbm_set_step_graph) against the same steps
launched one by one: bit-identical populations, macros and particle records
after every call of a sequence of calls of uneven lengths, with host
synchronisation and read-back between them, including calls that start
inside a Zou/He density ramp (the replayed launches read the step index
from device memory, set before each replay), odd start phases and a change
of an argument the graph captured (which forces a new capture)."""
import numpy as np
import pytest

from helpers import c1_ops, channel_flags_2d, make_particle, parabolic_inlet, seeded_discs
from lbm_mirror_amd import abi

pytestmark = pytest.mark.gpu


def _c1(st, graph, ramp=40):
    from lbm_mirror_amd.lattice import Lattice
    nx, ny = 96, 24
    lat = Lattice(abi.D2Q9, st, (nx, ny))
    lat.set_flags(channel_flags_2d(nx, ny))
    ops = c1_ops()
    ops[4] = abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.8,
                         ramp=(1.0, -0.002 / ramp, 0.998, ramp))
    for fl, op in ops.items():
        lat.set_operator(fl, op)
    lat.set_boundary_velocity(parabolic_inlet(nx, ny, 0.05))
    lat.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    lat.set_step_graph(graph)
    return lat


def _c3(st, graph, coupling):
    from lbm_mirror_amd.lattice import Lattice
    nx, ny = 160, 64
    lat = Lattice(abi.D2Q9, st, (nx, ny))
    lat.set_flags(channel_flags_2d(nx, ny))
    coll = {abi.COUPLING_PSM: abi.COLL_PSM, abi.COUPLING_HLBM: abi.COLL_HLBM, abi.COUPLING_FPLBM: abi.COLL_FPLBM}[coupling]
    ops = {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK), 2: abi.op_desc(collision=coll, tau=0.7),
           3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.7,
                          ramp=(1.0, 0.001 / 24, 1.001, 24)),
           4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.7, rho=1.0)}
    for fl, op in ops.items():
        lat.set_operator(fl, op)
    lat.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    from lbm_mirror_amd.lattice import spheroid_mass
    parts = []
    for (x, y, r) in seeded_discs(5, (nx, ny), 5.0, 8.0, 31, margin=4.0, gap=3.0):
        m, i = spheroid_mass(2, st, r, 1.5)
        parts.append(make_particle([x, y], r, 1.5, 2, st, mass=m, inertia=i))
    prm = abi.particle_params(k_n=1e3, zeta_n=0.03, walls=True, sub_steps=8, body_acceleration=(0.0, -1e-5),
                              max_radius=8.0, coupling=coupling)
    lat.particles_set(parts, prm)
    lat.set_step_graph(graph)
    return lat, prm


def _same(a, b):
    assert np.array_equal(a.get_populations(), b.get_populations())
    ra, ua = a.get_macros()
    rb, ub = b.get_macros()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)


CALLS = [1, 37, 100, 5, 64, 66, 3, 130]


# ramp 400: every replay of the first calls runs inside the density ramp
@pytest.mark.parametrize("ramp", [40, 400])
@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_channel_graph_replays_match_launches(hip, st, ramp):
    g, e = _c1(st, 32, ramp), _c1(st, 0, ramp)
    t = 0
    for n in CALLS:
        g.step(n, t0=t)
        e.step(n, t0=t)
        g.sync()
        _same(g, e)
        t += n
    # a captured argument changes (periodicity): the next replay captures anew
    for lat in (g, e):
        if t % 2:
            lat.step(1, t0=t)
    t += t % 2
    for lat in (g, e):
        lat.set_periodic(1)
    for n in (70, 41):
        g.step(n, t0=t)
        e.step(n, t0=t)
        _same(g, e)
        t += n
    g.close()
    e.close()


@pytest.mark.parametrize("coupling", [abi.COUPLING_PSM, abi.COUPLING_HLBM, abi.COUPLING_FPLBM])
def test_particle_graph_replays_match_launches(hip, coupling):
    st = abi.F32
    (g, _), (e, _) = _c3(st, 16, coupling), _c3(st, 0, coupling)
    t = 0
    for n in CALLS[:6]:
        g.step(n, t0=t)
        e.step(n, t0=t)
        _same(g, e)
        for a, b in zip(g.particles_get(), e.particles_get()):
            for fld in ("position", "position_old", "rotation", "force", "torque", "acceleration"):
                assert list(getattr(a, fld)) == list(getattr(b, fld)), fld
        assert np.array_equal(g.get_porosity(), e.get_porosity())
        t += n
    g.close()
    e.close()


def test_step_graph_setting_is_checked(hip):
    from lbm_mirror_amd import LbmError
    lat = _c1(abi.F32, 0)
    with pytest.raises(LbmError):
        lat.set_step_graph(3)
    lat.set_step_graph(4096)
    lat.close()
