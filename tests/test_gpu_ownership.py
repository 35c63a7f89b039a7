"""Which particle writes a cell that several particles' boxes or solid
fractions share: the highest index, since the reference's gathers run in
particle order and each overwrites the cell (psm.hpp:304-424,
hlbm.hpp:200-260, fplbm.hpp:204-224).  The GPU decides this inside each
gather block from the boxes that overlap its particle's (up to 1024
particles), tests every higher-index particle where a block's overlap list
overflows, and runs the per-cell owner pass above 1024 particles.  Each path
against the oracle, for the three couplings and both storages."""
import numpy as np
import pytest

from helpers import channel_flags_2d, make_particle
from lbm_mirror_amd import abi

pytestmark = pytest.mark.gpu


def _discs(case, rng):
    if case == "dense":  # 400 freely overlapping discs
        ext = (128, 96)
        n = 400
        r = rng.uniform(1.5, 4.5, n)
        xy = np.stack([rng.uniform(6, ext[0] - 6, n), rng.uniform(6, ext[1] - 6, n)], 1)
    elif case == "overflow":  # particle 0's box holds 300 higher-index discs
        ext = (128, 96)
        n = 301
        r = np.concatenate([[24.0], rng.uniform(0.8, 1.6, n - 1)])
        xy = np.concatenate([[[64.0, 48.0]], np.stack([rng.uniform(42, 86, n - 1), rng.uniform(26, 70, n - 1)], 1)])
    else:  # "owner_pass": more than 1024 particles
        ext = (160, 120)
        n = 1100
        r = rng.uniform(1.0, 3.0, n)
        xy = np.stack([rng.uniform(5, ext[0] - 5, n), rng.uniform(5, ext[1] - 5, n)], 1)
    return ext, xy, r


@pytest.mark.parametrize("coupling", [abi.COUPLING_PSM, abi.COUPLING_HLBM, abi.COUPLING_FPLBM])
@pytest.mark.parametrize("case", ["dense", "overflow", "owner_pass"])
@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_overlapping_particles_write_like_the_oracle(hip, orc, st, case, coupling):
    from lbm_mirror_amd.lattice import Lattice
    rng = np.random.default_rng(11)
    ext, xy, rad = _discs(case, rng)
    N = ext[0] * ext[1]
    g, o = Lattice(abi.D2Q9, st, ext), orc.Sim(abi.D2Q9, st, ext)
    D, Q, c, w, _ = orc.lattice(abi.D2Q9)
    f0 = np.outer(w, np.ones(N)) * (1 + 0.05 * rng.uniform(-1, 1, size=(Q, N)))
    u0 = 0.02 * rng.uniform(-1, 1, size=(2, N))
    coll = {abi.COUPLING_PSM: abi.COLL_PSM, abi.COUPLING_HLBM: abi.COLL_HLBM,
            abi.COUPLING_FPLBM: abi.COLL_FPLBM}[coupling]
    for s in (g, o):
        s.set_flags(channel_flags_2d(*ext))
        s.set_operator(1, abi.op_desc(boundary=abi.BC_BOUNCE_BACK))
        for fl in (2, 3, 4):
            s.set_operator(fl, abi.op_desc(collision=coll, tau=0.7))
        s.set_populations(f0)
        s.set_macros(np.ones(N), u0)
    parts = []
    for (x, y), r in zip(xy, rad):
        v = rng.uniform(-0.01, 0.01, size=2)
        parts.append(make_particle([x, y], r, 1.5, 2, st, vel=list(v / 16), orc=orc))
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, sub_steps=16, eps=1.5, max_radius=float(rad.max()),
                              coupling=coupling, psm_torque=True)
    for s in (g, o):
        s.particles_set(parts, prm)
    o.match_gpu_arithmetic()  # FP64 sums, the engine's ramp sine
    Fg, Tg = g.psm_couple()
    Fo, To, A = o.psm_couple()
    tol = 1e-12 if st == abi.F64 else 1e-5
    assert np.all(np.abs(Fg - Fo) <= tol * A[:, None] + 1e-300)
    assert np.all(np.abs(Tg - To) <= tol * A[:, None] * 64 + 1e-300)
    # porosity: the owner's value, bit for bit (the same ramp sine on both
    # sides); a wrong owner shows as another disc's value
    pg, po = g.get_porosity(), o.get_porosity()
    assert np.array_equal(pg, po) and np.sum(po < 1) > 0
    ffg, ffo = g.get_force_field(), o.get_force_field()
    fscale = max(np.max(np.abs(ffo)), 1e-300)
    assert np.max(np.abs(ffg - ffo)) <= (1e-13 if st == abi.F64 else 1e-5) * fscale
    g.close()
