"""HlbmOneParticleStamp where post-move boxes overlap (hlbm.hpp:264-308): the
reference stamps particle after particle, so a shared cell sums the terms
D += 1 - B, N += u_p (1 - B) in particle order.  The GPU's sums after a step
against that serial order restated here in float32 from the GPU's own
porosity and particle records: bit for bit up to 1024 particles (dense
overlaps, and a block whose overlap list overflows), to rounding above (the
atomic path); and against the oracle's run to the porosity's ulps."""
import numpy as np
import pytest

from helpers import channel_flags_2d, make_particle
from lbm_mirror_amd import abi

pytestmark = pytest.mark.gpu
f32 = np.float32


def _serial_stamp(por, parts, ext, sub_steps):
    """The reference's stamp loop in float32, particles in index order."""
    nx, ny = ext
    N = nx * ny
    n = np.zeros((2, N), f32)
    d = np.zeros(N, f32)
    ts = f32(1) / f32(sub_steps)
    flip_all = (f32(1) - por.astype(f32)).astype(f32)
    for p in parts:
        pos = [f32(p.position[k]) for k in range(2)]
        old = [f32(p.position_old[k]) for k in range(2)]
        r = f32(p.radius)
        vp = [f32(f32(pos[k] - old[k]) / ts) for k in range(2)]
        lo, hi = [], []
        for k, e in enumerate((nx, ny)):
            lower = f32(pos[k] - r) if pos[k] >= r else f32(0)
            up1 = f32(f32(pos[k] + r) + f32(1))
            h = min(int(up1) if up1 > 0 else 0, e)
            l = min(int(lower) if lower > 0 else 0, h)
            lo.append(l)
            hi.append(h)
        xs, ys = np.meshgrid(np.arange(lo[0], hi[0]), np.arange(lo[1], hi[1]), indexing="ij")
        c = (xs + nx * ys).reshape(-1)
        fl = flip_all[c]
        d[c] = (d[c] + fl).astype(f32)
        for k in range(2):
            n[k, c] = (n[k, c] + (f32(vp[k]) * fl).astype(f32)).astype(f32)
    return n, d


def _case(kind, rng):
    if kind == "dense":  # 60 freely overlapping discs
        ext = (96, 64)
        r = rng.uniform(3.0, 6.0, 60)
        xy = np.stack([rng.uniform(8, ext[0] - 8, 60), rng.uniform(8, ext[1] - 8, 60)], 1)
    elif kind == "overflow":  # particle 0's box holds 300 more
        ext = (128, 96)
        r = np.concatenate([[24.0], rng.uniform(0.8, 1.6, 300)])
        xy = np.concatenate([[[64.0, 48.0]], np.stack([rng.uniform(42, 86, 300), rng.uniform(26, 70, 300)], 1)])
    else:  # "atomic": more than 1024 particles
        ext = (160, 120)
        r = rng.uniform(1.0, 3.0, 1100)
        xy = np.stack([rng.uniform(5, ext[0] - 5, 1100), rng.uniform(5, ext[1] - 5, 1100)], 1)
    return ext, xy, r


@pytest.mark.parametrize("kind", ["dense", "overflow", "atomic"])
def test_hlbm_stamp_sums_in_particle_order(hip, orc, kind):
    from lbm_mirror_amd.lattice import Lattice
    rng = np.random.default_rng(17)
    ext, xy, rad = _case(kind, rng)
    N = ext[0] * ext[1]
    sub = 8
    g, o = Lattice(abi.D2Q9, abi.F32, ext), orc.Sim(abi.D2Q9, abi.F32, ext)
    D, Q, c, w, _ = orc.lattice(abi.D2Q9)
    f0 = np.outer(w, np.ones(N)) * (1 + 0.02 * rng.uniform(-1, 1, size=(Q, N)))
    parts = [make_particle([x, y], r, 1.5, 2, abi.F32, vel=list(rng.uniform(-0.05, 0.05, 2) / sub), orc=orc)
             for (x, y), r in zip(xy, rad)]
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, sub_steps=sub, max_radius=float(rad.max()),
                              body_acceleration=(0.0, -1e-4), coupling=abi.COUPLING_HLBM)
    for s in (g, o):
        s.set_flags(channel_flags_2d(*ext))
        s.set_operator(1, abi.op_desc(boundary=abi.BC_BOUNCE_BACK))
        for fl in (2, 3, 4):
            s.set_operator(fl, abi.op_desc(collision=abi.COLL_HLBM, tau=0.7))
        s.set_populations(f0)
        s.particles_set(parts, prm)
        s.step(1)
    ng, dg = g.get_hlbm_sums()
    nh, dh = _serial_stamp(g.get_porosity(), g.particles_get(), ext, sub)
    assert np.count_nonzero(dh > 1) > 0  # cells that several particles stamp
    if kind == "atomic":
        assert np.max(np.abs(dg - dh)) <= 1e-6 * max(1.0, np.max(dh))
        assert np.max(np.abs(ng - nh)) <= 1e-6 * max(1e-30, np.max(np.abs(nh)))
    else:
        assert np.array_equal(dg, dh.astype(np.float64)) and np.array_equal(ng, nh.astype(np.float64))
    no, do = o.get_hlbm()
    assert np.max(np.abs(dg - do)) <= 1e-5 * max(1.0, np.max(do))
    assert np.max(np.abs(ng - no)) <= 1e-5 * max(1e-30, np.max(np.abs(no))) + 1e-9
    g.close()
