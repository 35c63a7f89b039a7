"""PSM scatter porosity kept without a per-cell reset: each gather sweeps its
particle's previous and current boxes and returns the cells that no current
box holds to porosity 1, which must leave exactly the field that PsmReset
followed by the gathers gives (psm.hpp:20-32, 304-424).  Checked bit for bit
against the oracle (which resets every cell) for discs that move several
cells per step, discs moved elsewhere by lbm_particles_set, a porosity
field set from outside between couplings, and runs of full steps."""
import numpy as np
import pytest

from helpers import channel_flags_2d, make_particle
from lbm_mirror_amd import abi

pytestmark = pytest.mark.gpu


def _setup(st, ext, rng):
    from lbm_mirror_amd.lattice import Lattice
    from oracle import oracle as orc
    N = ext[0] * ext[1]
    g, o = Lattice(abi.D2Q9, st, ext), orc.Sim(abi.D2Q9, st, ext)
    _, Q, _, w, _ = orc.lattice(abi.D2Q9)
    f0 = np.outer(w, np.ones(N)) * (1 + 0.05 * rng.uniform(-1, 1, size=(Q, N)))
    for s in (g, o):
        s.set_flags(channel_flags_2d(*ext))
        s.set_operator(1, abi.op_desc(boundary=abi.BC_BOUNCE_BACK))
        for fl in (2, 3, 4):
            s.set_operator(fl, abi.op_desc(collision=abi.COLL_PSM, tau=0.7))
        s.set_populations(f0)
    return g, o


def _discs(rng, ext, n, vmax, st, orc):
    parts = []
    for _ in range(n):
        r = rng.uniform(2.5, 7.0)
        x = rng.uniform(r + 2, ext[0] - r - 2)
        y = rng.uniform(r + 2, ext[1] - r - 2)
        v = rng.uniform(-vmax, vmax, size=2)
        # heavy discs: the fluid's force changes no position bit, so GPU and
        # oracle boxes stay equal although their force sums differ in order
        parts.append(make_particle([x, y], r, 1.5, 2, st, vel=list(v), mass=1e12, inertia=1e12))
    return parts


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_sweep_matches_reset_across_moves_and_outside_fields(hip, orc, st):
    rng = np.random.default_rng(5)
    ext = (120, 72)
    g, o = _setup(st, ext, rng)
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, sub_steps=1, eps=1.5, max_radius=7.0)
    o.match_gpu_arithmetic()
    for rnd in range(5):
        # new positions every round: boxes far from the previous ones
        parts = _discs(rng, ext, 14, 3.0, st, orc)  # up to 3 cells per move
        for s in (g, o):
            s.particles_set(parts, prm)
        for k in range(3):
            g.psm_couple()
            o.psm_couple()
            assert np.array_equal(g.get_porosity(), o.get_porosity()), f"round {rnd} coupling {k}"
            # the discs move between couplings (integrate only; the boxes shift)
            for s in (g, o):
                s.integrate()
        if rnd % 2:
            # a porosity field set from outside: the next coupling starts over
            junk = rng.uniform(0, 1, size=ext[0] * ext[1])
            for s in (g, o):
                s.set_porosity(junk)
            g.psm_couple()
            o.psm_couple()
            assert np.array_equal(g.get_porosity(), o.get_porosity()), f"round {rnd} after set_porosity"
    g.close()


@pytest.mark.parametrize("st", [abi.F64, abi.F32])
def test_sweep_over_full_steps_with_moving_discs(hip, orc, st):
    """Full steps (step graphs replay inside) with discs moving up to 0.05
    cells per step (their boxes shift every 20 steps or more; faster discs
    would drive the fluid past the lattice's speed of sound), against the
    oracle's reset-then-gather steps."""
    rng = np.random.default_rng(9)
    ext = (160, 64)
    g, o = _setup(st, ext, rng)
    for s in (g, o):
        s.set_periodic(1)  # a closed channel: periodic x, walls in y
    parts = _discs(rng, ext, 10, 0.05, st, orc)
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, sub_steps=1, eps=1.5, max_radius=7.0, walls=False)
    for s in (g, o):
        s.particles_set(parts, prm)
    o.match_gpu_arithmetic()
    for n in (1, 2, 37, 70, 150):
        g.step(n)
        o.step(n)
        assert np.array_equal(g.get_porosity(), o.get_porosity()), f"after {n} more steps"
        pg = np.array([[q.position[0], q.position[1]] for q in g.particles_get()])
        po = np.array([[q.position[0], q.position[1]] for q in o.particles_get()])
        tol = 1e-11 if st == abi.F64 else 1e-5
        assert np.max(np.abs(pg - po)) <= tol * np.max(np.abs(po))
    g.close()
