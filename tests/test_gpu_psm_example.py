"""The particle-laden channel driver ported as a C++ example
(lbm_mirror_amd/examples/moving_particles_2d_psm.cpp, after
examples/moving_poiseulle_particles_2d_psm_gpu/sim.cpp): its per-step
particle lines and VTK files byte for byte against the same run driven from
Python through the C ABI, and the disc's trajectory against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import make_particle
from lbm_mirror_amd import abi, output

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm_mirror_amd", "_build", "moving_particles_2d_psm")
NY = 256
NX = NY * 20
MACROS = "velocity,density,porosity,force"  # the driver's MacroStruct order (sim.cpp:55-60)


def _flags():
    f = np.ones((NY, NX), np.uint8)
    f[1:-1, 1:-1] = 2
    f[1:-1, 0] = 3
    f[1:-1, NX - 1] = 4
    return f.reshape(-1)


def _ops():
    return {1: abi.op_desc(boundary=abi.BC_BOUNCE_BACK),
            2: abi.op_desc(collision=abi.COLL_PSM, tau=0.8),
            3: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=1, tau=0.8,
                           ramp=(1.0, 0.01 / 8, 1.01, 8)),
            4: abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_BGK, east=0, tau=0.8, rho=1.0)}


def _setup(s, orc, sub_steps):
    s.set_flags(_flags())
    for fl, op in _ops().items():
        s.set_operator(fl, op)
    s.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    r = float(np.float32(NY) * np.float32(0.1))
    p = make_particle([0.25 * NX, 0.5 * NY], r, 1.0, 2, abi.F32, orc=orc)
    prm = abi.particle_params(k_n=1.0e-2, zeta_n=0.3, walls=True, sub_steps=sub_steps, max_radius=r,
                              coupling=abi.COUPLING_PSM)
    return p, prm


def test_psm_channel_example_matches_abi_run_and_oracle(hip, orc, tmp_path):
    assert os.path.exists(EXE), "moving_particles_2d_psm example not built"
    steps, every, sub = 40, 16, 16
    home = tmp_path / "home"
    home.mkdir()
    r = subprocess.run([EXE, "--name", "chan", "--time_steps", str(steps), "--vtk_every", str(every)],
                       capture_output=True, text=True, timeout=180, env=dict(os.environ, HOME=str(home)))
    assert r.returncode == 0, r.stderr
    out = home / ".lbm" / "data" / "chan"

    from lbm_mirror_amd.lattice import Lattice
    lat = Lattice(abi.D2Q9, abi.F32, (NX, NY))
    p, prm = _setup(lat, orc, sub)
    mine = tmp_path / "py"
    output.write_vtk(lat, str(mine), "initial_state", 0, "info,dfs")
    lat.particles_set([p], prm)
    lines = []
    for i in range(steps):
        lat.step(1, t0=i)
        if i % every == 0:
            output.write_vtk(lat, str(mine), "m", i, MACROS)
        q = lat.particles_get()[0]
        lines.append("Particle: %g, %g; %g, %g;%d" % (q.position[0], q.position[1], q.acceleration[0],
                                                       q.acceleration[1], i))
    assert r.stdout.splitlines() == lines

    def files(d):
        return {os.path.relpath(os.path.join(a, f), d): open(os.path.join(a, f), "rb").read()
                for a, _, fs in os.walk(d) for f in fs}
    got = files(str(out))
    assert sorted(got) == sorted(["vtk/initial_state_0.vtk"] + [f"vtk/m_{t}.vtk" for t in range(0, steps, every)])
    assert got == files(str(mine))

    # the same driver in the oracle: the disc's path
    o = orc.Sim(abi.D2Q9, abi.F32, (NX, NY))
    po, prm_o = _setup(o, orc, sub)
    o.particles_set([po], prm_o)
    o.step(steps)
    a, b = q, o.particles_get()[0]
    for d in range(2):
        # FP32 positions near 1280: the first steps' displacements are below
        # their resolution, so the positions agree to a few ulps
        assert abs(a.position[d] - b.position[d]) <= 8 * float(np.spacing(np.float32(b.position[d])))
        # within 40 steps the inlet's pressure front (~c_s t = 23 cells) has not
        # reached the disc 1254 cells downstream: both see (near) zero force
        assert abs(a.acceleration[d] - b.acceleration[d]) <= 1e-9 + 1e-3 * abs(b.acceleration[d])
    # the flow: the inlet ramp's front, density and velocity against the oracle
    rg, ug = lat.get_macros()
    ro, uo = o.get_macros()
    assert np.max(np.abs(uo)) > 1e-4
    assert np.max(np.abs(rg - ro)) <= 1e-6
    assert np.max(np.abs(ug - uo)) <= 1e-5 * np.max(np.abs(uo))
    lat.close()
