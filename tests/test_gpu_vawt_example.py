"""The rotor driver ported as a C++ example (lbm_mirror_amd/examples/vawt_3d.cpp,
after examples/vawt_3d_hlbm_gpu/sim.cpp): its command line, output directory,
particle.csv and VTK series against the same run driven from Python through
the C ABI, byte for byte."""
import os
import subprocess

import numpy as np
import pytest

from lbm_mirror_amd import abi, output
from lbm_mirror_amd.lattice import Lattice, spheroid_mass
from lbm_mirror_amd.warmstart import init_equilibrium_field

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm_mirror_amd", "_build", "vawt_3d")
DIM = 32
MEMBERS = "velocity,density,force,porosity"
# the header of sim.cpp:515-518
HEADER = ("step,theta,omega,alpha,tau,drag,x,y,z,vx,vy,vz,fx,fy,fz,rot_x,rot_y,rot_z,"
          "omega_x,omega_y,omega_z,alpha_x,alpha_y,alpha_z,tau_x,tau_y,tau_z")


def _g(x):
    # std::ostream's default for a double: %g with 6 significant digits
    return "%g" % x


def _python_run(out_dir, steps, vtk_every, axis, stable):
    zf = abi.OPF_ZOUHE_NORMAL_ONLY if stable else 0
    tau = 3 * 0.1 + 0.5
    lat = Lattice(abi.D3Q27, abi.F32, (DIM, DIM, DIM))
    f = np.ones((DIM, DIM, DIM), np.uint8)
    f[1:-1, 1:-1, 1:-1] = 2
    f[1:-1, 1:-1, 0] = 3
    f[1:-1, 1:-1, DIM - 1] = 4
    lat.set_flags(f.reshape(-1))
    lat.set_operator(1, abi.op_desc(boundary=abi.BC_BOUNCE_BACK))
    lat.set_operator(2, abi.op_desc(collision=abi.COLL_HLBM, tau=tau))
    lat.set_operator(3, abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=1, tau=tau,
                                    ramp=(1.0, 0.001 / 256, 1.001, 256), flags=zf))
    lat.set_operator(4, abi.op_desc(boundary=abi.BC_ZOUHE_PRESSURE_X, collision=abi.COLL_HLBM, east=0, tau=tau,
                                    rho=1.0, flags=zf))
    N = DIM ** 3
    u = np.zeros((3, DIM, DIM, DIM))
    u[0, 1:-1, 1:-1, 1:-1] = 0.01
    init_equilibrium_field(lat, np.ones(N), u.reshape(3, -1))
    r = 0.15 * DIM
    p = abi.Particle()
    for d, x in enumerate((0.3 * DIM, 0.5 * DIM, 0.5 * DIM)):
        p.position[d] = p.position_old[d] = x
    p.radius, p.density = r, 1.0
    p.mass, p.inertia = spheroid_mass(3, abi.F32, r, 1.0)
    prm = abi.particle_params(k_n=0.0, zeta_n=0.0, walls=False, sub_steps=16, max_radius=r,
                              coupling=abi.COUPLING_HLBM, spin_axis=axis)
    lat.particles_set([p], prm)
    # the axis normalised in float, summed in order (the example's projection)
    a = np.array(axis, np.float32)
    dot = np.float32(0)
    for x in a:
        dot = np.float32(dot + x * x)
    hat = (a / np.sqrt(dot)).astype(np.float64)
    rows = [HEADER]
    for t in range(steps):
        if vtk_every and t % vtk_every == 0:
            output.write_vtk(lat, out_dir, "t", t, MEMBERS)
        lat.step(1, t0=t)
        q = lat.particles_get()[0]
        rot, rot0 = np.array(q.rotation[:3]), np.array(q.rotation_old[:3])
        pos, pos0 = np.array(q.position[:3]), np.array(q.position_old[:3])
        th = om = al = tq = 0.0
        for k in range(3):
            th += rot[k] * hat[k]
            om += (rot[k] - rot0[k]) * 16.0 * hat[k]
            al += q.angular_acceleration[k] * hat[k]
            tq += q.torque[k] * hat[k]
        vals = [th, om, al, tq, q.force[0], *pos, *((pos - pos0) * 16.0), *q.force[:3], *rot,
                *((rot - rot0) * 16.0), *q.angular_acceleration[:3], *q.torque[:3]]
        rows.append(",".join([str(t)] + [_g(v) for v in vals]))
    output.write_vtk(lat, out_dir, "t", steps, MEMBERS)
    q = lat.particles_get()[0]
    lat.close()
    return "\n".join(rows) + "\n", q


def _files(d):
    out = {}
    for root, _, files in os.walk(d):
        for f in files:
            out[os.path.relpath(os.path.join(root, f), d)] = open(os.path.join(root, f), "rb").read()
    return out


@pytest.mark.parametrize("axis,stable", [((0.0, 1.0, 0.0), False), ((0.3, 1.0, -0.2), True)])
def test_vawt_example_matches_abi_run(hip, tmp_path, axis, stable):
    assert os.path.exists(EXE), "vawt_3d example not built"
    steps, every = 40, 10
    home = tmp_path / "home"
    home.mkdir()
    env = dict(os.environ, HOME=str(home))
    cmd = [EXE, "--name", "rotor", "--time_steps", str(steps), "--vtk_every", str(every),
           "--axis_x", str(axis[0]), "--axis_y", str(axis[1]), "--axis_z", str(axis[2]),
           "--stable_faces", "true" if stable else "false"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    out = home / ".lbm" / "data" / "rotor"
    assert out.is_dir()
    csv = (out / "particle.csv").read_text()
    assert csv.splitlines()[0] == HEADER and len(csv.splitlines()) == steps + 1
    mine = tmp_path / "py"
    want, q = _python_run(str(mine), steps, every, axis, stable)
    assert csv == want
    got = _files(str(out))
    del got["particle.csv"]
    assert sorted(got) == sorted(f"vtk/t_{t}.vtk" for t in list(range(0, steps, every)) + [steps])
    assert got == _files(str(mine))
    # pinned at its start, turning about the axis only
    assert list(q.position[:3]) == [float(np.float32(x)) for x in (0.3 * DIM, 0.5 * DIM, 0.5 * DIM)]
    rot = np.array(q.rotation[:3])
    a = np.array(axis) / np.linalg.norm(axis)
    # F32 state: the axis is normalised and the rotation kept in float
    assert np.linalg.norm(rot - np.dot(rot, a) * a) <= 1e-4 * max(np.linalg.norm(rot), 1e-30)

