"""The C++ driver written against include/lbm_mirror_amd/component.hpp (the
component vocabulary of the reference, bound per flag) gives the oracle's
channel flow bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from lbm_mirror_amd import abi
from helpers import channel_flags_2d, c1_ops, parabolic_inlet

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm_mirror_amd", "_build", "poiseuille_2d")


def test_cpp_component_driver_matches_oracle(hip, orc):
    assert os.path.exists(EXE), "C++ driver not built"
    nx, ny, steps = 64, 18, 401
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "m.bin")
        r = subprocess.run([EXE, str(nx), str(ny), str(steps), "0.05", "0.8", out],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        data = np.fromfile(out, dtype=np.float64)
    rho_g, u_g = data[: nx * ny], data[nx * ny:].reshape(2, nx * ny)
    o = orc.Sim(abi.D2Q9, abi.F64, (nx, ny))
    o.set_flags(channel_flags_2d(nx, ny))
    for fl, op in c1_ops(0.8, 0.05).items():
        o.set_operator(fl, op)
    o.set_boundary_velocity(parabolic_inlet(nx, ny, 0.05))
    o.init_equilibrium(1.0, (0.0, 0.0))
    o.step(steps)
    rho_o, u_o = o.get_macros()
    assert np.array_equal(rho_g, rho_o) and np.array_equal(u_g, u_o)
