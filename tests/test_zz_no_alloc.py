"""lbm_residual_noalloc through the C ABI: the same three FP64 values, bit for
bit, as lbm_residual on two calls in a row (the second on a smaller lattice,
both through the one static scratch array), and its size limit."""
import ctypes as C

import numpy as np
import pytest

from lbm_mirror_amd import abi, check
from helpers import duct_flags_3d, c2_ops, channel_flags_2d, c1_ops, parabolic_inlet

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime library not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _duct(ext):
    from lbm_mirror_amd.lattice import Lattice
    g = Lattice(abi.D3Q19, abi.F32, ext)
    g.set_flags(duct_flags_3d(*ext))
    g.set_periodic(1)
    for fl, op in c2_ops(0.6, 1e-5).items():
        g.set_operator(fl, op)
    g.init_equilibrium()
    g.set_residual(True)
    return g


def _channel(ext):
    from lbm_mirror_amd.lattice import Lattice
    g = Lattice(abi.D2Q9, abi.F64, ext)
    g.set_flags(channel_flags_2d(*ext))
    for fl, op in c1_ops(0.8, 0.05).items():
        g.set_operator(fl, op)
    g.set_boundary_velocity(parabolic_inlet(*ext, 0.05))
    g.init_equilibrium(1.0, (0.0, 0.0))
    g.set_residual(True)
    return g


def test_noalloc_residual_equals_residual(hip):
    h = _hip()
    d_out = C.c_void_p()
    assert h.hipMalloc(C.byref(d_out), 3 * 8) == 0
    try:
        # the larger lattice first (more block sums), then a smaller one of
        # another lattice kind and precision through the same scratch
        cases = [_duct((200, 40, 36)), _channel((300, 41))]
        for g, steps in zip(cases, (25, 31)):
            g.step(steps)
            ref = g.residual()
            assert h.hipMemset(d_out, 0xff, 3 * 8) == 0  # NaN pattern: the call must write all three
            check(g.L.lbm_residual_noalloc(g._h, d_out))
            g.sync()
            got = np.zeros(3)
            assert h.hipMemcpy(got.ctypes.data_as(C.c_void_p), d_out, 3 * 8, 2) == 0  # device to host
            assert ref[0] > 0 and ref[2] > 0
            assert got.tobytes() == np.array(ref).tobytes(), (got, ref)
            # and the other way round: lbm_residual after the no-alloc call
            assert g.residual() == ref
    finally:
        h.hipFree(d_out)


def test_noalloc_residual_limits(hip):
    from lbm_mirror_amd.lattice import Lattice
    h = _hip()
    d_out = C.c_void_p()
    assert h.hipMalloc(C.byref(d_out), 3 * 8) == 0
    try:
        g = _duct((40, 8, 8))
        assert g.L.lbm_residual_noalloc(g._h, None) == abi.ERR_INVALID_ARG
        g.set_residual(False)
        assert g.L.lbm_residual_noalloc(g._h, d_out) == abi.ERR_INVALID_STATE
        # one cell over the maximum (8192 blocks of 2048 cells)
        big = Lattice(abi.D2Q9, abi.F32, (16_777_217, 1))
        assert big.L.lbm_residual_noalloc(big._h, d_out) == abi.ERR_NOT_SUPPORTED
        assert b"exceed" in big.L.lbm_last_error()
        big.close()
    finally:
        h.hipFree(d_out)
