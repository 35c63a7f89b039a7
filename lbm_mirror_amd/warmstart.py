"""Learned warm start: the convolutional encoder-decoder behind lbm_net_*.

``Net`` wraps one network on one GPU: its on-device training set, Adam
training, inference, and the raw forward / loss-gradient calls used to check
the kernels.  ``warm_start`` initialises batch members from predictions
(equilibrium of the predicted rho, u; solid cells rho 1, u 0).
"""
import ctypes as C

import numpy as np

from . import abi, check, load_library


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class Net:
    FEATURES = 8
    OUTPUTS = 3

    def __init__(self, height, width, widths=(32, 64, 96, 160), batch=16, capacity=0, seed=1, device=0):
        self.L = load_library()
        self.cfg = abi.net_config(height, width, widths, batch, capacity, seed)
        self.height, self.width, self.widths = height, width, tuple(widths)
        self.batch, self.capacity = batch, capacity
        h = C.c_void_p()
        check(self.L.lbm_net_create(C.byref(h), C.byref(self.cfg), device))
        self._h = h
        n = C.c_uint64()
        check(self.L.lbm_net_param_count(self._h, C.byref(n)))
        self.param_count = n.value

    def close(self):
        if getattr(self, "_h", None):
            self.L.lbm_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ----------------------------------------------------------
    def get_params(self):
        p = np.zeros(self.param_count, dtype=np.float32)
        check(self.L.lbm_net_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.param_count:
            raise ValueError("parameter count mismatch")
        check(self.L.lbm_net_set_params(self._h, _fp(a)))

    # ---- data ------------------------------------------------------------------
    def features(self, flags, u_max, tau):
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        out = np.zeros(f.size * self.FEATURES, dtype=np.float32)
        check(self.L.lbm_net_features(self._h, _u8(f), u_max, tau, _fp(out)))
        return out.reshape(self.height, self.width, self.FEATURES)

    def add_sample(self, flags, u_max, tau, rho, u):
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        check(self.L.lbm_net_add_sample(self._h, _u8(f), u_max, tau, _dp(r), _dp(uu)))

    def clear_samples(self):
        check(self.L.lbm_net_clear_samples(self._h))

    def sample_count(self):
        n = C.c_uint32()
        check(self.L.lbm_net_sample_count(self._h, C.byref(n)))
        return n.value

    # ---- training / inference ----------------------------------------------------
    def train(self, steps, lr=1e-3, seed=0):
        losses = np.zeros(max(1, steps), dtype=np.float32)
        check(self.L.lbm_net_train(self._h, steps, lr, seed, _fp(losses)))
        return losses[:steps]

    def predict(self, flags_list, u_max, tau):
        k = len(flags_list)
        N = self.height * self.width
        f = np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.uint8).reshape(-1) for x in flags_list]))
        um = np.ascontiguousarray(u_max, dtype=np.float64)
        ta = np.ascontiguousarray(tau, dtype=np.float64)
        rho = np.zeros(k * N)
        u = np.zeros(k * 2 * N)
        check(self.L.lbm_net_predict(self._h, k, _u8(f), _dp(um), _dp(ta), _dp(rho), _dp(u)))
        return rho.reshape(k, N), u.reshape(k, 2, N)

    def correct_flux(self, flags_list, u_max, u):
        """Rescale each column of predicted velocities u (count, 2, N) to the
        inlet's flux (lbm_net_correct_flux); returns the corrected copy."""
        k = len(flags_list)
        f = np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.uint8).reshape(-1) for x in flags_list]))
        um = np.ascontiguousarray(u_max, dtype=np.float64)
        out = np.ascontiguousarray(np.array(u, dtype=np.float64).reshape(k, 2, -1))
        check(self.L.lbm_net_correct_flux(self._h, k, _u8(f), _dp(um), _dp(out)))
        return out

    def forward_raw(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        k = x.size // (self.height * self.width * self.FEATURES)
        y = np.zeros(k * self.height * self.width * self.OUTPUTS, dtype=np.float32)
        check(self.L.lbm_net_forward_raw(self._h, k, _fp(x), _fp(y)))
        return y.reshape(k, self.height, self.width, self.OUTPUTS)

    def loss_grad_raw(self, x, t, m):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = np.ascontiguousarray(m, dtype=np.float32)
        k = x.size // (self.height * self.width * self.FEATURES)
        loss = C.c_double()
        g = np.zeros(self.param_count, dtype=np.float32)
        check(self.L.lbm_net_loss_grad_raw(self._h, k, _fp(x), _fp(t), _fp(m), C.byref(loss), _fp(g)))
        return loss.value, g

    def output_scale(self):
        m = np.zeros(3, dtype=np.float32)
        s = np.zeros(3, dtype=np.float32)
        check(self.L.lbm_net_get_output_scale(self._h, _fp(m), _fp(s)))
        return m, s

    def set_output_scale(self, mean, stdev):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(stdev, dtype=np.float32)
        check(self.L.lbm_net_set_output_scale(self._h, _fp(m), _fp(s)))

    def last_ms(self):
        m = C.c_double()
        check(self.L.lbm_net_last_ms(self._h, C.byref(m)))
        return m.value


def init_equilibrium_field(lattice, rho, u):
    """Start a Lattice (or batch member) from per-cell (rho, u)."""
    r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    if r.size != lattice.N or uu.size != lattice.D * lattice.N:
        raise ValueError("field sizes do not match the lattice")
    check(lattice.L.lbm_init_equilibrium_field(lattice._h, _dp(r), _dp(uu)))
