"""ctypes wrapper of the CPU oracle (oracle/lbm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product never imports this module.
"""
import ctypes as C
import os

import numpy as np

from lbm_mirror_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liblbm_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        u64p = C.POINTER(C.c_uint64)
        vp = C.c_void_p
        L.orc_lattice.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), dp, C.POINTER(C.c_int)]
        L.orc_equilibrium.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_rho_u.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_cell_op.argtypes = [C.c_int, C.c_int, C.POINTER(abi.OpDesc), C.c_uint64, dp,
                                  dp, dp, C.c_double, dp, dp, dp]
        L.orc_porosity.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double, dp]
        L.orc_aabb.argtypes = [C.c_int, C.c_int, dp, C.c_double, u64p, u64p, u64p]
        L.orc_spheroid_mass.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp, dp]
        L.orc_sim_create.argtypes = [C.c_int, C.c_int, u64p]
        L.orc_sim_create.restype = vp
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_set_flags.argtypes = [vp, C.POINTER(C.c_uint8)]
        L.orc_sim_set_periodic.argtypes = [vp, C.c_uint32]
        L.orc_sim_set_operator.argtypes = [vp, C.c_uint8, C.POINTER(abi.OpDesc)]
        L.orc_sim_set_boundary_velocity.argtypes = [vp, dp]
        L.orc_sim_set_force_field.argtypes = [vp, dp]
        L.orc_sim_get_force_field.argtypes = [vp, dp]
        L.orc_sim_set_porosity.argtypes = [vp, dp]
        L.orc_sim_get_porosity.argtypes = [vp, dp]
        L.orc_sim_set_populations.argtypes = [vp, C.c_uint64, dp]
        L.orc_sim_get_populations.argtypes = [vp, C.c_uint64, dp]
        L.orc_sim_init_equilibrium.argtypes = [vp, C.c_uint64, C.c_double, dp]
        L.orc_sim_get_macros.argtypes = [vp, dp, dp]
        L.orc_sim_set_macros.argtypes = [vp, dp, dp]
        L.orc_sim_step.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_int]
        L.orc_particles_set.argtypes = [vp, C.POINTER(abi.Particle), C.c_uint32,
                                        C.POINTER(abi.ParticleParams)]
        L.orc_particles_get.argtypes = [vp, C.POINTER(abi.Particle), C.c_uint32]
        L.orc_psm_couple.argtypes = [vp, C.c_uint64, dp, dp, dp]
        L.orc_dem_contact.argtypes = [vp]
        L.orc_particle_integrate.argtypes = [vp]
        L.orc_cell_list_overflow.argtypes = [vp]
        L.orc_cell_hlbm.argtypes = [C.c_int, C.c_int, C.POINTER(abi.OpDesc), dp, C.c_double, dp, C.c_double,
                                    dp, dp]
        L.orc_sim_get_hlbm.argtypes = [vp, dp, dp]
        L.orc_sim_set_sum64.argtypes = [vp, C.c_int]
        L.orc_sim_set_series_sin.argtypes = [vp, C.c_int]
        L.orc_series_sin.argtypes = [C.c_double]
        L.orc_series_sin.restype = C.c_double
        L.orc_cell_list_overflow.restype = C.c_uint64
        L.orc_max_threads.restype = C.c_int
        L.orc_narrow.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, C.c_double, dp, dp, dp]
        L.orc_narrow.restype = C.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lattice(lat):
    D, Q = C.c_int(), C.c_int()
    c = (C.c_int * 81)()
    w = (C.c_double * 27)()
    opp = (C.c_int * 27)()
    lib().orc_lattice(lat, C.byref(D), C.byref(Q), c, w, opp)
    q = Q.value
    return (D.value, q, np.array(c[: q * 3]).reshape(q, 3), np.array(w[:q]), np.array(opp[:q]))


def equilibrium(lat, st, rho, u):
    uu = np.zeros(3)
    uu[: len(u)] = u
    eq = np.zeros(27)
    lib().orc_equilibrium(lat, st, rho, _dp(uu), _dp(eq))
    return eq[: abi.LATTICE_DQ[lat][1]]


def rho_u(lat, st, f):
    ff = np.zeros(27)
    ff[: len(f)] = f
    r = np.zeros(1)
    u = np.zeros(3)
    lib().orc_rho_u(lat, st, _dp(ff), _dp(r), _dp(u))
    return r[0], u[: abi.LATTICE_DQ[lat][0]]


def cell_op(lat, st, op, f, t=0, cell_force=None, bc_u=None, porosity=1.0):
    """Apply one flag's (boundary, collision) to one cell's populations."""
    D, Q = abi.LATTICE_DQ[lat]
    ff = np.zeros(27)
    ff[:Q] = f
    cf = np.zeros(3)
    bu = np.zeros(3)
    if cell_force is not None:
        cf[: len(cell_force)] = cell_force
    if bc_u is not None:
        bu[: len(bc_u)] = bc_u
    r = np.zeros(1)
    u = np.zeros(3)
    fo = np.zeros(3)
    lib().orc_cell_op(lat, st, C.byref(op), t, _dp(ff), _dp(cf), _dp(bu), porosity,
                      _dp(r), _dp(u), _dp(fo))
    return ff[:Q], r[0], u[:D], fo[:D]


def cell_hlbm(lat, st, op, f, porosity, nsum, dsum):
    """The HLBM collision (hlbm.hpp:62-105) on one cell with given porosity
    and particle sums N (D values) and D."""
    D, Q = abi.LATTICE_DQ[lat]
    ff = np.zeros(27)
    ff[:Q] = f
    nn = np.zeros(3)
    nn[: len(nsum)] = nsum
    r = np.zeros(1)
    u = np.zeros(3)
    lib().orc_cell_hlbm(lat, st, C.byref(op), _dp(ff), porosity, _dp(nn), dsum, _dp(r), _dp(u))
    return ff[:Q], r[0], u[:D]


def porosity(D, st, rel, r, eps):
    rr = np.zeros(3)
    rr[: len(rel)] = rel
    out = np.zeros(1)
    lib().orc_porosity(D, st, _dp(rr), r, eps, _dp(out))
    return out[0]


def aabb(D, st, pos, r, extent):
    pp = np.zeros(3)
    pp[: len(pos)] = pos
    ext = (C.c_uint64 * 3)(*list(extent) + [1] * (3 - len(extent)))
    a = (C.c_uint64 * 3)()
    b = (C.c_uint64 * 3)()
    lib().orc_aabb(D, st, _dp(pp), r, ext, a, b)
    return list(a[:D]), list(b[:D])


def spheroid_mass(D, st, r, rho):
    m, i = C.c_double(), C.c_double()
    lib().orc_spheroid_mass(D, st, r, rho, C.byref(m), C.byref(i))
    return m.value, i.value


def narrow(D, st, pos_lo, r_lo, pos_hi, r_hi):
    a = np.zeros(3)
    b = np.zeros(3)
    a[:len(pos_lo)] = pos_lo
    b[:len(pos_hi)] = pos_hi
    depth = np.zeros(1)
    n = np.zeros(3)
    pt = np.zeros(3)
    hit = lib().orc_narrow(D, st, _dp(a), r_lo, _dp(b), r_hi, _dp(depth), _dp(n), _dp(pt))
    return bool(hit), depth[0], n[:D], pt[:D]


def series_sin(x):
    """The engine's series sine of the porosity ramp (oracle restatement)."""
    return lib().orc_series_sin(float(x))


def max_threads():
    return lib().orc_max_threads()


class Sim:
    """The reference's two-buffer simulation on the CPU (oracle)."""

    def __init__(self, lat, storage, extent):
        self.lat, self.storage = lat, storage
        self.D, self.Q = abi.LATTICE_DQ[lat]
        ext = list(extent) + [1] * (3 - len(extent))
        self.extent = tuple(ext[: self.D])
        self.N = int(np.prod(self.extent))
        self._h = lib().orc_sim_create(lat, storage, (C.c_uint64 * 3)(*ext))
        if not self._h:
            raise ValueError("oracle does not support this lattice/storage")
        self.t = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_sim_destroy(h)
            self._h = None

    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        assert f.size == self.N
        lib().orc_sim_set_flags(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8)))

    def set_periodic(self, mask):
        lib().orc_sim_set_periodic(self._h, mask)

    def set_operator(self, flag, op):
        lib().orc_sim_set_operator(self._h, flag, C.byref(op))

    def set_boundary_velocity(self, u):
        a = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        lib().orc_sim_set_boundary_velocity(self._h, _dp(a))

    def set_force_field(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        lib().orc_sim_set_force_field(self._h, _dp(a))

    def get_force_field(self):
        a = np.zeros(self.D * self.N)
        lib().orc_sim_get_force_field(self._h, _dp(a))
        return a.reshape(self.D, self.N)

    def set_porosity(self, p):
        a = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        lib().orc_sim_set_porosity(self._h, _dp(a))

    def get_porosity(self):
        a = np.zeros(self.N)
        lib().orc_sim_get_porosity(self._h, _dp(a))
        return a

    def set_populations(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert a.size == self.Q * self.N
        lib().orc_sim_set_populations(self._h, self.t, _dp(a))

    def get_populations(self):
        a = np.zeros(self.Q * self.N)
        lib().orc_sim_get_populations(self._h, self.t, _dp(a))
        return a.reshape(self.Q, self.N)

    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        uu = np.zeros(3)
        uu[: len(u)] = u
        lib().orc_sim_init_equilibrium(self._h, self.t, rho, _dp(uu))

    def get_macros(self):
        r = np.zeros(self.N)
        u = np.zeros(self.D * self.N)
        lib().orc_sim_get_macros(self._h, _dp(r), _dp(u))
        return r, u.reshape(self.D, self.N)

    def set_macros(self, rho, u):
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        lib().orc_sim_set_macros(self._h, _dp(r), _dp(uu))

    def step(self, n=1, threads=0):
        lib().orc_sim_step(self._h, self.t, n, threads)
        self.t += n

    # particles
    def particles_set(self, parts, params):
        arr = (abi.Particle * len(parts))(*parts)
        self.n_particles = len(parts)
        lib().orc_particles_set(self._h, arr, len(parts), C.byref(params))

    def particles_get(self):
        arr = (abi.Particle * self.n_particles)()
        lib().orc_particles_get(self._h, arr, self.n_particles)
        return list(arr)

    def psm_couple(self):
        n = self.n_particles
        f = np.zeros(3 * n)
        t = np.zeros(3 * n)
        a = np.zeros(n)
        lib().orc_psm_couple(self._h, self.t, _dp(f), _dp(t), _dp(a))
        return f.reshape(n, 3), t.reshape(n, 3), a

    def dem_contact(self):
        lib().orc_dem_contact(self._h)

    def integrate(self):
        lib().orc_particle_integrate(self._h)

    def overflow(self):
        return lib().orc_cell_list_overflow(self._h)

    def set_particle_sums_fp64(self, on=True):
        """Particle force/torque sums in FP64, rounded to T once (the GPU's
        arithmetic), instead of the reference's serial T sums (default)."""
        lib().orc_sim_set_sum64(self._h, 1 if on else 0)

    def set_series_sin(self, on=True):
        """The porosity ramp's sine from the engine's series instead of the C
        library's sin (the reference's)."""
        lib().orc_sim_set_series_sin(self._h, 1 if on else 0)

    def match_gpu_arithmetic(self):
        """Both switches the GPU comparison tests use: FP64 particle sums and
        the engine's ramp sine."""
        self.set_particle_sums_fp64(True)
        self.set_series_sin(True)

    def get_hlbm(self):
        """HLBM particle sums: N (D x N) and D (N)."""
        n = np.zeros(self.D * self.N)
        d = np.zeros(self.N)
        lib().orc_sim_get_hlbm(self._h, _dp(n), _dp(d))
        return n.reshape(self.D, self.N), d
