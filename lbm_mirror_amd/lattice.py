"""Lattice: one GPU lattice behind the C ABI (include/lbm_mirror_amd/lbm.h).

The reference drives its components from a per-cell `switch(info)` inside a
SYCL kernel (examples/poiseulle_2d_gpu/poiseulle_2d_gpu.cpp:193-222).  Here
the same mapping flag -> (boundary component, collision component) is
registered once with ``set_flag`` and ``step`` advances whole steps on the GPU.
"""
import ctypes as C

import numpy as np

from . import abi, check, load_library


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Lattice:
    def __init__(self, lattice, storage, extent, device=0, _handle=None):
        self.L = load_library()
        self.lattice, self.storage = lattice, storage
        self.D, self.Q = abi.LATTICE_DQ[lattice]
        ext = list(extent) + [1] * (3 - len(extent))
        self.extent = tuple(ext[: self.D])
        self.N = int(np.prod(self.extent))
        self._owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            check(self.L.lbm_ctx_create(C.byref(h), lattice, storage, (C.c_uint64 * 3)(*ext), device))
            self._h = h
        else:
            self._h = _handle
        self.t = 0
        self.n_particles = 0

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                self.L.lbm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up -----------------------------------------------------------
    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if f.size != self.N:
            raise ValueError(f"flags: expected {self.N} cells, got {f.size}")
        check(self.L.lbm_set_flags(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8))))

    def get_flags(self):
        f = np.zeros(self.N, dtype=np.uint8)
        check(self.L.lbm_get_flags(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8))))
        return f

    def set_periodic(self, mask):
        check(self.L.lbm_set_periodic(self._h, mask))

    def set_operator(self, flag, op):
        check(self.L.lbm_set_operator(self._h, flag, C.byref(op)))

    def set_flag(self, flag, boundary=None, collision=None):
        """Register what the drivers' switch does for ``flag``: an optional
        boundary component, then an optional collision component."""
        from .components import combine
        self.set_operator(flag, combine(boundary, collision))

    def set_boundary_velocity(self, u):
        a = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if a.size != self.D * self.N:
            raise ValueError("boundary velocity: expected D*N values")
        check(self.L.lbm_set_boundary_velocity(self._h, _dp(a)))

    def set_force_field(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        check(self.L.lbm_set_force_field(self._h, _dp(a)))

    def get_force_field(self):
        a = np.zeros(self.D * self.N)
        check(self.L.lbm_get_force_field(self._h, _dp(a)))
        return a.reshape(self.D, self.N)

    def set_porosity(self, p):
        a = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        check(self.L.lbm_set_porosity(self._h, _dp(a)))

    def get_porosity(self):
        a = np.zeros(self.N)
        check(self.L.lbm_get_porosity(self._h, _dp(a)))
        return a

    def get_hlbm_sums(self):
        """HLBM particle sums N (D x N) and D (N) after the last step."""
        n = np.zeros(self.D * self.N)
        d = np.zeros(self.N)
        check(self.L.lbm_get_hlbm_sums(self._h, _dp(n), _dp(d)))
        return n.reshape(self.D, self.N), d

    def population_count(self, layout=abi.LAYOUT_SOA):
        n = C.c_uint64()
        check(self.L.lbm_population_count(self._h, layout, C.byref(n)))
        return n.value

    def set_populations(self, f, layout=abi.LAYOUT_SOA):
        """SoA f[i, cell] (default) or the reference's chunk layout with its
        ghost ring, flat (lbm_set_populations_layout)."""
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        if a.size != self.population_count(layout):
            raise ValueError(f"populations: expected {self.population_count(layout)} values")
        check(self.L.lbm_set_populations_layout(self._h, _dp(a), layout))

    def get_populations(self, layout=abi.LAYOUT_SOA):
        """(Q, N) SoA by default; the chunk layout as a flat array of
        (ghost-widened cells) x Q doubles."""
        a = np.zeros(self.population_count(layout))
        check(self.L.lbm_get_populations_layout(self._h, _dp(a), layout))
        return a.reshape(self.Q, self.N) if layout == abi.LAYOUT_SOA else a

    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        uu = np.zeros(3)
        uu[: len(u)] = u
        check(self.L.lbm_init_equilibrium(self._h, rho, _dp(uu)))

    def get_macros(self):
        r = np.zeros(self.N)
        u = np.zeros(self.D * self.N)
        check(self.L.lbm_get_macros(self._h, _dp(r), _dp(u)))
        return r, u.reshape(self.D, self.N)

    def set_macros(self, rho, u):
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        check(self.L.lbm_set_macros(self._h, _dp(r), _dp(uu)))

    # ---- stepping ---------------------------------------------------------
    def step(self, n=1, t0=None):
        t = self.t if t0 is None else t0
        check(self.L.lbm_step(self._h, t, n))
        self.t = t + n

    def sync(self):
        check(self.L.lbm_sync(self._h))

    def set_residual(self, enable=True):
        check(self.L.lbm_set_residual(self._h, 1 if enable else 0))

    def residual(self):
        """(r, sum |u_t - u_{t-1}|^2, sum |u_t|^2) in FP64 (lbm_residual)."""
        out = np.zeros(3)
        check(self.L.lbm_residual(self._h, _dp(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def set_step_graph(self, steps):
        """Steps per hipGraph replay inside step() (0: launch every step)."""
        check(self.L.lbm_set_step_graph(self._h, steps))

    def last_step_ms(self):
        m = C.c_double()
        check(self.L.lbm_last_step_ms(self._h, C.byref(m)))
        return m.value

    def autotune(self, rounds=8):
        """Time the fused step's launch shapes on this lattice and keep the
        fastest (lbm_autotune); returns the chosen shape and its times."""
        t = abi.StepTuning()
        check(self.L.lbm_autotune(self._h, rounds, C.byref(t)))
        return {"block_x": t.block_x, "form": [t.form[0], t.form[1]], "us": [t.us[0], t.us[1]]}

    def set_step_tuning(self, block_x, form=(0, 0)):
        t = abi.StepTuning()
        t.block_x = block_x
        t.form[0], t.form[1] = form
        check(self.L.lbm_set_step_tuning(self._h, C.byref(t)))

    def get_step_tuning(self):
        t = abi.StepTuning()
        check(self.L.lbm_get_step_tuning(self._h, C.byref(t)))
        return {"block_x": t.block_x, "form": [t.form[0], t.form[1]]}

    # ---- particles --------------------------------------------------------
    def particles_set(self, parts, params):
        arr = (abi.Particle * max(1, len(parts)))(*parts)
        check(self.L.lbm_particles_set(self._h, arr, len(parts), C.byref(params)))
        self.n_particles = len(parts)

    def particles_get(self):
        arr = (abi.Particle * max(1, self.n_particles))()
        check(self.L.lbm_particles_get(self._h, arr, self.n_particles))
        return list(arr)[: self.n_particles]

    def psm_couple(self):
        check(self.L.lbm_psm_couple(self._h, self.t))
        n = self.n_particles
        f = np.zeros(3 * n)
        t = np.zeros(3 * n)
        check(self.L.lbm_particle_reduce(self._h, _dp(f), _dp(t)))
        return f.reshape(n, 3), t.reshape(n, 3)

    def dem_contact(self):
        check(self.L.lbm_dem_contact(self._h))

    def integrate(self):
        check(self.L.lbm_particle_integrate(self._h))

    def overflow(self):
        o = C.c_uint64()
        check(self.L.lbm_cell_list_overflow(self._h, C.byref(o)))
        return o.value


def spheroid_mass(dims, storage, radius, density):
    L = load_library()
    m, i = C.c_double(), C.c_double()
    check(L.lbm_spheroid_mass(dims, storage, radius, density, C.byref(m), C.byref(i)))
    return m.value, i.value


class Batch:
    """Many independent lattices of uneven size stepped by one launch per
    step (lbm_batch_*).  ``members[i]`` is a Lattice view of member i for
    set-up and read-back."""

    def __init__(self, lattice, storage, extents, device=0):
        self.L = load_library()
        self.lattice, self.storage = lattice, storage
        ext = []
        for e in extents:
            e = list(e) + [1] * (3 - len(e))
            ext.extend(e[:3])
        arr = (C.c_uint64 * len(ext))(*ext)
        h = C.c_void_p()
        check(self.L.lbm_batch_create(C.byref(h), lattice, storage, len(extents), arr, device))
        self._h = h
        self.members = []
        for i, e in enumerate(extents):
            mh = self.L.lbm_batch_member(self._h, i)
            self.members.append(Lattice(lattice, storage, e, _handle=C.c_void_p(mh)))
        self.t = 0

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            self.L.lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n, t0=None):
        t = self.t if t0 is None else t0
        check(self.L.lbm_batch_step(self._h, t, n))
        self.t = t + n

    def run(self, tol, check_every=100, max_steps=100000, t0=None):
        t = self.t if t0 is None else t0
        M = len(self.members)
        steps = (C.c_uint64 * M)()
        res = (C.c_double * M)()
        check(self.L.lbm_batch_run(self._h, t, tol, check_every, max_steps, steps, res))
        return np.array(steps[:], dtype=np.uint64), np.array(res[:])

    def autotune(self):
        t = abi.BatchTuning()
        check(self.L.lbm_batch_autotune(self._h, C.byref(t)))
        return {"block": t.block, "rows": t.rows, "grid": t.grid, "ms_per_step": t.ms_per_step}

    def set_tuning(self, block, rows, grid):
        t = abi.BatchTuning()
        t.block, t.rows, t.grid = block, rows, grid
        check(self.L.lbm_batch_set_tuning(self._h, C.byref(t)))

    def stats(self):
        ms = C.c_double()
        up = C.c_uint64()
        check(self.L.lbm_batch_stats(self._h, C.byref(ms), C.byref(up)))
        return ms.value, up.value
