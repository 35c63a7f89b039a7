"""Result output in the reference's formats: legacy-ASCII VTK (write_vtk.hpp),
one CSV per member (write_csv.hpp) and Forst binary snapshots of a struct of
chunks (abstract/forst.hpp, forst_chunk.hpp).

``members`` names the fields of the struct being written, in file order, with
the reference drivers' keys: velocity, density, force, porosity, info, dfs
(e.g. "velocity,density" for the MacroStruct of cavity_2d_gpu.cpp:47-50).

The functions taking a Lattice read its fields back from the GPU; the
``*_fields`` variants take host arrays (batch members, post-processing).
A Forst snapshot with "info,dfs,velocity,density" (plus force/porosity when
used) restores a lattice that continues bit for bit (``read_forst``).
"""
import ctypes as C
import os

import numpy as np

from . import abi, check, load_library

MACROS = "velocity,density"


def _s(x):
    return os.fsencode(x) if not isinstance(x, bytes) else x


def write_vtk(lattice, out_dir, file_name, t, members=MACROS):
    """write_vtk_file: <out_dir>/vtk/<file_name>_<t>.vtk."""
    check(lattice.L.lbm_write_vtk(lattice._h, _s(out_dir), _s(file_name), t, _s(members)))
    return os.path.join(out_dir, "vtk", f"{file_name}_{t}.vtk")


def write_csv(lattice, out_dir, file_name, t, members=MACROS):
    """write_csv_file: <out_dir>/csv/<file_name>_<member>_<t>.csv per member."""
    check(lattice.L.lbm_write_csv(lattice._h, _s(out_dir), _s(file_name), t, _s(members)))
    return [os.path.join(out_dir, "csv", f"{file_name}_{m.strip()}_{t}.csv") for m in members.split(",") if m.strip()]


def write_forst(lattice, path, members):
    check(lattice.L.lbm_write_forst(lattice._h, _s(path), _s(members)))


def read_forst(lattice, path, members):
    check(lattice.L.lbm_read_forst(lattice._h, _s(path), _s(members)))


class Fields:
    """Host field set (lbm_field_set).  Arrays are SoA float64 (velocity and
    force (D, N), dfs (Q, N), density and porosity (N,)) and uint8 info (N,);
    any may be None when not listed."""

    def __init__(self, lattice_kind, storage, extent, **arrays):
        self.kind, self.storage = lattice_kind, storage
        self.D, self.Q = abi.LATTICE_DQ[lattice_kind]
        ext = list(extent) + [1] * (3 - len(extent))
        self.extent = tuple(ext)
        self.N = int(np.prod(ext[: self.D]))
        self.arrays = {}
        for k in ("velocity", "density", "force", "porosity", "dfs", "info"):
            a = arrays.get(k)
            if a is not None:
                dt = np.uint8 if k == "info" else np.float64
                self.arrays[k] = np.ascontiguousarray(a, dtype=dt).reshape(-1)

    def allocate(self, members):
        """Zeroed buffers for the listed members (for read_forst_fields)."""
        size = {"velocity": self.D, "force": self.D, "dfs": self.Q, "density": 1, "porosity": 1, "info": 1}
        for m in [x.strip() for x in members.split(",") if x.strip()]:
            if m not in size:
                raise KeyError(m)
            self.arrays[m] = np.zeros(size[m] * self.N, dtype=np.uint8 if m == "info" else np.float64)
        return self

    def _struct(self):
        fs = abi.FieldSet()
        fs.lattice, fs.storage = self.kind, self.storage
        fs.extent = (C.c_uint64 * 3)(*self.extent)
        for k, a in self.arrays.items():
            t = C.c_uint8 if k == "info" else C.c_double
            setattr(fs, k, a.ctypes.data_as(C.POINTER(t)))
        return fs

    def get(self, key):
        a = self.arrays[key]
        if key in ("velocity", "force"):
            return a.reshape(self.D, self.N)
        if key == "dfs":
            return a.reshape(self.Q, self.N)
        return a


def write_vtk_fields(fields, out_dir, file_name, t, members=MACROS):
    L = load_library()
    check(L.lbm_write_vtk_fields(C.byref(fields._struct()), _s(out_dir), _s(file_name), t, _s(members)))
    return os.path.join(out_dir, "vtk", f"{file_name}_{t}.vtk")


def write_csv_fields(fields, out_dir, file_name, t, members=MACROS):
    L = load_library()
    check(L.lbm_write_csv_fields(C.byref(fields._struct()), _s(out_dir), _s(file_name), t, _s(members)))


def write_forst_fields(fields, path, members):
    L = load_library()
    check(L.lbm_write_forst_fields(C.byref(fields._struct()), _s(path), _s(members)))


def read_forst_fields(fields, path, members):
    L = load_library()
    check(L.lbm_read_forst_fields(C.byref(fields._struct()), _s(path), _s(members)))
    return fields
