"""Writes the result-writer fixtures under tests/golden/writers/ with the
reference's own writers (write_vtk_file, write_csv_file and the Forst codec,
built from the reference sources into oracle/_ref/libref_harness.so by
oracle/ref/Makefile).

For each (shape, storage, struct) case: inputs.npz holds the field values
(SoA float64, info uint8), and the reference's output files sit beside it.
Inputs mix seeded random values with the special cases of %g formatting:
signed zero, subnormals, exact ties at the sixth digit, values that need the
exponent form, +-inf, NaN and doubles beyond float range.

Run from the repository root: python tests/golden/make_writer_golden.py
"""
import ctypes as C
import os
import shutil
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "writers")
SHAPES = {0: ((5, 3), 2, 9), 1: ((3, 4, 2), 3, 19)}
SPECIAL = [0.0, -0.0, 5e-324, -1.2e-310, 1e-45, 3e-39, 0.5, 123456.5, 1234565.0, 999999.5, 1e6, 1e-5,
           1.0000005, -2.5e-7, 1e300, -1e300, np.inf, -np.inf, np.nan, -np.nan, 3.4028235e38, 1e16, 0.1, 1.0 / 3]


def inputs(shape, seed):
    ext, D, Q = SHAPES[shape]
    N = int(np.prod(ext))
    rng = np.random.default_rng(seed)

    def fld(k, scale):
        a = rng.standard_normal(k * N) * scale * 10.0 ** rng.integers(-3, 4, k * N)
        n = min(len(SPECIAL), a.size)
        a[:n] = np.array(SPECIAL[:n])[rng.permutation(n)] if seed % 2 else np.array(SPECIAL[:n])
        return a

    return {"velocity": fld(D, 0.05), "density": 1.0 + fld(1, 1e-3), "force": fld(D, 1e-5),
            "porosity": rng.uniform(0, 1, N), "info": rng.integers(0, 256, N).astype(np.uint8),
            "dfs": fld(Q, 0.1)}


def main():
    lib = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    if not os.path.exists(lib):
        sys.exit("build oracle/_ref first (make -C oracle/ref)")
    L = C.CDLL(lib)
    L.ref_write_fields.restype = C.c_int64
    dp = C.POINTER(C.c_double)
    shutil.rmtree(OUT, ignore_errors=True)
    for shape in SHAPES:
        for st, stn in ((0, "f64"), (1, "f32")):
            for full, fn in ((1, "full"), (0, "macro")):
                case = f"s{shape}_{stn}_{fn}"
                d = os.path.join(OUT, case)
                os.makedirs(d)
                x = inputs(shape, 100 * shape + 10 * st + full)
                np.savez(os.path.join(d, "inputs.npz"), **x)
                buf = np.zeros(1 << 20, np.uint8)
                ptrs = [x[k].ctypes.data_as(dp) for k in ("velocity", "density", "force", "porosity")]
                for what in (0, 1, 2):
                    r = L.ref_write_fields(shape, st, full, *ptrs, x["info"].ctypes.data_as(C.POINTER(C.c_uint8)),
                                           x["dfs"].ctypes.data_as(dp), d.encode(), b"m", C.c_uint64(7), what,
                                           buf.ctypes.data_as(C.POINTER(C.c_uint8)), C.c_uint64(buf.size), None)
                    if r < 0:
                        sys.exit(f"reference writer failed: {case} what={what}")
                    if what == 2:
                        buf[:r].tofile(os.path.join(d, "m.forst"))
                print(case, sorted(os.listdir(d)))


if __name__ == "__main__":
    main()
