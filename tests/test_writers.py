"""Result writers (VTK, CSV, Forst) on host field sets, byte for byte against
the reference's own writers: the committed fixtures in tests/golden/writers
(made by tests/golden/make_writer_golden.py from write_vtk.hpp, write_csv.hpp
and abstract/forst.hpp built from the reference sources), the reference-built
writers themselves where oracle/_ref is present, and the known answers of the
reference's Forst tests (modules/core/tests/forst.cpp:205-282).  No GPU: the
*_fields entry points format host arrays."""
import ctypes as C
import os

import numpy as np
import pytest

from lbm_mirror_amd import LbmError, abi, output

GOLD = os.path.join(os.path.dirname(__file__), "golden", "writers")
SHAPES = {0: (abi.D2Q9, (5, 3)), 1: (abi.D3Q19, (3, 4, 2))}
FULL = "velocity,density,force,porosity,info,dfs"
CASES = sorted(os.listdir(GOLD))


def _fields(case):
    shape, st = int(case[1]), abi.F64 if "_f64_" in case else abi.F32
    kind, ext = SHAPES[shape]
    x = np.load(os.path.join(GOLD, case, "inputs.npz"))
    return output.Fields(kind, st, ext, **{k: x[k] for k in x.files}), ("full" in case)


def _tree(d):
    out = {}
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith((".vtk", ".csv")):
                out[os.path.relpath(os.path.join(root, f), d)] = open(os.path.join(root, f), "rb").read()
    return out


@pytest.mark.parametrize("case", CASES)
def test_vtk_and_csv_match_reference_fixtures(case, tmp_path):
    f, full = _fields(case)
    members = FULL if full else output.MACROS
    output.write_vtk_fields(f, str(tmp_path), "m", 7, members)
    output.write_csv_fields(f, str(tmp_path), "m", 7, members)
    got, ref = _tree(str(tmp_path)), _tree(os.path.join(GOLD, case))
    assert sorted(got) == sorted(ref)
    for name in ref:
        assert got[name] == ref[name], name


@pytest.mark.parametrize("case", CASES)
def test_forst_matches_reference_fixture_and_round_trips(case, tmp_path):
    f, full = _fields(case)
    members = FULL if full else output.MACROS
    p = str(tmp_path / "m.forst")
    output.write_forst_fields(f, p, members)
    assert open(p, "rb").read() == open(os.path.join(GOLD, case, "m.forst"), "rb").read()
    back = output.Fields(f.kind, f.storage, f.extent).allocate(members)
    output.read_forst_fields(back, p, members)
    with np.errstate(over="ignore"):
        narrow32 = {k: f.arrays[k].astype(np.float32).astype(np.float64) for k in f.arrays if k != "info"}
    narrow = (lambda k: narrow32[k]) if f.storage == abi.F32 else (lambda k: f.arrays[k])
    for k in [m for m in members.split(",")]:
        want = f.arrays[k] if k == "info" else narrow(k)
        assert np.array_equal(back.arrays[k], want, equal_nan=k != "info"), k


def test_against_reference_build(tmp_path):
    """Fresh random values through the reference-built writers (skipped where
    oracle/_ref is absent)."""
    lib = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "libref_harness.so")
    if not os.path.exists(lib):
        pytest.skip("oracle/_ref not built")
    L = C.CDLL(lib)
    L.ref_write_fields.restype = C.c_int64
    dp = C.POINTER(C.c_double)
    rng = np.random.default_rng(31)
    for shape, (kind, ext) in SHAPES.items():
        D, Q = abi.LATTICE_DQ[kind]
        N = int(np.prod(ext))
        for st in (abi.F64, abi.F32):
            x = {"velocity": rng.standard_normal(D * N) * 10.0 ** rng.integers(-12, 12, D * N),
                 "density": 1 + 1e-4 * rng.standard_normal(N), "force": rng.standard_normal(D * N) * 1e-6,
                 "porosity": rng.uniform(0, 1, N), "info": rng.integers(1, 6, N).astype(np.uint8),
                 "dfs": rng.uniform(0, 0.5, Q * N)}
            rd, md = tmp_path / f"r{shape}{st}", tmp_path / f"m{shape}{st}"
            buf = np.zeros(1 << 20, np.uint8)
            ptrs = [x[k].ctypes.data_as(dp) for k in ("velocity", "density", "force", "porosity")]
            n = 0
            for what in (0, 1, 2):
                n = L.ref_write_fields(shape, st, 1, *ptrs, x["info"].ctypes.data_as(C.POINTER(C.c_uint8)),
                                       x["dfs"].ctypes.data_as(dp), str(rd).encode(), b"q", C.c_uint64(123), what,
                                       buf.ctypes.data_as(C.POINTER(C.c_uint8)), C.c_uint64(buf.size), None)
                assert n >= 0
            f = output.Fields(kind, st, ext, **x)
            output.write_vtk_fields(f, str(md), "q", 123, FULL)
            output.write_csv_fields(f, str(md), "q", 123, FULL)
            output.write_forst_fields(f, str(md / "q.forst"), FULL)
            assert _tree(str(md)) == _tree(str(rd))
            assert open(md / "q.forst", "rb").read() == bytes(buf[:n])


def test_forst_struct_of_chunks_is_its_raw_bytes(tmp_path):
    """forst.cpp:205-262: velocity (x, y) and density x*dim_y + y on 16x8 float
    chunks encode to 16*8*3*4 bytes, velocity then density, no framing."""
    nx, ny = 16, 8
    xs, ys = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    xs, ys = xs.reshape(-1).astype(float), ys.reshape(-1).astype(float)
    f = output.Fields(abi.D2Q9, abi.F32, (nx, ny), velocity=np.stack([xs, ys]), density=xs * ny + ys)
    p = str(tmp_path / "m.forst")
    output.write_forst_fields(f, p, "velocity,density")
    b = open(p, "rb").read()
    assert len(b) == nx * ny * 3 * 4
    vel = np.frombuffer(b[:nx * ny * 8], np.float32).reshape(ny, nx, 2)
    assert vel[5, 3, 1] == 5.0 and vel[5, 3, 0] == 3.0
    rho = np.frombuffer(b[nx * ny * 8:], np.float32).reshape(ny, nx)
    assert rho[5, 3] == 3 * ny + 5


def test_forst_ghost_ring_is_part_of_the_encoding(tmp_path):
    """forst.cpp:264-282: a ghost-1 FixedArray<Float32,9> chunk on 8x4 encodes
    (8+2)*(4+2)*9*4 bytes; interior cell (0,0) sits at stored (1,1)."""
    dfs = np.zeros((9, 32))
    dfs[3, 0] = 7.0
    f = output.Fields(abi.D2Q9, abi.F32, (8, 4), dfs=dfs)
    p = str(tmp_path / "d.forst")
    output.write_forst_fields(f, p, "dfs")
    b = np.fromfile(p, np.float32)
    assert b.size == 10 * 6 * 9
    assert b[(1 + 1 * 10) * 9 + 3] == 7.0 and np.count_nonzero(b) == 1


def test_forst_truncated_and_trailing(tmp_path):
    f = output.Fields(abi.D2Q9, abi.F64, (4, 3), density=np.arange(12.0), velocity=np.zeros((2, 12)))
    p = str(tmp_path / "m.forst")
    output.write_forst_fields(f, p, "density,velocity")
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-4])
    back = output.Fields(f.kind, f.storage, f.extent).allocate("density,velocity")
    with pytest.raises(LbmError, match="Truncated"):
        output.read_forst_fields(back, p, "density,velocity")
    # bytes past the struct are ignored, as the reference's decode does
    open(p, "wb").write(raw + b"\x01\x02\x03")
    output.read_forst_fields(back, p, "density,velocity")
    assert np.array_equal(back.get("density"), np.arange(12.0))


def test_member_errors(tmp_path):
    f = output.Fields(abi.D2Q9, abi.F32, (4, 3), density=np.ones(12))
    with pytest.raises(LbmError, match="unknown member"):
        output.write_vtk_fields(f, str(tmp_path), "m", 0, "density,pressure")
    with pytest.raises(LbmError, match="twice"):
        output.write_csv_fields(f, str(tmp_path), "m", 0, "density,density")
    with pytest.raises(LbmError, match="no buffer"):
        output.write_vtk_fields(f, str(tmp_path), "m", 0, "density,velocity")
    g = output.Fields(abi.D2Q9, abi.F16S, (4, 3), dfs=np.zeros((9, 12)))
    with pytest.raises(LbmError, match="16-bit"):
        output.write_forst_fields(g, str(tmp_path / "x"), "dfs")


def test_large_field_is_split_across_workers(tmp_path):
    """Enough cells for every worker thread: the pieces are joined in cell
    order (checked against a one-line-per-cell rendering in numpy)."""
    n = (200, 90)
    N = n[0] * n[1]
    rng = np.random.default_rng(5)
    rho = 1 + rng.standard_normal(N) * 1e-3
    f = output.Fields(abi.D2Q9, abi.F64, n, density=rho)
    output.write_csv_fields(f, str(tmp_path), "big", 3, "density")
    lines = open(tmp_path / "csv" / "big_density_3.csv").read().splitlines()
    assert len(lines) == N
    assert [float(v) for v in lines] == [float("%g" % v) for v in rho]
    assert lines == ["%g" % v for v in rho]
