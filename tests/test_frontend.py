"""The unit/config front end (converter.hpp, abstract/args.hpp, config.hpp,
environment.hpp) in both of this engine's forms — the Python module
lbm_mirror_amd/frontend.py and the C++ header
include/lbm_mirror_amd/frontend.hpp (driven by tests/frontend_check.cpp) —
against the reference's own code: the committed outcomes in
tests/golden/frontend.json (made by tests/golden/make_frontend_golden.py
from the reference build) and, where oracle/_ref is built, fresh cases
through it.  Conversions are compared bit for bit, parse and load outcomes
value for value with the reference's error category and message text."""
import json
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from lbm_mirror_amd import abi, frontend as fe

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = json.load(open(os.path.join(HERE, "golden", "frontend.json")))
SPEC = fe.ArgSpec({"name": (fe.STR, "default_name"), "time_steps": (fe.U64, 256), "sub_steps": (fe.U64, 16),
                   "axis_x": (fe.F64, 0.0), "visc": (fe.F32, 0.0), "shift": (fe.I64, 0),
                   "coupling": (fe.STR, ""), "verbose": (fe.BOOL, False)}, [fe.U64, fe.F64])
WHAT = ["meter_si_to_lbm", "meter_lbm_to_si", "second_si_to_lbm", "velocity_si_to_lbm", "acceleration_si_to_lbm",
        "kinematic_viscosity_si_to_lbm", "kinematic_viscosity_si_to_tau", "kinematic_viscosity_lbm_to_tau",
        "delta_x", "delta_t", "delta_v", "delta_a", "kinematic_viscosity_si_to_tau"]


def _same(a, b):
    a, b = float(a), float(b)
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b and math.copysign(1, a) == math.copysign(1, b)


def _py_convert(st, dx, dt, what, v):
    c = fe.Converter(dx, dt, np.float64 if st == 0 else np.float32)
    f = getattr(c, WHAT[what])
    return f() if what in (8, 9, 10, 11) else f(v)


@pytest.fixture(scope="module")
def cpp(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fe") / "frontend_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "frontend_check.cpp"), "-o", exe], check=True)

    def run(lines, whole=False):
        r = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
        if whole:  # one case whose result may hold newlines
            return r.stdout[:-1]
        return r.stdout.split("\n")[:len(lines)]
    return run


def _args_expect(case):
    if not case["ok"]:
        cat, msg = case["error"].split("|", 1)
        return ("err", cat, msg)
    nums = [float.fromhex(x) for x in case["nums"]]
    name, coupling, program = case["text"].split("\n")
    return ("ok", nums, name, coupling, program)


def _py_args(argv):
    try:
        r = SPEC.parse(argv)
    except fe.FrontendError as e:
        return ("err", e.category, e.message)
    a = r["args"]
    nums = [a["time_steps"], a["sub_steps"], a["axis_x"], a["visc"], a["shift"], 1.0 if a["verbose"] else 0.0,
            r["positionals"][0], r["positionals"][1]]
    return ("ok", [float(x) for x in nums], a["name"], a["coupling"], r["program"])


def _cpp_args_line(out):
    f = out.split("\x1f")
    if f[0] == "err":
        return ("err", f[1], f[2])
    return ("ok", [float.fromhex(x) for x in f[1:9]], f[9], f[10], f[11])


def _match(got, want):
    if got[0] != want[0]:
        return False
    if got[0] == "err":
        return got[1:] == want[1:]
    return all(_same(a, b) for a, b in zip(got[1], want[1])) and got[2:] == want[2:]


def test_converter_matches_reference(cpp):
    lines = [f"conv\t{st}\t{dx!r}\t{dt!r}\t{what}\t{v!r}" for st, dx, dt, what, v, _ in GOLD["convert"]]
    cout = cpp(lines)
    for (st, dx, dt, what, v, want), c in zip(GOLD["convert"], cout):
        w = float.fromhex(want)
        assert _same(_py_convert(st, dx, dt, what, v), w), (st, dx, dt, WHAT[what], v)
        assert _same(float.fromhex(c), w), (st, dx, dt, WHAT[what], v, c)


def test_converter_known_answers():
    # tests/converter.cpp: delta_x = delta_t = 0.1, 1 m -> 10 cells
    assert fe.Converter(0.1, 0.1).meter_si_to_lbm(1.0) == 10.0
    # the drivers' tau from nu = 0.1 with delta_x = delta_t = 1 (poiseulle_particle_3d_hlbm_gpu/sim.cpp:295)
    assert fe.Converter(1.0, 1.0, np.float32).kinematic_viscosity_si_to_tau(0.1) == float(np.float32(0.8))
    # tau = 3 nu dt / dx^2 + 1/2, the standard BGK relation, to rounding
    c = fe.Converter(1e-3, 1e-4)
    assert abs(c.kinematic_viscosity_si_to_tau(1e-6) - (3 * 1e-6 * 1e-4 / 1e-6 + 0.5)) < 1e-15
    m = fe.lbm_meta(fe.Converter(1.0, 1.0), 0.1, 0.01, 9.6)
    assert abs(m["Re"] - 0.01 * 9.6 / 0.1) < 1e-15 and m["tau"] == 0.8


@pytest.mark.parametrize("i", range(len(GOLD["args"])))
def test_parse_args_matches_reference(cpp, i):
    case = GOLD["args"][i]
    want = _args_expect(case)
    assert _match(_py_args(case["argv"]), want), (case["argv"], _py_args(case["argv"]), want)
    got = _cpp_args_line(cpp(["args\t" + "\x1f".join(case["argv"])])[0])
    assert _match(got, want), (case["argv"], got, want)


def test_parse_args_known_answers():
    """tests/args.cpp:33-58: interleaved named and positional arguments."""
    spec = fe.ArgSpec({"dx": (fe.F64, 0.0), "dt": (fe.F64, 0.0)}, [fe.U64, fe.U64])
    r = spec.parse(["prog", "--dx", "0.25", "2", "--dt", "0.5", "9"])
    assert r["program"] == "prog" and r["args"] == {"dx": 0.25, "dt": 0.5} and r["positionals"] == [2, 9]
    with pytest.raises(fe.FrontendError, match="nope"):
        fe.ArgSpec({"dx": (fe.F64, 0.0)}).parse(["prog", "--nope", "1"])


def test_load_config_matches_reference(cpp, tmp_path):
    lines, cases = [], []
    for k, case in enumerate(GOLD["config"]):
        if case["doc"] is None:
            path = "/nonexistent/lbm_config.json"
        else:
            path = str(tmp_path / f"c{k}.json")
            with open(path, "w") as fh:
                fh.write(case["doc"])
        lines.append(f"cfg\t{case['st']}\t{path}")
        cases.append((case, path))
    cout = cpp(lines)
    for (case, path), c in zip(cases, cout):
        dtype = np.float64 if case["st"] == 0 else np.float32
        if case["ok"]:
            want = [float.fromhex(x) for x in case["values"]]
            got = fe.load_lbm_config(path, dtype)
            assert all(_same(got[k], w) for k, w in zip(fe.CONFIG_KEYS, want)), (case["doc"], got, want)
            f = c.split("\x1f")
            assert f[0] == "ok" and all(_same(float.fromhex(x), w) for x, w in zip(f[1:], want)), (case["doc"], c)
        else:
            cat, msg = case["error"].split("|", 1)
            with pytest.raises(fe.FrontendError) as e:
                fe.load_lbm_config(path, dtype)
            assert (e.value.category, e.value.message) == (cat, msg), case["doc"]
            assert c.split("\x1f") == ["err", cat, msg], (case["doc"], c)


def test_environment_matches_reference(cpp, monkeypatch):
    for case in GOLD["env"]:
        home = tempfile.mkdtemp()
        try:
            os.makedirs(os.path.join(home, ".lbm"))
            if case["content"] is not None:
                with open(os.path.join(home, ".lbm", "data_dir_location.txt"), "w") as fh:
                    fh.write(case["content"])
            monkeypatch.setenv("HOME", home)
            want = [p.replace("{HOME}", home) for p in case["paths"]]
            e = fe.init_lbm_env(case["name"])
            assert [e["lbm_dir"], e["data_dir"], e["name_dir"], fe.output_directory()] == want, case
            out = cpp(["env\t" + case["name"]], whole=True)
            assert out.split("\x1f") == ["ok"] + want, (case, out)
        finally:
            shutil.rmtree(home, ignore_errors=True)
    monkeypatch.delenv("HOME")
    with pytest.raises(fe.FrontendError, match="home dir"):
        fe.init_lbm_env("x")


def test_against_reference_build_fresh_cases(tmp_path):
    """Random conversions and argv through the live reference build (skipped
    where oracle/_ref is absent)."""
    import ctypes as C
    lib = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    if not os.path.exists(lib):
        pytest.skip("oracle/_ref not built")
    L = C.CDLL(lib)
    L.ref_convert.restype = C.c_double
    L.ref_convert.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double]
    rng = np.random.default_rng(77)
    for _ in range(300):
        st = int(rng.integers(0, 2))
        dx, dt = float(10.0 ** rng.uniform(-6, 1)), float(10.0 ** rng.uniform(-7, 1))
        v = float(rng.standard_normal() * 10.0 ** rng.integers(-9, 5))
        what = int(rng.integers(0, 13))
        assert _same(_py_convert(st, dx, dt, what, v), L.ref_convert(st, dx, dt, what, v))
    words = ["--visc", "--axis_x", "--time_steps", "--shift", "--verbose", "--name", "3", "-1", "0.5", "x", "1e5",
             "", "true", "--", "7", "2.5e-3", "-0", "0x10"]
    for _ in range(300):
        argv = ["p"] + [str(rng.choice(words)) for _ in range(int(rng.integers(0, 7)))]
        arr = (C.c_char_p * len(argv))(*[w.encode() for w in argv])
        nums = (C.c_double * 8)()
        txt = C.create_string_buffer(1024)
        err = C.create_string_buffer(1024)
        r = L.ref_parse_args(len(argv), C.cast(arr, C.POINTER(C.c_char_p)), nums, txt, 1024, err, 1024)
        want = _args_expect({"ok": r == 0, "nums": [float(x).hex() for x in nums], "text": txt.value.decode(),
                             "error": err.value.decode()})
        assert _match(_py_args(argv), want), (argv, _py_args(argv), want)


@pytest.mark.parametrize("example,bad", [
    ("vawt_3d", (["--time_steps", "-3"], ["--axis_q", "1"], ["--stable_faces", "maybe"])),
    ("moving_particles_2d_psm", (["--sub_steps", "-1"], ["--radius", "2"], ["--vtk_every", ""])),
])
def test_examples_reject_bad_arguments(tmp_path, example, bad):
    """The example drivers' command lines go through the argument parser
    before any device call: a negative count, an unknown name or an empty
    number ends the program with status 1 and a message, without touching a
    GPU or creating the output directory."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lbm_mirror_amd", "_build", example)
    if not os.path.exists(exe):
        pytest.skip(f"{example} example not built")
    env = dict(os.environ, HOME=str(tmp_path))
    for argv in bad:
        r = subprocess.run([exe, *argv], capture_output=True, text=True, timeout=60, env=env)
        assert r.returncode == 1 and r.stderr, argv
    assert not os.path.exists(tmp_path / ".lbm")
