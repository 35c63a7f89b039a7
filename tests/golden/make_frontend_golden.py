"""Writes tests/golden/frontend.json: the reference's own converter,
parse_args, load_lbm_config and init_lbm_env/output_directory (built from
converter.hpp, abstract/args.hpp, config.hpp and environment.hpp into
oracle/_ref/libref_harness.so by oracle/ref/Makefile) on batteries of inputs.

The argument schema is the harness's RefArgs: named name (string, default
"default_name"), time_steps (u64, 256), sub_steps (u64, 16), axis_x (f64),
visc (f32), shift (i64), coupling (string), verbose (bool); positionals
(u64, f64).  Environment cases run with HOME set to a scratch directory,
recorded as "{HOME}".

Run from the repository root: python tests/golden/make_frontend_golden.py
"""
import ctypes as C
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "frontend.json")

ARGS = [
    ["p", "3", "2.5"], ["p", "--nope", "1", "3", "2"], ["p", "--time_steps"], ["p", "--time_steps", "1.5", "1", "2"],
    ["p", "--visc", "abc", "1", "2"], ["p", "1"], ["p", "1", "2", "3"], ["p", "--verbose", "maybe", "1", "2"],
    ["p", "--shift", "-5", "--name", "x", "1", "-2e3"], ["p", "--time_steps", "-1", "1", "2"],
    ["p", "--axis_x", "1e-3x", "1", "2"], ["p", "--axis_x", " 3", "1", "2"], ["p", "--axis_x", "inf", "1", "2"],
    ["p", "-5", "2"], ["p", "--coupling", "", "1", "2"], ["p", "--time_steps", " 4", "1", "2"],
    ["p", "--time_steps", "+4", "1", "2"], ["p", "--time_steps", "18446744073709551615", "1", "2"],
    ["p", "--time_steps", "18446744073709551616", "1", "2"], ["p", "--shift", "-9223372036854775808", "1", "2"],
    ["p", "--shift", "9223372036854775808", "1", "2"], ["p", "--visc", "0.1", "--axis_x", "0x1p-3", "7", "1e-310"],
    ["p", "--visc", "1e39", "1", "2"], ["p", "--verbose", "yes", "--verbose", "off", "1", "2"],
    ["p", "--verbose", "", "1", "2"], ["p", "--axis_x", "", "1", "2"], ["p", "--name", "--coupling", "1", "2"],
    ["p", "--coupling", "fplbm", "--sub_steps", "0", "0", "-0"], ["p", "--", "x", "1", "2"],
    ["p", "--axis_x", "nan", "1", "2"], ["p", "--axis_x", "1.5 ", "1", "2"], ["p", "1", "--shift"],
    ["./sim", "--time_steps", "00012", "12", ".5"], ["p", "--shift", "-", "1", "2"],
]

CONFIGS = [
    '{"delta_x": 0.1, "kinematic_viscosity": 1e-3, "delta_t": 0.01}', '{}', ' {"delta_x":2} ', '{"delta_x":2} x',
    '{"delta_x":"a"}', '{"delta_x":+1.5e2,"other":{"a":[1,2,{"b":null}],"c":"\\u00e9\\ud83d\\ude00"},"delta_t":.5}',
    '{"delta_x":1,}', '{"delta_x" 1}', '', '{"delta_x":-}', '{"delta\\u005fx":7}', '[1]', '{"delta_x":1e400}',
    '{"kinematic_viscosity":0.1,"kinematic_viscosity":0.2}', '{"a":' + '[' * 70 + ']' * 70 + '}',
    '{"a":' + '[' * 62 + ']' * 62 + ',"delta_t":3}', '{"delta_x":true}', '{"x":"\\q"}', '{"x":"abc', '{"delta_x":0x10}',
    '{"delta_x":1.}', '{"delta_x": 3', '{"a":tru}', '{"a":nul}', '{"delta_x":1}\n\t ', '{"delta_x":1e}',
    '{"x":"\\ud800"}', '{"x":"\\u12"}', '{"x":"\\uzzzz"}', '{"a":{},"b":[],"delta_t":2.5e-3}', '{"a":[1,]}',
    '{"a":false,"b":true,"c":null,"delta_x":-0.0}', '{"delta_x":1}{', '{"delta_x" :  \r\n 1 , "delta_t":2}',
    '{"delta_x":0.1,"kinematic_viscosity":1.5e-5,"delta_t":1e-4}', '{"delta_x":3.4028235e38,"delta_t":1e-46}',
]


def main():
    lib = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    if not os.path.exists(lib):
        sys.exit("build oracle/_ref first (make -C oracle/ref)")
    L = C.CDLL(lib)
    L.ref_convert.restype = C.c_double
    L.ref_convert.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double]
    rng = np.random.default_rng(5)
    conv = []
    for st in (0, 1):
        for _ in range(40):
            dx = float(10.0 ** rng.uniform(-4, 0))
            dt = float(10.0 ** rng.uniform(-5, 0))
            v = float(rng.uniform(-2, 2) * 10.0 ** rng.integers(-6, 3))
            for what in range(13):
                conv.append([st, dx, dt, what, v, L.ref_convert(st, dx, dt, what, v).hex()])
    conv.append([0, 0.1, 0.1, 0, 1.0, L.ref_convert(0, 0.1, 0.1, 0, 1.0).hex()])  # converter.cpp known answer

    args = []
    for words in ARGS:
        arr = (C.c_char_p * len(words))(*[w.encode() for w in words])
        nums = (C.c_double * 8)()
        txt = C.create_string_buffer(1024)
        err = C.create_string_buffer(1024)
        r = L.ref_parse_args(len(words), C.cast(arr, C.POINTER(C.c_char_p)), nums, txt, 1024, err, 1024)
        args.append({"argv": words, "ok": r == 0, "nums": [float(x).hex() for x in nums],
                     "text": txt.value.decode(), "error": err.value.decode()})

    d = tempfile.mkdtemp()
    configs = []
    L.ref_load_config.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_char_p, C.c_uint64]
    for i, doc in enumerate(CONFIGS):
        p = os.path.join(d, f"c{i}.json")
        with open(p, "w") as fh:
            fh.write(doc)
        for st in (0, 1):
            out = (C.c_double * 3)()
            err = C.create_string_buffer(256)
            r = L.ref_load_config(p.encode(), st, out, err, 256)
            configs.append({"doc": doc, "st": st, "ok": r == 0, "values": [float(x).hex() for x in out],
                            "error": err.value.decode()})
    out = (C.c_double * 3)()
    err = C.create_string_buffer(256)
    r = L.ref_load_config(b"/nonexistent/lbm_config.json", 0, out, err, 256)
    configs.append({"doc": None, "st": 0, "ok": r == 0, "values": [0.0.hex()] * 3, "error": err.value.decode()})

    env = []
    old = os.environ.get("HOME")
    for case in ("default", "located", "located_newline", "empty_located"):
        home = tempfile.mkdtemp()
        os.environ["HOME"] = home
        os.makedirs(os.path.join(home, ".lbm"))
        content = {"default": None, "located": "/data/lbm_runs", "located_newline": "/data/lbm_runs\n",
                   "empty_located": ""}[case]
        if content is not None:
            with open(os.path.join(home, ".lbm", "data_dir_location.txt"), "w") as fh:
                fh.write(content)
        for name in ("poiseulle_particle_3d_hlbm_gpu", "sub/dir", "/abs"):
            buf = C.create_string_buffer(4096)
            err = C.create_string_buffer(256)
            r = L.ref_lbm_env(name.encode(), buf, 4096, err, 256)
            env.append({"case": case, "content": content, "name": name, "ok": r == 0,
                        "paths": buf.value.decode().replace(home, "{HOME}").split("\x1f")})
    if old is None:
        del os.environ["HOME"]
    else:
        os.environ["HOME"] = old

    with open(OUT, "w") as fh:
        json.dump({"convert": conv, "args": args, "config": configs, "env": env}, fh, indent=0)
    print(f"{OUT}: {len(conv)} conversions, {len(args)} argv, {len(configs)} configs, {len(env)} env")


if __name__ == "__main__":
    main()
