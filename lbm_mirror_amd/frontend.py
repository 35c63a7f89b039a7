"""Unit and configuration front end of the reference drivers, so their
parameter sets are consumed as they are:

  Converter          converter.hpp:15-93      SI <-> lattice units, tau from nu
  lbm_meta           lbm.hpp:36-65            the numbers print_lbm_meta shows
  parse_args         abstract/args.hpp:43-340 "--name value" pairs + positionals
  load_lbm_config    config.hpp:14-51         {delta_x, kinematic_viscosity, delta_t}
  init_lbm_env       environment.hpp:9-56     ~/.lbm/data/<name> (or data_dir_location.txt)
  output_directory   environment.hpp:58-89

Arithmetic is done in the reference's type T (numpy float32 or float64
scalars, one rounding per operation, in the reference's operation order), so
a tau or a lattice velocity is the value the reference's driver computes, bit
for bit.  Number parsing goes through the C library's strtod, as the
reference's does.  Errors raise FrontendError carrying the reference's error
category and message text.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_libc = C.CDLL(None)
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]

# reference error categories (abstract/error.hpp) and their lbm_status codes
NOT_FOUND, INVALID_STATE, OUT_OF_MEMORY = "Not found", "Invalid state", "Out of memory"
JSON_MALFORMED, JSON_TOO_DEEP = "Malformed JSON", "JSON nesting is too deep"
_STATUS = {NOT_FOUND: abi.ERR_NOT_FOUND, INVALID_STATE: abi.ERR_INVALID_STATE,
           OUT_OF_MEMORY: abi.ERR_OUT_OF_MEMORY, JSON_MALFORMED: abi.ERR_INVALID_ARG,
           JSON_TOO_DEEP: abi.ERR_INVALID_ARG}


class FrontendError(Exception):
    def __init__(self, category, message=""):
        super().__init__(f"{category}: {message}" if message else category)
        self.category, self.message = category, message
        self.status = _STATUS.get(category, abi.ERR_INVALID_ARG)


def _strtod(text):
    """C strtod: (value, characters consumed)."""
    buf = C.create_string_buffer(text.encode())
    end = C.c_void_p()
    v = _libc.strtod(buf, C.byref(end))
    return v, (end.value or C.addressof(buf)) - C.addressof(buf)


# ---- converter.hpp ----------------------------------------------------------------
class Converter:
    """converter<T>{delta_x, delta_t}: SI metres / seconds per lattice unit."""

    INV_CS2 = 3.0  # df_info<T, Desc>::inv_cs2 for D2Q9, D3Q19 and D3Q27

    def __init__(self, delta_x, delta_t, dtype=np.float64):
        self.T = np.dtype(dtype).type
        self.dx = self.T(delta_x)
        self.dt = self.T(delta_t)

    def _v(self, x):
        return self.T(x)

    def delta_x(self):
        return float(self.dx * self.T(1.0))

    def delta_t(self):
        return float(self.dt * self.T(1.0))

    def delta_v(self):
        return float((self.dx / self.dt) * self.T(1.0))

    def delta_a(self):
        return float((self.dx / (self.dt * self.dt)) * self.T(1.0))

    def meter_si_to_lbm(self, m):
        return float(self._v(m) / self.dx)

    def meter_lbm_to_si(self, m):
        return float(self._v(m) * self.dx)

    def second_si_to_lbm(self, s):
        return float(self._v(s) / self.dt)

    def velocity_si_to_lbm(self, v):
        return float((self._v(v) / self.dx) * self.dt)

    def acceleration_si_to_lbm(self, a):
        return float((self._v(a) / self.dx) * (self.dt * self.dt))

    def kinematic_viscosity_si_to_lbm(self, nu):
        return float((self._v(nu) / (self.dx * self.dx)) * self.dt)

    def kinematic_viscosity_lbm_to_tau(self, nu_lbm, lattice=abi.D2Q9):
        return float(self.T(self.INV_CS2) * self._v(nu_lbm) + self.T(0.5))

    def kinematic_viscosity_si_to_tau(self, nu, lattice=abi.D2Q9):
        return self.kinematic_viscosity_lbm_to_tau(self.kinematic_viscosity_si_to_lbm(nu), lattice)


def lbm_meta(conv, kin_vis_si, char_vel, char_len):
    """The quantities print_lbm_meta reports (lbm.hpp:36-65), in T: Re =
    v L / nu; Ma as the reference forms it, v * sqrt(inv_cs2) with v the SI
    characteristic velocity; the lattice viscosity and tau."""
    T = conv.T
    nu = T(kin_vis_si)
    return {"Re": float((T(char_vel) * T(char_len)) / nu),
            "Ma": float(T(char_vel) * T(np.sqrt(T(Converter.INV_CS2)))),
            "delta_x": conv.delta_x(), "delta_t": conv.delta_t(), "delta_v": conv.delta_v(),
            "delta_a": conv.delta_a(), "kinematic_viscosity_lbm": conv.kinematic_viscosity_si_to_lbm(nu),
            "tau": conv.kinematic_viscosity_si_to_tau(nu)}


# ---- abstract/args.hpp ------------------------------------------------------------
F64, F32, U64, I64, STR, BOOL = "f64", "f32", "u64", "i64", "str", "bool"
_TRUE = ("true", "1", "yes", "on")
_FALSE = ("false", "0", "no", "off")


def _decode(kind, view):
    if kind == STR:
        return view
    if kind == BOOL:
        if view in _TRUE:
            return True
        if view in _FALSE:
            return False
        raise FrontendError(INVALID_STATE, f"Expected a boolean, got '{view}'")
    if not view:
        raise FrontendError(INVALID_STATE, "Can't decode empty data.")
    if kind in (F64, F32):
        v, n = _strtod(view)
        if n == 0 or n != len(view.encode()):
            raise FrontendError(INVALID_STATE, f"Expected a number, got '{view}'")
        if kind == F32:
            with np.errstate(over="ignore"):  # static_cast<float>: out of range -> inf
                return float(np.float32(v))
        return v
    # integers: std::from_chars — optional '-' for signed, digits only, in range
    body = view[1:] if (kind == I64 and view.startswith("-")) else view
    ok = body.isascii() and body.isdigit()
    if ok:
        v = int(view)
        ok = (0 <= v < 2 ** 64) if kind == U64 else (-2 ** 63 <= v < 2 ** 63)
    if not ok:
        raise FrontendError(INVALID_STATE, f"Expected an integer, got '{view}'")
    return v


class ArgSpec:
    """sch::Args<Struct<Member<...>...>, Tuple<...>>: named members
    {name: (kind, default)} and positional kinds, kinds one of F64, F32, U64,
    I64, STR, BOOL."""

    def __init__(self, named, positionals=()):
        self.named = dict(named)
        self.positionals = list(positionals)

    def parse(self, argv):
        """parse_args (args.hpp:287-340): returns {"program", "args",
        "positionals"}; raises FrontendError as the reference reports."""
        if len(argv) == 0:
            raise FrontendError(INVALID_STATE, "No arguments at all, not even argv[0]")
        out = {"program": argv[0], "args": {k: d for k, (_, d) in self.named.items()},
               "positionals": [None] * len(self.positionals)}
        pos = 0
        i = 1
        while i < len(argv):
            view = argv[i]
            if view.startswith("--"):
                name = view[2:]
                i += 1
                if i >= len(argv):
                    raise FrontendError(INVALID_STATE, f"Missing value for --{name}")
                if name not in self.named:
                    raise FrontendError(NOT_FOUND, f"Unknown argument --{name}")
                out["args"][name] = _decode(self.named[name][0], argv[i])
                i += 1
                continue
            if pos >= len(self.positionals):
                raise FrontendError(INVALID_STATE, f"Too many positional arguments, did not expect '{view}'")
            out["positionals"][pos] = _decode(self.positionals[pos], view)
            pos += 1
            i += 1
        if pos != len(self.positionals):
            raise FrontendError(INVALID_STATE,
                                f"Expected {len(self.positionals)} positional arguments, got {pos}")
        return out


def parse_args(spec, argv):
    return spec.parse(argv)


# ---- config.hpp: the reference's JSON grammar (abstract/json.hpp:96-420) -------------
class _Json:
    def __init__(self, src, max_depth=64):
        self.s, self.p, self.depth, self.max_depth = src, 0, 0, max_depth

    def ws(self):
        while self.p < len(self.s) and self.s[self.p] in " \t\n\r":
            self.p += 1

    def at_end(self):
        return self.p >= len(self.s)

    def peek(self):
        self.ws()
        if self.at_end():
            raise FrontendError(JSON_MALFORMED, "Unexpected end of input")
        return self.s[self.p]

    def expect(self, c):
        if self.peek() != c:
            raise FrontendError(JSON_MALFORMED, "Unexpected character")
        self.p += 1

    def consume_if(self, c):
        self.ws()
        if not self.at_end() and self.s[self.p] == c:
            self.p += 1
            return True
        return False

    def enter(self):
        self.depth += 1
        if self.depth > self.max_depth:
            raise FrontendError(JSON_TOO_DEEP)

    def leave(self):
        if self.depth > 0:
            self.depth -= 1

    def hex4(self):
        if self.p + 4 > len(self.s):
            raise FrontendError(JSON_MALFORMED, "Truncated \\u escape")
        v = 0
        for ch in self.s[self.p:self.p + 4]:
            if not (ch.isascii() and ch in "0123456789abcdefABCDEF"):
                raise FrontendError(JSON_MALFORMED, "Bad hex digit")
            v = (v << 4) | int(ch, 16)
        self.p += 4
        return v

    def string(self):
        self.expect('"')
        out = []
        while True:
            if self.at_end():
                raise FrontendError(JSON_MALFORMED, "Unterminated string")
            c = self.s[self.p]
            self.p += 1
            if c == '"':
                return "".join(out)
            if c != "\\":
                out.append(c)
                continue
            if self.at_end():
                raise FrontendError(JSON_MALFORMED, "Dangling escape")
            e = self.s[self.p]
            self.p += 1
            simple = {'"': '"', "\\": "\\", "/": "/", "b": "\b", "f": "\f", "n": "\n", "r": "\r", "t": "\t"}
            if e in simple:
                out.append(simple[e])
            elif e == "u":
                cp = self.hex4()
                if 0xD800 <= cp <= 0xDBFF:
                    if self.p + 1 < len(self.s) and self.s[self.p] == "\\" and self.s[self.p + 1] == "u":
                        self.p += 2
                        lo = self.hex4()
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00) if 0xDC00 <= lo <= 0xDFFF else 0xFFFD
                    else:
                        cp = 0xFFFD
                elif 0xDC00 <= cp <= 0xDFFF:
                    cp = 0xFFFD
                out.append(chr(cp))
            else:
                raise FrontendError(JSON_MALFORMED, "Unknown escape")

    def number_token(self):
        self.ws()
        s, start = self.s, self.p
        n = len(s)
        if self.p < n and s[self.p] in "-+":
            self.p += 1
        while self.p < n and "0" <= s[self.p] <= "9":
            self.p += 1
        if self.p < n and s[self.p] == ".":
            self.p += 1
            while self.p < n and "0" <= s[self.p] <= "9":
                self.p += 1
        if self.p < n and s[self.p] in "eE":
            self.p += 1
            if self.p < n and s[self.p] in "-+":
                self.p += 1
            while self.p < n and "0" <= s[self.p] <= "9":
                self.p += 1
        if self.p == start:
            raise FrontendError(JSON_MALFORMED, "Expected a number")
        return s[start:self.p]

    def boolean(self):
        self.ws()
        if self.s.startswith("true", self.p):
            self.p += 4
        elif self.s.startswith("false", self.p):
            self.p += 5
        else:
            raise FrontendError(JSON_MALFORMED, "Expected a boolean")

    def null(self):
        self.ws()
        if self.s.startswith("null", self.p):
            self.p += 4
            return True
        return False

    def skip_value(self):
        c = self.peek()
        self.enter()
        if c in "{[":
            close = "}" if c == "{" else "]"
            self.p += 1
            if not self.consume_if(close):
                while True:
                    if c == "{":
                        self.string()
                        self.expect(":")
                    self.skip_value()
                    if self.consume_if(close):
                        break
                    self.expect(",")
        elif c == '"':
            self.string()
        elif self.null():
            pass
        elif c in "tf":
            self.boolean()
        else:
            self.number_token()
        self.leave()

    def real(self):
        tok = self.number_token()
        v, n = _strtod(tok)
        if n == 0:
            raise FrontendError(JSON_MALFORMED, "Bad floating point")
        return v


CONFIG_KEYS = ("delta_x", "kinematic_viscosity", "delta_t")


def load_lbm_config(path, dtype=np.float64):
    """load_lbm_config<T, Desc> (config.hpp:22-50): the JSON object's
    delta_x, kinematic_viscosity and delta_t in T; members not given stay 0,
    others are skipped; trailing content is an error."""
    try:
        with open(path, "rb") as fh:
            raw = fh.read()
    except OSError:
        raise FrontendError(NOT_FOUND, "Couldn't open file") from None
    src = raw.decode("utf-8", errors="surrogateescape")
    T = np.dtype(dtype).type
    out = {k: 0.0 for k in CONFIG_KEYS}
    p = _Json(src)
    p.enter()
    p.expect("{")
    if not p.consume_if("}"):
        while True:
            key = p.string()
            p.expect(":")
            if key in out:
                with np.errstate(over="ignore"):
                    out[key] = float(T(p.real()))
            else:
                p.skip_value()
            if p.consume_if("}"):
                break
            p.expect(",")
    p.leave()
    p.ws()
    if not p.at_end():
        raise FrontendError(JSON_MALFORMED, "Trailing content after the JSON value")
    return out


# ---- environment.hpp ------------------------------------------------------------------
def _lbm_dir():
    home = os.environ.get("HOME")
    if home is None:
        raise FrontendError(NOT_FOUND, "Couldn't find home dir")
    return os.path.join(home, ".lbm")


def _located_data_dir(lbm_dir):
    loc = os.path.join(lbm_dir, "data_dir_location.txt")
    if os.path.isfile(loc):
        try:
            with open(loc, "rb") as fh:
                return fh.read().decode("utf-8", errors="surrogateescape")
        except OSError:
            return None
    return None


def init_lbm_env(name):
    """init_lbm_env (environment.hpp:18-56): lbm_dir = $HOME/.lbm, data_dir =
    lbm_dir/data unless lbm_dir/data_dir_location.txt names another (its
    whole content), name_dir = data_dir/name."""
    lbm_dir = _lbm_dir()
    data_dir = os.path.join(lbm_dir, "data")
    loc = _located_data_dir(lbm_dir)
    if loc is not None:
        data_dir = loc
    return {"name": name, "lbm_dir": lbm_dir, "data_dir": data_dir, "name_dir": os.path.join(data_dir, name)}


def output_directory():
    """output_directory (environment.hpp:63-89)."""
    lbm_dir = _lbm_dir()
    loc = _located_data_dir(lbm_dir)
    return loc if loc is not None else os.path.join(lbm_dir, "data")
