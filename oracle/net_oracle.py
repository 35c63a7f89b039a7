"""CPU restatement of the learned warm-start network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
network kernels (lbm_mirror_amd/csrc/net.hip); never part of the product path.

The network has no counterpart in the reference (it is new with this engine),
so this oracle is pinned by mathematics rather than by reference vectors:
  * the 3x3 convolution against scipy.signal.correlate2d summed over channel
    pairs (a textbook routine, different formulation);
  * the backward pass against central finite differences of the loss in the
    float64 mode (no rounding anywhere);
  * pooling / upsampling against reshape-means and np.repeat.
Those checks live in tests/test_net_oracle.py.

Two modes:
  "bf16"  rounds to bf16 exactly where the GPU kernels round (inputs,
          weights, every stored activation and activation gradient), with
          FP32 arithmetic elsewhere; accumulation order differs from the GPU,
          so comparisons use a relative tolerance.
  "exact" float64 throughout, no rounding (for finite differences).
"""
import numpy as np

FEATURES = 8
OUTPUTS = 3


def bf16(x):
    """Round-to-nearest-even to bf16, returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def layer_specs(widths):
    w = widths
    return [(8, w[0], True, 0), (w[0], w[0], True, 0), (w[0], w[1], True, 1), (w[1], w[1], True, 1),
            (w[1], w[2], True, 2), (w[2], w[2], True, 2), (w[2], w[3], True, 3), (w[3], w[3], True, 3),
            (w[3] + w[2], w[2], True, 2), (w[2] + w[1], w[1], True, 1), (w[1] + w[0], w[0], True, 0),
            (w[0], 3, False, 0)]


def param_count(widths):
    return sum(9 * ci * co + co for ci, co, _, _ in layer_specs(widths))


def unpack(params, widths):
    """Flat parameters -> list of (W[9][cin][cout], b[cout])."""
    out, off = [], 0
    for ci, co, _, _ in layer_specs(widths):
        n = 9 * ci * co
        out.append((params[off:off + n].reshape(9, ci, co), params[off + n:off + n + co]))
        off += n + co
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in layers])


def features(flags, H, W, u_max, tau):
    """Per-cell network input (what lbm_net_features returns), float32 (H, W, 8):
    fluid indicator, u_max/0.08, (tau-1/2)/0.7, inlet profile shape, x/(nx-1),
    y/(ny-1), log1p of the 1D hydraulic pressure estimate to the outlet (sum of
    12 nu Q / h^3 over the columns downstream, over u_max^2), log((ny-2)/h)."""
    fl = np.asarray(flags).reshape(H, W)
    y = np.arange(H, dtype=np.float64)
    prof = 4.0 * (y - 0.5) * (H - 1.5 - y) / ((H - 2.0) * (H - 2.0))
    nu = (tau - 0.5) / 3.0
    S = 0.0
    for yy in range(1, H - 1):
        S += 4.0 * (yy - 0.5) * (H - 1.5 - yy) / ((H - 2.0) * (H - 2.0))
    h = (fl >= 2).sum(axis=0).astype(np.float64)
    h[h == 0] = 1.0
    press = np.zeros(W)
    acc = 0.0
    for x in range(W - 1, -1, -1):
        press[x] = 12.0 * nu * S * acc / u_max
        acc += 1.0 / (h[x] * h[x] * h[x])
    hfull = float(H - 2) if H > 2 else 1.0
    f = np.zeros((H, W, 8), dtype=np.float32)
    f[..., 0] = (fl >= 2).astype(np.float32)
    f[..., 1] = np.float32(u_max / 0.08)
    f[..., 2] = np.float32((tau - 0.5) / 0.7)
    f[..., 3] = prof.astype(np.float32)[:, None]
    f[..., 4] = (np.arange(W, dtype=np.float64) / (W - 1)).astype(np.float32)[None, :]
    f[..., 5] = (y / (H - 1)).astype(np.float32)[:, None]
    f[..., 6] = np.log1p(press).astype(np.float32)[None, :]
    f[..., 7] = np.log(hfull / h).astype(np.float32)[None, :]
    return f


def rho_unit(u_max):
    """Inlet dynamic pressure in density units: rho - 1 = 3 p, p / u_max^2."""
    return 3.0 * u_max * u_max


def targets(rho, u, H, W, u_max):
    t = np.zeros((H, W, 3), dtype=np.float32)
    N = H * W
    t[..., 0] = (np.asarray(u[:N]) / u_max).reshape(H, W)
    t[..., 1] = (np.asarray(u[N:2 * N]) / u_max).reshape(H, W)
    t[..., 2] = ((np.asarray(rho) - 1.0) / rho_unit(u_max)).reshape(H, W)
    return t


class _Mode:
    def __init__(self, name):
        self.name = name
        self.dt = np.float64 if name == "exact" else np.float32

    def r(self, x):
        """Storage rounding of an activation / activation gradient."""
        return x.astype(np.float64) if self.name == "exact" else bf16(x)


def _shifted(x):
    """x (B,H,W,C) -> (B,H,W,9,C): tap t = 3*(dy+1)+(dx+1) holds x[y+dy, x+dx] (0 outside)."""
    B, H, W, C = x.shape
    p = np.zeros((B, H + 2, W + 2, C), dtype=x.dtype)
    p[:, 1:-1, 1:-1] = x
    return np.stack([p[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)], axis=3)


def conv3x3(x, w, b=None, dt=np.float32):
    """out[p][co] = sum_{tap,ci} x[p+off(tap)][ci] * w[tap][ci][co] (+ b)."""
    B, H, W, C = x.shape
    cols = _shifted(x.astype(dt)).reshape(B * H * W, 9 * C)
    out = cols @ w.astype(dt).reshape(9 * C, -1)
    if b is not None:
        out = out + b.astype(dt)
    return out.reshape(B, H, W, -1)


def conv3x3_wgrad(x, d, dt=np.float32):
    B, H, W, C = x.shape
    cols = _shifted(x.astype(dt)).reshape(B * H * W, 9 * C)
    return (cols.T @ d.astype(dt).reshape(B * H * W, -1)).reshape(9, C, -1)


def conv3x3_dgrad(d, w, dt=np.float32):
    """Input gradient: correlation of d with the flipped, transposed kernel."""
    wt = w[::-1].transpose(0, 2, 1)  # wt[8-tap][co][ci] = w[tap][ci][co]
    return conv3x3(d, np.ascontiguousarray(wt), None, dt)


def pool(x, dt=np.float32):
    B, H, W, C = x.shape
    a = x.astype(dt)
    s = ((a[:, 0::2, 0::2] + a[:, 0::2, 1::2]) + a[:, 1::2, 0::2]) + a[:, 1::2, 1::2]
    return s * dt(0.25)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(g, dt=np.float32):
    a = g.astype(dt)
    return ((a[:, 0::2, 0::2] + a[:, 0::2, 1::2]) + a[:, 1::2, 0::2]) + a[:, 1::2, 1::2]


def forward(params, widths, x, mode="bf16", keep=False):
    """x (B,H,W,8) -> y (B,H,W,3); with keep, also the stored activations."""
    md = _Mode(mode)
    dt = md.dt
    L = unpack(np.asarray(params), widths)
    Wq = [(w.astype(dt) if mode == "exact" else bf16(w), b.astype(dt)) for w, b in L]

    def cv(i, inp, relu=True):
        o = conv3x3(inp, Wq[i][0], Wq[i][1], dt)
        if relu:
            o = np.maximum(o, 0)
        return md.r(o)

    A = {}
    A["x"] = md.r(x)
    A["a0"] = cv(0, A["x"])
    A["e0"] = cv(1, A["a0"])
    A["p1"] = md.r(pool(A["e0"], dt))
    A["a1"] = cv(2, A["p1"])
    A["e1"] = cv(3, A["a1"])
    A["p2"] = md.r(pool(A["e1"], dt))
    A["a2"] = cv(4, A["p2"])
    A["e2"] = cv(5, A["a2"])
    A["p3"] = md.r(pool(A["e2"], dt))
    A["a3"] = cv(6, A["p3"])
    A["bt"] = cv(7, A["a3"])
    A["c2"] = np.concatenate([up(A["bt"]), A["e2"]], axis=3)
    A["d2"] = cv(8, A["c2"])
    A["c1"] = np.concatenate([up(A["d2"]), A["e1"]], axis=3)
    A["d1"] = cv(9, A["c1"])
    A["c0"] = np.concatenate([up(A["d1"]), A["e0"]], axis=3)
    A["d0"] = cv(10, A["c0"])
    y = conv3x3(A["d0"], Wq[11][0], Wq[11][1], dt)
    if keep:
        return y, A, Wq
    return y


def loss_grad(params, widths, x, t, m, mode="bf16"):
    """Masked MSE over the 3 outputs and its gradient (flat, parameter order)."""
    md = _Mode(mode)
    dt = md.dt
    y, A, Wq = forward(params, widths, x, mode, keep=True)
    mk = (np.asarray(m) > 0)[..., None]
    nf = max(1.0, float(mk.sum()))
    d = (y - t.astype(dt)) * mk
    loss = float(np.sum(d.astype(np.float64) ** 2) / (3.0 * nf))
    if mode == "exact":
        gy = 2.0 * d / (3.0 * nf)
    else:
        inv = np.float32(1.0 / (3.0 * nf))
        gy = bf16((np.float32(2.0) * d.astype(np.float32)) * inv)
    grads = [None] * 12

    def bw(i, inp, dpre, mask_in):
        gw = conv3x3_wgrad(inp, dpre, dt)
        gb = dpre.astype(dt).sum(axis=(0, 1, 2))
        grads[i] = (gw, gb)
        if mask_in is None:
            return None
        dx = conv3x3_dgrad(dpre, Wq[i][0], dt)
        if mask_in:
            dx = dx * (inp > 0)
        return md.r(dx)

    w = widths
    g_d0 = bw(11, A["d0"], gy, True)
    g_c0 = bw(10, A["c0"], g_d0, True)
    g_d1 = md.r(up_bwd(g_c0[..., :w[1]], dt))
    g_c1 = bw(9, A["c1"], g_d1, True)
    g_d2 = md.r(up_bwd(g_c1[..., :w[2]], dt))
    g_c2 = bw(8, A["c2"], g_d2, True)
    g_bt = md.r(up_bwd(g_c2[..., :w[3]], dt))
    g_a3 = bw(7, A["a3"], g_bt, True)
    g_p3 = bw(6, A["p3"], g_a3, False)

    def pool_back(gpool, gskip, e):
        add = np.where(e > 0, up(gpool).astype(dt) * dt(0.25), dt(0))
        return md.r(gskip.astype(dt) + add)

    g_e2 = pool_back(g_p3, g_c2[..., w[3]:], A["e2"])
    g_a2 = bw(5, A["a2"], g_e2, True)
    g_p2 = bw(4, A["p2"], g_a2, False)
    g_e1 = pool_back(g_p2, g_c1[..., w[2]:], A["e1"])
    g_a1 = bw(3, A["a1"], g_e1, True)
    g_p1 = bw(2, A["p1"], g_a1, False)
    g_e0 = pool_back(g_p1, g_c0[..., w[1]:], A["e0"])
    g_a0 = bw(1, A["a0"], g_e0, True)
    bw(0, A["x"], g_a0, None)
    flat = pack([(gw, gb) for gw, gb in grads])
    return loss, flat
