"""Pins the network oracle (oracle/net_oracle.py) without a GPU:
bf16 rounding against torch's bfloat16 cast, the convolution against
scipy.signal.correlate2d, pooling / upsampling against reshape-means, and the
backward pass against central finite differences in float64."""
import numpy as np
import pytest
import torch
from scipy.signal import correlate2d

from oracle import net_oracle as no


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0, 3.0e38])])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(no.bf16(x), ref)
    # ties go to even
    assert no.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0
    assert no.bf16(np.float32([1.0 + 3 * 2 ** -8]))[0] == np.float32(1.0 + 2 ** -6)


def test_conv3x3_matches_correlate2d():
    rng = np.random.default_rng(1)
    B, H, W, Ci, Co = 2, 7, 5, 3, 4
    x = rng.standard_normal((B, H, W, Ci))
    w = rng.standard_normal((9, Ci, Co))
    b = rng.standard_normal(Co)
    got = no.conv3x3(x, w, b, np.float64)
    for bb in range(B):
        for co in range(Co):
            ref = np.full((H, W), b[co])
            for ci in range(Ci):
                ref += correlate2d(x[bb, :, :, ci], w[:, ci, co].reshape(3, 3), mode="same")
            assert np.allclose(got[bb, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_pool_and_upsample():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 4, 3))
    p = no.pool(x, np.float64)
    ref = x.reshape(2, 3, 2, 2, 2, 3).mean(axis=(2, 4))
    assert np.allclose(p, ref, rtol=1e-14)
    u = no.up(p)
    assert u.shape == x.shape
    assert np.array_equal(u[:, ::2, ::2], p) and np.array_equal(u[:, 1::2, 1::2], p)
    # up_bwd is the adjoint of up: <up(a), g> = <a, up_bwd(g)>
    g = rng.standard_normal(x.shape)
    assert np.isclose(np.sum(no.up(p) * g), np.sum(p * no.up_bwd(g, np.float64)), rtol=1e-12)


def _tiny_problem(seed, H=8, W=8, widths=(8, 8, 16, 8), B=2):
    rng = np.random.default_rng(seed)
    specs = no.layer_specs(widths)
    layers = [(rng.standard_normal((9, ci, co)) * np.sqrt(2.0 / (9 * ci)), rng.standard_normal(co) * 0.1)
              for ci, co, _, _ in specs]
    params = no.pack(layers)
    x = rng.uniform(0, 1, (B, H, W, 8))
    t = rng.standard_normal((B, H, W, 3))
    m = (rng.uniform(0, 1, (B, H, W)) > 0.3).astype(np.float64)
    return params, x, t, m


def test_backward_matches_finite_differences():
    widths = (8, 8, 16, 8)
    params, x, t, m = _tiny_problem(3, widths=widths)
    loss, g = no.loss_grad(params, widths, x, t, m, mode="exact")
    rng = np.random.default_rng(4)
    specs = no.layer_specs(widths)
    offs, off = [], 0
    for ci, co, _, _ in specs:
        offs.append((off, off + 9 * ci * co, off + 9 * ci * co + co))
        off += 9 * ci * co + co
    picks = []
    for a, b, c in offs:
        picks += list(rng.integers(a, b, 4)) + [int(rng.integers(b, c))]
    eps = 1e-6
    for i in picks:
        p1 = params.copy()
        p1[i] += eps
        p2 = params.copy()
        p2[i] -= eps
        l1, _ = no.loss_grad(p1, widths, x, t, m, mode="exact")
        l2, _ = no.loss_grad(p2, widths, x, t, m, mode="exact")
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_input_gradient_path_is_exercised():
    """A wrong data-gradient (e.g. unflipped taps) changes first-layer grads:
    flipping the kernel in the dgrad must matter for this input."""
    widths = (8, 8, 16, 8)
    params, x, t, m = _tiny_problem(5, widths=widths)
    _, g = no.loss_grad(params, widths, x, t, m, mode="exact")
    saved = no.conv3x3_dgrad
    try:
        no.conv3x3_dgrad = lambda d, w, dt: no.conv3x3(d, np.ascontiguousarray(w.transpose(0, 2, 1)), None, dt)
        _, g_bad = no.loss_grad(params, widths, x, t, m, mode="exact")
    finally:
        no.conv3x3_dgrad = saved
    n0 = 9 * 8 * 8 + 8
    assert np.linalg.norm(g_bad[:n0] - g[:n0]) > 1e-3 * np.linalg.norm(g[:n0])


def test_bf16_mode_close_to_exact():
    widths = (8, 8, 16, 8)
    params, x, t, m = _tiny_problem(6, widths=widths)
    le, ge = no.loss_grad(params, widths, x, t, m, mode="exact")
    lb, gb = no.loss_grad(params.astype(np.float32), widths, x.astype(np.float32), t.astype(np.float32), m, mode="bf16")
    # bf16 keeps 8 significant bits: each stored activation / gradient moves
    # by up to 2^-9 relative, compounded over 12 layers forward and back
    assert abs(lb - le) < 2e-2 * le
    assert np.linalg.norm(gb - ge) < 0.15 * np.linalg.norm(ge)


def test_param_count_of_default_widths():
    # 915,552 weights + 899 biases
    assert no.param_count((32, 64, 96, 160)) == 916451


def test_features_closed_form():
    H, W = 16, 24
    fl = np.full((H, W), 2, np.uint8)
    fl[0] = fl[-1] = 1
    f = no.features(fl, H, W, 0.04, 0.85)
    assert f.shape == (H, W, 8)
    assert np.all(f[0, :, 0] == 0) and np.all(f[5, :, 0] == 1)
    assert np.isclose(f[3, 3, 1], 0.5) and np.isclose(f[3, 3, 2], 0.5)
    # profile: zero at the half-way walls y = 1/2 and ny - 3/2, maximum 1 between
    y = np.arange(H)
    prof = f[:, 0, 3]
    assert np.isclose(prof.max(), np.max(4 * (y - 0.5) * (H - 1.5 - y) / (H - 2.0) ** 2))
    assert f[0, 0, 4] == 0 and f[0, -1, 4] == 1 and f[-1, 0, 5] == 1
    # straight channel: the hydraulic channel is the plane-Poiseuille drop
    # 12 nu Q L / h^3 from x to the outlet (Q = mean inlet flux), no blockage
    u_max, tau = 0.04, 0.85
    nu, h = (tau - 0.5) / 3, H - 2
    Q = u_max * np.sum(4 * (y[1:-1] - 0.5) * (H - 1.5 - y[1:-1]) / (H - 2.0) ** 2)
    dp = 12 * nu * Q * (W - 1 - np.arange(W)) / h ** 3
    assert np.allclose(np.expm1(f[0, :, 6].astype(np.float64)), dp / u_max ** 2, rtol=1e-6)
    assert np.all(f[..., 7] == 0)
    # a blocked column raises the estimate upstream of it only
    fl2 = fl.copy()
    fl2[1:9, 12] = 1
    f2 = no.features(fl2, H, W, u_max, tau)
    assert np.all(f2[0, :12, 6] > f[0, :12, 6]) and np.array_equal(f2[0, 12:, 6], f[0, 12:, 6])
    assert np.isclose(f2[0, 12, 7], np.log(14 / 6))
