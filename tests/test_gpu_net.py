"""GPU checks of the learned warm start (lbm_net_*, lbm_init_equilibrium_field):
the bf16 MFMA forward and backward kernels against the numpy oracle
(oracle/net_oracle.py, itself pinned by finite differences), determinism of
training, and what the trained network gains over a cold start."""
import numpy as np
import pytest

from lbm_mirror_amd import abi
from oracle import net_oracle as no

pytestmark = pytest.mark.gpu


def _random_params(widths, seed):
    rng = np.random.default_rng(seed)
    return no.pack([(rng.standard_normal((9, ci, co)) * np.sqrt(2.0 / (9 * ci)), 0.1 * rng.standard_normal(co))
                    for ci, co, _, _ in no.layer_specs(widths)]).astype(np.float32)


def _x(rng, B, H, W):
    x = rng.uniform(0, 1, (B, H, W, 8)).astype(np.float32)
    x[..., 6:] = 0
    return x


@pytest.mark.parametrize("H,W,widths,B", [(32, 32, (16, 16, 24, 32), 2), (24, 40, (8, 24, 40, 56), 3)])
def test_forward_matches_oracle(hip, H, W, widths, B):
    from lbm_mirror_amd.warmstart import Net
    net = Net(H, W, widths, batch=B, seed=11)
    assert net.param_count == no.param_count(widths)
    net.set_params(_random_params(widths, 21))
    params = net.get_params()
    rng = np.random.default_rng(1)
    x = _x(rng, B, H, W)
    y = net.forward_raw(x)
    ref = no.forward(params, widths, x, mode="bf16")
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"forward {H}x{W} widths {widths}: rel L2 {err:.2e}")
    # same rounding points, different FP32 accumulation order: rare 1-ulp
    # bf16 flips propagate, nothing more
    assert err < 1e-2


@pytest.mark.parametrize("H,W,widths,B", [(32, 32, (16, 16, 24, 32), 2), (24, 40, (8, 24, 40, 56), 3)])
def test_loss_gradient_matches_oracle(hip, H, W, widths, B):
    from lbm_mirror_amd.warmstart import Net
    net = Net(H, W, widths, batch=B, seed=12)
    net.set_params(_random_params(widths, 22))
    params = net.get_params()
    rng = np.random.default_rng(2)
    x = _x(rng, B, H, W)
    t = rng.standard_normal((B, H, W, 3)).astype(np.float32)
    m = (rng.uniform(0, 1, (B, H, W)) > 0.25).astype(np.float32)
    loss, g = net.loss_grad_raw(x, t, m)
    lref, gref = no.loss_grad(params, widths, x, t, m, mode="bf16")
    assert abs(loss - lref) < 1e-3 * lref
    off = 0
    for i, (ci, co, _, _) in enumerate(no.layer_specs(widths)):
        n = 9 * ci * co + co
        a, b = g[off:off + n], gref[off:off + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"layer {i}: grad rel L2 {err:.2e}")
        assert err < 3e-2, (i, err)
        off += n


def test_features_match_oracle(hip):
    from lbm_mirror_amd.warmstart import Net
    H, W = 16, 24
    net = Net(H, W, (8, 8, 8, 8), batch=1)
    rng = np.random.default_rng(3)
    fl = rng.integers(1, 5, (H, W)).astype(np.uint8)
    got = net.features(fl, 0.037, 0.93)
    ref = no.features(fl, H, W, 0.037, 0.93)
    assert np.array_equal(got[..., :6], ref[..., :6])
    # log / log1p of the host C library vs numpy: equal to float32 rounding
    assert np.allclose(got[..., 6:], ref[..., 6:], rtol=2e-7, atol=1e-7)


def test_training_is_deterministic_and_fits(hip):
    from lbm_mirror_amd.warmstart import Net
    H, W, widths = 32, 32, (16, 16, 24, 32)
    rng = np.random.default_rng(4)
    nets = [Net(H, W, widths, batch=4, capacity=8, seed=5) for _ in range(2)]
    for k in range(8):
        fl = np.full((H, W), 2, np.uint8)
        fl[0] = fl[-1] = 1
        um, tau = float(rng.uniform(0.01, 0.08)), float(rng.uniform(0.55, 1.2))
        yy = np.arange(H)
        prof = np.clip(4 * (yy - 0.5) * (H - 1.5 - yy) / (H - 2.0) ** 2, 0, None)
        u = np.zeros((2, H, W))
        u[0] = um * prof[:, None] * (fl == 2)
        rho = 1.0 + 0.001 * um * np.linspace(1, 0, W)[None, :] * np.ones((H, 1))
        for n in nets:
            n.add_sample(fl, um, tau, rho, u)
    l0 = nets[0].train(150, lr=2e-3, seed=9)
    l1 = nets[1].train(150, lr=2e-3, seed=9)
    assert np.array_equal(l0, l1)
    assert np.array_equal(nets[0].get_params(), nets[1].get_params())
    print(f"loss {l0[0]:.4f} -> {l0[-1]:.5f}")
    assert l0[-10:].mean() < 0.1 * l0[:5].mean()


def test_init_equilibrium_field(hip, orc):
    from lbm_mirror_amd.lattice import Lattice
    from lbm_mirror_amd.warmstart import init_equilibrium_field
    rng = np.random.default_rng(6)
    for st in (abi.F64, abi.F32):
        lat = Lattice(abi.D2Q9, st, (12, 7))
        rho = 1 + 0.01 * rng.standard_normal(lat.N)
        u = 0.05 * rng.standard_normal((2, lat.N))
        init_equilibrium_field(lat, rho, u)
        f = lat.get_populations()
        for x in range(0, lat.N, 5):
            eq = orc.equilibrium(abi.D2Q9, st, rho[x], (u[0, x], u[1, x]))
            assert np.array_equal(f[:, x], eq)

