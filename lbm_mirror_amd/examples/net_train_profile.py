"""A short training run of the 256x256 warm-start network on synthetic
samples (straight channel, random fields), for kernel traces of the bf16
MFMA training kernels.  Prints the device time per training step."""
import os
import sys

R = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, R)
sys.path.insert(0, os.path.join(R, "tests"))
import numpy as np
from lbm_mirror_amd.warmstart import Net
from helpers import channel_flags_2d
H = W = 256
net = Net(H, W, (32, 64, 96, 160), batch=16, capacity=32, seed=1)
rng = np.random.default_rng(0)
for k in range(32):
    fl = channel_flags_2d(W, H)
    u = rng.standard_normal((2, H * W)) * 0.01
    rho = 1 + rng.standard_normal(H * W) * 1e-4
    net.add_sample(fl, 0.05, 0.8, rho, u)
net.train(30, 1e-3, 0)
print("train ms/step", net.last_ms() / 30, flush=True)
