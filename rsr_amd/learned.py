"""Learned shading: the MLP surrogate of the heavy GGX + environment shader.

The network, its training and its inference all run in the HIP library
(rgl_mlp_* and rgl_gpu_set_learned_shading in include/rgl_hip.h).  This
module holds what a caller needs around them: the parameter layout and
initialisation, training-data collection from seeded scenes, the training
loop, and PSNR.  No counterpart exists in the reference, whose shaders are
analytic.
"""
from __future__ import annotations

import math

import numpy as np

from . import scenes

FEAT, BANDS, ENC, K1, HID, OUT = 13, 4, 117, 128, 64, 3
SAMPLE = 16                      # floats per sample: 13 inputs + 3 targets

# fp32 parameter block (matches rgl_mlp.h)
W1 = 0
B1 = W1 + HID * K1
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * HID
W4 = B3 + HID
B4 = W4 + OUT * HID
NORM = B4 + OUT
TRAINABLE = NORM
PARAMS = NORM + 4

# world-space normalisation of positions: the config scenes sit in a
# 40-unit cube centred 25 units in front of the camera
POS_CENTER = (0.0, 0.0, -25.0)
POS_HALF_EXTENT = 30.0


def layers(p: np.ndarray):
    """Views (W1 [64][128], b1, W2, b2, W3, b3, W4 [3][64], b4) of a block."""
    return (p[W1:B1].reshape(HID, K1), p[B1:W2], p[W2:B2].reshape(HID, HID), p[B2:W3],
            p[W3:B3].reshape(HID, HID), p[B3:W4], p[W4:B4].reshape(OUT, HID), p[B4:NORM])


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases; W1's padding columns are zero."""
    rng = np.random.RandomState(seed)
    p = np.zeros(PARAMS, np.float32)
    w1, _, w2, _, w3, _, w4, _ = layers(p)
    w1[:, :ENC] = rng.uniform(-1, 1, (HID, ENC)) * math.sqrt(6.0 / ENC)
    w2[:] = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(6.0 / HID)
    w3[:] = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(6.0 / HID)
    w4[:] = rng.uniform(-1, 1, (OUT, HID)) * math.sqrt(3.0 / HID)
    p[B4:NORM] = 0.5
    p[NORM:NORM + 3] = POS_CENTER
    p[NORM + 3] = 1.0 / POS_HALF_EXTENT
    return p


def psnr(a: np.ndarray, b: np.ndarray, mask: np.ndarray | None = None) -> float:
    """PSNR in dB between two 0x00RRGGBB images (over `mask` if given)."""
    ca = np.stack([(a >> 16) & 255, (a >> 8) & 255, a & 255], -1).astype(np.float64)
    cb = np.stack([(b >> 16) & 255, (b >> 8) & 255, b & 255], -1).astype(np.float64)
    if mask is not None:
        ca, cb = ca[mask], cb[mask]
    mse = float(np.mean((ca - cb) ** 2))
    return float("inf") if mse == 0.0 else 10.0 * math.log10(255.0 ** 2 / mse)


def training_scene(seed: int, width: int = 1920, height: int = 1080):
    """A seeded random scene for training data: the heavy shader (same lights,
    environment and texture as the evaluation scene) over a fresh instance
    layout; the evaluation layout (seed 2) is not among the training seeds."""
    return scenes.scene_heavy(width, height, instances=256, seed=seed, rough_seed=seed + 7919)


def collect_samples(gpu, seeds, per_scene: int, width: int = 1920, height: int = 1080, device="cuda:0"):
    """Draws each training scene with StoreTrainingSamples; returns a
    (n, 16) float32 device tensor of samples."""
    import torch
    from .rgl import DeviceScene
    cap = per_scene * len(seeds)
    out = torch.empty((cap, SAMPLE), dtype=torch.float32, device=device)
    count = torch.zeros(1, dtype=torch.int32, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    total = 0
    for k, seed in enumerate(seeds):
        sc = training_scene(seed, width, height)
        ds = DeviceScene(sc, device, prim=False)
        # a first frame counts the covered pixels, so that the keep rate
        # fills the scene's share without truncation (the buffer fills in
        # completion order, which would favour the top of the image)
        ds.record(gpu)
        gpu.run_and_wait(stream)
        covered = int(np.count_nonzero(ds.depth.cpu().numpy() < 1.0))
        keep = min(1.0, 0.97 * per_scene / max(covered, 1))
        count.zero_()
        ds.record(gpu)
        gpu.gl_store_training_samples(out[total:].data_ptr(), per_scene, count.data_ptr(), keep, 1000 + seed)
        gpu.run_and_wait(stream)
        got = min(int(count.item()), per_scene)
        total += got
    return out[:total]


def train(gpu, samples, params, opt, epochs: float, batch: int = 8192, lr: float = 2e-3, lr_end: float = 1e-4,
          seed: int = 1, chunk: int = 64, log=None):
    """Adam over `epochs` passes of hash-drawn batches with a cosine learning
    rate from lr to lr_end; returns the last batch loss."""
    import torch
    n = samples.shape[0]
    total = max(1, int(epochs * n / batch))
    stream = torch.cuda.current_stream(samples.device).cuda_stream
    step, loss = 0, float("nan")
    while step < total:
        k = min(chunk, total - step)
        t = (step + k / 2) / total
        rate = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * t))
        loss = gpu.mlp_train(samples.data_ptr(), n, params.data_ptr(), opt.data_ptr(), k, batch, rate, seed, step, stream)
        step += k
        if log:
            log(step, total, loss)
    return loss


# ---------------------------------------------------------------------------
# numpy restatement of the network (test infrastructure for the gradient and
# forward checks): every product operand rounded to bf16 exactly where the
# device rounds it, accumulation in float64
# ---------------------------------------------------------------------------
def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def encode(samples: np.ndarray, params: np.ndarray) -> np.ndarray:
    """(n, 16) samples -> (n, 128) float32 frequency encoding."""
    x = samples[:, :FEAT].astype(np.float32).copy()
    c = params[NORM:NORM + 3].astype(np.float32)
    inv = np.float32(params[NORM + 3])
    x[:, 6:9] = (x[:, 6:9] - c) * inv
    out = np.zeros((len(x), K1), np.float32)
    for i in range(FEAT):
        xi = x[:, i]
        out[:, 9 * i] = xi
        sn = np.sin(np.pi * xi.astype(np.float64)).astype(np.float32)
        cs = np.cos(np.pi * xi.astype(np.float64)).astype(np.float32)
        for k in range(BANDS):
            if k:
                sn, cs = np.float32(2) * sn * cs, (cs - sn) * (cs + sn)
            out[:, 9 * i + 1 + 2 * k] = sn
            out[:, 9 * i + 2 + 2 * k] = cs
    return out


def ref_forward(samples: np.ndarray, params: np.ndarray):
    w1, b1, w2, b2, w3, b3, w4, b4 = [np.asarray(t, np.float64) for t in layers(params)]
    X = encode(samples, params)
    q = lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)
    H1 = np.maximum(q(X) @ q(w1).T + b1, 0)
    H2 = np.maximum(q(H1) @ q(w2).T + b2, 0)
    H3 = np.maximum(q(H2) @ q(w3).T + b3, 0)
    Y = q(H3) @ q(w4).T + b4
    return Y, (X, H1, H2, H3)


def ref_gradient(samples: np.ndarray, params: np.ndarray):
    """Mean squared error over all outputs and its gradient (trainable block)."""
    w1, b1, w2, b2, w3, b3, w4, b4 = [np.asarray(t, np.float64) for t in layers(params)]
    q = lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)
    Y, (X, H1, H2, H3) = ref_forward(samples, params)
    T = samples[:, 13:16].astype(np.float64)
    n = len(samples)
    loss = float(np.mean((Y - T) ** 2))
    dY = np.asarray(((Y - T) * (2.0 / (3.0 * n))).astype(np.float32), np.float64)
    g = np.zeros(TRAINABLE, np.float64)
    gw1, gb1, gw2, gb2, gw3, gb3, gw4, gb4 = layers(g)
    gw4[:] = q(dY).T @ q(H3)
    gb4[:] = dY.sum(0)
    dZ3 = np.asarray(((q(dY) @ q(w4)) * (H3 > 0)).astype(np.float32), np.float64)
    gw3[:] = q(dZ3).T @ q(H2)
    gb3[:] = dZ3.sum(0)
    dZ2 = np.asarray(((q(dZ3) @ q(w3)) * (H2 > 0)).astype(np.float32), np.float64)
    gw2[:] = q(dZ2).T @ q(H1)
    gb2[:] = dZ2.sum(0)
    dZ1 = np.asarray(((q(dZ2) @ q(w2)) * (H1 > 0)).astype(np.float32), np.float64)
    gw1[:] = q(dZ1).T @ q(X)
    gb1[:] = dZ1.sum(0)
    return loss, g


def exact_loss(samples: np.ndarray, params64: np.ndarray) -> float:
    """The same network in float64 with no rounding (finite-difference target)."""
    w1, b1, w2, b2, w3, b3, w4, b4 = layers(params64)
    X = encode(samples, params64.astype(np.float32)).astype(np.float64)
    H1 = np.maximum(X @ w1.T + b1, 0)
    H2 = np.maximum(H1 @ w2.T + b2, 0)
    H3 = np.maximum(H2 @ w3.T + b3, 0)
    Y = H3 @ w4.T + b4
    return float(np.mean((Y - samples[:, 13:16]) ** 2))


def exact_gradient(samples: np.ndarray, params64: np.ndarray) -> np.ndarray:
    """Analytic gradient of exact_loss (backpropagation in float64)."""
    w1, b1, w2, b2, w3, b3, w4, b4 = layers(params64)
    X = encode(samples, params64.astype(np.float32)).astype(np.float64)
    H1 = np.maximum(X @ w1.T + b1, 0)
    H2 = np.maximum(H1 @ w2.T + b2, 0)
    H3 = np.maximum(H2 @ w3.T + b3, 0)
    Y = H3 @ w4.T + b4
    n = len(samples)
    dY = (Y - samples[:, 13:16]) * (2.0 / (3.0 * n))
    g = np.zeros(TRAINABLE, np.float64)
    gw1, gb1, gw2, gb2, gw3, gb3, gw4, gb4 = layers(g)
    gw4[:] = dY.T @ H3; gb4[:] = dY.sum(0)
    d3 = (dY @ w4) * (H3 > 0); gw3[:] = d3.T @ H2; gb3[:] = d3.sum(0)
    d2 = (d3 @ w3) * (H2 > 0); gw2[:] = d2.T @ H1; gb2[:] = d2.sum(0)
    d1 = (d2 @ w2) * (H1 > 0); gw1[:] = d1.T @ X; gb1[:] = d1.sum(0)
    return g
