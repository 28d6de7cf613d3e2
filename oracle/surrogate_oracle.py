"""FP64 restatement of the learned surrogate (numpy): forward pass, mixture
loss and backpropagated gradients.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The network is the library's (include/mlgpu.h, mlgpu_surrogate): 5 -> 256 ->
256 -> 256 -> 256 -> 57 with ReLU, flat parameters W0 b0 ... W4 b4 (W
row-major out x in).  With ``bf16=True`` the forward pass rounds the
weights, the features and every hidden activation to bf16 (round to nearest
even) as the library stores them, and computes everything else in FP64, so
the library's bf16 / fp32 results can be compared against it with
tolerances of bf16 rounding.  The 57 outputs are the survival logit s and,
per component k, (w_k, mu_k[3], log sigma_k[3]); the loss of a photon with
target (d, y) is softplus(s) - d s - d log sum_k pi_k N(y; mu_k,
diag sigma_k^2), pi = softmax(w), log sigma clamped to [-9, 6].
"""
from __future__ import annotations

import numpy as np

SIZES = [(256, 5), (256, 256), (256, 256), (256, 256), (57, 256)]
N_COMP = 8
LOG_SIGMA_MIN, LOG_SIGMA_MAX = -9.0, 6.0
NUM_PARAMS = sum(o * i + o for o, i in SIZES)


def bf16(x) -> np.ndarray:
    """Round float values to bf16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    layers, o = [], 0
    for out, inp in SIZES:
        W = params[o:o + out * inp].reshape(out, inp)
        o += out * inp
        b = params[o:o + out]
        o += out
        layers.append((W, b))
    return layers


def pack(grads) -> np.ndarray:
    return np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])


def forward(params: np.ndarray, features: np.ndarray, bf16_storage: bool = True):
    """Outputs (n x 57) and the hidden activations (post-ReLU)."""
    q = bf16 if bf16_storage else (lambda a: np.asarray(a, dtype=np.float64))
    # the library keeps biases in fp32
    qb = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if bf16_storage else q
    layers = unpack(np.asarray(params, dtype=np.float64))
    h = q(features)
    hs = [h]
    for i, (W, b) in enumerate(layers):
        z = h @ q(W).T + qb(b)
        if i < len(layers) - 1:
            h = q(np.maximum(z, 0.0))
            hs.append(h)
        else:
            return z, hs
    raise AssertionError


E4M3_MAX = 448.0


def e4m3(x) -> np.ndarray:
    """Round values to OCP fp8 e4m3 (4 exponent bits of bias 7, 3 mantissa
    bits, subnormals down to 2^-9, largest finite 448) to nearest even,
    returned as float64; magnitudes beyond 448 saturate to 448."""
    a = np.asarray(x, dtype=np.float64)
    m = np.abs(a)
    _, e = np.frexp(m)                       # m = f 2^e, f in [0.5, 1)
    ulp = np.ldexp(1.0, np.maximum(e - 1, -6) - 3)
    r = np.round(m / ulp) * ulp              # numpy rounds halves to even
    return np.copysign(np.minimum(r, E4M3_MAX), a)


def fp8_row_shift(m) -> np.ndarray:
    """The largest s with m 2^s <= 448 = 1.75 2^8 (m > 0; 0 for m = 0),
    clamped to [-100, 100]: the power-of-two row scale of the fp8 weights."""
    m = np.asarray(m, dtype=np.float64)
    f, e = np.frexp(m)
    s = np.where(2.0 * f <= 1.75, 8, 7) - (e - 1)
    return np.where(m > 0, np.clip(s, -100, 100), 0).astype(np.int64)


def forward_fp8(params: np.ndarray, features: np.ndarray) -> np.ndarray:
    """Outputs (n x 57) of the fp8 inference path (mlgpu_surrogate_set_inference
    MLGPU_SURROGATE_FP8): layer 0 on bf16 features and weights; every later
    layer on e4m3 weights W[r] 2^s rounded and scaled back by 2^-s (s per
    row, fp8_row_shift of the row's largest fp32 magnitude) and e4m3
    activations relu(z) clamped to 448; fp32 biases added after the sums."""
    layers = unpack(np.asarray(params, dtype=np.float32).astype(np.float64))
    qb = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    W0, b0 = layers[0]
    h = e4m3(np.clip(bf16(features) @ bf16(W0).T + qb(b0), 0.0, E4M3_MAX))
    for i, (W, b) in enumerate(layers[1:]):
        shift = fp8_row_shift(np.abs(W).max(axis=1))[:, None]
        Wq = np.ldexp(e4m3(np.ldexp(W, shift)), -shift)
        z = h @ Wq.T + qb(b)
        if i < len(layers) - 2:
            h = e4m3(np.clip(z, 0.0, E4M3_MAX))
        else:
            return z
    raise AssertionError


def mixture_loss(out: np.ndarray, targets: np.ndarray, survival_scale: float = 1.0):
    """Per-photon loss and its gradient wrt the 57 outputs, the gradient wrt
    the survival logit times survival_scale."""
    n = len(out)
    s = out[:, 0]
    d = targets[:, 0]
    loss = np.maximum(s, 0) + np.log1p(np.exp(-np.abs(s))) - d * s
    g = np.zeros_like(out)
    g[:, 0] = survival_scale * (1.0 / (1.0 + np.exp(-s)) - d)
    w = out[:, 1::7][:, :N_COMP]
    mu = np.stack([out[:, 2 + j::7][:, :N_COMP] for j in range(3)], axis=2)      # n x K x 3
    raw = np.stack([out[:, 5 + j::7][:, :N_COMP] for j in range(3)], axis=2)
    ls = np.clip(raw, LOG_SIGMA_MIN, LOG_SIGMA_MAX)
    y = targets[:, None, 1:4]
    z = (y - mu) * np.exp(-ls)
    lse_w = np.logaddexp.reduce(w, axis=1)
    a = w - lse_w[:, None] - 1.5 * np.log(2 * np.pi) - (0.5 * z * z + ls).sum(axis=2)
    lse_a = np.logaddexp.reduce(a, axis=1)
    det = d > 0.5
    loss = loss - np.where(det, lse_a, 0.0)
    r = np.exp(a - lse_a[:, None])
    pi = np.exp(w - lse_w[:, None])
    gw = np.where(det[:, None], pi - r, 0.0)
    gmu = np.where(det[:, None, None], -r[:, :, None] * z * np.exp(-ls), 0.0)
    inside = (raw >= LOG_SIGMA_MIN) & (raw <= LOG_SIGMA_MAX)
    gls = np.where(det[:, None, None] & inside, r[:, :, None] * (1.0 - z * z), 0.0)
    for k in range(N_COMP):
        g[:, 1 + 7 * k] = gw[:, k]
        g[:, 2 + 7 * k:5 + 7 * k] = gmu[:, k]
        g[:, 5 + 7 * k:8 + 7 * k] = gls[:, k]
    return loss, g


def loss_and_gradients(params: np.ndarray, features: np.ndarray, targets: np.ndarray, bf16_storage: bool = True,
                       survival_scale: float = 1.0):
    """Mean loss over the batch and its gradient wrt the flat parameters
    (with survival_scale != 1: the backpropagation of the outputs' gradient
    whose survival component is scaled)."""
    n = len(features)
    out, hs = forward(params, features, bf16_storage)
    loss, g = mixture_loss(out, targets, survival_scale)
    q = bf16 if bf16_storage else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float64))
    grads = [None] * len(layers)
    g = g / n
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads[i] = (g.T @ hs[i], g.sum(axis=0))
        if i > 0:
            g = (g @ q(W)) * (hs[i] > 0)
    return loss.mean(), pack(grads)
