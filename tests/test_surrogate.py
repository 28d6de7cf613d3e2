"""The learned surrogate: its FP64 oracle (oracle/surrogate_oracle.py) pinned
by finite differences and closed forms on the CPU; the library's bf16
matrix-core forward pass and gradients against the oracle on the GPU."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import surrogate_oracle as SO


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, (n, 5))
    t = np.zeros((n, 4))
    t[:, 0] = rng.uniform(size=n) < 0.6
    t[:, 1:] = rng.normal(0, 0.5, (n, 3))
    return f, t


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for out, inp in SO.SIZES:
        parts.append(rng.normal(0, scale * np.sqrt(2.0 / inp), out * inp))
        parts.append(rng.normal(0, 0.05, out))
    return np.concatenate(parts)


def test_oracle_loss_closed_form():
    """One photon on the mean of identical unit components: the mixture term
    is -log N(0; 0, I) = 1.5 log(2 pi); survival logit 0 with d = 1 gives log 2."""
    out = np.zeros((1, 57))
    y = np.array([[1.0, 0.3, -0.2, 0.7]])
    for k in range(8):
        out[0, 2 + 7 * k:5 + 7 * k] = y[0, 1:]
    loss, g = SO.mixture_loss(out, y)
    assert np.isclose(loss[0], np.log(2.0) + 1.5 * np.log(2 * np.pi), rtol=1e-15)
    assert np.isclose(g[0, 0], -0.5)
    for k in range(8):
        assert np.allclose(g[0, 1 + 7 * k:5 + 7 * k], 0.0)          # weights and means at the optimum
        assert np.allclose(g[0, 5 + 7 * k:8 + 7 * k], 1.0 / 8.0)    # d/dlog sigma of 0.5 z^2 + log sigma, z = 0
    # an undetected photon only has the survival term
    loss0, g0 = SO.mixture_loss(out, np.array([[0.0, 5.0, 5.0, 5.0]]))
    assert np.isclose(loss0[0], np.log(2.0)) and np.allclose(g0[0, 1:], 0.0)


def test_oracle_gradients_match_finite_differences():
    p = _params(1)
    f, t = _batch(6, 2)
    loss, g = SO.loss_and_gradients(p, f, t, bf16_storage=False)
    rng = np.random.default_rng(3)
    # coordinates in every layer, weights and biases
    idx = []
    o = 0
    for out, inp in SO.SIZES:
        idx += list(o + rng.integers(0, out * inp, 6)) + list(o + out * inp + rng.integers(0, out, 3))
        o += out * inp + out
    for i in idx:
        e = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (SO.loss_and_gradients(pp, f, t, False)[0] - SO.loss_and_gradients(pm, f, t, False)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_bf16_rounding():
    # bf16 keeps 7 mantissa bits: neighbours of 1 are 1 and 1 + 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.14159, 0.0, 1.0 + 2 ** -8 + 2 ** -10,
                  1.0 + 3 * 2 ** -8])
    r = SO.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0               # tie to even
    assert r[2] == 1.0 + 2 ** -7                     # below the tie
    assert r[5] == 1.0 + 2 ** -7                     # above the tie
    assert r[6] == 1.0 + 2 ** -6                     # tie to even (upwards)
    assert abs(r[3] + 3.140625) < 1e-12 and r[4] == 0.0


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.gpu
def test_gpu_surrogate_forward_matches_oracle(cuda):
    from merlict_development_kit_amd.surrogate import Surrogate
    s = Surrogate(seed=5)
    p = _params(7)
    s.set_params(torch.tensor(p, dtype=torch.float32, device=cuda))
    np.testing.assert_array_equal(s.params().cpu().numpy(), p.astype(np.float32))
    f, _ = _batch(1000, 8)  # not a multiple of the 64-row tiles
    got = s.forward(torch.tensor(f, dtype=torch.float32, device=cuda)).cpu().numpy().astype(np.float64)
    want, _ = SO.forward(p.astype(np.float32).astype(np.float64), f.astype(np.float32))
    err = np.abs(got - want)
    scale = np.abs(want).mean()
    print(f"forward: max abs err {err.max():.3e}, mean abs err {err.mean():.3e}, mean |out| {scale:.3e}")
    assert err.mean() < 1e-2 * scale and err.max() < 0.1 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("n", [512, 1000, 4096])
def test_gpu_surrogate_gradients_match_oracle(cuda, n):
    """bf16 matrix-core backpropagation vs FP64 backpropagation of the same
    bf16-rounded network: every layer's weight and bias gradient agrees in
    direction (cosine > 0.999) and size (relative L2 error < 3 %)."""
    from merlict_development_kit_amd.surrogate import Surrogate
    s = Surrogate(seed=6)
    p = _params(9)
    s.set_params(torch.tensor(p, dtype=torch.float32, device=cuda))
    f, t = _batch(n, 10)
    loss, g = s.gradients(torch.tensor(f, dtype=torch.float32, device=cuda),
                          torch.tensor(t, dtype=torch.float32, device=cuda))
    g = g.cpu().numpy().astype(np.float64)
    want_loss, want = SO.loss_and_gradients(p.astype(np.float32).astype(np.float64), f.astype(np.float32), t)
    assert abs(loss.item() - want_loss) < 1e-2 * abs(want_loss)
    o = 0
    for li, (out, inp) in enumerate(SO.SIZES):
        for name, size in (("W", out * inp), ("b", out)):
            a, b = g[o:o + size], want[o:o + size]
            o += size
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
            print(f"layer {li} {name}: cosine {cos:.6f}, relative error {rel:.2e}")
            assert cos > 0.999 and rel < 3e-2, (li, name, cos, rel)


@pytest.mark.gpu
def test_gpu_surrogate_training_lowers_the_loss(cuda):
    """Adam steps of the library fit a smooth synthetic map (detected iff
    x^2 + y^2 < 0.5, position and time smooth functions of the inputs)."""
    from merlict_development_kit_amd.surrogate import Surrogate
    s = Surrogate(seed=11)
    gen = torch.Generator(device=cuda)
    gen.manual_seed(1)

    def batch(n):
        f = torch.rand((n, 5), generator=gen, device=cuda) * 2 - 1
        t = torch.zeros((n, 4), device=cuda)
        t[:, 0] = (f[:, 0] ** 2 + f[:, 1] ** 2 < 0.5).float()
        t[:, 1] = 0.8 * f[:, 2] + 0.1 * f[:, 0]
        t[:, 2] = 0.8 * f[:, 3] - 0.1 * f[:, 1]
        t[:, 3] = 0.3 * f[:, 4] + 0.2 * (f[:, 0] ** 2 + f[:, 1] ** 2)
        return f.contiguous(), t.contiguous()

    fe, te = batch(8192)
    first, _ = s.gradients(fe, te)
    loss = torch.zeros(1, device=cuda)
    for step in range(300):
        f, t = batch(4096)
        s.train_step(f, t, 1e-3, loss)
    last, _ = s.gradients(fe, te)
    print(f"loss {first.item():.3f} -> {last.item():.3f}")
    assert last.item() < first.item() - 2.0


@pytest.mark.gpu
def test_gpu_trained_surrogate_gains_over_untrained(cuda):
    """What training buys: on the FACT-like telescope, the hybrid's per-pixel
    histogram error against the plain propagator drops from that of the
    untrained network to a small fraction of it, and the trained network
    hands few photons back to the propagator (a photon falls back when its
    draws keep landing off the sensor lattice), far fewer than the untrained
    one."""
    from merlict_development_kit_amd import (ArrivalBinning, GpuScene, PropagationConfig, SceneDescription,
                                             histogram_arrivals)
    from merlict_development_kit_amd import surrogate as SG
    desc = SceneDescription.fact(light_guides=True)
    gs = GpuScene(desc)
    cells = SG.HexCells(desc, cuda)
    dom = SG.fact_domain()
    trained = SG.Surrogate(seed=1)
    dom, _ = SG.train(gs, desc, trained, dom, 100_000_000, seed=3)
    untrained = SG.Surrogate(seed=1)
    n = 2_000_000
    ph = dom.source(10.0).generate(n, seed=99, device=cuda)
    binning = ArrivalBinning(1440, 1, dom.time_offset - 50 * dom.time_scale, 100 * dom.time_scale)
    cfg = PropagationConfig(5, seed=7)
    plain = histogram_arrivals(gs.propagate(ph, cfg), binning).cpu().numpy().astype(np.float64)[:, 0]

    def l1(sur):
        counts, fallback = SG.propagate_hybrid(gs, sur, dom, cells, ph, binning, cfg, max_entropy=1e9, seed=5)
        h = counts.cpu().numpy().astype(np.float64)[:, 0]
        return np.abs(h - plain).sum() / plain.sum(), fallback

    e_untrained, f_untrained = l1(untrained)
    e_trained, f_trained = l1(trained)
    print(f"per-pixel L1: untrained {e_untrained:.3f}, trained {e_trained:.3f}; fallbacks {f_untrained}, {f_trained}")
    assert f_trained < 0.01 * n and f_trained < 0.1 * f_untrained
    assert e_trained < 0.35 * e_untrained and e_trained < 0.15
