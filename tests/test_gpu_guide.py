"""Path-guiding network on the GPU (bf16 matrix cores, fp32 accumulation) against the FP64
oracle run with the same bf16 rounding points, plus training and unbiasedness checks."""
import math

import numpy as np
import pytest

import oracle_mlp as M
from raccoongpurender_amd import guide as G
from raccoongpurender_amd import render_scenes as R
from raccoongpurender_amd.guiding_study import luminance_image
from raccoongpurender_amd.render import build_lane

pytestmark = pytest.mark.gpu


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _samples(rng, n):
    s = np.zeros(n, G.SAMPLE_DTYPE)
    s["p"] = rng.random((n, 3))
    s["n"] = _unit(rng, n)
    s["wo"] = _unit(rng, n)
    s["wi"] = _unit(rng, n)
    s["weight"] = rng.uniform(0.5, 1.5, n)
    return s


def _x(s):
    return M.encode(s["p"].astype(np.float64), s["n"].astype(np.float64), s["wo"].astype(np.float64))


def test_forward_matches_oracle(ctx):
    rng = np.random.default_rng(1)
    g = G.Guide(ctx, seed=3)
    w = g.params().astype(np.float64)
    s = _samples(rng, 333)
    out = g.forward(s)
    ref = M.forward(w, _x(s), emulate=True)
    rel = np.abs(out - ref) / (1.0 + np.abs(ref))
    print("forward rel err q99 %.3g max %.3g" % (np.quantile(rel, 0.99), rel.max()), flush=True)
    assert np.quantile(rel, 0.99) < 1e-4 and rel.max() < 2e-2


def test_gradients_match_oracle(ctx):
    rng = np.random.default_rng(2)
    g = G.Guide(ctx, seed=4)
    w = g.params().astype(np.float64)
    s = _samples(rng, 512)
    gg = g.gradients(s).astype(np.float64)
    _, gr = M.loss_grad(w, _x(s), s["wi"].astype(np.float64), s["weight"].astype(np.float64), emulate=True)
    cos = gg @ gr / (np.linalg.norm(gg) * np.linalg.norm(gr))
    rel = np.linalg.norm(gg - gr) / np.linalg.norm(gr)
    print(f"gradient cosine {cos:.6f} relative error {rel:.3g}", flush=True)
    assert cos > 0.9999 and rel < 1e-2
    # per layer as well
    offs = [0, 6144, 6208, 10304, 10368, 14464, 14528, 18624, 18688]
    for a, b in zip(offs[:-1], offs[1:]):
        if np.linalg.norm(gr[a:b]) > 1e-9:
            assert np.linalg.norm(gg[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b]) < 3e-2, (a, b)


def test_adam_step_matches_closed_form(ctx):
    rng = np.random.default_rng(3)
    g = G.Guide(ctx, seed=5, minibatch=256, lr=1e-3)
    w0 = g.params().astype(np.float64)
    s = _samples(rng, 256)
    grad = g.gradients(s).astype(np.float64)
    g.train(s)  # one minibatch: gradient + one Adam step (t = 1)
    w1 = g.params().astype(np.float64)
    m, v = np.zeros_like(w0), np.zeros_like(w0)
    expect = w0.copy()
    M.adam(expect, m, v, grad, 1e-3, 1)
    assert np.abs(w1 - expect).max() < 2e-6


def test_training_learns_a_vmf_target(ctx):
    rng = np.random.default_rng(4)
    mu = np.array([0.3, 0.8, -0.52])
    mu /= np.linalg.norm(mu)
    out_t = np.zeros(64)
    out_t[0:3] = mu
    out_t[3] = math.log(20.0)
    out_t[4] = 10.0  # one dominant lobe
    for k in range(1, 8):
        out_t[5 * k + 4] = -30.0

    def draw(n):
        s = np.zeros(n, G.SAMPLE_DTYPE)
        s["p"] = (0.4, 0.6, 0.5)
        s["n"] = (0.0, 1.0, 0.0)
        s["wo"] = (0.0, 0.6, 0.8)
        s["wi"] = M.sample(out_t, rng.random((n, 3)))
        s["weight"] = 1.0
        return s

    g = G.Guide(ctx, seed=6, capacity=1 << 16, minibatch=4096, lr=3e-3)
    test = draw(4096)

    def mean_log_q():
        out = g.forward(test[:1]).astype(np.float64)
        return float(np.log(M.pdf(np.repeat(out, len(test), 0), test["wi"].astype(np.float64))).mean())

    before = mean_log_q()
    for _ in range(12):
        g.train(draw(1 << 16))
    after = mean_log_q()
    print(f"mean log q: {before:.3f} -> {after:.3f} (target entropy bound ~0.16)", flush=True)
    assert after > before + 1.5 and after > -0.5


def test_guided_render_is_unbiased(ctx):
    sd = R.material_test_scene(48, 32, volume=True)
    lo, hi = sd["points"].min(0) - 1e-3, sd["points"].max(0) + 1e-3
    means = {"plain": [], "guided": []}
    for seed in range(3):
        for mode in ("plain", "guided"):
            wf = build_lane(ctx, sd, 48 * 32 * 64, clamp=1e30, seed=1000 * seed + 17)
            if mode == "guided":
                gd = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=seed)
                G.attach(wf, gd, sampling=True, training=True)
            for _ in range(30):
                wf.step(4)
                if mode == "guided":
                    G.train_pass(wf, 1 << 15)
            y, cnt = luminance_image(wf)
            means[mode].append(float((y * cnt).sum() / cnt.sum()))
    mp, mg = np.mean(means["plain"]), np.mean(means["guided"])
    sp = np.std(means["plain"], ddof=1) / math.sqrt(3)
    sg = np.std(means["guided"], ddof=1) / math.sqrt(3)
    z = abs(mp - mg) / math.sqrt(sp ** 2 + sg ** 2 + (1e-4 * mp) ** 2)
    print(f"plain {means['plain']} guided {means['guided']} z={z:.2f}", flush=True)
    assert z < 4.0 and abs(mp - mg) < 0.02 * mp
