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


def test_forward_with_scene_classes_matches_oracle(ctx):
    # batch jobs feed their scene as a one-hot class; the network must see it as the oracle does
    rng = np.random.default_rng(21)
    g = G.Guide(ctx, seed=3)
    # the class columns of the first layer start at zero (every class the same function until
    # training tells them apart): give them weights so that the comparison exercises them
    w = g.params()
    w0 = w[:64 * 96].reshape(64, 96)
    w0[:, 78:86] = rng.uniform(-0.25, 0.25, (64, 8)).astype(np.float32)
    g.set_params(w)
    w = g.params().astype(np.float64)
    s = _samples(rng, 320)
    s["scene_class"] = rng.integers(0, 9, len(s))
    out = g.forward(s)
    x = M.encode(s["p"].astype(np.float64), s["n"].astype(np.float64), s["wo"].astype(np.float64),
                 s["scene_class"].astype(np.uint32))
    ref = M.forward(w, x, emulate=True)
    rel = np.abs(out - ref) / (1.0 + np.abs(ref))
    print("forward (classes) rel err q99 %.3g max %.3g" % (np.quantile(rel, 0.99), rel.max()), flush=True)
    assert np.quantile(rel, 0.99) < 1e-4 and rel.max() < 2e-2
    # the class changes the output of an otherwise identical row
    one = np.repeat(s[:1], 2)
    one["scene_class"] = (0, 3)
    o = g.forward(one)
    assert np.abs(o[0] - o[1]).max() > 1e-4


def test_single_render_samples_are_in_the_box_with_no_class(ctx):
    sd = R.material_test_scene(32, 24)
    lo, hi = sd["points"].min(0) - 1e-3, sd["points"].max(0) + 1e-3
    wf = build_lane(ctx, sd, 32 * 24 * 8, clamp=1e30, seed=3)
    gd = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=3)
    G.attach(wf, gd, sampling=True, training=True)
    wf.step(12)
    s = gd.buffered_samples()
    print(f"{len(s)} samples buffered", flush=True)
    assert len(s) > 100
    assert (s["scene_class"] == 0).all()
    assert (s["p"] >= 0).all() and (s["p"] <= 1).all()
    assert np.isfinite(s["weight"]).all() and (s["weight"] > 0).all()
    assert np.allclose(np.linalg.norm(s["wi"], axis=1), 1.0, atol=1e-4)


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


def test_selection_gradients_match_oracle(ctx):
    # learn_select: rows with a known density of the other strategies add the one-sample mixture
    # term; the selection logit (output 40) and, scaled, the lobes get its gradient
    rng = np.random.default_rng(12)
    g = G.Guide(ctx, seed=4, learn_select=True)
    w = g.params().astype(np.float64)
    s = _samples(rng, 512)
    s["other_pdf"] = np.where(rng.random(512) < 0.75, rng.uniform(0.02, 3.0, 512), 0.0)
    gg = g.gradients(s).astype(np.float64)
    _, gr = M.loss_grad(w, _x(s), s["wi"].astype(np.float64), s["weight"].astype(np.float64), emulate=True,
                        other_pdf=s["other_pdf"].astype(np.float64))
    cos = gg @ gr / (np.linalg.norm(gg) * np.linalg.norm(gr))
    rel = np.linalg.norm(gg - gr) / np.linalg.norm(gr)
    sel = np.r_[M.OFFSETS["W3"] + 40 * 64 + np.arange(64), M.OFFSETS["B3"] + 40]
    rs = np.linalg.norm(gg[sel] - gr[sel]) / np.linalg.norm(gr[sel])
    print(f"gradient cosine {cos:.6f} relative error {rel:.3g}; selection output {rs:.3g}", flush=True)
    assert np.linalg.norm(gr[sel]) > 1e-6
    assert cos > 0.9999 and rel < 1e-2 and rs < 3e-2
    # without learn_select the same rows give the plain gradient (no selection term)
    g0 = G.Guide(ctx, seed=4)
    g0.set_params(w.astype(np.float32))
    _, gr0 = M.loss_grad(w, _x(s), s["wi"].astype(np.float64), s["weight"].astype(np.float64), emulate=True)
    gg0 = g0.gradients(s).astype(np.float64)
    assert np.abs(gg0[sel]).max() == 0.0
    assert np.linalg.norm(gg0 - gr0) / np.linalg.norm(gr0) < 1e-2


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


def _guided_lane(ctx, sd, n, seed, guided, flags=0, passes=12, learn_select=False):
    lo, hi = sd["points"].min(0) - 1e-3, sd["points"].max(0) + 1e-3
    wf = build_lane(ctx, sd, n, clamp=1e30, seed=seed, flags=flags)
    if guided:
        gd = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=seed, learn_select=learn_select)
        G.attach(wf, gd, sampling=True, training=True)
        wf._guide = gd
    for _ in range(passes):
        wf.step(4)
        if guided:
            G.train_pass(wf, 1 << 15)
    return wf


def _pixel_means(wf):
    rgb, cnt = wf.accumulation()
    assert cnt.min() > 0
    return rgb / cnt[:, None]


@pytest.mark.parametrize("guided", [False, True])
def test_lambertian_furnace_is_exact_with_guiding(ctx, guided):
    # Every primary ray hits a convex lambertian sphere (R = 0.5) under a constant sky of 0.5:
    # each sample's expectation is exactly 0.25 whatever the guide has learned, provided the
    # one-sample mixture pdf is the density the three strategies actually sample.
    sd = R.furnace_scene("lambertian", 32, 32, value=0.5, fill=True)
    wf = _guided_lane(ctx, sd, 32 * 32 * 32, seed=5, guided=guided)
    m = _pixel_means(wf)[:, 0]
    se = m.std(ddof=1) / math.sqrt(m.size)
    z = abs(m.mean() - 0.25) / se
    print(f"guided={guided} mean {m.mean():.6f} se {se:.3g} z={z:.2f}", flush=True)
    assert z < 5.0


def test_furnace_is_exact_with_a_learned_selection_probability(ctx):
    # Each vertex mixes the guide with its own learned probability; sampling and the mixture pdf
    # read the same per-slot value, so the convex lambertian furnace stays exact (0.25).
    sd = R.furnace_scene("lambertian", 32, 32, value=0.5, fill=True)
    wf = _guided_lane(ctx, sd, 32 * 32 * 32, seed=15, guided=True, learn_select=True)
    m = _pixel_means(wf)[:, 0]
    se = m.std(ddof=1) / math.sqrt(m.size)
    z = abs(m.mean() - 0.25) / se
    wf.step(4)  # fill the training buffer again (the last pass consumed it)
    samples = wf._guide.buffered_samples()
    print(f"learned selection: mean {m.mean():.6f} se {se:.3g} z={z:.2f}; "
          f"{(samples['other_pdf'] > 0).mean():.3f} of the samples carry the other strategies' pdf", flush=True)
    assert z < 5.0
    # the samples a path emits carry the density of the other strategies at wi
    assert len(samples) > 100 and (samples["other_pdf"] > 0).mean() > 0.9 and np.isfinite(samples["other_pdf"]).all()


def test_ward_furnace_guided_matches_plain(ctx):
    # A convex Ward sphere: paths end after one bounce, so the completed samples are an
    # unbiased set in both modes and the two pixel means must agree within noise.
    sd = R.furnace_scene("ward", 32, 32, value=0.5, fill=True)
    mp = _pixel_means(_guided_lane(ctx, sd, 32 * 32 * 32, seed=7, guided=False)).mean(1)
    mg = _pixel_means(_guided_lane(ctx, sd, 32 * 32 * 32, seed=8, guided=True)).mean(1)
    d = mg - mp
    z = abs(d.mean()) / (d.std(ddof=1) / math.sqrt(d.size))
    print(f"plain {mp.mean():.6f} guided {mg.mean():.6f} z={z:.2f}", flush=True)
    assert z < 5.0


def test_guided_render_close_to_plain(ctx):
    # All materials and a medium. With consistent normals the guided and plain estimators
    # target the same image; the fixed-length run excludes the paths still in flight, which
    # favours short paths differently in the two modes, so the check is a relative one.
    sd = R.material_test_scene(48, 32, volume=True)
    means = {}
    for guided in (False, True):
        wf = _guided_lane(ctx, sd, 48 * 32 * 64, seed=17, guided=guided, flags=4, passes=30)
        y, cnt = luminance_image(wf)
        means[guided] = float((y * cnt).sum() / cnt.sum())
    print(f"plain {means[False]:.6f} guided {means[True]:.6f}", flush=True)
    assert abs(means[True] - means[False]) < 0.01 * means[False]


def test_guide_config_and_hooks_reject_bad_values(ctx):
    from raccoongpurender_amd import _lib

    for kw in (dict(weight_cap=-1.0), dict(log_kappa_max=12.5), dict(log_kappa_max=-3.0), dict(select_p=1.0)):
        with pytest.raises(_lib.RgpuError) as e:
            G.Guide(ctx, **kw)
        assert e.value.code == -1, kw
    g = G.Guide(ctx, seed=2, weight_cap=5.0, log_kappa_max=10.0)
    assert len(g.buffered_samples()) == 0  # nothing rendered into it yet
    with pytest.raises(_lib.RgpuError):
        g.buffered_samples(rows=1)
    # the cap limits what a sample contributes to the gradient
    rng = np.random.default_rng(8)
    s = _samples(rng, 64)
    s["weight"] = 50.0
    capped = g.gradients(s).astype(np.float64)
    s["weight"] = 5.0
    at_cap = g.gradients(s).astype(np.float64)
    assert np.allclose(capped, at_cap, rtol=1e-6, atol=1e-9)


def test_trained_guide_lowers_error_on_the_cornell_box(ctx):
    # What the trained network gains: on the Cornell box (a small ceiling light and a black sky,
    # so plain sampling finds the light only by chance) a guide trained for 200 passes and then
    # frozen renders 1024 spp with far lower error than plain sampling at the same spp, both
    # scored against a 2^18-spp plain render. Measured: about 26x; required: 4x.
    from raccoongpurender_amd import guiding_study as S

    sd = R.cornell_box(32, 32)
    sd["max_depth"] = 5
    ref, _, _ = S.render_reference(ctx, sd, 1 << 18, 1 << 18)
    lo, hi = sd["points"].min(0) - 1e-3, sd["points"].max(0) + 1e-3
    gd = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=3)
    tr = build_lane(ctx, sd, 1 << 16, clamp=1e30, seed=11)
    G.attach(tr, gd, sampling=True, training=True)
    passes = 0
    for _ in range(4000):
        tr.step(4)
        passes += G.train_pass(tr, 1 << 15) > 0
        if passes >= 200:
            break
    err = {}
    for guided in (False, True):
        wf = build_lane(ctx, sd, 1 << 16, clamp=1e30, seed=21)
        if guided:
            G.attach(wf, gd, sampling=True, training=False)
        while wf.stat() < 1024 * 32 * 32:
            wf.step(8)
        y, _ = luminance_image(wf)
        err[guided] = S.rel_mse(y, ref)
        wf.close()
    print(f"{passes} training passes; relMSE plain {err[False]:.3g} guided {err[True]:.3g} "
          f"({err[False] / err[True]:.1f}x lower)", flush=True)
    assert passes >= 200 and err[True] * 4.0 < err[False]


def test_guide_lifetime_and_detach_reattach(ctx):
    # A guide refuses to be destroyed while a wavefront holds it; a detached lane renders plain,
    # keeps its per-slot guiding buffers and starts recording afresh on the next attach.
    from raccoongpurender_amd import _lib

    sd = R.material_test_scene(32, 24)
    lo, hi = sd["points"].min(0) - 1e-3, sd["points"].max(0) + 1e-3
    wf = build_lane(ctx, sd, 32 * 24 * 8, clamp=1e30, seed=4)
    a = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=3)
    b = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=5)
    G.attach(wf, a, sampling=True, training=True)
    wf.step(6)
    with pytest.raises(_lib.RgpuError) as e:
        a.close()
    assert "attached" in str(e.value) and a.h is not None
    G.attach(wf, None)  # detach: renders plain, collects nothing
    n_a = len(a.buffered_samples())
    wf.step(4)
    assert len(a.buffered_samples()) == n_a
    a.close()  # no wavefront holds it now
    G.attach(wf, b, sampling=True, training=True)  # re-attach reuses the lane's buffers
    wf.step(8)
    s = b.buffered_samples()
    print(f"{n_a} samples before the detach, {len(s)} after re-attaching", flush=True)
    assert n_a > 100 and len(s) > 100
    assert (s["p"] >= 0).all() and (s["p"] <= 1).all()
    assert np.isfinite(s["weight"]).all() and (s["weight"] > 0).all()
    # attaching another guide moves the lane: the old one can then be destroyed
    c = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=6)
    G.attach(wf, c, sampling=True, training=False)
    b.close()
    wf.close()  # destroying the lane releases its hold
    c.close()


def test_inference_lobes_match_the_forward_pass(ctx):
    # The lobe records the bounce's inference writes for its lambertian and Ward vertices (weights
    # staged in LDS, chained sin/cos encoding, one-instruction head) against the tested forward
    # pass on the same vertices, encoded here from the lane's state, and the head in float64.
    sd = R.material_test_scene(48, 32, volume=True)
    lo = (sd["points"].min(0) - 1e-3).astype(np.float32)
    hi = (sd["points"].max(0) + 1e-3).astype(np.float32)
    wf = build_lane(ctx, sd, 48 * 32 * 16, clamp=1e30, seed=9)
    gd = G.Guide(ctx, lo, hi, capacity=1 << 15, minibatch=4096, seed=2)
    G.attach(wf, gd, sampling=True, training=True)
    for _ in range(3):  # a trained network, not the initial one
        wf.step(4)
        G.train_pass(wf, 1 << 15)
    wf.step(1)
    st = wf.state()
    counts, _ = wf.queue_counts()
    slots = np.concatenate([st["items"][1, :counts[1]], st["items"][4, :counts[4]]]).astype(np.int64)
    assert len(slots) > 1000
    wf.step(1)  # its inference runs over these two queues
    lob = G.vertex_lobes(wf, 0, len(slots)).astype(np.float64)  # in the same queue order
    ray, ext = st["rays"][slots], st["ext"][slots]
    inv = np.float32(1.0) / (hi - lo)
    s = np.zeros(len(slots), G.SAMPLE_DTYPE)
    P = ray["ro"] + ray["rd"] * ext["tmin"][:, None]
    s["p"] = (P - lo) * inv
    wo = -ray["rd"]
    ng = ext["Ng"]
    d = ng[:, 0] * wo[:, 0] + ng[:, 1] * wo[:, 1] + ng[:, 2] * wo[:, 2]
    s["n"] = np.where((d < 0)[:, None], -ng, ng)
    s["wo"] = wo
    out = gd.forward(s).astype(np.float64)
    mu, kappa, w = G.raw_to_lobes(out)
    lw = np.log(w)
    logwc = lw + np.log(kappa) - math.log(2.0 * math.pi) - np.log1p(-np.exp(-2.0 * kappa))
    err = {"mu": np.abs(lob[:, :, :3] - mu).max(), "kappa": (np.abs(lob[:, :, 3] - kappa) / kappa).max(),
           "w": (np.abs(lob[:, :, 4] - w) / (w + 1e-7)).max(),
           "log(w C)": np.abs(lob[:, :, 5] - logwc).max()}
    print(f"{len(slots)} vertices: " + ", ".join(f"{k} {v:.3g}" for k, v in err.items()), flush=True)
    # hardware rcp/rsq/exp/log, about one float ulp each; exp's argument is scaled by log2(e) in
    # float, which adds |x| * 6e-8 relative for weights far below the largest (w ~ e^-30: 2e-6)
    assert err["mu"] < 2e-6 and err["kappa"] < 2e-6 and err["w"] < 2e-5 and err["log(w C)"] < 2e-5
    G.attach(wf, None)
    wf.close()
    gd.close()
