"""The FP64 model of the path-guiding network (oracle/oracle_mlp.cpp), pinned by finite
differences, the vMF-mixture normalisation integral, sampling histograms and Adam's closed
form."""
import math

import numpy as np

import oracle_mlp as M


def _rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=scale, size=M.PARAM_COUNT) / np.sqrt(8.0)


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_encoding_values():
    x = M.encode(np.array([[0.0, 0.5, 0.25]]), np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))[0]
    for k in range(12):
        assert abs(x[6 * k + 0]) < 1e-12 and abs(x[6 * k + 1] - 1.0) < 1e-12  # p = 0
        assert abs(x[6 * k + 2] - math.sin(2 ** k * math.pi * 0.5)) < 1e-12
        assert abs(x[6 * k + 5] - math.cos(2 ** k * math.pi * 0.25)) < 1e-12
    assert list(x[72:78]) == [0, 0, 1, 1, 0, 0] and not x[78:].any()


def test_scene_class_is_one_hot_after_the_directions():
    # class code c = 1..8 (a batch job's scene) sets feature 77 + c and nothing else; 0 and
    # out-of-range codes leave the encoding of a single render
    rng = np.random.default_rng(11)
    p, n, wo = rng.random((10, 3)), _unit(rng, 10), _unit(rng, 10)
    base = M.encode(p, n, wo)
    cls = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9])
    x = M.encode(p, n, wo, cls)
    assert np.array_equal(x[:, :78], base[:, :78])
    for i, c in enumerate(cls):
        expect = np.zeros(18)
        if 1 <= c <= 8:
            expect[c - 1] = 1.0
        assert np.array_equal(x[i, 78:], expect), c


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = _rand_params(1)
    n = 4
    x = M.encode(rng.random((n, 3)), _unit(rng, n), _unit(rng, n))
    wi = _unit(rng, n)
    wt = rng.uniform(0.5, 2.0, n)
    loss, g = M.loss_grad(params, x, wi, wt)
    idx = rng.choice(M.PARAM_COUNT, 300, replace=False)
    h = 1e-6
    ok = 0
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (M.loss_grad(p1, x, wi, wt)[0] - M.loss_grad(p2, x, wi, wt)[0]) / (2 * h)
        ok += abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i]))
    assert ok >= 297  # a parameter on a ReLU kink may disagree
    assert np.abs(g).max() > 1e-4


def test_selection_gradient_matches_finite_differences():
    # The learned selection probability c = clamp(sigmoid(out[40]), 0.05, 0.95) adds, for rows
    # with a known density p_o of the other strategies, -w log(c q + (1 - c) p_o): checked by
    # central differences over the whole loss, on rows with and without p_o.
    rng = np.random.default_rng(3)
    params = _rand_params(2)
    n = 6
    x = M.encode(rng.random((n, 3)), _unit(rng, n), _unit(rng, n))
    wi = _unit(rng, n)
    wt = rng.uniform(0.5, 2.0, n)
    po = np.array([0.3, 0.0, 1.7, 0.05, 0.0, 0.9])
    loss, g = M.loss_grad(params, x, wi, wt, other_pdf=po)
    loss0, g0 = M.loss_grad(params, x, wi, wt)
    loss_z, g_z = M.loss_grad(params, x, wi, wt, other_pdf=np.zeros(n))
    assert loss_z == loss0 and np.array_equal(g_z, g0)  # p_o = 0: no selection term
    assert loss != loss0  # the mixture term of the four rows with p_o > 0
    # output 40's weights and bias get a gradient only from the selection term
    sel = np.r_[M.OFFSETS["W3"] + 40 * 64 + np.arange(64), M.OFFSETS["B3"] + 40]
    assert np.abs(g0[sel]).max() == 0.0 and np.abs(g[sel]).max() > 1e-4
    idx = np.concatenate([rng.choice(M.PARAM_COUNT, 200, replace=False), sel[::8]])
    h = 1e-6
    ok = 0
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (M.loss_grad(p1, x, wi, wt, other_pdf=po)[0] - M.loss_grad(p2, x, wi, wt, other_pdf=po)[0]) / (2 * h)
        ok += abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i]))
    assert ok >= len(idx) - 3  # a parameter on a ReLU kink may disagree


def test_selection_gradient_moves_toward_the_better_strategy():
    # One row, everything but the selection bias fixed: where the guide's density at the sampled
    # direction beats the other strategies' (q > p_o) the loss falls as c rises, and the other
    # way round; gradient descent on the bias moves c accordingly.
    params = np.zeros(M.PARAM_COUNT)
    b40 = M.OFFSETS["B3"] + 40
    params[M.OFFSETS["B3"] + 3] = 2.0  # one sharp-ish lobe along +z (bias-only network)
    params[M.OFFSETS["B3"] + 2] = 1.0
    params[M.OFFSETS["B3"] + 4] = 5.0
    x = M.encode(np.array([[0.5, 0.5, 0.5]]), np.array([[0.0, 0.0, 1.0]]), np.array([[0.0, 0.0, 1.0]]))
    wi = np.array([[0.0, 0.0, 1.0]])
    q = M.pdf(M.forward(params, x), wi)[0]
    for po, sign in ((q / 10.0, -1.0), (q * 10.0, 1.0)):
        _, g = M.loss_grad(params, x, wi, np.ones(1), other_pdf=np.array([po]))
        assert np.sign(g[b40]) == sign, (po, q, g[b40])


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = math.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def _head(rng, max_logk=3.0):
    out = rng.normal(size=(1, M.N_OUT))
    out[0, 3:40:5] = rng.uniform(-3, max_logk, 8)
    return out


def test_vmf_mixture_integrates_to_one():
    rng = np.random.default_rng(2)
    d = _fib_sphere(400000)
    for _ in range(4):
        out = _head(rng)
        q = M.pdf(np.repeat(out, len(d), 0), d)
        assert abs(q.mean() * 4 * math.pi - 1.0) < 2e-3


def test_vmf_mixture_sampling_follows_pdf():
    rng = np.random.default_rng(3)
    out = _head(rng, max_logk=2.5)
    n = 200000
    w = M.sample(out[0], rng.random((n, 3)))
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-9)
    # equal-area bins in (z, phi)
    nz, nph = 16, 32
    bz = np.clip(((w[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    bp = ((np.arctan2(w[:, 1], w[:, 0]) + math.pi) / (2 * math.pi) * nph).astype(int) % nph
    counts = np.bincount(bz * nph + bp, minlength=nz * nph)
    d = _fib_sphere(2000000)
    q = M.pdf(np.repeat(out, len(d), 0), d)
    dz = np.clip(((d[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    dp = ((np.arctan2(d[:, 1], d[:, 0]) + math.pi) / (2 * math.pi) * nph).astype(int) % nph
    mass = np.bincount(dz * nph + dp, weights=q, minlength=nz * nph) * 4 * math.pi / len(d)
    expect = mass * n
    m = expect > 10
    chi2 = ((counts[m] - expect[m]) ** 2 / expect[m]).sum()
    dof = m.sum() - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    w = rng.normal(size=50)
    g = rng.normal(size=50)
    w0 = w.copy()
    m, v = np.zeros(50), np.zeros(50)
    M.adam(w, m, v, g, 1e-3, 1)
    # bias-corrected first step: m_hat = g, v_hat = g^2
    assert np.allclose(w, w0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    M.adam(w, m, v, g, 1e-3, 2)
    mh = (0.9 * 0.1 * g + 0.1 * g) / (1 - 0.9 ** 2)
    vh = (0.999 * 0.001 * g * g + 0.001 * g * g) / (1 - 0.999 ** 2)
    assert np.allclose(w, w0 - 1e-3 * g / (np.abs(g) + 1e-8) - 1e-3 * mh / (np.sqrt(vh) + 1e-8), atol=1e-12)


def test_emulated_rounding_is_close_to_exact():
    rng = np.random.default_rng(5)
    params = _rand_params(6)
    x = M.encode(rng.random((16, 3)), _unit(rng, 16), _unit(rng, 16))
    a = M.forward(params, x)
    b = M.forward(params, x, emulate=True)
    assert np.abs(a - b).max() < 0.05 * (1 + np.abs(a).max())
    assert M.bf16(1.0 + 2 ** -9) == 1.0 and M.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7  # ties to even


def test_zero_length_mean_gives_a_unit_lobe():
    """A row whose hidden units are all zero leaves a lobe's mean outputs at their biases, which
    can be exactly zero: that lobe must still be a normalised vMF about a unit mean (+z), so that
    its samples are directions and the mixture integrates to one."""
    out = np.zeros(M.N_OUT)
    out[3::5][:8] = np.log(50.0)  # sharp lobes: a zero-length mean would shrink every sample
    out[4::5][:8] = np.linspace(-1.0, 1.0, 8)
    out[5:8] = (0.3, -0.2, 0.9)  # lobe 1 keeps a mean of its own
    rng = np.random.default_rng(3)
    w = M.sample(out, rng.random((20000, 3)))
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-12)
    d = rng.normal(size=(400000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q = M.pdf(np.tile(out, (len(d), 1)), d)
    assert abs(4.0 * np.pi * q.mean() - 1.0) < 0.02
    # the lobe's mean is +z: its density peaks there
    q_z, q_mz = M.pdf(np.tile(out, (2, 1)), np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]))
    assert q_z > 100.0 * q_mz
