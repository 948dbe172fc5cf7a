"""The learned guide's device kernels against the CPU restatement
(oracle/guide_oracle.py, itself pinned by finite differences and
normalisation in test_oracle_guide.py).

BF16 weights and activations make the device and the oracle round at the same
points; they still differ in FP32 accumulation order, which can flip a BF16
rounding here and there, so network outputs and gradients are compared by
relative L2 error (1e-2 and 2e-2), the mixture arithmetic on the same raw
outputs at 1e-4.
"""
import numpy as np
import pytest

from oracle import guide_oracle as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def guide(gpu):
    from pathtracer_amd import render as R, scenes
    sc = R.Scene(scenes.cornell(width=8, height=8, spp_x=1, spp_y=1))
    g = R.Guide(sc, seed=3)
    yield g
    g.close()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _samples(rng, n):
    x = rng.uniform(0, 1, (n, 6))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    pb = rng.uniform(0.05, 1.0, n)
    wt = rng.uniform(0.0, 2.0, n)
    wt[rng.uniform(size=n) < 0.1] = 0.0
    return np.concatenate([x, w, pb[:, None], wt[:, None]], 1).astype(np.float32)


def test_params_roundtrip(guide):
    p = G.init_like_device(11)
    guide.set_params(p)
    np.testing.assert_array_equal(guide.params(), p)


@pytest.mark.parametrize("seed", [1, 2])
def test_forward_matches_oracle(guide, seed):
    rng = np.random.default_rng(seed)
    p = G.init_like_device(seed)
    guide.set_params(p)
    x = rng.uniform(0, 1, (1000, 6)).astype(np.float32)
    x[:7] = [0, 0, 0, 0, 0, 0]  # ends of the encoding
    x[7:14] = 1
    raw, mix = guide.forward(x)
    ref, _, _ = G.forward(p, x)
    assert _rel(raw, ref) < 1e-2
    # the processed mixture is the oracle's arithmetic on the device's raw outputs
    pm = G.processed(raw.astype(np.float64))
    np.testing.assert_allclose(mix[:, :32:4], pm[:, :32:4], rtol=1e-4, atol=1e-5)    # kappa mu x
    np.testing.assert_allclose(mix[:, 32:41], pm[:, 32:41], rtol=1e-4, atol=1e-6)   # pi, c
    np.testing.assert_allclose(mix[:, 3:32:4], pm[:, 3:32:4], rtol=1e-4, atol=1e-4)  # log-weights


def test_mixture_pdf_and_samples(guide):
    from pathtracer_amd import render as R
    rng = np.random.default_rng(3)
    guide.set_params(G.init_like_device(3))
    x = rng.uniform(0, 1, (2000, 6)).astype(np.float32)
    _, mix = guide.forward(x)
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u = rng.uniform(size=(2000, 3)).astype(np.float32)
    pdf, sd, sp = R.guide_eval(mix, d, u)
    np.testing.assert_allclose(pdf, G.processed_pdf(mix, d), rtol=2e-4)
    assert np.allclose(np.linalg.norm(sd, axis=1), 1, atol=1e-5)
    np.testing.assert_allclose(sp, G.processed_pdf(mix, sd), rtol=2e-4)
    # the sampled lobe and the angle to its mean are the oracle's (the azimuth
    # frame is free)
    od = G.sample(mix, u)
    m64 = mix.astype(np.float64)
    cdf = np.cumsum(m64[:, 32:40], 1)
    lobe = np.minimum((u[:, :1] >= cdf[:, :7]).sum(1), 7)
    mu = np.stack([m64[np.arange(2000), 4 * lobe + k] for k in range(3)], 1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    np.testing.assert_allclose((sd * mu).sum(1), (od * mu).sum(1), atol=2e-4)


def test_gradient_matches_oracle(guide):
    rng = np.random.default_rng(4)
    p = G.init_like_device(4)
    guide.set_params(p)
    s = _samples(rng, 3000)
    loss, g = guide.gradient(s)
    ref_loss, ref = G.loss_and_grad(p, s, emulate=True)
    assert abs(loss - ref_loss) < 1e-3 * abs(ref_loss)
    for l in range(5):
        o = G.w_offset(l)
        sl = slice(o, o + G.K[l] * G.N[l])
        assert _rel(g[sl], ref[sl]) < 2e-2, ("W", l, _rel(g[sl], ref[sl]))
        bl = slice(G.b_offset(l), G.b_offset(l) + G.N[l])
        assert _rel(g[bl], ref[bl]) < 2e-2, ("b", l, _rel(g[bl], ref[bl]))
    # a batch that is not a multiple of the 16-sample tile
    loss7, g7 = guide.gradient(s[:7])
    ref7, r7 = G.loss_and_grad(p, s[:7], emulate=True)
    assert abs(loss7 - ref7) < 1e-3 * abs(ref7)
    assert _rel(g7, r7) < 2e-2


def test_adam_step_matches_oracle(guide):
    rng = np.random.default_rng(5)
    p = G.init_like_device(5)
    guide.set_params(p)
    s = _samples(rng, 500)
    _, g = guide.gradient(s)
    guide.train_step(s, lr=1e-3)
    p1 = guide.params()
    ref, _, _ = G.adam(p.astype(np.float64), 0.0, 0.0, g.astype(np.float64), 1, lr=1e-3)
    np.testing.assert_allclose(p1, ref, rtol=1e-6, atol=1e-9)


def test_training_learns_a_directional_target(guide):
    """Radiance arriving from direction A where x0 < 0.5 and from B elsewhere:
    uniformly sampled directions weighted by F/q; after training the guide puts
    most of its density near the right direction in each half."""
    rng = np.random.default_rng(6)
    guide.set_params(G.init_like_device(6))
    A, B = np.array([0.0, 1.0, 0.0]), np.array([0.6, -0.48, 0.64])
    kt = 30.0

    def batch(n):
        x = rng.uniform(0, 1, (n, 6))
        w = rng.normal(size=(n, 3))
        w /= np.linalg.norm(w, axis=1, keepdims=True)
        t = np.where((x[:, 0] < 0.5)[:, None], A, B)
        f = kt / (2 * np.pi * (1 - np.exp(-2 * kt))) * np.exp(kt * ((w * t).sum(1) - 1))
        q = 1 / (4 * np.pi)
        return np.concatenate([x, w, np.full((n, 1), q), (f / q)[:, None]], 1).astype(np.float32)

    first = guide.train_step(batch(8192), lr=5e-3)
    for _ in range(300):
        last = guide.train_step(batch(8192), lr=5e-3)
    assert last < 0.7 * first, (first, last)
    x = rng.uniform(0, 1, (400, 6)).astype(np.float32)
    _, mix = guide.forward(x)
    from pathtracer_amd import render as R
    t = np.where((x[:, 0] < 0.5)[:, None], A, B).astype(np.float32)
    pdf_t, _, _ = R.guide_eval(mix, t)
    assert np.median(pdf_t) > 10 / (4 * np.pi), np.median(pdf_t)
    # the other direction is not favoured
    o = np.where((x[:, 0] < 0.5)[:, None], B, A).astype(np.float32)
    pdf_o, _, _ = R.guide_eval(mix, o)
    assert np.median(pdf_o) < 0.2 * np.median(pdf_t)
