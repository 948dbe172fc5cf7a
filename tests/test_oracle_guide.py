"""Pins the CPU restatement of the learned guide (oracle/guide_oracle.py).

The guide has no counterpart in the reference, so every check here is one
the mathematics fixes: the analytic gradient against central finite
differences (float64), the vMF mixture's normalisation over the sphere and
the distribution of its samples, the encoding's support, and the parameter
layout against the C header.
"""
import os
import re

import numpy as np
import pytest

from oracle import guide_oracle as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _samples(rng, n, target=None):
    x = rng.uniform(0, 1, (n, 6))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    pb = rng.uniform(0.05, 1.0, n)
    wt = rng.uniform(0.0, 2.0, n)
    wt[rng.uniform(size=n) < 0.1] = 0.0
    return np.concatenate([x, w, pb[:, None], wt[:, None]], 1)


def test_layout_matches_header():
    text = open(os.path.join(ROOT, "include", "pathtracer_amd.h")).read()
    assert int(re.search(r"#define PT_GUIDE_NPARAMS (\d+)", text).group(1)) == G.NPARAMS == 27497
    assert int(re.search(r"#define PT_GUIDE_OUTPUTS (\d+)", text).group(1)) == G.OUT


def test_gradient_matches_finite_differences():
    """float64 analytic gradient of the mean loss vs central differences, on
    parameters of every layer and every output group (logits, selection,
    concentrations, the three mean components) and biases."""
    rng = np.random.default_rng(1)
    p = G.init_like_device(3).astype(np.float64)
    s = _samples(rng, 64)
    loss, g = G.loss_and_grad(p, s, emulate=False)
    assert np.isfinite(loss) and loss != 0.0
    idx = []
    for l in range(5):
        o = G.w_offset(l)
        idx += list(rng.integers(o, o + G.K[l] * G.N[l], 6))
        idx += list(G.b_offset(l) + rng.integers(0, G.N[l], 3))
    # one weight into each raw output of the last layer
    for out in range(G.OUT):
        idx.append(G.w_offset(4) + int(rng.integers(G.K[4])) * G.N[4] + out)
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += h
        lp, _ = G.loss_and_grad(pp, s, emulate=False)
        pp[i] -= 2 * h
        lm, _ = G.loss_and_grad(pp, s, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def test_emulated_gradient_close_to_float64():
    """BF16 rounding (2^-8 relative per value, through five layers and the
    1/p terms of the loss) changes the gradient by some percent."""
    rng = np.random.default_rng(2)
    p = G.init_like_device(4)
    s = _samples(rng, 256)
    l64, g64 = G.loss_and_grad(p.astype(np.float64), s, emulate=False)
    l16, g16 = G.loss_and_grad(p, s, emulate=True)
    assert abs(l16 - l64) < 2e-2 * abs(l64)
    assert np.linalg.norm(g16 - g64) < 0.1 * np.linalg.norm(g64)


def _sphere_quadrature(n=400_000, seed=0):
    """Stratified uniform directions and their solid angle (4 pi / n each)."""
    rng = np.random.default_rng(seed)
    u = (np.arange(n) + rng.uniform(size=n)) / n
    v = rng.uniform(size=n)
    z = 1 - 2 * u
    r = np.sqrt(np.maximum(0, 1 - z * z))
    ph = 2 * np.pi * v
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1), 4 * np.pi / n


def test_mixture_normalised_over_sphere():
    rng = np.random.default_rng(3)
    p = G.init_like_device(5)
    x = rng.uniform(0, 1, (4, 6))
    raw, _, _ = G.forward(p, x)
    raw = raw.astype(np.float64)
    raw[:, 9:17] = rng.uniform(-3, 2, (4, 8))  # kappa up to ~31
    dirs, dw = _sphere_quadrature()
    mix = G.processed(raw)
    for i in range(4):
        integral = G.guide_pdf(np.repeat(raw[i:i + 1], dirs.shape[0], 0), dirs).sum() * dw
        assert abs(integral - 1) < 2e-3, integral
        # the processed form is the same density
        a = G.processed_pdf(np.repeat(mix[i:i + 1], 1000, 0), dirs[:1000])
        b = G.guide_pdf(np.repeat(raw[i:i + 1], 1000, 0), dirs[:1000])
        np.testing.assert_allclose(a, b, rtol=1e-9)


def test_samples_follow_the_pdf():
    """Histogram of sampled directions on a 12 x 24 (cos theta, phi) grid vs
    the pdf integrated over each cell."""
    rng = np.random.default_rng(4)
    raw = np.zeros((1, G.OUT))
    raw[0, 0:8] = rng.normal(size=8)
    raw[0, 9:17] = rng.uniform(-3, 1.2, 8)  # kappa 0.3 .. 15
    raw[0, 17:41] = rng.normal(size=24)
    mix = G.processed(raw)
    n = 400_000
    u = rng.uniform(size=(n, 3))
    d = G.sample(np.repeat(mix, n, 0), u)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-9)
    nz, nphi = 12, 24

    def cell(v):
        iz = np.minimum(((v[:, 2] + 1) / 2 * nz).astype(int), nz - 1)
        ip = np.minimum(((np.arctan2(v[:, 1], v[:, 0]) + np.pi) / (2 * np.pi) * nphi).astype(int), nphi - 1)
        return iz * nphi + ip
    hist = np.bincount(cell(d), minlength=nz * nphi) / n
    dirs, dw = _sphere_quadrature(800_000, 1)
    pdf = G.guide_pdf(np.repeat(raw, dirs.shape[0], 0), dirs)
    expect = np.bincount(cell(dirs), weights=pdf * dw, minlength=nz * nphi)
    big = expect > 2e-3
    assert np.all(np.abs(hist[big] - expect[big]) < 5 * np.sqrt(expect[big] / n) + 2e-3 * expect[big])
    assert abs(hist.sum() - 1) < 1e-12


def test_encoding_support_and_sum():
    x = np.linspace(0.1, 0.9, 101)[:, None].repeat(6, 1)
    e = G.encode(x, np.float64).reshape(-1, 6, 32)
    nz = (e > 0).sum(-1)
    assert nz.min() >= 5 and nz.max() <= 6  # |b - y| < 3
    s = e.sum(-1)
    assert np.all(np.abs(s - s.mean()) < 0.02 * s.mean())  # nearly constant away from the ends


def test_adam_first_step_is_signed_lr():
    p = np.zeros(5)
    g = np.array([1e-3, -2.0, 5.0, 0.0, 1e-12])
    p1, m, v = G.adam(p, np.zeros(5), np.zeros(5), g, 1, lr=1e-3)
    np.testing.assert_allclose(p1[:3], -1e-3 * np.sign(g[:3]), rtol=1e-4)
    assert p1[3] == 0.0


# ---- the C oracle's mixture (the guided estimator's pdf and sampler) ----

def _raw_mixture(rng, n, kappa_r=(-1.0, 4.0)):
    raw = rng.normal(size=(n, G.OUT))
    raw[:, 9:17] = rng.uniform(*kappa_r, (n, 8))
    return raw


def test_c_mixture_pdf_matches_numpy_and_integrates_to_one():
    """or_guide_pdf on the processed mixture equals the numpy oracle's mixture
    pdf (float64) and integrates to 1 over the sphere (Fibonacci points)."""
    from oracle import oracle_py as O
    rng = np.random.default_rng(7)
    raw = _raw_mixture(rng, 64)
    mix = O.processed_mixture(raw)
    w = rng.normal(size=(64, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    pc, _ = O.guide_eval(mix, dirs=w)
    pn = G.guide_pdf(raw, w)
    np.testing.assert_allclose(pc, pn, rtol=2e-5)
    # quadrature on 2^18 Fibonacci points for moderate concentrations
    raw = _raw_mixture(rng, 4, kappa_r=(-2.0, 1.0))
    mix = O.processed_mixture(raw)
    m = 1 << 18
    k = np.arange(m) + 0.5
    z = 1 - 2 * k / m
    ph = np.pi * (1 + 5 ** 0.5) * k
    r = np.sqrt(1 - z * z)
    dirs = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    for row in mix:
        p, _ = O.guide_eval(np.repeat(row[None], m, 0), dirs=dirs)
        assert abs(p.astype(np.float64).mean() * 4 * np.pi - 1.0) < 1e-3


def test_c_mixture_sampler_inverts_the_vmf_cdf():
    """or_guide_sample: unit directions; the lobe the weights' CDF picks from
    u0; cos(angle to that lobe's mean) = 1 + log(u1 + (1 - u1) e^{-2 kappa}) /
    kappa, the vMF's inverse cosine CDF; the sample mean equals the closed form
    sum_i pi_i (coth kappa_i - 1 / kappa_i) mu_i."""
    from oracle import oracle_py as O
    rng = np.random.default_rng(8)
    raw = _raw_mixture(rng, 1)
    mix = O.processed_mixture(raw)
    pi, kappa, mu, _ = G.mixture(raw)
    n = 400_000
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    _, w = O.guide_eval(np.repeat(mix, n, 0), u=u)
    w = w.astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(w, axis=1), 1.0, atol=2e-6)
    cdf = np.cumsum(mix[0, 32:40].astype(np.float32))
    lobe = np.minimum(np.searchsorted(cdf, u[:, 0], side="right"), 7)
    kap = kappa[0, lobe]
    ct = 1 + np.log(u[:, 1] + (1 - u[:, 1]) * np.exp(-2 * kap)) / kap
    got = np.einsum("ij,ij->i", w, mu[0, lobe])
    assert np.abs(got - np.clip(ct, -1, 1)).max() < 2e-4
    mean = (pi[0, :, None] * (1 / np.tanh(kappa[0]) - 1 / kappa[0])[:, None] * mu[0]).sum(0)
    se = w.std(0) / np.sqrt(n)
    assert np.all(np.abs(w.mean(0) - mean) < 5 * se + 1e-4), (w.mean(0), mean)
