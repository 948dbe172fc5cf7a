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
