"""CPU restatement of the learned path-guiding network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device's
guide kernels (pathtracer_amd/csrc/pt_guide.hip); the product never imports
it.  The guide has no counterpart in the reference, so this module is pinned
by mathematics alone (tests/test_oracle_guide.py): the analytic gradient by
central finite differences in float64, the vMF mixture by its normalisation
over the sphere and by the distribution of its samples, the encoding by its
support and partition-of-unity-like sum.

Model (the same as pathtracer_amd/csrc/pt_guide.h):
  x (6 values in [0,1]) -> one-blob encoding, 32 bins per value, feature
  (v, b) = bump(b - (32 x_v - 0.5)), bump(t) = max(0, 1 - t^2/9)^2
  -> 192 -> 64 -> 64 -> 64 -> 64 (ReLU after each) -> 41 raw outputs:
  o[0..7] lobe logits, o[8] selection logit, o[9..16] concentrations r,
  o[17..40] mean directions (x of lobes 0..7, y, z).
  pi = softmax, kappa = exp(1.5 + 6.5 tanh(r / 6.5)), mu = normalize,
  c = C_LO + C_SPAN sigmoid(s) (the guide's default range [0.05, 0.5]).
  p_g(w) = sum_i pi_i kappa_i / (2 pi (1 - exp(-2 kappa_i))) exp(kappa_i (mu_i.w - 1)).
Loss of a sample (x, w, p_bsdf, weight F/q), averaged over the batch:
  -weight * (log p_g(w) + log(c p_g(w) + (1 - c) p_bsdf)).

emulate=True rounds as the device does: weights, layer inputs and output
gradients to BF16 (round to nearest even), products accumulated in FP32,
biases and bias gradients in FP32.
"""
from __future__ import annotations

import numpy as np

IN, BINS, H, OUT, LOBES = 6, 32, 64, 41, 8
K = [IN * BINS, H, H, H, H]
N = [H, H, H, H, OUT]
NPARAMS = sum(k * n + n for k, n in zip(K, N))
C_LO, C_SPAN = 0.05, 0.45


def w_offset(l):
    return sum(K[i] * N[i] + N[i] for i in range(l))


def b_offset(l):
    return w_offset(l) + K[l] * N[l]


def unpack(params):
    """Views W_l (K_l x N_l) and b_l (N_l) of the flat parameter vector."""
    Ws, bs = [], []
    for l in range(5):
        o = w_offset(l)
        Ws.append(params[o:o + K[l] * N[l]].reshape(K[l], N[l]))
        bs.append(params[o + K[l] * N[l]:o + K[l] * N[l] + N[l]])
    return Ws, bs


def bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def encode(x, dtype=np.float32):
    """(n, 6) -> (n, 192): feature v*32 + b."""
    x = np.asarray(x, dtype)
    y = dtype(32.0) * x - dtype(0.5)
    b = np.arange(BINS, dtype=dtype)
    t = b[None, None, :] - y[:, :, None]
    u = np.maximum(dtype(0), dtype(1) - t * t * dtype(1.0 / 9.0))
    return (u * u).reshape(x.shape[0], IN * BINS)


def forward(params, x, emulate=True):
    """Raw outputs (n, 41) and the per-layer inputs / pre-activations."""
    dt = np.float32 if emulate else np.float64
    Ws, bs = unpack(np.asarray(params, dt))
    h = encode(x, dt)
    xs, zs = [], []
    for l in range(5):
        hin = bf16(h) if emulate else h
        W = bf16(Ws[l]) if emulate else Ws[l]
        z = (hin.astype(dt) @ W.astype(dt)) + bs[l]
        xs.append(hin)
        zs.append(z)
        h = np.maximum(z, 0)
    return zs[-1], xs, zs


def mixture(raw):
    """pi (n,8), kappa (n,8), mu (n,8,3), c (n,), plus the raw pieces."""
    raw = np.asarray(raw, np.float64)
    a = raw[:, 0:8]
    s = raw[:, 8]
    r = raw[:, 9:17]
    m = np.stack([raw[:, 17:25], raw[:, 25:33], raw[:, 33:41]], -1)
    a = a - a.max(1, keepdims=True)
    e = np.exp(a)
    pi = e / e.sum(1, keepdims=True)
    kappa = np.exp(1.5 + 6.5 * np.tanh(r / 6.5))
    ln = np.maximum(np.linalg.norm(m, axis=-1), 1e-12)
    mu = m / ln[..., None]
    sg = 1.0 / (1.0 + np.exp(-s))
    c = C_LO + C_SPAN * sg
    return pi, kappa, mu, c


def log_norm(kappa):
    return np.log(kappa) - np.log(2 * np.pi) - np.log1p(-np.exp(-2 * kappa))


def vmf(kappa, mu, w):
    """vMF densities (n, 8) at directions w (n, 3)."""
    d = np.einsum("nij,nj->ni", mu, w)
    return np.exp(log_norm(kappa) + kappa * (d - 1.0))


def guide_pdf(raw, w):
    pi, kappa, mu, _ = mixture(raw)
    return (pi * vmf(kappa, mu, np.asarray(w, np.float64))).sum(1)


def processed(raw):
    """The renderer's 48-float form: (kappa mu, log(pi C(kappa)) - kappa) x 8, pi x 8, c."""
    pi, kappa, mu, c = mixture(raw)
    n = raw.shape[0]
    out = np.zeros((n, 48))
    for i in range(LOBES):
        out[:, 4 * i:4 * i + 3] = kappa[:, i:i + 1] * mu[:, i]
        with np.errstate(divide="ignore"):
            out[:, 4 * i + 3] = np.log(pi[:, i]) + log_norm(kappa[:, i]) - kappa[:, i]
        out[:, 32 + i] = pi[:, i]
    out[:, 40] = c
    return out


def processed_pdf(mix, w):
    """Guide pdf from the processed form (what the renderer evaluates)."""
    mix = np.asarray(mix, np.float64)
    w = np.asarray(w, np.float64)
    p = np.zeros(mix.shape[0])
    for i in range(LOBES):
        p += np.exp(mix[:, 4 * i] * w[:, 0] + mix[:, 4 * i + 1] * w[:, 1] + mix[:, 4 * i + 2] * w[:, 2] +
                    mix[:, 4 * i + 3])
    return p


def sample(mix, u):
    """Directions from the processed form and uniforms u (n, 3): lobe by pi
    with u0, cos(theta) = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa about
    mu, phi = 2 pi u2 (any frame about mu)."""
    mix = np.asarray(mix, np.float64)
    u = np.asarray(u, np.float64)
    n = mix.shape[0]
    cdf = np.cumsum(mix[:, 32:40], 1)
    lobe = np.minimum((u[:, :1] >= cdf[:, :7]).sum(1), 7)
    m = np.stack([mix[np.arange(n), 4 * lobe + k] for k in range(3)], 1)
    kappa = np.linalg.norm(m, axis=1)
    mu = m / kappa[:, None]
    ct = 1.0 + np.log(u[:, 1] + (1.0 - u[:, 1]) * np.exp(-2 * kappa)) / kappa
    ct = np.clip(ct, -1, 1)
    st = np.sqrt(np.maximum(0, 1 - ct * ct))
    ph = 2 * np.pi * u[:, 2]
    a = np.where((np.abs(mu[:, 0]) > 0.9)[:, None], np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t1 = np.cross(mu, a)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(mu, t1)
    return t1 * (st * np.cos(ph))[:, None] + t2 * (st * np.sin(ph))[:, None] + mu * ct[:, None]


def loss_and_grad(params, samples, emulate=True):
    """Mean loss and its gradient (flat, NPARAMS) over samples (n, 11)."""
    samples = np.asarray(samples, np.float64)
    n = samples.shape[0]
    x, w, pb, wt = samples[:, :6], samples[:, 6:9], samples[:, 9], samples[:, 10]
    raw, xs, zs = forward(params, x, emulate)
    raw64 = raw.astype(np.float64)
    pi, kappa, mu, c = mixture(raw64)
    d = np.einsum("nij,nj->ni", mu, w)
    with np.errstate(divide="ignore"):
        lj = np.log(pi) + log_norm(kappa) + kappa * (d - 1)          # log(pi_i v_i)
    lmax = lj.max(1, keepdims=True)
    log_pg = (lmax + np.log(np.exp(lj - lmax).sum(1, keepdims=True)))[:, 0]
    pg = np.exp(log_pg)
    pm = c * pg + (1 - c) * pb
    with np.errstate(divide="ignore"):
        log_pm = np.where(pb > 0, np.log(np.where(pm > 0, pm, 1.0)), np.log(c) + log_pg)
    rho = np.exp(np.log(c) + log_pg - log_pm)
    live = (wt > 0) & np.isfinite(log_pg) & np.isfinite(log_pm)
    loss = np.where(live, -wt * (log_pg + log_pm), 0.0).sum() / n
    coef = np.where(live, -wt * (1 + rho) / n, 0.0)               # times d log p_g
    resp = np.exp(lj - log_pg[:, None])
    dout = np.zeros((n, OUT))
    dout[:, 0:8] = coef[:, None] * (resp - pi)
    s = raw64[:, 8]
    sg = 1 / (1 + np.exp(-s))
    dsel = np.where(live, -wt / n * (rho / c - (1 - rho) / (1 - c)), 0.0)
    dout[:, 8] = dsel * C_SPAN * sg * (1 - sg)
    e2 = np.exp(-2 * kappa)
    dlv_dk = 1 / kappa - 2 * e2 / (1 - e2) + d - 1
    r = raw64[:, 9:17]
    dout[:, 9:17] = coef[:, None] * resp * dlv_dk * kappa * (1 - np.tanh(r / 6.5) ** 2)
    m = np.stack([raw64[:, 17:25], raw64[:, 25:33], raw64[:, 33:41]], -1)
    ln = np.maximum(np.linalg.norm(m, axis=-1), 1e-12)
    gm = coef[:, None] * resp * kappa / ln
    for k in range(3):
        dout[:, 17 + 8 * k:25 + 8 * k] = gm * (w[:, None, k] - d * mu[:, :, k])
    dt = np.float32 if emulate else np.float64
    Ws, _ = unpack(np.asarray(params, dt))
    grad = np.zeros(NPARAMS, np.float64)
    dz = dout.astype(dt)
    for l in range(4, -1, -1):
        dzr = bf16(dz) if emulate else dz
        o = w_offset(l)
        grad[o:o + K[l] * N[l]] = (xs[l].astype(dt).T @ dzr.astype(dt)).reshape(-1)
        grad[o + K[l] * N[l]:o + K[l] * N[l] + N[l]] = dz.sum(0)
        if l == 0:
            break
        W = bf16(Ws[l]) if emulate else Ws[l]
        dx = dzr.astype(dt) @ W.astype(dt).T
        dz = dx * (zs[l - 1] > 0)
    return float(loss), grad


def adam(params, m, v, grad, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def init_like_device(seed):
    """Not the device's initial weights (those are seeded on the host):
    a He-uniform draw of the same shapes, for tests that set parameters."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    for l in range(5):
        a = np.sqrt(6.0 / K[l]) * (0.3 if l == 4 else 1.0)
        o = w_offset(l)
        p[o:o + K[l] * N[l]] = rng.uniform(-a, a, K[l] * N[l])
        p[o + K[l] * N[l]:o + K[l] * N[l] + N[l]] = rng.uniform(-0.1, 0.1, N[l])
    return p
