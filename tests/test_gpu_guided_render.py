"""Guided rendering (pt_render_guided) on the slit-lit room.

1. With a kinds mask that guides nothing, the guided estimator is the NEE+MIS
   estimator path for path (it draws no extra random numbers), so the two
   films agree to float-atomic order.
2. Unbiased: guided and plain NEE+MIS films agree in 6x6-pixel block means to
   within their estimated noise (z-test over the blocks).
3. What the trained network gains: at equal samples per pixel, the guided
   film's relMSE against a high-spp plain reference is below the plain
   film's and below that of the same method with an untrained network.
   The relMSE rejects the 0.1% largest per-pixel errors: a handful of
   fireflies from the slit light move the full relMSE by 30% between seeds
   of the same method, and training is not bit-reproducible (its samples are
   collected in arrival order), so three films of each method are averaged.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = H = 36


@pytest.fixture(scope="module")
def room(gpu):
    from pathtracer_amd import render as R, scenes
    return R.Scene(scenes.guiding_room(width=W, height=H, spp_x=16, spp_y=16))


def _plain(room, spp_side, seed):
    img, st = room.render(room.cfg(spp_x=spp_side, spp_y=spp_side, seed=seed))
    return img.astype(np.float64), st


def _guided(room, spp_side, seed, gseed=1, **gk):
    from pathtracer_amd import render as R, _lib as L
    g = R.Guide(room, seed=gseed)
    gc = L.GuideCfg()
    for k, v in gk.items():
        setattr(gc, k, v)
    img, st, gs = g.render(room.cfg(spp_x=spp_side, spp_y=spp_side, seed=seed), gc)
    g.close()
    return img.astype(np.float64), st, gs


def test_unguided_kinds_equal_plain(room):
    from pathtracer_amd import _lib as L
    plain, _ = _plain(room, 4, 3)
    img, _, gs = _guided(room, 4, 3, kinds=1 << L.MAT_EMITTER, flags=L.GUIDE_NO_TRAIN)
    assert gs["steps"] == 0
    np.testing.assert_allclose(img, plain, rtol=1e-5, atol=1e-6)


def _blocks(img, b=6):
    h, w, _ = img.shape
    return img.reshape(h // b, b, w // b, b, 3).mean((1, 3))


def test_guided_is_unbiased(room):
    """Four independent films of each method: per 6x6 block, the difference of
    the method means over its standard error (from the replicate spread)."""
    P = np.stack([_blocks(_plain(room, 24, 11 + k)[0]) for k in range(4)])
    runs = [_guided(room, 24, 21 + k, gseed=1 + k, train_spp=96) for k in range(4)]
    assert all(r[2]["steps"] > 0 and r[2]["samples"] > 0 and np.isfinite(r[2]["last_loss"]) for r in runs)
    Gm = np.stack([_blocks(r[0]) for r in runs])
    var = P.var(0, ddof=1) / 4 + Gm.var(0, ddof=1) / 4
    z = (Gm.mean(0) - P.mean(0)) / np.sqrt(var.mean(-1, keepdims=True) + 1e-12)
    assert np.mean(z ** 2) < 2.5, np.mean(z ** 2)
    # the whole-image means
    pm, gm = P.mean((1, 2, 3)), Gm.mean((1, 2, 3))
    zi = (gm.mean() - pm.mean()) / np.sqrt(pm.var(ddof=1) / 4 + gm.var(ddof=1) / 4)
    assert abs(zi) < 4, zi


def _relmse(img, ref, keep=0.999):
    """relMSE of the films scaled to unit mean (the room is dim), over the
    pixels whose error is below the 99.9th percentile."""
    mu = ref.mean()
    e = (((img - ref) / mu) ** 2 / ((ref / mu) ** 2 + 1e-2)).mean(-1).ravel()
    k = int(keep * e.size)
    return float(np.partition(e, k)[:k].mean())


def test_trained_guide_reduces_error(gpu):
    """At 256x256 and 256 spp (32 of them while training): guided below the
    plain tracer and below the same method with an untrained network.  (On
    six seeds each: plain 0.297-0.316, guided 0.247-0.281, untrained
    0.340-0.367.)"""
    from pathtracer_amd import render as R, scenes, _lib as L
    room = R.Scene(scenes.guiding_room(width=256, height=256, spp_x=16, spp_y=16))
    ref = sum(_plain(room, 64, 900 + k)[0] for k in range(8)) / 8   # 32768 spp
    plain = np.mean([_relmse(_plain(room, 16, s)[0], ref) for s in (21, 22, 23)])
    guided = np.mean([_relmse(_guided(room, 16, s, gseed=s)[0], ref) for s in (31, 32, 33)])
    untrained = np.mean([_relmse(_guided(room, 16, s, gseed=s, flags=L.GUIDE_NO_TRAIN)[0], ref)
                         for s in (41, 42, 43)])
    print(f"relMSE (0.1% outliers rejected) at 256 spp: plain {plain:.4g}, guided {guided:.4g}, "
          f"untrained guide {untrained:.4g}")
    assert guided < 0.95 * plain
    assert guided < 0.9 * untrained


def _constant_guide(scene, mix_raw):
    """A guide whose network outputs mix_raw at every point: every weight zero,
    the output layer's biases (the last 41 parameters) set to it."""
    from pathtracer_amd import render as R, _lib as L
    g = R.Guide(scene, seed=1)
    p = np.zeros(L.GUIDE_NPARAMS, np.float32)
    p[-L.GUIDE_OUTPUTS:] = mix_raw
    g.set_params(p)
    return g


def _mix_raw(seed, c_logit):
    rng = np.random.default_rng(seed)
    raw = np.zeros(41, np.float32)
    raw[0:8] = rng.normal(size=8)
    raw[8] = c_logit
    raw[9:17] = rng.uniform(0.0, 4.0, 8)                 # kappa ~ 4.5 .. 160
    d = rng.normal(size=(8, 3))
    d[0] = (-1.0, 0.1, 0.0)                               # towards the slit wall
    raw[17:25], raw[25:33], raw[33:41] = d[:, 0], d[:, 1], d[:, 2]
    return raw


@pytest.mark.parametrize("scene_name,verts,c_logit", [("guiding_room", 2, 0.0), ("guiding_room", 16, 1.5),
                                                      ("cornell", 8, -0.5)])
def test_guided_estimator_per_sample_against_oracle(gpu, scene_name, verts, c_logit):
    """The guided estimator (k_shade's guided variant) against the oracle's
    restatement, sample for sample, with a network whose output is one fixed
    mixture: technique choice, vMF sampling, the one-sample MIS pdf in the
    throughput, NEE's power heuristic and the next vertex's emitter MIS.  The
    device's processed mixture first matches the oracle's (numpy) processing
    of the same raw outputs."""
    from oracle import oracle_py as O
    from pathtracer_amd import render as R, scenes, _lib as L
    from tests.helpers import check_samples
    spec = getattr(scenes, scene_name)(width=40, height=40, spp_x=2, spp_y=2)
    sc = R.Scene(spec)
    raw = _mix_raw(17 + verts, c_logit)
    g = _constant_guide(sc, raw)
    x = np.random.default_rng(0).uniform(0, 1, (64, 6)).astype(np.float32)
    graw, gmix = g.forward(x)
    np.testing.assert_allclose(graw, np.repeat(raw[None], 64, 0), rtol=0, atol=0)
    omix = O.processed_mixture(raw[None].astype(np.float64))[0]
    np.testing.assert_allclose(gmix[:, :41], np.repeat(omix[None, :41], 64, 0), rtol=2e-5, atol=2e-6)
    kinds = (1 << L.MAT_LAMBERTIAN) | (1 << L.MAT_CONDUCTOR) | (1 << L.MAT_DISNEY)
    gc = L.GuideCfg()
    gc.flags = L.GUIDE_NO_TRAIN
    gc.kinds = kinds
    gc.guide_verts = verts
    cfg = sc.cfg(flags=L.RENDER_KEEP_SAMPLES, estimator=L.EST_NEE_MIS)
    g.render(cfg, gc)
    gs = sc.samples(cfg)
    g.close()
    o = O.OracleScene(spec)
    o.set_guide(omix, kinds, verts)
    _, _, os_, _ = o.render(o.cfg(estimator=O.EST_GUIDED), want_samples=True)
    check_samples(gs, os_, f"guided {scene_name} verts {verts} c_logit {c_logit}")
    # the guide was used: the guided film differs from the plain tracer's samples
    cfg_p = sc.cfg(flags=L.RENDER_KEEP_SAMPLES, estimator=L.EST_NEE_MIS)
    sc.render(cfg_p)
    assert not np.allclose(sc.samples(cfg_p), gs)


@pytest.mark.parametrize("spp_begin", [0, 1])
def test_pixel_shared_query_matches_per_path_query(gpu, spp_begin):
    """The camera vertices of a pixel share one network query, at the hit of
    the pixel's first sample in the chunk (the default), against one query
    per path (PT_GUIDE_PATH_QUERIES), with a randomly initialised network:
    each pixel's first sample of the pass is the same path in both, sample
    for sample; chunks of 37 pixels exercise the chunk offsets, spp_begin 1
    a pass that starts past sample 0.  The other samples use their pixel's
    shared mixture instead of their own, so some of them differ."""
    from pathtracer_amd import render as R, scenes, _lib as L
    sc = R.Scene(scenes.guiding_room(width=24, height=24, spp_x=2, spp_y=2))
    g = R.Guide(sc, seed=5)
    out = []
    for extra in (0, L.GUIDE_PATH_QUERIES):
        gc = L.GuideCfg()
        gc.flags = L.GUIDE_NO_TRAIN | extra
        gc.guide_verts = 2
        cfg = sc.cfg(flags=L.RENDER_KEEP_SAMPLES, estimator=L.EST_NEE_MIS, spp_begin=spp_begin,
                     wavefront=37 * (4 - spp_begin), seed=7)
        g.render(cfg, gc)
        out.append(sc.samples(cfg))
    g.close()
    shared, per_path = out
    assert shared.shape == (24 * 24, 4 - spp_begin, 3)
    np.testing.assert_allclose(shared[:, 0], per_path[:, 0], rtol=1e-6, atol=1e-7)
    assert np.abs(shared[:, 0]).sum() > 0
    assert not np.allclose(shared[:, 1:], per_path[:, 1:], rtol=1e-5, atol=1e-7)
