"""Batched rendering (pt_batch_*): many renders of uneven resolution, spp,
depth and scene size in one wavefront.

A path sample's radiance depends only on its render's scene, configuration
and (pixel, sample) index, never on where in the wavefront it runs, and the
film sums a pixel's samples in a fixed order (in sample order in a batch, by
a fixed lane tree in a single render).  So every film of a batch must equal
the film of the same render done alone with pt_render, up to that order and
the order of the float atomics that add shadow-ray contributions (rtol 1e-5).  This
holds for eager launches and for every replay of the captured chunk graph.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-6


def _close(a, b):
    return np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(a), np.abs(b)) + ATOL)


def _mixed_renders(estimator):
    """Renders that differ in everything a batch must keep apart."""
    from pathtracer_amd import render as R, scenes, _lib as L
    rng = np.random.default_rng(5)
    out = []
    for i in range(3):
        spec = scenes.batch_instance(rng, i, tri_range=(500, 4000), res_choices=(16, 24, 40), spp_range=(1, 12),
                                     depth_range=(3, 9), estimator=estimator)
        out.append((R.Scene(spec), {}))
    corn = R.Scene(scenes.cornell(width=20, height=12, spp_x=2, spp_y=3, estimator=estimator))
    out.append((corn, {}))
    out.append((corn, dict(seed=99, max_depth=5, sampler=L.SAMPLER_HALTON_PERM)))  # same scene, other settings
    out.append((corn, dict(seed=7, sampler=L.SAMPLER_HALTON_PERM, spp_begin=1, spp_count=3)))
    fur = R.Scene(scenes.furnace(width=8, height=8, spp_x=3, spp_y=3))
    out.append((fur, dict(estimator=estimator)))
    return [(s, s.cfg(**dict(dict(estimator=estimator), **kw))) for s, kw in out]


def _alone(renders):
    imgs, nan = [], 0
    for s, cfg in renders:
        img, st = s.render(cfg)
        imgs.append(img)
        nan += st["nan_samples"]
    return imgs, nan


@pytest.mark.parametrize("estimator", [1, 0, 4])
def test_batch_matches_single_renders(gpu, estimator):
    from pathtracer_amd import render as R, _lib as L
    renders = _mixed_renders(estimator)
    ref, nan = _alone(renders)
    b = R.Batch()
    for s, cfg in renders:
        b.add(s, cfg)
    # a small wavefront: chunks span several renders and renders span chunks
    for wavefront, flags in ((700, L.BATCH_NO_GRAPH), (700, 0), (1 << 20, 0)):
        st = b.render(wavefront=wavefront, flags=flags | L.RENDER_COUNTERS)
        assert st["paths"] == sum(c.width * c.height * (c.spp_count or c.spp_x * c.spp_y - c.spp_begin)
                                  for _, c in renders)
        # the reservoir estimator gives NaN samples under an environment light
        # (the oracle's restatement too); the film drops them (framebuffer.cpp:27-58)
        assert st["nan_samples"] == nan
        for i, r in enumerate(ref):
            got = b.film(i)
            assert _close(got, r), (estimator, wavefront, flags, i, float(np.abs(got - r).max()))
    b.close()


def test_batch_graph_replays(gpu):
    """The chunk graph is captured once and replayed for every chunk of every
    render call; each replay sequence must give the eager result, with eager
    work and synchronisation in between."""
    from pathtracer_amd import render as R, _lib as L
    renders = _mixed_renders(1)
    b = R.Batch()
    for s, cfg in renders:
        b.add(s, cfg)
    b.render(wavefront=900, flags=L.BATCH_NO_GRAPH)
    eager = [b.film(i) for i in range(len(renders))]
    for rep in range(3):
        b.render(wavefront=900)
        for i, e in enumerate(eager):
            assert _close(b.film(i), e), (rep, i)
        renders[3][0].render(renders[3][1])  # unrelated eager work on another stream
    # a changed configuration (seed) replays the same graph with new tables
    c = renders[0][1]
    c.seed = 1234
    b.set_cfg(0, c)
    b.render(wavefront=900)
    alone, _ = renders[0][0].render(c)
    assert _close(b.film(0), alone)
    b.close()


def test_batch_accumulate(gpu):
    from pathtracer_amd import render as R, scenes, _lib as L
    s = R.Scene(scenes.cornell(width=16, height=16, spp_x=2, spp_y=2))
    c0 = s.cfg(spp_begin=0, spp_count=2)
    c1 = s.cfg(spp_begin=2, spp_count=2, flags=L.RENDER_ACCUMULATE)
    full, _ = s.render(s.cfg())
    b = R.Batch()
    b.add(s, c0)
    b.render()
    b.set_cfg(0, c1)
    b.render()
    assert _close(b.film(0), full)
    b.close()


def test_batch_errors(gpu):
    from pathtracer_amd import render as R, scenes, _lib as L
    s = R.Scene(scenes.cornell(width=8, height=8, spp_x=1, spp_y=1))
    b = R.Batch()
    with pytest.raises(L.PTError):
        b.render()  # empty
    with pytest.raises(L.PTError):
        b.add(s, s.cfg(flags=L.RENDER_KEEP_SAMPLES))
    b.add(s, s.cfg(estimator=0))
    b.add(s, s.cfg(estimator=1))
    with pytest.raises(L.PTError, match="one estimator"):
        b.render()
    with pytest.raises(L.PTError):
        b.film(5)
    b.close()


def test_autotune(gpu):
    """The tuner returns values from its candidate sets, and a render with the
    tuned launch parameters (same BVH) gives the same image."""
    from pathtracer_amd import render as R, scenes, _lib as L
    s = R.Scene(scenes.cornell(width=64, height=64, spp_x=2, spp_y=2))
    cfg = s.cfg()
    before, _ = s.render(cfg)
    t = s.autotune(cfg, flags=L.TUNE_WAVEFRONT | L.TUNE_GRIDS, spp_div=2)
    assert t["wavefront"] in (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)
    assert t["trav_blocks_per_cu"] in (4, 6, 8)
    assert t["shade_blocks_per_cu"] in (8, 16, 32, 64)
    after, _ = s.render(cfg)
    assert _close(before, after)
    t = s.autotune(cfg, flags=L.TUNE_LEAF, spp_div=2)
    assert t["leaf_size"] in (1, 2, 4, 8)
    relit, _ = s.render(cfg)
    # another BVH can only change ties between equal hits: the means agree
    assert abs(relit.mean() - before.mean()) < 1e-3 * before.mean()

    b = R.Batch()
    renders = _mixed_renders(1)
    for r, c in renders:
        b.add(r, c)
    b.render()
    films = [b.film(k) for k in range(len(renders))]
    bt = b.autotune(spp_div=2)
    assert bt["wavefront"] in (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)
    assert bt["leaf_size"] in (1, 2, 4, 8)
    # the batch's scenes stay committed with the chosen leaf size: the same
    # films up to ties between equal hits
    b.render()
    for k in range(len(renders)):
        f = b.film(k)
        assert abs(f.mean() - films[k].mean()) <= 1e-3 * films[k].mean() + 1e-6
    b.close()
