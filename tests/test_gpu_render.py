"""Whole-path parity of the GPU wavefront integrator against the oracle.

1. Per sample: GPU and oracle draw the same sample streams, so each path
   sample's radiance must agree (1e-3 relative) except for the rare path whose
   branch decision flips at a floating-point boundary: at least 99.9% of the
   samples agree and the others carry at most 5e-4 of the film's radiance
   (tests/helpers.py check_samples).
2. Closed form: white furnace (a*E unbiased, 2*a*E reference estimator).
3. Images: GPU mean vs oracle mean, per-pixel z-test and relMSE.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import check_samples

pytestmark = pytest.mark.gpu


def _counts(spec, **cfg_kw):
    """Stats of a counting render (PT_RENDER_COUNTERS) of the same frame."""
    from pathtracer_amd import render as R, _lib as L
    g = R.Scene(spec)
    _, st = g.render(g.cfg(flags=L.RENDER_COUNTERS, **cfg_kw))
    return st


def _per_sample(spec, **cfg_kw):
    from pathtracer_amd import render as R, _lib as L
    g = R.Scene(spec)
    o = O.OracleScene(spec)
    cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES, **cfg_kw)
    gimg, st = g.render(cfg)
    gs = g.samples(cfg)
    ocfg = o.cfg(**cfg_kw)
    oimg, _, osmp, cnt = o.render(ocfg, want_samples=True)
    return gimg, gs, oimg, osmp, st


@pytest.mark.parametrize("estimator", [0, 1])
def test_cornell_per_sample(gpu, estimator):
    from pathtracer_amd import scenes
    spec = scenes.cornell(width=48, height=48, spp_x=2, spp_y=2)
    gimg, gs, oimg, os_, st = _per_sample(spec, estimator=estimator, rr_depth=0)
    check_samples(gs, os_, f"cornell est {estimator}")
    assert st["nan_samples"] == 0


@pytest.mark.parametrize("estimator", [0, 1, 2, 3, 4])
def test_integrators_per_sample(gpu, estimator):
    """The reference NEE, NEE+MIS, maxdepth_naive, roulette_naive and
    maxdepth_reservoir_di estimators on the Cornell box (quad light) and on a
    room lit by a sphere light (SphereLight, light.hpp:215-306)."""
    from pathtracer_amd import scenes
    from tests.test_oracle_render import _sphere_lit_room
    for spec in (scenes.cornell(width=32, height=32, spp_x=2, spp_y=2), _sphere_lit_room()):
        _, gs, _, os_, st = _per_sample(spec, estimator=estimator, rr_depth=0, spp_x=2, spp_y=2)
        check_samples(gs, os_, f"{spec.name} est {estimator}")
        st = _counts(spec, estimator=estimator, rr_depth=0, spp_x=2, spp_y=2)
        if estimator in (2, 3):
            assert st["shadow_rays"] == 0
        else:
            assert st["shadow_rays"] > 0


def test_cornell_smooth_normals_per_sample(gpu):
    """Per-corner shading normals (pt_scene_add_mesh_normals) on the gold icosphere."""
    from pathtracer_amd import scenes
    spec = scenes.cornell(width=40, height=40, spp_x=2, spp_y=2, smooth=True)
    _, gs, _, os_, st = _per_sample(spec, rr_depth=0)
    check_samples(gs, os_, "cornell smooth")
    assert st["nan_samples"] == 0


@pytest.mark.parametrize("sampler", [1, 2, 3, 4])
def test_cornell_per_sample_samplers(gpu, sampler):
    from pathtracer_amd import scenes
    spec = scenes.cornell(width=32, height=32, spp_x=2, spp_y=2)
    _, gs, _, os_, _ = _per_sample(spec, sampler=sampler, rr_depth=3)
    check_samples(gs, os_, f"cornell sampler {sampler}")


@pytest.mark.parametrize("estimator", [0, 1])
def test_motion_blur_and_lens_per_sample(gpu, estimator):
    """A camera moving from pos to pos1 while the shutter is open (time uniform
    in [0, shutter)), with a thin lens: every sample's time and lens point are
    drawn from the same stream on both sides."""
    from pathtracer_amd import scenes
    spec = scenes.cornell(width=40, height=40, spp_x=2, spp_y=2)
    spec.camera = dict(spec.camera, pos1=(0.3, 1.2, 3.6), shutter=0.8, focal_dist=3.0, aperture=3.0 / 0.08)
    _, gs, _, os_, st = _per_sample(spec, estimator=estimator, rr_depth=3)
    check_samples(gs, os_, f"motion blur est {estimator}")
    static = scenes.cornell(width=40, height=40, spp_x=2, spp_y=2)
    static.camera = dict(spec.camera, shutter=0.0)
    _, gs0, _, _, _ = _per_sample(static, estimator=estimator, rr_depth=3)
    assert np.mean(np.any(gs0 != gs, axis=-1)) > 0.1      # the shutter changes the paths


def test_glass_hdri_per_sample(gpu):
    """A small stand-in of the glass/HDRI workload; the full one is in
    tests/test_gpu_workloads.py."""
    from pathtracer_amd import scenes
    spec = scenes.glass_hdri(width=32, height=32, spp_x=2, spp_y=2, subdiv=4, env_w=128, env_h=64, max_depth=8)
    for est in (0, 1):
        _, gs, _, os_, _ = _per_sample(spec, estimator=est, rr_depth=0)
        check_samples(gs, os_, f"glass small est {est}")


def test_disney_hair_grid_per_sample(gpu):
    from pathtracer_amd import scenes
    spec = scenes.disney_hair_grid(width=48, height=48, spp_x=2, spp_y=2, max_depth=6)
    for est in (0, 1):
        _, gs, _, os_, st = _per_sample(spec, estimator=est, rr_depth=0)
        check_samples(gs, os_, f"disney grid est {est}")
        assert st["nan_samples"] == 0


@pytest.mark.parametrize("estimator,expect", [(0, 1.0), (1, 0.5)])
def test_white_furnace(gpu, estimator, expect):
    from pathtracer_amd import render as R, scenes
    spec = scenes.furnace(albedo=0.5, env_value=1.0)
    g = R.Scene(spec)
    img, st = g.render(g.cfg(estimator=estimator, max_depth=64, spp_x=32, spp_y=32))
    # centre pixels see only the sphere
    c = img[3:5, 3:5].mean()
    assert abs(c - expect) < 0.02 * expect


def test_mis_matches_reference_expectation(gpu):
    """Where the reference estimator is unbiased (area lights only), the NEE+MIS
    estimator must converge to the same image."""
    from pathtracer_amd import render as R, scenes
    spec = scenes.cornell(width=24, height=24, spp_x=32, spp_y=32)
    g = R.Scene(spec)
    ref, _ = g.render(g.cfg(estimator=0, rr_depth=0, seed=1))
    mis, _ = g.render(g.cfg(estimator=1, rr_depth=0, seed=2))
    # per-channel image means agree to 1% and pixels agree on average
    assert np.allclose(ref.mean(axis=(0, 1)), mis.mean(axis=(0, 1)), rtol=0.01)
    assert abs(np.mean(mis - ref)) < 0.01 * np.mean(ref)


def test_cornell_image_statistics(gpu):
    """GPU at many samples vs oracle at fewer: per-pixel z-test and relMSE of the means."""
    from pathtracer_amd import render as R, scenes
    spec = scenes.cornell(width=32, height=32, spp_x=16, spp_y=16)
    g = R.Scene(spec)
    gimg, _ = g.render(g.cfg(rr_depth=3, seed=101))
    o = O.OracleScene(spec)
    oimg, ovar, _, _ = o.render(o.cfg(rr_depth=0, seed=202), want_var=True)
    n = 256
    # both means carry sampling noise of about the oracle's variance
    se = np.sqrt(2.0 * np.maximum(ovar, 1e-12) / n)
    z = (gimg - oimg) / se
    assert np.abs(z).mean() < 1.0
    assert (np.abs(z) > 5).mean() < 0.01


@pytest.mark.parametrize("which", ["absorption", "scatter", "scatter_hg"])
def test_volume_per_sample(gpu, which):
    """maxdepth_volume (single scattering in a homogeneous medium) per sample."""
    from pathtracer_amd import scenes
    if which == "absorption":
        spec = scenes.absorption_only_medium(width=40, height=40, spp_x=2, spp_y=2)
    else:
        spec = scenes.single_scatter_medium(width=40, height=40, spp_x=2, spp_y=2, g=0.6 if which == "scatter_hg" else 0.0)
    _, gs, _, os_, st = _per_sample(spec)
    check_samples(gs, os_, f"volume {which}")
    assert st["nan_samples"] == 0


@pytest.mark.parametrize("estimator", [0, 1])
@pytest.mark.parametrize("kk", [False, True])
def test_hair_curves_per_sample(gpu, estimator, kk):
    """Hair on Bezier curves.  Marschner: the curve's tangent frames the hair
    sample (HairMarschner::scatter), the normal frames f and the pdf.
    Kajiya-Kay: the curve's tangent is the fibre axis of f."""
    from pathtracer_amd import scenes
    spec = scenes.hair_curves(width=40, height=40, spp_x=2, spp_y=2, estimator=estimator, kajiya_kay=kk)
    _, gs, _, os_, st = _per_sample(spec, estimator=estimator, rr_depth=0)
    moved = None
    if kk:
        # exponent-40 cones on fibres 0.008-0.03 wide: a 1e-6 camera move
        # changes about 10% of the oracle's own samples here
        sp2 = scenes.hair_curves(width=40, height=40, spp_x=2, spp_y=2, estimator=estimator, kajiya_kay=kk)
        sp2.camera = dict(sp2.camera, pos=tuple(np.asarray(sp2.camera["pos"]) + 1e-6))
        o2 = O.OracleScene(sp2)
        _, _, moved, _ = o2.render(o2.cfg(estimator=estimator, rr_depth=0), want_samples=True)
    check_samples(gs, os_, f"{'kajiya-kay' if kk else 'marschner'} curves est {estimator}", self_moved=moved)
    assert st["nan_samples"] == 0


@pytest.mark.parametrize("estimator", [0, 1])
def test_sorted_shading_same_samples(gpu, estimator):
    """PT_RENDER_SORT_HITS only changes which lane shades which path: every
    sample's radiance is the same as in queue order (up to the order of the
    float atomics of a vertex's several shadow rays)."""
    from pathtracer_amd import render as R, scenes, _lib as L
    for spec in (scenes.cornell(width=48, height=48, spp_x=2, spp_y=2),
                 scenes.disney_hair_grid(width=64, height=64, spp_x=2, spp_y=2)):
        g = R.Scene(spec)
        cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES, estimator=estimator)
        g.render(cfg)
        a = g.samples(cfg)
        cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES | L.RENDER_SORT_HITS, estimator=estimator)
        g.render(cfg)
        b = g.samples(cfg)
        np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("which", ["cornell", "glass_hdri"])
def test_image_4096spp_against_oracle(gpu, which):
    """Whole images at 4096 spp: the GPU film against the oracle's film of the
    same configuration and seed (the reference's NEE estimator, no Russian
    roulette, so every sample is the same path): relMSE <= 1e-3 and no
    per-pixel bias by z-test against the oracle's per-pixel sample variance."""
    from pathtracer_amd import render as R, scenes
    if which == "cornell":
        spec = scenes.cornell(width=32, height=32, spp_x=64, spp_y=64)
    else:
        spec = scenes.glass_hdri(width=24, height=24, spp_x=64, spp_y=64, subdiv=4, env_w=128, env_h=64,
                                 max_depth=8)
    g = R.Scene(spec)
    gimg, st = g.render(g.cfg(estimator=0, rr_depth=0, seed=5))
    o = O.OracleScene(spec)
    oimg, ovar, _, _ = o.render(o.cfg(estimator=0, rr_depth=0, seed=5), want_var=True)
    gimg, oimg = gimg.astype(np.float64), oimg.astype(np.float64)
    rel = float(np.mean((gimg - oimg) ** 2 / (oimg ** 2 + 1e-2)))
    assert rel <= 1e-3, rel
    se = np.sqrt(np.maximum(ovar.astype(np.float64), 1e-12) / 4096)
    z = (gimg - oimg) / se
    assert np.abs(z).mean() < 0.5, np.abs(z).mean()
    assert st["nan_samples"] == 0
