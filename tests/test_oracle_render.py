"""Pins the oracle's integrator (CPU only) with closed forms.

* Quad light over a Lambertian floor, one bounce: radiance = albedo * Le * F,
  F the point-to-parallel-rectangle form factor -- pins NEE sampling, the
  geometric term, pdf = 1/(A*N) and the one-sided emission (maxdepth_nee.hpp:147-231).
* White furnace: a*E for an unbiased estimator; 2*a*E for the reference
  estimator, which counts the environment through NEE and again through the
  BSDF ray that misses (maxdepth_nee.hpp:26-37, 107-145).
* The reference estimator and NEE+MIS agree on an area-lit scene.
* The mixture-pdf integrators (maxdepth_naive.hpp, roulette_naive.hpp): the
  same form factor through BSDF/shape-pdf mixture sampling (pins the rect
  pdf d^2/(cos A), the mixture weights and the roulette 1/q), the depth limit
  (bg_color for naive, 0 for roulette) and agreement with the reference
  estimator on a scene lit by a sphere light (pins the cone pdf
  1/(2 pi (1 - cos theta_max)) and its sampler).
* The reservoir integrator (maxdepth_reservoir_di.hpp): over the floor, its
  estimate is the closed form of what it computes -- the selection keeps
  candidate i with probability w_i / W and scales it by W / M, so the
  expectation is E_x[|c(x)| c(x)] for one candidate c = f Le A G (the
  vector length |c| = sqrt(3) c for grey values), times the probability
  (1 - F) that the continuation ray misses the light, which would make the
  vertex return 0.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from pathtracer_amd import scenes


def _corner_form_factor(a, b, h):
    X, Y = a / h, b / h
    return (X / np.sqrt(1 + X * X) * np.arctan(Y / np.sqrt(1 + X * X)) +
            Y / np.sqrt(1 + Y * Y) * np.arctan(X / np.sqrt(1 + Y * Y))) / (2 * np.pi)


def floor_under_light(A=1.0, B=0.6, h=1.5, albedo=0.6, Le=5.0):
    s = scenes.SceneSpec(name="floor")
    fl = s.add_material(1, [albedo] * 3)
    lm = s.add_material(2, [1.0, 1.0, 1.0, Le])
    s.add_rect(1, 0.0, -20, 20, -20, 20, fl)
    s.add_rect(1, h, -A / 2, A / 2, -B / 2, B / 2, lm)
    s.camera = dict(fov_deg=0.2, aspect=1, focal_dist=1, aperture=0, up=(0, 0, 1), front=(0, -1, 0),
                    pos=(0, 0.5, 0))
    s.render = dict(width=1, height=1, spp_x=64, spp_y=64, max_depth=1, sampler=0, estimator=0, rr_depth=0,
                    light_samples=4, seed=1, bg=(0, 0, 0))
    F = 4 * _corner_form_factor(A / 2, B / 2, h)
    return s, albedo * Le * F


@pytest.mark.parametrize("estimator", [0, 1])
def test_quad_light_form_factor(estimator):
    spec, expect = floor_under_light()
    o = O.OracleScene(spec)
    img, var, _, _ = o.render(o.cfg(estimator=estimator), want_var=True)
    se = np.sqrt(var[0, 0, 0] / 4096)
    assert abs(img[0, 0, 0] - expect) < max(4 * se, 2e-3 * expect)


@pytest.mark.parametrize("estimator,factor", [(0, 2.0), (1, 1.0)])
def test_white_furnace(estimator, factor):
    spec = scenes.furnace(albedo=0.5, env_value=1.0)
    o = O.OracleScene(spec)
    img, var, _, _ = o.render(o.cfg(estimator=estimator, max_depth=64, spp_x=32, spp_y=32), want_var=True)
    c = img[3:5, 3:5, 0].mean()
    assert abs(c - 0.5 * factor) < 0.01 * factor


def test_depth_cutoff_returns_zero():
    # max_depth == 0 returns 0 with no environment or background (maxdepth_nee.hpp:16-17)
    spec = scenes.furnace()
    o = O.OracleScene(spec)
    img, _, _, _ = o.render(o.cfg(estimator=0, max_depth=0, spp_x=1, spp_y=1))
    assert np.all(img == 0)


def test_reference_and_mis_agree_on_area_lights():
    spec = scenes.cornell(width=16, height=16, spp_x=16, spp_y=16)
    o = O.OracleScene(spec)
    a, va, _, _ = o.render(o.cfg(estimator=0, rr_depth=0), want_var=True)
    b, vb, _, _ = o.render(o.cfg(estimator=1, rr_depth=0, seed=99), want_var=True)
    assert np.allclose(a.mean((0, 1)), b.mean((0, 1)), rtol=0.02)


def test_russian_roulette_keeps_expectation():
    spec = scenes.cornell(width=16, height=16, spp_x=16, spp_y=16)
    o = O.OracleScene(spec)
    a, _, _, _ = o.render(o.cfg(rr_depth=0))
    b, _, _, _ = o.render(o.cfg(rr_depth=2, seed=5))
    assert np.allclose(a.mean((0, 1)), b.mean((0, 1)), rtol=0.03)


@pytest.mark.parametrize("estimator", [2, 3])
def test_mixture_estimators_form_factor(estimator):
    spec, expect = floor_under_light()
    o = O.OracleScene(spec)
    img, var, _, _ = o.render(o.cfg(estimator=estimator, max_depth=2, spp_x=128, spp_y=128), want_var=True)
    se = np.sqrt(var[0, 0, 0] / (128 * 128))
    assert abs(img[0, 0, 0] - expect) < max(4 * se, 2e-3 * expect)


def _sphere_under_sky(albedo=0.7, E=2.0):
    s = scenes.SceneSpec(name="sky")
    m = s.add_material(1, [albedo] * 3)
    s.add_sphere((0, 0, 0), 1.0, m)
    s.camera = dict(fov_deg=30, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1), pos=(0, 0, 5))
    s.render = dict(width=8, height=8, spp_x=2, spp_y=2, max_depth=1, sampler=0, estimator=2, rr_depth=0,
                    light_samples=4, seed=1, bg=(E, E, E))
    return s


def test_naive_depth_limit_returns_background():
    """maxdepth_naive returns bg_color at depth 0 and on misses: a convex
    Lambertian sphere under a uniform background gives a*E at every depth
    (cosine sampling: f cos / pdf = a), E where the ray misses."""
    spec = _sphere_under_sky()
    o = O.OracleScene(spec)
    for depth in (1, 3):
        img, _, smp, _ = o.render(o.cfg(max_depth=depth), want_samples=True)
        v = smp[..., 0]
        on = np.isclose(v, 0.7 * 2.0, rtol=1e-5)
        off = v == np.float32(2.0)
        assert np.all(on | off) and on.any() and off.any()


def test_roulette_depth_limit_returns_zero():
    spec = _sphere_under_sky()
    o = O.OracleScene(spec)
    _, _, smp1, _ = o.render(o.cfg(estimator=3, max_depth=1), want_samples=True)
    v = smp1[..., 0]
    assert np.all((v == 0) | (v == np.float32(2.0)))        # sphere paths end at the limit with 0
    # one bounce more: the sphere vertex survives roulette (q = clamp(a/pi, 0.5, 1) = 0.5) with
    # weight 1/q, so its samples are 0 or 2 a E, half of each; misses are E as before
    cfg = o.cfg(estimator=3, max_depth=2, spp_x=8, spp_y=8)
    _, _, smp2, _ = o.render(cfg, want_samples=True)
    _, _, ref, _ = o.render(o.cfg(estimator=2, max_depth=2, spp_x=8, spp_y=8), want_samples=True)
    r, n = smp2[..., 0], ref[..., 0]
    miss = n == np.float32(2.0)
    assert np.array_equal(r[miss], n[miss])
    hit = ~miss
    assert np.all((r[hit] == 0) | np.isclose(r[hit], 2 * 1.4, rtol=1e-5))
    frac = np.mean(r[hit] > 0)
    assert abs(frac - 0.5) < 4 * np.sqrt(0.25 / hit.sum())


def _sphere_lit_room():
    s = scenes.SceneSpec(name="sphere_light")
    w = s.add_material(1, [0.6, 0.6, 0.6])
    r = s.add_material(1, [0.7, 0.2, 0.1])
    lm = s.add_material(2, [1.0, 0.9, 0.7, 6.0])
    s.add_rect(1, 0.0, -3, 3, -3, 3, w)
    s.add_rect(2, -2.0, -3, 3, 0, 4, r)
    s.add_sphere((0.0, 0.6, 0.0), 0.6, w)
    s.add_sphere((1.2, 2.0, 0.5), 0.35, lm)
    s.camera = dict(fov_deg=45, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, -0.3, -1),
                    pos=(0, 1.5, 4.5))
    s.render = dict(width=12, height=12, spp_x=16, spp_y=16, max_depth=4, sampler=0, estimator=0, rr_depth=0,
                    light_samples=4, seed=3, bg=(0, 0, 0))
    return s


@pytest.mark.parametrize("estimator", [2, 3])
def test_mixture_estimators_agree_with_reference_on_a_sphere_light(estimator):
    """Where the reference estimator is unbiased (area lights, no environment,
    no background: the naive depth limit adds 0) the mixture integrators
    converge to the same image."""
    spec = _sphere_lit_room()
    o = O.OracleScene(spec)
    a, va, _, _ = o.render(o.cfg(estimator=0, spp_x=32, spp_y=32), want_var=True)
    b, vb, _, _ = o.render(o.cfg(estimator=estimator, spp_x=64, spp_y=64, seed=11), want_var=True)
    ma, mb = a.mean((0, 1)), b.mean((0, 1))
    se = np.sqrt(va.mean((0, 1)) / (32 * 32 * 144) + vb.mean((0, 1)) / (64 * 64 * 144))
    assert np.all(np.abs(ma - mb) < 5 * se + 0.01 * ma), (ma, mb, se)


def test_reservoir_closed_form():
    A, B, h, albedo, Le = 1.0, 0.6, 1.5, 0.6, 5.0
    spec, _ = floor_under_light(A, B, h, albedo, Le)
    o = O.OracleScene(spec)
    n = 192
    img, var, _, _ = o.render(o.cfg(estimator=4, max_depth=1, spp_x=n, spp_y=n), want_var=True)
    # E_x over the light of c(x)^2, c = (a/pi) Le A h^2 / d^4 (cos_i = cos_l = h/d)
    g, wts = np.polynomial.legendre.leggauss(64)
    u = g * A / 2
    w = g * B / 2
    U, Wg = np.meshgrid(u, w, indexing="ij")
    WW = np.outer(wts, wts) / 4.0
    c = albedo / np.pi * Le * (A * B) * h * h / (U * U + h * h + Wg * Wg) ** 2
    e_c2 = float(np.sum(WW * c * c))
    F = 4 * _corner_form_factor(A / 2, B / 2, h)
    expect = (1 - F) * np.sqrt(3.0) * e_c2
    se = np.sqrt(var[0, 0, 0] / (n * n))
    assert abs(img[0, 0, 0] - expect) < max(4 * se, 3e-3 * expect), (img[0, 0, 0], expect, se)


# ---------------------------------------------------------------- maxdepth_volume

def test_volume_absorption_only_matches_transmittance():
    """absorption_only_medium (scenes_medium.hpp:6-31): with sigma_s = 0 a sample
    returns the sphere's emission when the free flight passes the hit, else 0,
    so its expectation is Le exp(-sigma_t t_hit) for its own camera ray."""
    from pathtracer_amd import scenes
    spec = scenes.absorption_only_medium(width=48, height=48, spp_x=8, spp_y=8)
    o = O.OracleScene(spec)
    cfg = o.cfg()
    img, _, smp, _ = o.render(cfg, want_samples=True)
    spp = 64
    pix = np.repeat(np.arange(48 * 48, dtype=np.uint32), spp)
    kk = np.tile(np.arange(spp, dtype=np.uint32), 48 * 48)
    rays = o.camera_rays(cfg, pix, kk)
    prim, t = o.intersect(rays)
    le = np.array([0.4, 2.32, 3.2])
    expect = np.where((prim >= 0)[:, None], le * np.exp(-1.5 * t.astype(np.float64))[:, None], 0.0)
    expect = expect.reshape(48 * 48, spp, 3).mean(1)
    hit = (prim.reshape(48 * 48, spp) >= 0).all(1)
    got = smp.reshape(48 * 48, spp, 3).mean(1)
    # each sample is Le or 0: binomial noise of the mean over all hit pixels
    p = (expect[hit] / le).mean(0)
    n = hit.sum() * spp
    assert np.all(np.abs((got[hit] / le).mean(0) - p) < 4 * np.sqrt(p * (1 - p) / n))
    miss = (prim.reshape(48 * 48, spp) < 0).all(1)
    assert miss.sum() > 100 and np.all(got[miss] == 0)


def _sphere_light_quadrature(p, wo, c, r, le, sig_t, g, occluders, n_th=96, n_ph=192):
    """The single-scattering source term at p from one sphere light, by
    quadrature over the sphere: sum over the surface of Le (one-sided, the
    outward normal towards p) |n . dir| HG(wo . wi) exp(-sig_t d) / d^2 dA,
    with the segment tested against the occluder spheres analytically."""
    th = (np.arange(n_th) + 0.5) / n_th * np.pi
    ph = (np.arange(n_ph) + 0.5) / n_ph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    nrm = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (r * r * np.sin(T) * (np.pi / n_th) * (2 * np.pi / n_ph)).reshape(-1)
    x = c + r * nrm
    dv = p - x
    d = np.linalg.norm(dv, axis=1)
    dirv = dv / d[:, None]
    front = (nrm * dirv).sum(1) > 0
    wi = -dirv
    cosg = (wo[None, :] * wi).sum(1)
    hg = (1 - g * g) / (4 * np.pi * (1 + g * g + 2 * g * cosg) ** 1.5)
    vis = np.ones(len(d), bool)
    for (oc, orad) in occluders:
        # segment x -> p against sphere (oc, orad)
        m = x - oc
        b = (m * dirv).sum(1)
        cc = (m * m).sum(1) - orad * orad
        disc = b * b - cc
        sq = np.sqrt(np.maximum(disc, 0))
        t1 = -b - sq
        t2 = -b + sq
        hitseg = (disc > 0) & (((t1 > 1e-3) & (t1 < d - 1e-3)) | ((t2 > 1e-3) & (t2 < d - 1e-3)))
        vis &= ~hitseg
    w = np.where(front & vis, (nrm * dirv).sum(1) * hg * np.exp(-sig_t * d) / (d * d) * dA, 0.0)
    return le[None, :] * w.sum()


@pytest.mark.parametrize("g", [0.0, 0.6])
def test_volume_single_scatter_matches_quadrature(g):
    """A thin camera cone (one pixel, fov 0.02 deg) through single_scatter_medium,
    missing both spheres: the estimator's expectation along the ray is
    (sigma_s / sigma_t) * integral of exp(-sigma_t t) L_s1(r(t)) dt, where L_s1
    sums the sphere lights' quadrature; the reference's estimator divides by
    sigma_t once more than the free-flight pdf calls for, so this is its
    expectation (not the physical sigma_s times the integral)."""
    from pathtracer_amd import scenes
    spec = scenes.single_scatter_medium(width=1, height=1, spp_x=256, spp_y=256, g=g)
    d0 = np.array([0.35, 0.55, 1.0])
    d0 /= np.linalg.norm(d0)
    spec.camera = dict(fov_deg=0.02, aspect=1.0, focal_dist=1.0, aperture=0.0, up=(0, 1, 0), front=tuple(d0),
                       pos=(0, 0, -3))
    o = O.OracleScene(spec)
    img, var, _, _ = o.render(o.cfg(), want_var=True)
    sig_t, sig_s = 0.8, 0.7
    lights = [((0.0, 0.0, 0.0), 1.0, np.array([0.4, 2.32, 3.2])), ((-3.0, 0.0, -1.5), 1.0, np.array([24.0, 10.0, 24.0]))]
    cam = np.array([0.0, 0.0, -3.0])
    ts = (np.arange(300) + 0.5) / 300 * 24.0   # exp(-0.8 * 24) is negligible
    dt = 24.0 / 300
    total = np.zeros(3)
    for t in ts:
        p = cam + t * d0
        ls = np.zeros(3)
        for li, (c, r, le) in enumerate(lights):
            occ = [(np.array(oc), orad) for j, (oc, orad, _) in enumerate(lights) if j != li]
            ls += _sphere_light_quadrature(p, -d0, np.array(c), r, le, sig_t, g, occ)[0]
        total += np.exp(-sig_t * t) * ls * dt
    expect = sig_s / sig_t * total
    n = 256 * 256
    err = np.abs(img[0, 0] - expect)
    assert np.all(err < 5 * np.sqrt(var[0, 0] / n) + 5e-3 * expect), (img[0, 0], expect, np.sqrt(var[0, 0] / n))


def test_reference_traversal_count_changes_no_sample():
    """OR_RENDER_REFERENCE_TRAVERSALS repeats the reference's second traversal
    of each continuation ray (maxdepth_nee.hpp:238-255): every sample is the
    same, and the closest-hit traversals per non-specular vertex double."""
    spec = scenes.cornell(width=24, height=24, spp_x=2, spp_y=2)
    o = O.OracleScene(spec)
    _, _, a, ca = o.render(o.cfg(estimator=0), want_samples=True)
    _, _, b, cb = o.render(o.cfg(estimator=0, flags=1 << 24), want_samples=True)
    assert np.array_equal(a, b)
    assert ca[2] == cb[2]                      # shadow rays unchanged
    assert 1.4 * ca[1] < cb[1] < 2.0 * ca[1]   # camera rays are traversed once, continuations twice


# ---- the guided estimator (a fixed vMF mixture in place of the network) ----

def guide_mixture(c_logit=0.0, seed=3, toward=(0.0, 1.0, 0.0)):
    """A processed mixture whose first lobe points along `toward` with
    kappa ~ 31, the others at random directions with kappa from 4 to 160;
    c = 0.05 + 0.45 sigmoid(c_logit)."""
    from oracle import guide_oracle as G
    rng = np.random.default_rng(seed)
    raw = np.zeros((1, G.OUT))
    raw[0, 0:8] = rng.normal(size=8)
    raw[0, 8] = c_logit
    raw[0, 9:17] = rng.uniform(0.0, 4.0, 8)
    raw[0, 9] = 2.0
    dirs = rng.normal(size=(8, 3))
    dirs[0] = toward
    raw[0, 17:25], raw[0, 25:33], raw[0, 33:41] = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    return O.processed_mixture(raw)[0]


GUIDED_KINDS = (1 << 1) | (1 << 3) | (1 << 7)   # Lambertian, conductor, Disney


@pytest.mark.parametrize("c_logit", [-1.0, 2.0])
def test_guided_form_factor(c_logit):
    """Direct light on a diffuse floor under a quad light: with guided
    directions drawn at the floor point and weighted by the one-sample MIS pdf
    (in the BSDF weight, NEE's power heuristic and the emitter hit's), the
    estimate is the closed-form albedo * Le * form factor."""
    spec, expect = floor_under_light()
    o = O.OracleScene(spec)
    o.set_guide(guide_mixture(c_logit), GUIDED_KINDS, 4)
    img, var, _, _ = o.render(o.cfg(estimator=O.EST_GUIDED, max_depth=1, spp_x=128, spp_y=128), want_var=True)
    se = np.sqrt(var[0, 0, 0] / (128 * 128))
    assert abs(img[0, 0, 0] - expect) < max(4 * se, 2e-3 * expect)


def test_guided_white_furnace():
    """Every vertex guided inside the white furnace (Lambertian albedo 0.5
    under a uniform environment of 1): the NEE+MIS value 0.5."""
    spec = scenes.furnace(albedo=0.5, env_value=1.0)
    o = O.OracleScene(spec)
    o.set_guide(guide_mixture(1.0, seed=4), GUIDED_KINDS, 64)
    img, _, _, _ = o.render(o.cfg(estimator=O.EST_GUIDED, max_depth=64, spp_x=32, spp_y=32))
    assert abs(img[3:5, 3:5, 0].mean() - 0.5) < 0.01


def test_guided_agrees_with_nee_mis_on_the_cornell_box():
    """Guided (every vertex, rough conductor and Lambertian walls) and plain
    NEE+MIS images of the Cornell box agree in their means."""
    spec = scenes.cornell(width=16, height=16, spp_x=16, spp_y=16)
    o = O.OracleScene(spec)
    o.set_guide(guide_mixture(0.5, seed=5), GUIDED_KINDS, 8)
    a, va, _, _ = o.render(o.cfg(estimator=1, rr_depth=0), want_var=True)
    b, vb, _, _ = o.render(o.cfg(estimator=O.EST_GUIDED, rr_depth=0, seed=77), want_var=True)
    ma, mb = a.mean((0, 1)), b.mean((0, 1))
    se = np.sqrt((va.mean((0, 1)) + vb.mean((0, 1))) / (256 * 256))
    assert np.all(np.abs(ma - mb) < 4 * se + 1e-3 * ma), (ma, mb, se)
