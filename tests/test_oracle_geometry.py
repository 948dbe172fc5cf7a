"""Pins the oracle's geometry, lights and camera (CPU only).

Moller-Trumbore known answer of tests/test_mesh_io.cpp:26-32; the median-split
mesh BVH (meshBVH.cpp:6-83) against brute force over every primitive; sphere
and quad intersections against their closed forms; QuadLight / SphereLight
sample and pdf closed forms (light.hpp:119-306); the environment CDF pdf
integrates to 1 and matches its sampler; camera rays at zero aperture pass
through the pinhole.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from pathtracer_amd import scenes
from tests.helpers import random_dirs, unit


def _scene_with(mesh=None, spheres=(), rects=(), env=None):
    s = scenes.SceneSpec()
    m = s.add_material(1, [0.5, 0.5, 0.5])
    e = s.add_material(2, [1.0, 0.9, 0.8, 7.0])
    if mesh is not None:
        s.add_mesh(mesh[0], mesh[1], m)
    for c, r, emissive in spheres:
        s.add_sphere(c, r, e if emissive else m)
    for args, emissive in rects:
        s.add_rect(*args, e if emissive else m)
    if env is not None:
        s.env = (env, 1.0)
    s.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1), pos=(0, 0, 5))
    return s


def test_mesh_io_known_answer():
    # tests/data/minimal.obj: (0,0,0), (1,0,0), (0,1,0); ray (.25,.25,-1) + t(0,0,1)
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.int32)
    o = O.OracleScene(_scene_with((pos, idx)))
    prim, t = o.intersect(np.array([[0.25, 0.25, -1, 0, 0, 1, 0.001, 100.0]], np.float32))
    assert prim[0] == 0
    assert t[0] == 1.0  # hit point (0.25, 0.25, 0)


def test_bvh_matches_brute_force():
    pos, idx = scenes.displaced_icosphere(3, 1.0, 0.1, 4)
    spec = _scene_with((pos, idx), spheres=[((1.5, 0, 0), 0.4, False)],
                       rects=[((1, -1.1, -2, 2, -2, 2), False), ((0, -1.5, -1, 1, -1, 1), False)])
    o = O.OracleScene(spec)
    rng = np.random.default_rng(2)
    n = 40_000
    org = rng.uniform(-3, 3, size=(n, 3))
    tgt = rng.uniform(-1.2, 1.2, size=(n, 3))
    d = unit(tgt - org)
    rays = np.concatenate([org, d, np.full((n, 1), 1e-4), rng.uniform(0.5, 10, size=(n, 1))], 1).astype(np.float32)
    p1, t1 = o.intersect(rays)
    p2, t2 = o.intersect(rays, brute=True)
    assert np.array_equal(p1 >= 0, p2 >= 0)
    h = p1 >= 0
    assert np.allclose(t1[h], t2[h], rtol=1e-6)
    assert np.mean(p1[h] == p2[h]) > 0.9999   # differing ids only on ties
    a, _ = o.intersect(rays, any_hit=True)
    assert np.array_equal(a >= 0, p2 >= 0)


def test_sphere_and_rect_closed_forms():
    o = O.OracleScene(_scene_with(spheres=[((0, 0, 0), 1.0, False)], rects=[((2, -3.0, -1, 1, -1, 1), False)]))
    rng = np.random.default_rng(4)
    n = 2000
    org = np.repeat(np.array([[0.0, 0.0, 5.0]]), n, 0)
    d = unit(np.concatenate([rng.uniform(-0.15, 0.15, (n, 2)), -np.ones((n, 1))], 1))
    rays = np.concatenate([org, d, np.full((n, 1), 1e-4), np.full((n, 1), 1e5)], 1).astype(np.float32)
    prim, t = o.intersect(rays)
    # |o + t d| = 1 for sphere hits (first root), z = -3 plane for rect hits
    b = np.sum(org * d, 1)
    c = np.sum(org * org, 1) - 1.0
    disc = b * b - c
    hit_s = disc >= 0
    t_exp = -b - np.sqrt(np.maximum(disc, 0))
    assert np.all(prim[hit_s] == 0)
    assert np.allclose(t[hit_s], t_exp[hit_s], rtol=1e-5)
    r = ~hit_s & (prim == 1)
    assert np.allclose(t[r], (5.0 + 3.0) / -d[r, 2], rtol=1e-5)


def test_quad_light_closed_forms():
    # AARectangle<Y> at y=2, x in [-0.5,0.5], z in [-0.25,0.25]: QuadLight normal is -Y
    spec = _scene_with(rects=[((1, 2.0, -0.5, 0.5, -0.25, 0.25), True)])
    o = O.OracleScene(spec)
    rng = np.random.default_rng(6)
    n = 1000
    u2 = rng.random((n, 2), np.float32)
    ref = np.repeat(np.array([[0.1, 0.0, 0.2]], np.float32), n, 0)
    p, nrm, pdf, le = o.light_sample(0, u2, ref)
    assert np.allclose(p[:, 0], -0.5 + u2[:, 0], atol=1e-6)
    assert np.allclose(p[:, 2], -0.25 + 0.5 * u2[:, 1], atol=1e-6)
    assert np.allclose(nrm, [0, -1, 0])
    d2 = np.sum((p - ref) ** 2, 1)
    cos = np.abs(p[:, 1] - ref[:, 1]) / np.sqrt(d2)
    assert np.allclose(pdf, d2 / (0.5 * cos), rtol=1e-5)   # area 0.5
    assert np.allclose(le, [7.0, 6.3, 5.6], rtol=1e-6)      # lit from below (one-sided)


def test_sphere_light_closed_forms():
    spec = _scene_with(spheres=[((0, 3, 0), 0.5, True)])
    o = O.OracleScene(spec)
    rng = np.random.default_rng(7)
    n = 200_000
    u2 = rng.random((n, 2), np.float32)
    ref = np.zeros((n, 3), np.float32)
    p, nrm, pdf, le = o.light_sample(0, u2, ref)
    assert np.allclose(np.linalg.norm(p - [0, 3, 0], axis=1), 0.5, atol=1e-5)
    # uniform over the whole sphere: the mean point is the centre
    assert np.allclose(p.mean(0), [0, 3, 0], atol=5e-3)
    cone = 1 / (2 * np.pi * (1 - np.sqrt(1 - 0.25 / 9)))
    assert np.allclose(pdf, cone, rtol=1e-5)                 # SphereLight::pdf_value: cone pdf
    facing = np.sum(nrm * (ref - p), 1) > 0
    assert np.all(np.any(le[facing] > 0, 1)) and np.all(le[~facing] == 0)


def test_env_cdf_pdf_normalised_and_matches_sampler():
    env = scenes.sky_hdri(64, 32, seed=2, sun_peak=50.0, sun_width_deg=8.0)
    o = O.OracleScene(_scene_with(spheres=[((0, 0, 0), 0.1, False)], env=env))
    nt, nph = 400, 800
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).reshape(-1)
    pdf = o.env_pdf(d[::5])
    assert abs(np.sum(pdf * dw[::5]) * 5 - 1.0) < 1e-2
    rng = np.random.default_rng(8)
    dirs, ps = o.env_sample(rng.random((100_000, 2), np.float32))
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-5)
    assert np.allclose(o.env_pdf(dirs[:500]), ps[:500], rtol=1e-3)
    # E[1/pdf] = 4*pi
    assert abs(np.mean(1 / ps) - 4 * np.pi) < 0.03 * 4 * np.pi


def test_pinhole_camera_rays():
    spec = scenes.cornell(width=16, height=16)
    o = O.OracleScene(spec)
    cfg = o.cfg()
    pix = np.arange(256, dtype=np.uint32)
    rays = o.camera_rays(cfg, pix, np.zeros(256, np.uint32))
    assert np.allclose(rays[:, :3], [0, 1.0, 3.9], atol=1e-6)         # zero aperture: all from the pinhole
    assert np.allclose(np.linalg.norm(rays[:, 3:6], axis=1), 1, atol=1e-6)
    # the image plane spans the vertical field of view (40 degrees)
    vy = rays[:, 4] / -rays[:, 5]
    assert abs(vy.max() - np.tan(np.radians(20)) * (15.9 / 16 * 2 - 1)) < 0.05


def test_thin_lens_focus():
    # points on the focal plane are sharp: every lens sample of one pixel meets there
    spec = scenes.glass_hdri(width=8, height=8, spp_x=4, spp_y=4, subdiv=1, env_w=16, env_h=8)
    spec.camera = dict(fov_deg=30, aspect=1, focal_dist=4.0, aperture=4.0 / 0.5, up=(0, 1, 0), front=(0, 0, -1),
                       pos=(0, 0, 4.0))
    o = O.OracleScene(spec)
    cfg = o.cfg()
    k = np.arange(16, dtype=np.uint32)
    rays = o.camera_rays(cfg, np.full(16, 27, np.uint32), k)
    # lens radius focal/aperture/2 = 0.25: origins spread on the lens square
    assert np.ptp(rays[:, 0]) > 0.05 and np.all(np.abs(rays[:, :2]) <= 0.25 + 1e-6)
    t = 4.0 / -rays[:, 5]
    hit = rays[:, :3] + t[:, None] * rays[:, 3:6]
    # k differs, so the pixel jitter differs; hits stay within one pixel footprint of each other
    pix = 2 * 4.0 * np.tan(np.radians(15)) / 8
    assert np.ptp(hit[:, 0]) <= pix * 1.01 and np.ptp(hit[:, 1]) <= pix * 1.01


def _smooth_scene(pos, idx, nrm, nidx, flags=0):
    s = scenes.SceneSpec()
    m = s.add_material(1, [0.5, 0.5, 0.5])
    s.add_mesh(pos, idx, m, flags, normals=nrm, nidx=nidx)
    s.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1), pos=(0, 0, 5))
    return s


def test_motion_blur_camera_moments():
    """A shutter moves the camera from pos (time 0) to pos1 (time 1), the time
    uniform in [0, shutter): with a pinhole every origin lies on the segment,
    its position along it has mean shutter/2 and variance shutter^2/12 (the
    first two moments of a uniform law), and the direction through a pixel
    does not depend on the time (the image plane moves with the camera)."""
    spec = scenes.cornell(width=16, height=16, spp_x=4, spp_y=4)
    p0, p1, shutter = np.array([0.0, 1.0, 3.9]), np.array([0.6, 1.3, 3.5]), 0.75
    spec.camera = dict(spec.camera, pos1=tuple(p1), shutter=shutter)
    o = O.OracleScene(spec)
    cfg = o.cfg()
    n = 256 * 16
    pix = np.repeat(np.arange(256, dtype=np.uint32), 16)
    k = np.tile(np.arange(16, dtype=np.uint32), 256)
    rays = o.camera_rays(cfg, pix, k)
    delta = p1 - p0
    s = (rays[:, :3] - p0) @ delta / (delta @ delta)
    off = rays[:, :3] - (p0 + s[:, None] * delta)
    assert np.abs(off).max() < 1e-5                        # on the segment
    assert s.min() >= -1e-6 and s.max() < shutter + 1e-6
    se_mean = shutter / np.sqrt(12 * n)
    assert abs(s.mean() - shutter / 2) < 4 * se_mean
    assert abs(s.var() - shutter ** 2 / 12) < 0.05 * shutter ** 2 / 12
    static = scenes.cornell(width=16, height=16, spp_x=4, spp_y=4)
    os_ = O.OracleScene(static)
    rays0 = os_.camera_rays(os_.cfg(), pix, k)
    assert np.allclose(rays[:, 3:6], rays0[:, 3:6], atol=2e-6)
    # zero shutter is the static camera
    spec.camera = dict(spec.camera, shutter=0.0)
    oz = O.OracleScene(spec)
    assert np.array_equal(oz.camera_rays(oz.cfg(), pix, k), rays0)


def test_smooth_normal_interpolation_corners():
    """MeshBVHNode::hitTriangle (meshBVH.cpp:116-130): the shading normal is
    normalize((1-u-v) n0 + u n1 + v n2) with u along p1-p0 and v along p2-p0,
    turned to the face normal's side, then against the ray (set_normal).  Rays
    hit the triangle at chosen barycentrics from both sides."""
    p = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], np.float32)
    nrm = unit([[0.2, 0.1, 1.0], [1.0, -0.3, 0.8], [-0.6, 1.0, 0.7]])
    o = O.OracleScene(_smooth_scene(p, np.array([[0, 1, 2]], np.int32), nrm, np.array([[0, 1, 2]], np.int32)))
    rng = np.random.default_rng(11)
    n = 2000
    u = rng.uniform(0.02, 0.96, n)
    v = rng.uniform(0.02, 0.98, n) * (1 - u)
    hit = p[0] + u[:, None] * (p[1] - p[0]) + v[:, None] * (p[2] - p[0])
    side = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    org = hit + side[:, None] * np.array([0.1, 0.2, 3.0])
    d = unit(hit - org)
    rays = np.concatenate([org, d, np.full((n, 1), 1e-4), np.full((n, 1), 100.0)], 1).astype(np.float32)
    got, inside = o.surface(rays)
    ns = (1 - u - v)[:, None] * nrm[0] + u[:, None] * nrm[1] + v[:, None] * nrm[2]
    ns = ns / np.linalg.norm(ns, axis=1, keepdims=True)           # all on the +z side here
    want = np.where((np.sum(d * ns, axis=1) < 0)[:, None], ns, -ns)
    np.testing.assert_allclose(got, want, atol=2e-5)
    assert np.array_equal(inside == 1, np.sum(d * ns, axis=1) < 0)


def test_smooth_normals_equal_to_face_normals_change_nothing():
    pos, idx = scenes.displaced_icosphere(2, 1.0, 0.1, 3)
    e1 = pos[idx[:, 1]] - pos[idx[:, 0]]
    e2 = pos[idx[:, 2]] - pos[idx[:, 0]]
    fn = unit(np.cross(e1, e2))
    nidx = np.arange(3 * len(idx), dtype=np.int32).reshape(-1, 3)
    flat = O.OracleScene(_scene_with((pos, idx)))
    smooth = O.OracleScene(_smooth_scene(pos, idx, np.repeat(fn, 3, axis=0), nidx))
    rng = np.random.default_rng(12)
    n = 5000
    org = rng.uniform(-3, 3, size=(n, 3))
    d = unit(rng.uniform(-0.8, 0.8, size=(n, 3)) - org)
    rays = np.concatenate([org, d, np.full((n, 1), 1e-4), np.full((n, 1), 100.0)], 1).astype(np.float32)
    a, ia = flat.surface(rays)
    b, ib = smooth.surface(rays)
    assert np.array_equal(ia, ib)
    np.testing.assert_allclose(a, b, atol=1e-6)


def test_smooth_icosphere_approaches_the_sphere_normal():
    v, f = scenes.icosphere(3)
    v = v.astype(np.float32)
    o_s = O.OracleScene(_smooth_scene(v, f.astype(np.int32), v, f.astype(np.int32)))
    o_f = O.OracleScene(_scene_with((v, f.astype(np.int32))))
    rng = np.random.default_rng(13)
    n = 4000
    org = unit(rng.normal(size=(n, 3))) * 4.0
    d = unit(rng.normal(scale=0.2, size=(n, 3)) - org)
    rays = np.concatenate([org, d, np.full((n, 1), 1e-4), np.full((n, 1), 100.0)], 1).astype(np.float32)
    prim, t = o_s.intersect(rays)
    h = prim >= 0
    p = rays[h, :3] + t[h, None] * rays[h, 3:6]
    true_n = unit(p)
    ns, _ = o_s.surface(rays[h])
    nf, _ = o_f.surface(rays[h])
    err_s = np.linalg.norm(ns - true_n, axis=1)   # rays arrive from outside: the normal turned against
    err_f = np.linalg.norm(nf - true_n, axis=1)   # the ray is the outward one
    assert err_s.max() < 0.02
    assert err_s.mean() < 0.25 * err_f.mean()


# ---------------------------------------------------------------- curves

def _curve_scene(cps, w0, w1):
    from pathtracer_amd import scenes, _lib as L
    s = scenes.SceneSpec(name="curve")
    m = s.add_material(L.MAT_LAMBERTIAN, [0.5, 0.5, 0.5])
    s.add_curve(cps, w0, w1, m)
    s.camera = dict(fov_deg=40.0, aspect=1.0, focal_dist=1.0, aperture=0.0, up=(0, 1, 0), front=(0, 0, -1),
                    pos=(0, 0, 5))
    s.render = dict(width=8, height=8, spp_x=1, spp_y=1, max_depth=2, sampler=0, estimator=1, rr_depth=0,
                    light_samples=1, seed=1, bg=(0, 0, 0))
    return s


def test_straight_curve_is_a_cylinder():
    """Collinear, evenly spaced control points make the Bezier a straight line
    (uniform in u), and equal widths a cylinder of that radius: rays
    perpendicular to the axis at offset b hit at t = D - sqrt(r^2 - b^2), with
    the radial normal and the axis as tangent; rays past the ends or beyond the
    radius miss (curve.hpp:47-130)."""
    cps = np.array([[0, -1, 0], [0, -1 / 3, 0], [0, 1 / 3, 0], [0, 1, 0]], np.float64)
    r = 0.2
    o = O.OracleScene(_curve_scene(cps, r, r))
    rng = np.random.default_rng(3)
    n = 2000
    b = rng.uniform(-0.99 * r, 0.99 * r, n)
    y = rng.uniform(-0.95, 0.95, n)
    D = 3.0
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = b
    rays[:, 1] = y
    rays[:, 2] = D
    rays[:, 5] = -1.0
    rays[:, 6] = 1e-4
    rays[:, 7] = 1e5
    prim, t = o.intersect(rays)
    assert np.all(prim >= 0)
    np.testing.assert_allclose(t, D - np.sqrt(r * r - b * b), rtol=0, atol=2e-5)
    # segment ids follow y: 8 segments over [-1, 1]
    np.testing.assert_array_equal(prim, np.clip(((y + 1) / 2 * 8).astype(int), 0, 7))
    nrm, inside = o.surface(rays)
    hitp = np.stack([b, y, D - t], 1)
    radial = hitp * [1, 0, 1]
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    np.testing.assert_allclose(nrm, radial, atol=1e-4)   # faces the ray (+z side)
    assert np.all(inside == 1)
    # misses: past the radius, beyond the ends
    rays[:, 0] = np.where(b >= 0, r * 1.01, -r * 1.01)
    prim, _ = o.intersect(rays)
    assert np.all(prim < 0)
    rays[:, 0] = 0.0
    rays[:, 1] = np.where(y >= 0, 1.05, -1.05)
    prim, _ = o.intersect(rays)
    assert np.all(prim < 0)


def test_curved_curve_hits_lie_on_the_tube():
    """A bent Bezier with tapering width: every hit point lies within the
    segment's radius of the Bezier polyline (the 8-segment approximation), on
    the surface (distance = radius to its segment's axis), and brute force
    agrees with the BVH-free scene test."""
    cps = np.array([[-1, -1, 0], [-0.2, 0.8, 0.3], [0.4, -0.6, -0.2], [1.0, 1.0, 0.1]], np.float64)
    w0, w1 = 0.15, 0.05
    o = O.OracleScene(_curve_scene(cps, w0, w1))
    rng = np.random.default_rng(4)
    n = 4000
    org = np.stack([rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n), np.full(n, 4.0)], 1)
    tgt = np.stack([rng.uniform(-1.2, 1.2, n), rng.uniform(-1.2, 1.2, n), rng.uniform(-0.3, 0.3, n)], 1)
    d = tgt - org
    rays = np.concatenate([org, d, np.full((n, 1), 1e-4), np.full((n, 1), 1e5)], 1).astype(np.float32)
    prim, t = o.intersect(rays)
    bp, bt = o.intersect(rays, brute=True)
    np.testing.assert_array_equal(prim, bp)
    np.testing.assert_allclose(t[prim >= 0], bt[prim >= 0], rtol=1e-6)
    hit = prim >= 0
    assert hit.sum() > 200
    u = np.linspace(0, 1, 9)
    B = ((1 - u) ** 3)[:, None] * cps[0] + (3 * u * (1 - u) ** 2)[:, None] * cps[1] + \
        (3 * u * u * (1 - u))[:, None] * cps[2] + (u ** 3)[:, None] * cps[3]
    p = rays[hit, :3].astype(np.float64) + t[hit, None] * rays[hit, 3:6]
    seg = prim[hit]
    pa, pb = B[seg], B[seg + 1]
    ax = (pb - pa) / np.linalg.norm(pb - pa, axis=1, keepdims=True)
    v = p - pa
    dist = np.linalg.norm(v - (v * ax).sum(1, keepdims=True) * ax, axis=1)
    rad = 0.5 * ((w0 + (w1 - w0) * seg / 8) + (w0 + (w1 - w0) * (seg + 1) / 8))
    np.testing.assert_allclose(dist, rad, atol=5e-5)
    proj = (v * ax).sum(1)
    assert np.all(proj >= -1e-5) and np.all(proj <= np.linalg.norm(pb - pa, axis=1) + 1e-5)
