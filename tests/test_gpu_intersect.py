"""GPU BVH traversal against the oracle: closest-hit primitive ids identical
and t within 1e-5 relative; any-hit agreement.  The oracle itself is pinned
against brute force in tests/test_oracle_geometry.py.  Where the ids differ,
the two hits must be a tie (same t) -- the reference's own tie rule depends
on traversal order (meshBVH.cpp:67-81)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import random_dirs

pytestmark = pytest.mark.gpu


def _rays_towards(rng, n, lo, hi, origin_box, tmin=1e-4, tmax=1e5):
    o = rng.uniform(origin_box[0], origin_box[1], size=(n, 3))
    target = rng.uniform(lo, hi, size=(n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3:6] = d
    r[:, 6] = tmin
    r[:, 7] = tmax
    return r


def _compare(gp, gt, cp, ct, tol=1e-5):
    same = gp == cp
    both = (gp >= 0) & (cp >= 0)
    assert np.all((gp >= 0) == (cp >= 0)), "hit/miss disagreement"
    rel = np.abs(gt[both] - ct[both]) / np.maximum(np.abs(ct[both]), 1e-6)
    assert np.all(rel < tol), f"t mismatch {rel.max()}"
    # differing ids only on exact (within tolerance) ties
    return int((~same).sum())


def test_cornell_closest_hit(gpu):
    from pathtracer_amd import render as R, scenes
    spec = scenes.cornell(ico_subdiv=4)
    g = R.Scene(spec)
    o = O.OracleScene(spec)
    rng = np.random.default_rng(11)
    rays = _rays_towards(rng, 1_000_000, [-1, 0, -1], [1, 2, 1], ([-0.9, 0.1, -0.9], [0.9, 1.9, 3.0]))
    gp, gt = g.intersect(rays)
    cp, ct = o.intersect(rays)
    ties = _compare(gp, gt, cp, ct)
    assert ties < 50
    ga, _ = g.intersect(rays, any_hit=True)
    ca, _ = o.intersect(rays, any_hit=True)
    assert np.array_equal(ga >= 0, ca >= 0)


def test_displaced_mesh_closest_hit(gpu):
    from pathtracer_amd import render as R, scenes
    spec = scenes.SceneSpec()
    m = spec.add_material(1, [0.5, 0.5, 0.5])
    pos, idx = scenes.displaced_icosphere(5, 1.0, 0.08, 2)     # 20,480 triangles
    spec.add_mesh(pos, idx, m)
    spec.add_sphere((2.0, 0.0, 0.0), 0.5, m)
    spec.add_rect(1, -1.2, -3, 3, -3, 3, m)
    spec.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1), pos=(0, 0, 5))
    g = R.Scene(spec)
    o = O.OracleScene(spec)
    rng = np.random.default_rng(5)
    rays = _rays_towards(rng, 2_000_000, [-1.2, -1.2, -1.2], [2.5, 1.2, 1.2], ([-4, -1, 3], [4, 4, 6]))
    # some rays from inside the sphere outwards and with finite tmax
    k = 200_000
    rays[:k, :3] = 0.0
    rays[:k, 3:6] = random_dirs(rng, k)
    rays[k:2 * k, 7] = rng.uniform(0.5, 6.0, size=k)
    gp, gt = g.intersect(rays)
    cp, ct = o.intersect(rays)
    ties = _compare(gp, gt, cp, ct)
    assert ties < 200
    ga, _ = g.intersect(rays, any_hit=True)
    ca, _ = o.intersect(rays, any_hit=True)
    assert np.array_equal(ga >= 0, ca >= 0)


def test_curves_closest_and_any_hit(gpu):
    """Bezier curves (8 cylinder segments each) among rects: ids, t, any hit."""
    from pathtracer_amd import render as R, scenes
    spec = scenes.hair_curves(n_curves=80)
    g = R.Scene(spec)
    o = O.OracleScene(spec)
    rng = np.random.default_rng(12)
    rays = _rays_towards(rng, 400_000, [-1.2, 0.0, -1.2], [1.2, 1.6, 1.2], ([-2, 0.2, -2], [2, 2.5, 4]))
    gp, gt = g.intersect(rays)
    cp, ct = o.intersect(rays)
    ties = _compare(gp, gt, cp, ct)
    assert ties < 20
    n_curve_hits = int((gp >= len(spec.rects)).sum())  # curve segments follow the two rects
    assert n_curve_hits > 10_000
    ga, _ = g.intersect(rays, any_hit=True)
    ca, _ = o.intersect(rays, any_hit=True)
    assert np.array_equal(ga >= 0, ca >= 0)


def test_env_sample_matches_oracle(gpu):
    """The bucketed CDF inversion finds the oracle's texel for every u,
    including u on and next to bucket bounds b/G and texel CDF values, and
    on a map with empty rows and columns (zero-probability intervals)."""
    from pathtracer_amd import render as R, scenes
    rng = np.random.default_rng(8)
    spec = scenes.glass_hdri(width=8, height=8, spp_x=1, spp_y=1, subdiv=1, env_w=256, env_h=128)
    rgb = spec.env[0].copy()
    rgb[40:48] = 0.0          # empty rows
    rgb[:, 100:130] = 0.0     # empty columns
    spec.env = (rgb, 1.0)
    g = R.Scene(spec)
    o = O.OracleScene(spec)
    n = 200_000
    u = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    for G, col in ((256, 0), (128, 1)):     # bounds of the conditional and marginal buckets
        b = np.arange(G, dtype=np.float32) / G
        for k, v in enumerate((b, np.nextafter(b, 0), np.nextafter(b, 1))):
            lo = 1000 + (3 * (G == 128) + k) * 256
            u[lo:lo + G, col] = v
    u = np.clip(u, 0, np.float32(1) - np.float32(2 ** -24))
    gd, gp = g.env_sample(u)
    od, op = o.env_sample(u)
    np.testing.assert_allclose(gp, op, rtol=2e-5, atol=0)
    np.testing.assert_allclose(gd, od, atol=2e-5)
    assert np.all(np.isfinite(gp))
    assert (gp > 0).mean() > 0.999   # pdf 0 only at the pole (u1 = 0: sin(theta) = 0)


def test_ten_million_rays(gpu):
    """Oracle level 2 at its full size: 10^7 seeded rays on the 20,480-triangle
    displaced mesh with a sphere and a rect, closest-hit ids identical up to
    exact ties and t within 1e-5 relative."""
    from pathtracer_amd import render as R, scenes
    spec = scenes.SceneSpec()
    m = spec.add_material(1, [0.5, 0.5, 0.5])
    pos, idx = scenes.displaced_icosphere(5, 1.0, 0.08, 2)
    spec.add_mesh(pos, idx, m)
    spec.add_sphere((2.0, 0.0, 0.0), 0.5, m)
    spec.add_rect(1, -1.2, -3, 3, -3, 3, m)
    spec.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1), pos=(0, 0, 5))
    g = R.Scene(spec)
    o = O.OracleScene(spec)
    rng = np.random.default_rng(17)
    total_ties = 0
    for k in range(5):   # 5 x 2M rays
        rays = _rays_towards(rng, 2_000_000, [-1.2, -1.2, -1.2], [2.5, 1.2, 1.2], ([-4, -1, 3], [4, 4, 6]))
        gp, gt = g.intersect(rays)
        cp, ct = o.intersect(rays)
        total_ties += _compare(gp, gt, cp, ct)
    assert total_ties < 1000, total_ties
