"""Any-hit traversal over the 64-byte quantised BVH nodes (PT_BUILD_ANY_HIT_QNODES).

The quantised child boxes are rounded outwards, so a shadow ray must find an
occluder exactly where the oracle does: any-hit results are compared with
the oracle's bit for bit, on scenes whose boxes are flat (axis-aligned rects),
tiny against their node (small spheres far from a huge one), far from the
origin, and on trees from both builds.  Renders over the quantised nodes
must equal renders over the 128-byte nodes sample for sample.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.test_gpu_intersect import _rays_towards

pytestmark = pytest.mark.gpu


def _any_hit_matches(spec, rays, build_flags, leaf=4, oracle=True, max_extra=0):
    """Any hit over the quantised nodes against any hit over the 128-byte
    nodes of the same tree: no hit is lost (the rounded-out boxes contain the
    full ones, and the slab test is monotone in the bounds), and at most
    `max_extra` rays gain one: a ray whose primitive test reports a hit just
    outside the primitive's float box, which the tight box culls.  With
    `oracle`, also equal to the oracle's."""
    from pathtracer_amd import render as R, _lib as L
    g = R.Scene(spec, leaf_size=leaf, build_flags=build_flags | L.BUILD_ANY_HIT_QNODES)
    ga, _ = g.intersect(rays, any_hit=True)
    f = R.Scene(spec, leaf_size=leaf, build_flags=build_flags | L.BUILD_ANY_HIT_FULL_NODES)
    fa, _ = f.intersect(rays, any_hit=True)
    lost = int(((fa >= 0) & (ga < 0)).sum())
    extra = int(((ga >= 0) & (fa < 0)).sum())
    assert lost == 0 and extra <= max_extra, (lost, extra)
    if oracle:
        ca, _ = O.OracleScene(spec).intersect(rays, any_hit=True)
        assert np.array_equal(ga >= 0, ca >= 0), int(((ga >= 0) != (ca >= 0)).sum())
    return ga


def _displaced_spec():
    from pathtracer_amd import scenes
    spec = scenes.SceneSpec()
    m = spec.add_material(1, [0.5, 0.5, 0.5])
    pos, idx = scenes.displaced_icosphere(5, 1.0, 0.08, 2)     # 20,480 triangles
    spec.add_mesh(pos, idx, m)
    spec.add_sphere((2.0, 0.0, 0.0), 0.5, m)
    spec.add_rect(1, -1.2, -3, 3, -3, 3, m)
    spec.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1),
                       pos=(0, 0, 5))
    return spec


@pytest.mark.parametrize("lbvh", [False, True])
@pytest.mark.parametrize("leaf", [1, 4])
def test_displaced_mesh_any_hit(gpu, lbvh, leaf):
    from pathtracer_amd import _lib as L
    rng = np.random.default_rng(31 + leaf)
    rays = _rays_towards(rng, 2_000_000, [-1.2, -1.2, -1.2], [2.5, 1.2, 1.2], ([-4, -1, 3], [4, 4, 6]))
    k = 200_000
    rays[:k, :3] = 0.0                           # from inside the mesh outwards
    d = rng.normal(size=(k, 3))
    rays[:k, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[k:2 * k, 7] = rng.uniform(0.5, 6.0, size=k)   # finite tmax: shadow-ray segments
    ga = _any_hit_matches(_displaced_spec(), rays, L.BUILD_DEVICE_LBVH if lbvh else 0, leaf)
    assert 0.2 < (ga >= 0).mean() < 0.95


def test_cornell_flat_boxes(gpu):
    """Axis-aligned rects have zero-thickness boxes; rays graze and end on them."""
    from pathtracer_amd import scenes
    spec = scenes.cornell(ico_subdiv=4)
    rng = np.random.default_rng(12)
    rays = _rays_towards(rng, 1_000_000, [-1, 0, -1], [1, 2, 1], ([-0.9, 0.1, -0.9], [0.9, 1.9, 3.0]))
    rays[:300_000, 7] = rng.uniform(0.1, 3.0, size=300_000)
    _any_hit_matches(spec, rays, 0)


def test_scales_far_from_origin(gpu):
    """Spheres of radius 0.01 to 30 around (5000, -3000, 2000): node extents
    from tiny to huge, far from the origin, against the per-axis scales
    (compared with the 128-byte nodes only: at these coordinates the sphere
    test itself is at the edge of float precision, and a few rays report a
    hit just outside a sphere's box)."""
    from pathtracer_amd import scenes
    rng = np.random.default_rng(8)
    spec = scenes.SceneSpec()
    m = spec.add_material(1, [0.5, 0.5, 0.5])
    c0 = np.array([5000.0, -3000.0, 2000.0])
    spec.add_sphere(tuple(c0 + [0, -40, 0]), 30.0, m)
    for _ in range(600):
        spec.add_sphere(tuple(c0 + rng.uniform(-10, 10, 3)), float(10 ** rng.uniform(-2, -0.5)), m)
    spec.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1),
                       pos=tuple(c0 + [0, 0, 30]))
    rays = _rays_towards(rng, 500_000, c0 - 10, c0 + 10, (c0 + [-20, -20, 20], c0 + [20, 20, 40]))
    rays[:100_000, 7] = rng.uniform(5.0, 40.0, size=100_000)
    ga = _any_hit_matches(spec, rays, 0, leaf=1, oracle=False, max_extra=5)
    assert (ga >= 0).mean() > 0.05


@pytest.mark.parametrize("which", ["cornell", "glass_hdri"])
def test_render_equals_full_nodes(gpu, which):
    """Shadow rays over the quantised nodes give every sample the radiance
    they give over the 128-byte nodes (the same visibility; up to the order
    of the float atomics of a vertex's several shadow rays)."""
    from pathtracer_amd import render as R, scenes, _lib as L
    if which == "cornell":
        spec = scenes.cornell(width=48, height=48, spp_x=2, spp_y=2)
    else:
        spec = scenes.glass_hdri(width=32, height=32, spp_x=2, spp_y=2, subdiv=5, env_w=128, env_h=64, max_depth=8)
    a = R.Scene(spec, build_flags=L.BUILD_ANY_HIT_FULL_NODES)
    b = R.Scene(spec, build_flags=L.BUILD_ANY_HIT_QNODES)
    for est in (0, 1):
        cfg = a.cfg(flags=L.RENDER_KEEP_SAMPLES, estimator=est)
        ia, _ = a.render(cfg)
        sa = a.samples(cfg)
        ib, _ = b.render(cfg)
        sb = b.samples(cfg)
        np.testing.assert_allclose(sb, sa, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(ib, ia, rtol=1e-5, atol=1e-6)
