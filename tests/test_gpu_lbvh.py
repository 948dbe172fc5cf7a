"""The on-device BVH build (PT_BUILD_DEVICE_LBVH) against the oracle.

A BVH only decides which primitives a ray tests, never the result: closest
hits must be the oracle's primitive ids and t (within 1e-5 relative; ids may
differ only on ties, whose winner depends on traversal order), any-hit must
agree, and renders must equal those over the host SAH tree sample for sample.
The cases cover the build's special shapes: one primitive, a handful, all
centroids coincident (equal Morton codes, so the tree is split by index),
every primitive type, every leaf size, a second commit of a grown scene on
the same handle (scratch reuse), and frames of a moving, deforming mesh
(pt_scene_update_mesh).
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.test_gpu_intersect import _compare, _rays_towards

pytestmark = pytest.mark.gpu


def _scene(spec, leaf=4):
    from pathtracer_amd import render as R, _lib as L
    return R.Scene(spec, leaf_size=leaf, build_flags=L.BUILD_DEVICE_LBVH)


def _check(spec, rays, leaf=4, max_ties=50):
    g = _scene(spec, leaf)
    o = O.OracleScene(spec)
    gp, gt = g.intersect(rays)
    cp, ct = o.intersect(rays)
    ties = _compare(gp, gt, cp, ct)
    assert ties <= max_ties, ties
    ga, _ = g.intersect(rays, any_hit=True)
    ca, _ = o.intersect(rays, any_hit=True)
    assert np.array_equal(ga >= 0, ca >= 0)
    return g, gp


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


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_displaced_mesh(gpu, leaf):
    rng = np.random.default_rng(5)
    rays = _rays_towards(rng, 1_000_000, [-1.2, -1.2, -1.2], [2.5, 1.2, 1.2], ([-4, -1, 3], [4, 4, 6]))
    rays[:100_000, :3] = 0.0                     # from inside the sphere outwards
    d = rng.normal(size=(100_000, 3))
    rays[:100_000, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[100_000:200_000, 7] = rng.uniform(0.5, 6.0, size=100_000)
    g, gp = _check(_displaced_spec(), rays, leaf, max_ties=200)
    assert (gp >= 0).mean() > 0.5
    assert g.info.bvh_nodes > 0 and 0 < g.info.bvh_depth <= 31


def test_cornell_and_curves(gpu):
    from pathtracer_amd import scenes
    rng = np.random.default_rng(11)
    rays = _rays_towards(rng, 500_000, [-1, 0, -1], [1, 2, 1], ([-0.9, 0.1, -0.9], [0.9, 1.9, 3.0]))
    _check(scenes.cornell(ico_subdiv=4), rays)
    spec = scenes.hair_curves(n_curves=80)
    rays = _rays_towards(rng, 400_000, [-1.2, 0.0, -1.2], [1.2, 1.6, 1.2], ([-2, 0.2, -2], [2, 2.5, 4]))
    _, gp = _check(spec, rays, max_ties=20)
    assert int((gp >= len(spec.rects)).sum()) > 10_000      # curve segments hit


def _spheres_spec(centres, radii):
    from pathtracer_amd import scenes
    spec = scenes.SceneSpec()
    m = spec.add_material(1, [0.5, 0.5, 0.5])
    for c, r in zip(centres, radii):
        spec.add_sphere(tuple(c), float(r), m)
    spec.camera = dict(fov_deg=40, aspect=1, focal_dist=1, aperture=0, up=(0, 1, 0), front=(0, 0, -1),
                       pos=(0, 0, 5))
    return spec


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17])
def test_few_primitives(gpu, n):
    rng = np.random.default_rng(n)
    c = rng.uniform(-1, 1, (n, 3))
    r = rng.uniform(0.1, 0.4, n)
    rays = _rays_towards(rng, 50_000, [-1.2] * 3, [1.2] * 3, ([-3, -3, 3], [3, 3, 5]))
    _, gp = _check(_spheres_spec(c, r), rays, leaf=1, max_ties=0)
    assert (gp >= 0).any()


def test_coincident_centroids(gpu):
    """Concentric spheres: one Morton code for all, the tree splits by index."""
    r = np.linspace(0.2, 1.0, 300)
    rays = _rays_towards(np.random.default_rng(3), 100_000, [-1] * 3, [1] * 3, ([-3, -3, 3], [3, 3, 5]))
    rays[:20_000, :3] = 0.0                      # from the centre: the innermost shell first
    _, gp = _check(_spheres_spec(np.zeros((300, 3)), r), rays, leaf=2, max_ties=0)
    assert np.all(gp[:20_000] == 0)


def test_recommit_grown_scene(gpu):
    """Commit, add primitives, commit again on the same handle."""
    from pathtracer_amd import _lib as L
    rng = np.random.default_rng(7)
    spec = _spheres_spec(rng.uniform(-1, 1, (40, 3)), rng.uniform(0.05, 0.2, 40))
    g = _scene(spec)
    extra = rng.uniform(-1, 1, (400, 3))
    for c in extra:
        spec.add_sphere(tuple(c), 0.08, 0)
        L.check(L.lib.pt_scene_add_sphere(g.h, L.ptr(np.asarray(c, np.float32)), ctypes.c_float(0.08), 0, 0),
                "pt_scene_add_sphere")
    b = L.BuildCfg()
    b.leaf_size, b.flags = 4, L.BUILD_DEVICE_LBVH
    L.check(L.lib.pt_scene_commit(g.h, ctypes.byref(b)), "pt_scene_commit")
    rays = _rays_towards(rng, 100_000, [-1.1] * 3, [1.1] * 3, ([-3, -3, 3], [3, 3, 5]))
    gp, gt = g.intersect(rays)
    cp, ct = O.OracleScene(spec).intersect(rays)
    assert _compare(gp, gt, cp, ct) == 0
    assert (gp >= 40).sum() > 1000


def test_render_equals_sah_tree(gpu):
    """Per-sample radiance over the device-built tree equals that over the host
    SAH tree (same hits, same arithmetic)."""
    from pathtracer_amd import render as R, scenes, _lib as L
    spec = scenes.cornell(width=64, height=64, spp_x=2, spp_y=2, ico_subdiv=4)
    a = R.Scene(spec)
    b = _scene(spec)
    cfg = a.cfg(flags=L.RENDER_KEEP_SAMPLES)
    ia, _ = a.render(cfg)
    sa = a.samples(cfg)
    ib, _ = b.render(cfg)
    sb = b.samples(cfg)
    same = np.all(np.abs(sa - sb) <= 1e-5 * np.maximum(np.abs(sa), np.abs(sb)) + 1e-7, axis=-1)
    assert same.mean() > 0.999, same.mean()
    np.testing.assert_allclose(ib, ia, rtol=2e-3, atol=2e-3)


def test_animated_mesh_frames(gpu):
    """Three frames of a moving, deforming mesh: pt_scene_update_mesh, a device
    re-build, then closest hits against the oracle of that frame's geometry."""
    from pathtracer_amd import _lib as L
    spec = _displaced_spec()
    pos0, idx, mat, flags = spec.meshes[0]
    g = _scene(spec)
    rng = np.random.default_rng(21)
    for f in range(3):
        a = 0.4 * (f + 1)
        rot = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        pos = (pos0 * (1 + 0.05 * np.sin(5 * pos0[:, 1:2] + f))) @ rot.T + [0.1 * f, 0, 0]
        pos = pos.astype(np.float32)
        g.update_mesh(0, pos)
        g.commit(build_flags=L.BUILD_DEVICE_LBVH)
        spec.meshes[0] = (pos, idx, mat, flags)
        rays = _rays_towards(rng, 300_000, [-1.2, -1.2, -1.2], [2.5, 1.2, 1.2], ([-4, -1, 3], [4, 4, 6]))
        gp, gt = g.intersect(rays)
        cp, ct = O.OracleScene(spec).intersect(rays)
        assert _compare(gp, gt, cp, ct) <= 60
        assert (gp >= 0).mean() > 0.5
