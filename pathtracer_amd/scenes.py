"""Scene descriptions and the seeded synthetic scenes of the benchmark workloads.

A SceneSpec is plain data (numpy arrays and parameter lists).  It is fed
unchanged to the GPU renderer (pathtracer_amd.render) and, in tests and the
CPU-baseline leg of bench.py, to the CPU oracle, so both sides see the same
scene.

Workloads (sizes as stated in each builder):
  cornell()       256x256, 16 spp Halton, depth 8: Cornell box of 20 triangles,
                  gold-conductor icosphere (1,280 triangles), Lambertian sphere,
                  0.5x0.5 ceiling light of radiance 15.
  glass_hdri()    1024x1024, 64 spp Sobol, depth 16: noise-displaced icosphere of
                  1,310,720 triangles, rough glass (IOR 1.5, alpha 0.05), copper
                  ground (alpha 0.3), 1024x512 synthetic sky + sun environment,
                  thin-lens depth of field (lens diameter 0.02).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib as L

GOLD_ETA, GOLD_K = (0.14, 0.43, 1.38), (4.54, 2.455, 1.914)  # commonMaterials.hpp:43-47
COPPER_ETA, COPPER_K = (0.200438, 0.924033, 1.10221), (3.91295, 2.45285, 2.14219)


@dataclasses.dataclass
class SceneSpec:
    materials: list = dataclasses.field(default_factory=list)   # (kind, params)
    meshes: list = dataclasses.field(default_factory=list)      # (pos f32 (n,3), idx i32 (m,3), mat, flags)
    spheres: list = dataclasses.field(default_factory=list)     # (center(3), r, mat, flags)
    rects: list = dataclasses.field(default_factory=list)       # (axis, k, d0min, d0max, d1min, d1max, mat, flags)
    env: tuple | None = None                                     # (rgb f32 (h,w,3) row y=0 is direction -Y, scale)
    camera: dict = dataclasses.field(default_factory=dict)
    render: dict = dataclasses.field(default_factory=dict)
    order: list = dataclasses.field(default_factory=list)        # insertion order of ("mesh"|"sphere"|"rect", index)
    mesh_normals: dict = dataclasses.field(default_factory=dict)  # mesh index -> (normals f32 (k,3), nidx i32 (m,3))
    curves: list = dataclasses.field(default_factory=list)      # (cps f32 (4,3), width0, width1, type, mat, flags)
    # global HomogeneousMedium: (sigma_a(3), sigma_s(3), sigma_scale, Le(3), Le_scale, g) or None
    medium: tuple | None = None
    name: str = ""
    # textures: ("image", rgb f32 (h,w,3), row 0 the bottom) or ("checker", c1, c2, scale)
    textures: list = dataclasses.field(default_factory=list)
    material_flags: dict = dataclasses.field(default_factory=dict)  # material index -> pt_material_desc.flags
    mesh_uvs: dict = dataclasses.field(default_factory=dict)        # mesh index -> (uv f32 (k,2), uvidx i32 (m,3))

    def add_material(self, kind, params, texture=None) -> int:
        """texture: a texture id (add_texture_*) the colour parameter comes from."""
        self.materials.append((kind, list(params)))
        if texture is not None:
            self.material_flags[len(self.materials) - 1] = L.mat_texture(texture)
        return len(self.materials) - 1

    def add_texture_image(self, rgb) -> int:
        """ImageTexture: rgb (h, w, 3), row 0 the image's bottom row."""
        self.textures.append(("image", np.ascontiguousarray(rgb, np.float32)))
        return len(self.textures) - 1

    def add_texture_checker(self, c1, c2, scale=10.0) -> int:
        """CheckerTexture(color1, color2, scale)."""
        self.textures.append(("checker", tuple(float(x) for x in c1), tuple(float(x) for x in c2), float(scale)))
        return len(self.textures) - 1

    def set_mesh_uv(self, mesh, uv, uvidx):
        """MeshData::uvs / uvIndices of mesh `mesh` (in adding order)."""
        self.mesh_uvs[mesh] = (np.ascontiguousarray(uv, np.float32), np.ascontiguousarray(uvidx, np.int32))

    def add_mesh(self, pos, idx, mat, flags=0, normals=None, nidx=None):
        self.meshes.append((np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(idx, np.int32), mat, flags))
        self.order.append(("mesh", len(self.meshes) - 1))
        if normals is not None:
            self.mesh_normals[len(self.meshes) - 1] = (np.ascontiguousarray(normals, np.float32),
                                                       np.ascontiguousarray(nidx, np.int32))

    def add_sphere(self, c, r, mat, flags=0):
        self.spheres.append((tuple(float(x) for x in c), float(r), mat, flags))
        self.order.append(("sphere", len(self.spheres) - 1))

    def add_rect(self, axis, k, d0min, d0max, d1min, d1max, mat, flags=0):
        self.rects.append((axis, k, d0min, d0max, d1min, d1max, mat, flags))
        self.order.append(("rect", len(self.rects) - 1))

    def add_curve(self, cps, width0, width1, mat, ctype=1, flags=0):
        """Curve(cps, width0, width1, type, material) (curve.hpp:13-140)."""
        self.curves.append((np.asarray(cps, np.float32).reshape(4, 3), float(width0), float(width1), ctype, mat,
                            flags))
        self.order.append(("curve", len(self.curves) - 1))

    @property
    def n_triangles(self):
        return sum(m[1].shape[0] for m in self.meshes)


def render_cfg(**kw) -> L.RenderCfg:
    c = L.RenderCfg()
    defaults = dict(width=64, height=64, spp_x=1, spp_y=1, max_depth=8, sampler=L.SAMPLER_HALTON_OWEN,
                    estimator=L.EST_REFERENCE_NEE, rr_depth=0, light_samples=4, seed=0,
                    bg=(0.0, 0.0, 0.0), spp_begin=0, spp_count=0, wavefront=0, flags=0)
    defaults.update(kw)
    for k, v in defaults.items():
        if k == "bg":
            for i in range(3):
                c.bg[i] = float(v[i])
        else:
            setattr(c, k, int(v))
    return c


def camera_desc(fov_deg, aspect, focal_dist, aperture, up, front, pos, pos1=None, shutter=0.0,
                normalize_dir=1, fov_rad=None) -> L.CameraDesc:
    c = L.CameraDesc()
    c.fov = math.radians(fov_deg) if fov_rad is None else fov_rad
    c.aspect = aspect
    c.focal_dist = focal_dist
    c.aperture = aperture
    for i in range(3):
        c.up[i], c.front[i], c.pos[i] = up[i], front[i], pos[i]
        c.pos1[i] = (pos1 if pos1 is not None else pos)[i]
    c.shutter = shutter
    c.normalize_dir = normalize_dir
    return c


# ---------------------------------------------------------------- geometry helpers

def quad_tris(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    pos = np.array([p0, p1, p2, p3], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return pos, idx


def merge(parts):
    pos, idx, off = [], [], 0
    for p, i in parts:
        pos.append(p)
        idx.append(i + off)
        off += p.shape[0]
    return np.concatenate(pos).astype(np.float32), np.concatenate(idx).astype(np.int32)


def icosphere(subdiv: int):
    """Unit icosphere: 20*4^subdiv triangles, shared vertices."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f


def value_noise(p: np.ndarray, seed: int, octaves: int = 4, freq: float = 2.0) -> np.ndarray:
    """Seeded 3-D value noise in [-1, 1] (lattice values, smoothstep trilinear), octaves summed
    with weights 1/2^i and normalised."""
    rng = np.random.default_rng(seed)
    lattice = rng.uniform(-1.0, 1.0, size=(64, 64, 64))
    out = np.zeros(p.shape[0])
    norm = 0.0
    for o in range(octaves):
        q = p * (freq * 2 ** o) + 17.0 * o
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3 - 2 * fr)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    val = lattice[(i0[:, 0] + dx) & 63, (i0[:, 1] + dy) & 63, (i0[:, 2] + dz) & 63]
                    wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                         (w[:, 2] if dz else 1 - w[:, 2])
                    acc += wt * val
        amp = 0.5 ** o
        out += amp * acc
        norm += amp
    return out / norm


def displaced_icosphere(subdiv, radius, amp, seed, center=(0, 0, 0)):
    v, f = icosphere(subdiv)
    r = radius * (1.0 + amp * value_noise(v, seed))
    pos = v * r[:, None] + np.asarray(center, np.float64)
    return pos.astype(np.float32), f.astype(np.int32)


def sky_hdri(w=1024, h=512, seed=1, zenith=0.3, horizon=1.0, sun_peak=5e4, sun_width_deg=0.5):
    """Synthetic equirect sky: gradient from horizon to zenith over the upper
    hemisphere, dim ground, plus a Gaussian sun lobe.  Texel (x, y) is
    direction theta = (y+0.5)/h*pi from -Y, phi = (x+0.5)/w*2pi with
    dir = (-sin th cos ph, -cos th, sin th sin ph) (texture.hpp:105-115)."""
    rng = np.random.default_rng(seed)
    y = (np.arange(h) + 0.5) / h * np.pi
    x = (np.arange(w) + 0.5) / w * 2 * np.pi
    th, ph = np.meshgrid(y, x, indexing="ij")
    d = np.stack([-np.sin(th) * np.cos(ph), -np.cos(th), np.sin(th) * np.sin(ph)], -1)
    elev = np.clip(d[..., 1], -1, 1)
    sky = np.where(elev >= 0, horizon + (zenith - horizon) * elev, 0.25 * horizon)
    tint = np.array([0.85, 0.92, 1.0])
    rgb = sky[..., None] * tint
    az = rng.uniform(0, 2 * np.pi)
    el = np.radians(rng.uniform(25, 50))
    sun = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    ang = np.arccos(np.clip(d @ sun, -1, 1))
    sig = np.radians(sun_width_deg)
    rgb = rgb + (sun_peak * np.exp(-0.5 * (ang / sig) ** 2))[..., None] * np.array([1.0, 0.95, 0.85])
    return rgb.astype(np.float32)


# ---------------------------------------------------------------- workloads

def cornell(width=256, height=256, spp_x=4, spp_y=4, max_depth=8, rr_depth=3,
            sampler=L.SAMPLER_HALTON_OWEN, estimator=L.EST_REFERENCE_NEE, ico_subdiv=3, smooth=False) -> SceneSpec:
    """Cornell box, side 2.0, made of 10 quads (20 triangles).  smooth: the
    icosphere carries per-vertex shading normals (the unit sphere's)."""
    s = SceneSpec(name="cornell")
    white = s.add_material(L.MAT_LAMBERTIAN, [0.73, 0.73, 0.73])
    red = s.add_material(L.MAT_LAMBERTIAN, [0.65, 0.05, 0.05])
    green = s.add_material(L.MAT_LAMBERTIAN, [0.12, 0.45, 0.15])
    light = s.add_material(L.MAT_EMITTER, [1.0, 1.0, 1.0, 15.0])
    gold = s.add_material(L.MAT_CONDUCTOR, [*GOLD_ETA, *GOLD_K, 0.2, 0.2])
    diffuse = s.add_material(L.MAT_LAMBERTIAN, [0.8, 0.8, 0.8])
    a, b = -1.0, 1.0
    # each wall in two halves
    white_parts = [
        quad_tris((a, 0, a), (0, 0, a), (0, 0, b), (a, 0, b)), quad_tris((0, 0, a), (b, 0, a), (b, 0, b), (0, 0, b)),
        quad_tris((a, 2, a), (a, 2, b), (0, 2, b), (0, 2, a)), quad_tris((0, 2, a), (0, 2, b), (b, 2, b), (b, 2, a)),
        quad_tris((a, 0, a), (a, 1, a), (b, 1, a), (b, 0, a)), quad_tris((a, 1, a), (a, 2, a), (b, 2, a), (b, 1, a)),
    ]
    red_parts = [quad_tris((a, 0, a), (a, 0, 0), (a, 2, 0), (a, 2, a)),
                 quad_tris((a, 0, 0), (a, 0, b), (a, 2, b), (a, 2, 0))]
    green_parts = [quad_tris((b, 0, a), (b, 2, a), (b, 2, 0), (b, 0, 0)),
                   quad_tris((b, 0, 0), (b, 2, 0), (b, 2, b), (b, 0, b))]
    s.add_mesh(*merge(white_parts), white)
    s.add_mesh(*merge(red_parts), red)
    s.add_mesh(*merge(green_parts), green)
    s.add_rect(1, 1.999, -0.25, 0.25, -0.25, 0.25, light)
    v, f = icosphere(ico_subdiv)
    s.add_mesh((v * 0.35 + np.array([-0.4, 0.35, -0.3])).astype(np.float32), f.astype(np.int32), gold,
               normals=v.astype(np.float32) if smooth else None, nidx=f.astype(np.int32) if smooth else None)
    s.add_sphere((0.45, 0.3, 0.3), 0.3, diffuse)
    s.camera = dict(fov_deg=40.0, aspect=width / height, focal_dist=1.0, aperture=0.0, up=(0, 1, 0),
                    front=(0, 0, -1), pos=(0, 1.0, 3.9))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth, sampler=sampler,
                    estimator=estimator, rr_depth=rr_depth, light_samples=4, seed=7, bg=(0, 0, 0))
    return s


def texture_image(w=64, h=48, seed=6) -> np.ndarray:
    """A procedural RGB image (row 0 the bottom): stripes over value noise."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    p = np.stack([x / w * 4.0, y / h * 4.0, np.zeros_like(x)], -1).reshape(-1, 3)
    n = value_noise(p, seed, octaves=3).reshape(h, w)
    stripes = 0.5 + 0.5 * np.sin(x * 0.6 + 3.0 * n)
    rgb = np.stack([0.15 + 0.7 * stripes, 0.1 + 0.5 * n, 0.2 + 0.6 * (1 - stripes) * n], -1)
    return np.clip(rgb, 0.0, 1.0).astype(np.float32)


def textured_room(width=64, height=64, spp_x=2, spp_y=2, max_depth=8, rr_depth=3,
                  sampler=L.SAMPLER_HALTON_OWEN, estimator=L.EST_REFERENCE_NEE) -> SceneSpec:
    """The Cornell box with ImageTexture and CheckerTexture materials: a
    checker floor rect, an image-textured back wall (rect), an image-textured
    sphere, a Disney mesh whose base colour is the image over its uvs, and a
    textured icosphere without uvs (the planar fallback)."""
    s = cornell(width, height, spp_x, spp_y, max_depth, rr_depth, sampler, estimator, ico_subdiv=2)
    s.name = "textured_room"
    img = s.add_texture_image(texture_image())
    chk = s.add_texture_checker((0.15, 0.2, 0.1), (0.85, 0.8, 0.75), 12.0)
    floor = s.add_material(L.MAT_LAMBERTIAN, [0, 0, 0], texture=chk)
    wall = s.add_material(L.MAT_LAMBERTIAN, [0, 0, 0], texture=img)
    disney = s.add_material(L.MAT_DISNEY, [0, 0, 0, 0.0, 0.4, 0.0, 0.5, 0.35, 0.0, 0.2, 0.0, 0.5, 0.3, 0.6, 1.5],
                            texture=img)
    s.add_rect(1, 0.001, -0.95, 0.95, -0.95, 0.95, floor)
    s.add_rect(2, -0.999, -0.9, 0.9, 0.05, 1.9, wall)
    s.add_sphere((0.1, 1.25, -0.5), 0.22, wall)
    pos = np.array([[-0.9, 0.6, -0.6], [-0.5, 0.6, -0.95], [-0.5, 1.4, -0.95], [-0.9, 1.4, -0.6]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    s.add_mesh(pos, idx, disney)
    s.set_mesh_uv(len(s.meshes) - 1, np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), idx)
    v, f = icosphere(2)
    s.add_mesh((v * 0.2 + np.array([0.55, 1.0, -0.3])).astype(np.float32), f.astype(np.int32), wall)
    return s


def glass_hdri(width=1024, height=1024, spp_x=8, spp_y=8, max_depth=16, subdiv=8, sampler=L.SAMPLER_SOBOL,
               estimator=L.EST_REFERENCE_NEE, rr_depth=3, env_w=1024, env_h=512, seed=1) -> SceneSpec:
    """Noise-displaced rough-glass icosphere on a copper plane under a sky."""
    s = SceneSpec(name="glass_hdri")
    glass = s.add_material(L.MAT_MFDIELECTRIC, [1.5, 0.05, 0.05])
    copper = s.add_material(L.MAT_CONDUCTOR, [*COPPER_ETA, *COPPER_K, 0.3, 0.3])
    pos, idx = displaced_icosphere(subdiv, 1.0, 0.08, seed)
    s.add_mesh(pos, idx, glass)
    g = 20.0
    y0 = -1.1
    s.add_mesh(*quad_tris((-g, y0, -g), (-g, y0, g), (g, y0, g), (g, y0, -g)), copper)
    s.env = (sky_hdri(env_w, env_h, seed), 1.0)
    focal = 4.0
    s.camera = dict(fov_deg=35.0, aspect=width / height, focal_dist=focal, aperture=focal / 0.02,
                    up=(0, 1, 0), front=(0, -0.1, -1), pos=(0, 0.4, 4.0))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth, sampler=sampler,
                    estimator=estimator, rr_depth=rr_depth, light_samples=4, seed=11, bg=(0, 0, 0))
    return s


EUMELANIN_SIGMA_A = (0.419, 0.697, 1.37)  # absorption per unit eumelanin concentration


def disney_hair_grid(width=1024, height=1024, spp_x=8, spp_y=16, max_depth=12, seed=3,
                     sampler=L.SAMPLER_STRATIFIED, estimator=L.EST_REFERENCE_NEE, rr_depth=3) -> SceneSpec:
    """Material-divergence workload: an 8x8 grid of analytic spheres with Disney
    principled BSDFs whose parameters are drawn uniform in [0,1], a ninth row of
    8 Marschner-hair spheres (sigma_a from eumelanin in [0,1.5], longitudinal
    roughness in [0.1,0.6], azimuthal roughness 0.3, IOR 1.55), one 4x4 quad
    light of radiance 8.  72 materials, 72 spheres, 1 light."""
    rng = np.random.default_rng(seed)
    s = SceneSpec(name="disney_hair_grid")
    r = 0.42
    for row in range(8):
        for col in range(8):
            base = rng.uniform(0, 1, 3)
            metallic, roughness, specular, spec_tint, sheen, clearcoat, gloss, subsurface, aniso, spec_trans = \
                rng.uniform(0, 1, 10)
            m = s.add_material(L.MAT_DISNEY, [*base, spec_trans, metallic, subsurface, specular, roughness,
                                              spec_tint, aniso, sheen, 0.5, clearcoat, gloss, 1.5])
            s.add_sphere((col - 3.5, row + 0.5, 0.0), r, m)
    for col in range(8):
        mel = rng.uniform(0, 1.5)
        beta_m = rng.uniform(0.1, 0.6)
        m = s.add_material(L.MAT_HAIR, [mel * EUMELANIN_SIGMA_A[0], mel * EUMELANIN_SIGMA_A[1],
                                        mel * EUMELANIN_SIGMA_A[2], 0.0, 1.55, beta_m, 0.3, math.radians(2.0)])
        s.add_sphere((col - 3.5, 8.5, 0.0), r, m)
    light = s.add_material(L.MAT_EMITTER, [1.0, 1.0, 1.0, 8.0])
    s.add_rect(1, 11.0, -2.0, 2.0, 2.0, 6.0, light)          # 4x4, facing down, in front of the grid
    floor = s.add_material(L.MAT_LAMBERTIAN, [0.5, 0.5, 0.5])
    s.add_rect(1, 0.0, -30, 30, -30, 30, floor)
    s.camera = dict(fov_deg=45.0, aspect=width / height, focal_dist=1.0, aperture=0.0, up=(0, 1, 0),
                    front=(0, 0, -1), pos=(0, 4.5, 12.0))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth, sampler=sampler,
                    estimator=estimator, rr_depth=rr_depth, light_samples=4, seed=17, bg=(0.02, 0.02, 0.03))
    return s


def furnace(albedo=0.5, env_value=1.0, max_depth=64, width=8, height=8, spp_x=8, spp_y=8) -> SceneSpec:
    """Lambertian sphere of albedo a under a uniform environment of radiance E.
    A convex object sees only the environment, so every camera ray that hits
    the sphere has expectation a*E under an unbiased estimator, and 2*a*E
    under the reference estimator, which counts the environment both through
    NEE and through the BSDF ray that misses (maxdepth_nee.hpp:26-37, 107-145);
    rays that miss see E.  Closed-form check of both estimators."""
    s = SceneSpec(name="furnace")
    m = s.add_material(L.MAT_LAMBERTIAN, [albedo] * 3)
    s.add_sphere((0, 0, 0), 1.0, m)
    s.env = (np.full((16, 32, 3), env_value, np.float32), 1.0)
    s.camera = dict(fov_deg=20.0, aspect=width / height, focal_dist=1.0, aperture=0.0, up=(0, 1, 0),
                    front=(0, 0, -1), pos=(0, 0, 5.0))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth,
                    sampler=L.SAMPLER_HALTON_OWEN, estimator=L.EST_NEE_MIS, rr_depth=0, light_samples=4, seed=3,
                    bg=(0, 0, 0))
    return s


def geodesic_sphere(freq: int):
    """Unit geodesic sphere: every icosahedron face cut into freq^2 triangles
    (20*freq^2 in all), vertices projected to the sphere and shared."""
    v0, f0 = icosphere(0)
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    keep = ii + jj <= freq
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((freq + 1, freq + 1), np.int64)
    lid[ii, jj] = np.arange(ii.size)
    # local triangles of one face: up (i,j)(i+1,j)(i,j+1), down (i+1,j)(i+1,j+1)(i,j+1)
    tri = []
    for i in range(freq):
        for j in range(freq - i):
            tri.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if i + j + 1 < freq:
                tri.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    tri = np.asarray(tri, np.int64)
    a, b, c = v0[f0[:, 0]], v0[f0[:, 1]], v0[f0[:, 2]]
    pts = (a[:, None, :] + (b - a)[:, None, :] * (ii[None, :, None] / freq) +
           (c - a)[:, None, :] * (jj[None, :, None] / freq)).reshape(-1, 3)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    faces = (tri[None, :, :] + (np.arange(20) * ii.size)[:, None, None]).reshape(-1, 3)
    # points on shared edges are computed once per face: merge them
    key = np.round(pts * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return pts[first], inv.reshape(-1)[faces]


BATCH_RES = (128, 256, 512, 1024)


def batch_instance(rng: np.random.Generator, index: int, tri_range=(1e4, 5e5), res_choices=BATCH_RES,
                   spp_range=(4, 256), depth_range=(8, 16), estimator=L.EST_NEE_MIS) -> SceneSpec:
    """One render of the batched workload: a noise-displaced geodesic sphere of
    N triangles (N log-uniform in tri_range) with a material drawn from
    {Lambertian, conductor, rough dielectric, Disney}, on a Lambertian floor,
    lit by 1-4 quad area lights of radiance U[2, 20]; resolution drawn from
    res_choices (square), spp log-uniform in spp_range, max depth uniform in
    depth_range."""
    s = SceneSpec(name=f"batch_{index}")
    n_tri = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
    freq = max(1, int(round((n_tri / 20.0) ** 0.5)))
    kind = int(rng.integers(4))
    if kind == 0:
        mat = s.add_material(L.MAT_LAMBERTIAN, list(rng.uniform(0.2, 0.9, 3)))
    elif kind == 1:
        a = float(rng.uniform(0.05, 0.5))
        mat = s.add_material(L.MAT_CONDUCTOR, [*GOLD_ETA, *GOLD_K, a, a] if rng.uniform() < 0.5
                             else [*COPPER_ETA, *COPPER_K, a, a])
    elif kind == 2:
        a = float(rng.uniform(0.0, 0.3))
        mat = s.add_material(L.MAT_MFDIELECTRIC, [float(rng.uniform(1.3, 1.8)), a, a])
    else:
        base = rng.uniform(0, 1, 3)
        metallic, roughness, specular, spec_tint, sheen, clearcoat, gloss, subsurface, aniso, spec_trans = \
            rng.uniform(0, 1, 10)
        mat = s.add_material(L.MAT_DISNEY, [*base, spec_trans, metallic, subsurface, specular, roughness,
                                            spec_tint, aniso, sheen, 0.5, clearcoat, gloss, 1.5])
    v, f = geodesic_sphere(freq)
    r = 1.0 + 0.08 * value_noise(v, int(rng.integers(1 << 30)))
    s.add_mesh((v * r[:, None]).astype(np.float32), f.astype(np.int32), mat)
    floor = s.add_material(L.MAT_LAMBERTIAN, [0.6, 0.6, 0.6])
    s.add_mesh(*quad_tris((-8, -1.1, -8), (-8, -1.1, 8), (8, -1.1, 8), (8, -1.1, -8)), floor)
    for _ in range(int(rng.integers(1, 5))):
        le = float(rng.uniform(2.0, 20.0))
        lm = s.add_material(L.MAT_EMITTER, [1.0, 1.0, 1.0, le])
        cx, cz = rng.uniform(-2.5, 2.5, 2)
        h = float(rng.uniform(0.5, 1.5)) / 2
        s.add_rect(1, float(rng.uniform(2.5, 4.0)), cx - h, cx + h, cz - h, cz + h, lm)
    res = int(rng.choice(res_choices))
    spp = int(round(np.exp(rng.uniform(np.log(spp_range[0]), np.log(spp_range[1])))))
    sx = max(d for d in range(1, int(spp ** 0.5) + 1) if spp % d == 0)
    depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
    az = float(rng.uniform(0, 2 * np.pi))
    pos = (4.2 * np.sin(az), 0.8, 4.2 * np.cos(az))
    front = (-np.sin(az), -0.15, -np.cos(az))
    s.camera = dict(fov_deg=40.0, aspect=1.0, focal_dist=1.0, aperture=0.0, up=(0, 1, 0), front=front, pos=pos)
    s.render = dict(width=res, height=res, spp_x=sx, spp_y=spp // sx, max_depth=depth,
                    sampler=L.SAMPLER_HALTON_OWEN, estimator=estimator, rr_depth=3, light_samples=4,
                    seed=1000 + index, bg=(0, 0, 0))
    return s


def batch64(seed=4, n=64, **kw):
    """The batched workload: n independent renders drawn from one seed."""
    rng = np.random.default_rng(seed)
    return [batch_instance(rng, i, **kw) for i in range(n)]


def guiding_room(width=1024, height=1024, spp_x=16, spp_y=16, max_depth=16, rr_depth=3, seed=5,
                 statue_freq=25, slit=0.05, estimator=L.EST_NEE_MIS) -> SceneSpec:
    """Indirect-dominated interior: a closed 4x3x3 room (x in [0,4], y in
    [0,3], z in [0,3]) of Lambertian walls (albedo 0.7), a Disney floor
    (roughness 0.3, metallic 0), a 12.5k-triangle noise-displaced glossy
    conductor statue, and one quad light of radiance 200 outside the left
    wall (0.4 x 0.4, 1.2 behind it), seen only through a vertical slit `slit`
    wide in that wall."""
    s = SceneSpec(name="guiding_room")
    wall = s.add_material(L.MAT_LAMBERTIAN, [0.7, 0.7, 0.7])
    floor = s.add_material(L.MAT_DISNEY, [0.6, 0.55, 0.5, 0.0, 0.0, 0.0, 0.5, 0.3, 0.0, 0.0, 0.0, 0.5, 0.0, 1.0,
                                          1.5])
    statue = s.add_material(L.MAT_CONDUCTOR, [*GOLD_ETA, *GOLD_K, 0.15, 0.15])
    light = s.add_material(L.MAT_EMITTER, [1.0, 1.0, 1.0, 200.0])
    X, Y, Z = 4.0, 3.0, 3.0
    s.add_rect(1, 0.0, 0.0, X, 0.0, Z, floor)          # floor
    s.add_rect(1, Y, 0.0, X, 0.0, Z, wall)             # ceiling
    s.add_rect(2, 0.0, 0.0, X, 0.0, Y, wall)           # back  (z = 0)
    s.add_rect(2, Z, 0.0, X, 0.0, Y, wall)             # front (z = 3)
    s.add_rect(0, X, 0.0, Y, 0.0, Z, wall)             # right (x = 4)
    z0, z1 = Z / 2 - slit / 2, Z / 2 + slit / 2
    y0, y1 = 0.6, 2.4                                   # the slit's extent
    s.add_rect(0, 0.0, 0.0, Y, 0.0, z0, wall)          # left wall (x = 0) around the slit
    s.add_rect(0, 0.0, 0.0, Y, z1, Z, wall)
    s.add_rect(0, 0.0, 0.0, y0, z0, z1, wall)
    s.add_rect(0, 0.0, y1, Y, z0, z1, wall)
    # outside, 1.2 behind the wall, emitting towards +x: the beam through the slit
    # lights a band of the right wall (z in ~[0.75, 2.25]) and parts of the floor
    # and ceiling; the rest of the room, and most of what the camera sees, is lit
    # only after two or more bounces
    s.add_rect(0, -1.2, 1.3, 1.7, 1.3, 1.7, light)
    v, f = geodesic_sphere(statue_freq)
    r = 0.55 * (1.0 + 0.08 * value_noise(v, seed))
    s.add_mesh((v * r[:, None] + np.array([2.5, 0.58, 1.2])).astype(np.float32), f.astype(np.int32), statue)
    pos = np.array([3.7, 1.7, 2.85])
    look = np.array([1.2, 0.9, 0.9])
    front = look - pos
    front = tuple(front / np.linalg.norm(front))
    s.camera = dict(fov_deg=60.0, aspect=width / height, focal_dist=1.0, aperture=0.0, up=(0, 1, 0),
                    front=front, pos=tuple(pos))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth,
                    sampler=L.SAMPLER_HALTON_OWEN, estimator=estimator, rr_depth=rr_depth, light_samples=4,
                    seed=23, bg=(0, 0, 0))
    return s


def _medium_scene(name, sigma_a, sigma_s, g, second_light, width, height, spp_x, spp_y):
    s = SceneSpec(name=name)
    m = s.add_material(L.MAT_EMITTER, [0.4, 2.32, 3.2, 1.0])
    s.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    if second_light:
        m2 = s.add_material(L.MAT_EMITTER, [24.0, 10.0, 24.0, 1.0])
        s.add_sphere((-3.0, 0.0, -1.5), 1.0, m2)
    s.medium = (sigma_a, sigma_s, 1.0, (0.0, 0.0, 0.0), 1.0, g)
    s.camera = dict(fov_deg=45.0, aspect=width / height, focal_dist=10.0, aperture=1000.0, up=(0, 1, 0),
                    front=(0, 0, 1), pos=(0, 0, -3))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=1,
                    sampler=L.SAMPLER_HALTON_OWEN, estimator=L.EST_VOLUME, rr_depth=0, light_samples=4, seed=3,
                    bg=(0, 0, 0))
    return s


def absorption_only_medium(width=128, height=128, spp_x=4, spp_y=4) -> SceneSpec:
    """absorption_only_medium (scenes/scenes_medium.hpp:6-31): an emissive
    sphere in a purely absorbing medium (sigma_a 1.5)."""
    return _medium_scene("absorption_only_medium", (1.5, 1.5, 1.5), (0.0, 0.0, 0.0), 0.0, False,
                         width, height, spp_x, spp_y)


def single_scatter_medium(width=128, height=128, spp_x=4, spp_y=4, g=0.0) -> SceneSpec:
    """single_scatter_medium (scenes/scenes_medium.hpp:33-61): two emissive
    spheres in a scattering medium (sigma_a 0.1, sigma_s 0.7, HG g)."""
    return _medium_scene("single_scatter_medium", (0.1, 0.1, 0.1), (0.7, 0.7, 0.7), g, True,
                         width, height, spp_x, spp_y)


def hair_curves(width=96, height=96, spp_x=4, spp_y=4, n_curves=60, seed=9, estimator=L.EST_NEE_MIS,
                max_depth=8, kajiya_kay=False) -> SceneSpec:
    """A tuft of hair Bezier curves (the curve primitive with its hair
    tangent) over a Lambertian floor under a quad light: Marschner hair of
    three melanin levels, or Kajiya-Kay fibres of three colours."""
    rng = np.random.default_rng(seed)
    s = SceneSpec(name="hair_curves")
    floor = s.add_material(L.MAT_LAMBERTIAN, [0.6, 0.6, 0.6])
    light = s.add_material(L.MAT_EMITTER, [1.0, 1.0, 1.0, 12.0])
    mats = []
    for mel in (0.3, 0.8, 1.3):
        if kajiya_kay:
            tint = np.exp(-mel * np.array(EUMELANIN_SIGMA_A))
            mats.append(s.add_material(L.MAT_KAJIYA_KAY, [*(0.6 * tint), *(0.3 * tint), 40.0]))
        else:
            mats.append(s.add_material(L.MAT_HAIR, [mel * EUMELANIN_SIGMA_A[0], mel * EUMELANIN_SIGMA_A[1],
                                                    mel * EUMELANIN_SIGMA_A[2], 0.0, 1.55, 0.3, 0.3,
                                                    math.radians(2.0)]))
    s.add_rect(1, 0.0, -3, 3, -3, 3, floor)
    s.add_rect(1, 3.0, -0.75, 0.75, -0.75, 0.75, light)
    for i in range(n_curves):
        x0, z0 = rng.uniform(-0.8, 0.8, 2)
        lean = rng.normal(size=2) * 0.3
        cps = np.array([[x0, 0.0, z0],
                        [x0 + lean[0] * 0.3, 0.5, z0 + lean[1] * 0.3],
                        [x0 + lean[0] * 0.8, 1.0, z0 + lean[1] * 0.8],
                        [x0 + lean[0] * 1.5, 1.4, z0 + lean[1] * 1.5]])
        s.add_curve(cps, 0.03, 0.008, mats[i % 3])
    s.camera = dict(fov_deg=40.0, aspect=width / height, focal_dist=1.0, aperture=0.0, up=(0, 1, 0),
                    front=(0, -0.25, -1), pos=(0, 1.2, 4.0))
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth,
                    sampler=L.SAMPLER_HALTON_OWEN, estimator=estimator, rr_depth=3, light_samples=4, seed=5,
                    bg=(0.05, 0.05, 0.06))
    return s
