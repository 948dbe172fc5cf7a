"""GPU renderer: lowers a SceneSpec onto the C ABI (the role of SceneInfo,
render/renderManager.hpp:35-158) and runs the wavefront kernels."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .scenes import SceneSpec, camera_desc, render_cfg


class Scene:
    """A committed device scene."""

    def __init__(self, spec: SceneSpec, leaf_size: int = 4, sah_bins: int = 16, build_flags: int = 0):
        self.spec = spec
        h = ctypes.c_void_p()
        L.check(L.lib.pt_scene_create(ctypes.byref(h)), "pt_scene_create")
        self.h = h
        self._keep = []
        for mi, (kind, params) in enumerate(spec.materials):
            mid = ctypes.c_uint32()
            md = L.material(kind, params, spec.material_flags.get(mi, 0))
            L.check(L.lib.pt_scene_add_material(h, ctypes.byref(md), ctypes.byref(mid)), "pt_scene_add_material")
        for tex in spec.textures:
            tid = ctypes.c_uint32()
            if tex[0] == "image":
                rgb = np.ascontiguousarray(tex[1], np.float32)
                L.check(L.lib.pt_scene_add_texture_image(h, L.ptr(rgb), rgb.shape[1], rgb.shape[0], ctypes.byref(tid)),
                        "pt_scene_add_texture_image")
            else:
                c1, c2 = (np.asarray(c, np.float32) for c in tex[1:3])
                L.check(L.lib.pt_scene_add_texture_checker(h, L.ptr(c1), L.ptr(c2), tex[3], ctypes.byref(tid)),
                        "pt_scene_add_texture_checker")
        n_mesh = 0
        for kind, i in spec.order:
            if kind == "mesh":
                pos, idx, mat, flags = spec.meshes[i]
                if i in spec.mesh_normals:
                    nrm, nidx = spec.mesh_normals[i]
                    L.check(L.lib.pt_scene_add_mesh_normals(h, L.ptr(pos), pos.shape[0], L.ptr(idx), idx.shape[0],
                                                            L.ptr(nrm), nrm.shape[0], L.ptr(nidx), mat, flags),
                            "pt_scene_add_mesh_normals")
                else:
                    L.check(L.lib.pt_scene_add_mesh(h, L.ptr(pos), pos.shape[0], L.ptr(idx), idx.shape[0], mat,
                                                    flags), "pt_scene_add_mesh")
                if i in spec.mesh_uvs:
                    uv, uvidx = spec.mesh_uvs[i]
                    L.check(L.lib.pt_scene_set_mesh_uv(h, n_mesh, L.ptr(uv), uv.shape[0], L.ptr(uvidx),
                                                       uvidx.shape[0]), "pt_scene_set_mesh_uv")
                n_mesh += 1
            elif kind == "sphere":
                c, r, mat, flags = spec.spheres[i]
                ca = np.asarray(c, np.float32)
                L.check(L.lib.pt_scene_add_sphere(h, L.ptr(ca), r, mat, flags), "pt_scene_add_sphere")
            elif kind == "curve":
                cps, w0, w1, ctype, mat, flags = spec.curves[i]
                cps = np.ascontiguousarray(cps, np.float32)
                L.check(L.lib.pt_scene_add_curve(h, L.ptr(cps), w0, w1, ctype, mat, flags), "pt_scene_add_curve")
            else:
                axis, k, a0, a1, b0, b1, mat, flags = spec.rects[i]
                L.check(L.lib.pt_scene_add_rect(h, axis, k, a0, a1, b0, b1, mat, flags), "pt_scene_add_rect")
        if spec.env is not None:
            rgb, scale = spec.env
            rgb = np.ascontiguousarray(rgb, np.float32)
            L.check(L.lib.pt_scene_set_env(h, L.ptr(rgb), rgb.shape[1], rgb.shape[0], scale), "pt_scene_set_env")
        cam = camera_desc(**spec.camera)
        L.check(L.lib.pt_scene_set_camera(h, ctypes.byref(cam)), "pt_scene_set_camera")
        if spec.medium is not None:
            sa, ss, scale, le, les, gph = spec.medium
            f3 = lambda v: np.ascontiguousarray(v, np.float32)
            L.check(L.lib.pt_scene_set_medium(h, L.ptr(f3(sa)), L.ptr(f3(ss)), scale, L.ptr(f3(le)), les, gph),
                    "pt_scene_set_medium")
        self.commit(leaf_size, sah_bins, build_flags)

    def commit(self, leaf_size: int = 4, sah_bins: int = 16, build_flags: int = 0):
        """pt_scene_commit: builds the BVH and uploads the scene."""
        b = L.BuildCfg()
        b.leaf_size, b.sah_bins, b.flags = leaf_size, sah_bins, build_flags
        L.check(L.lib.pt_scene_commit(self.h, ctypes.byref(b)), "pt_scene_commit")
        info = L.SceneInfo()
        L.check(L.lib.pt_scene_get_info(self.h, ctypes.byref(info)), "pt_scene_get_info")
        self.info = info

    def update_mesh(self, mesh: int, pos: np.ndarray):
        """pt_scene_update_mesh: new vertex positions for mesh `mesh` (add order);
        takes effect at the next commit()."""
        pos = np.ascontiguousarray(pos, np.float32)
        L.check(L.lib.pt_scene_update_mesh(self.h, mesh, L.ptr(pos), pos.shape[0]), "pt_scene_update_mesh")

    def close(self):
        if self.h:
            L.lib.pt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cfg(self, **overrides) -> L.RenderCfg:
        kw = dict(self.spec.render)
        kw.update(overrides)
        return render_cfg(**kw)

    def render(self, cfg: L.RenderCfg | None = None, **overrides):
        """Returns (image (h, w, 3) per-pixel mean, stats dict)."""
        cfg = cfg or self.cfg(**overrides)
        st = L.Stats()
        L.check(L.lib.pt_render(self.h, ctypes.byref(cfg), ctypes.byref(st)), "pt_render")
        img = np.empty((cfg.height, cfg.width, 3), np.float32)
        L.check(L.lib.pt_film_read(self.h, L.ptr(img)), "pt_film_read")
        return img, st.as_dict()

    def render_device(self, cfg: L.RenderCfg) -> dict:
        st = L.Stats()
        L.check(L.lib.pt_render(self.h, ctypes.byref(cfg), ctypes.byref(st)), "pt_render")
        return st.as_dict()

    def samples(self, cfg: L.RenderCfg) -> np.ndarray:
        spp = cfg.spp_x * cfg.spp_y
        sppc = cfg.spp_count or (spp - cfg.spp_begin)
        n = cfg.width * cfg.height * sppc * 3
        out = np.empty(n, np.float32)
        L.check(L.lib.pt_film_read_samples(self.h, L.ptr(out), n), "pt_film_read_samples")
        return out.reshape(cfg.height * cfg.width, sppc, 3)

    def autotune(self, cfg: L.RenderCfg, flags: int = 0, spp_div: int = 8, repeats: int = 1) -> dict:
        """pt_autotune: picks wavefront / grid sizes (and the BVH leaf size) for cfg's frame."""
        tc = L.AutotuneCfg(flags, spp_div, repeats)
        t = L.Tuning()
        L.check(L.lib.pt_autotune(self.h, ctypes.byref(cfg), ctypes.byref(tc), ctypes.byref(t)), "pt_autotune")
        return t.as_dict()

    def set_tuning(self, wavefront=0, trav_blocks_per_cu=0, shade_blocks_per_cu=0):
        t = L.Tuning(wavefront, trav_blocks_per_cu, shade_blocks_per_cu, 0)
        L.check(L.lib.pt_scene_set_tuning(self.h, ctypes.byref(t)), "pt_scene_set_tuning")

    def intersect(self, rays: np.ndarray, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        L.check(L.lib.pt_intersect(self.h, L.ptr(rays), n, int(any_hit), L.ptr(prim), L.ptr(t)), "pt_intersect")
        return prim, t

    def env_sample(self, u2: np.ndarray):
        """pt_env_sample: importance-sampled environment directions and pdfs."""
        u2 = np.ascontiguousarray(u2, np.float32)
        n = u2.shape[0]
        d = np.empty((n, 3), np.float32)
        pdf = np.empty(n, np.float32)
        L.check(L.lib.pt_env_sample(self.h, L.ptr(u2), n, L.ptr(d), L.ptr(pdf)), "pt_env_sample")
        return d, pdf

    def light_sample(self, light: int, u2: np.ndarray, ref: np.ndarray):
        u2 = np.ascontiguousarray(u2, np.float32)
        ref = np.ascontiguousarray(ref, np.float32)
        n = u2.shape[0]
        p, nrm, le = (np.empty((n, 3), np.float32) for _ in range(3))
        pdf = np.empty(n, np.float32)
        L.check(L.lib.pt_light_sample(self.h, light, L.ptr(u2), L.ptr(ref), n, L.ptr(p), L.ptr(nrm), L.ptr(pdf),
                                      L.ptr(le)), "pt_light_sample")
        return p, nrm, pdf, le


def sample(sampler: int, seed: int, dims, index, pixel=None, spp: int = 1) -> np.ndarray:
    dims = np.ascontiguousarray(dims, np.uint32)
    index = np.ascontiguousarray(index, np.uint64)
    pix = None if pixel is None else np.ascontiguousarray(pixel, np.uint32)
    out = np.empty(dims.shape[0], np.float32)
    L.check(L.lib.pt_sample(sampler, seed, spp, L.ptr(dims), L.ptr(index), L.ptr(pix), dims.shape[0], L.ptr(out)),
            "pt_sample")
    return out


def bsdf_eval(kind: int, params, wo, wi, nrm, inside=None):
    m = L.material(kind, params)
    wo, wi, nrm = (np.ascontiguousarray(a, np.float32) for a in (wo, wi, nrm))
    n = wo.shape[0]
    ins = None if inside is None else np.ascontiguousarray(inside, np.uint8)
    f = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    L.check(L.lib.pt_bsdf_eval(ctypes.byref(m), L.ptr(wo), L.ptr(wi), L.ptr(nrm), L.ptr(ins), n, L.ptr(f),
                               L.ptr(pdf)), "pt_bsdf_eval")
    return f, pdf


def bsdf_sample(kind: int, params, wo, nrm, uc, u2, aux=None, inside=None):
    m = L.material(kind, params)
    wo, nrm, uc, u2 = (np.ascontiguousarray(a, np.float32) for a in (wo, nrm, uc, u2))
    n = wo.shape[0]
    ins = None if inside is None else np.ascontiguousarray(inside, np.uint8)
    ax = None if aux is None else np.ascontiguousarray(aux, np.float32)
    wi, f = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    flags = np.empty(n, np.uint32)
    ok = np.empty(n, np.uint8)
    L.check(L.lib.pt_bsdf_sample(ctypes.byref(m), L.ptr(wo), L.ptr(nrm), L.ptr(ins), L.ptr(uc), L.ptr(u2), L.ptr(ax),
                                 n, L.ptr(wi), L.ptr(f), L.ptr(pdf), L.ptr(flags), L.ptr(ok)), "pt_bsdf_sample")
    return wi, f, pdf, flags, ok


class Batch:
    """Many renders in one wavefront (pt_batch_*): each entry is a committed
    Scene and its RenderCfg; every render gets its own film."""

    def __init__(self):
        h = ctypes.c_void_p()
        L.check(L.lib.pt_batch_create(ctypes.byref(h)), "pt_batch_create")
        self.h = h
        self.cfgs = []
        self.scenes = []  # the scenes must outlive the batch

    def add(self, scene: Scene, cfg: L.RenderCfg) -> int:
        idx = ctypes.c_uint32()
        L.check(L.lib.pt_batch_add(self.h, scene.h, ctypes.byref(cfg), ctypes.byref(idx)), "pt_batch_add")
        self.cfgs.append(cfg)
        self.scenes.append(scene)
        return idx.value

    def set_cfg(self, index: int, cfg: L.RenderCfg):
        L.check(L.lib.pt_batch_set_cfg(self.h, index, ctypes.byref(cfg)), "pt_batch_set_cfg")
        self.cfgs[index] = cfg

    def render(self, wavefront: int = 0, flags: int = 0) -> dict:
        bc = L.BatchCfg(wavefront, flags)
        st = L.Stats()
        L.check(L.lib.pt_batch_render(self.h, ctypes.byref(bc), ctypes.byref(st)), "pt_batch_render")
        return st.as_dict()

    def film(self, index: int) -> np.ndarray:
        if not 0 <= index < len(self.cfgs):  # the library reports the error
            L.check(L.lib.pt_batch_film_read(self.h, index, L.ptr(np.empty(3, np.float32))), "pt_batch_film_read")
        c = self.cfgs[index]
        img = np.empty((c.height, c.width, 3), np.float32)
        L.check(L.lib.pt_batch_film_read(self.h, index, L.ptr(img)), "pt_batch_film_read")
        return img

    def autotune(self, flags: int = 0, spp_div: int = 8, repeats: int = 1) -> dict:
        tc = L.AutotuneCfg(flags, spp_div, repeats)
        t = L.Tuning()
        L.check(L.lib.pt_batch_autotune(self.h, ctypes.byref(tc), ctypes.byref(t)), "pt_batch_autotune")
        return t.as_dict()

    def set_tuning(self, wavefront=0, trav_blocks_per_cu=0, shade_blocks_per_cu=0):
        t = L.Tuning(wavefront, trav_blocks_per_cu, shade_blocks_per_cu, 0)
        L.check(L.lib.pt_batch_set_tuning(self.h, ctypes.byref(t)), "pt_batch_set_tuning")

    def close(self):
        if self.h:
            L.lib.pt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Guide:
    """The learned path-guiding network of a committed Scene (pt_guide_*)."""

    def __init__(self, scene: Scene, seed: int = 1):
        h = ctypes.c_void_p()
        L.check(L.lib.pt_guide_create(scene.h, seed, ctypes.byref(h)), "pt_guide_create")
        self.h = h
        self.scene = scene

    def params(self) -> np.ndarray:
        p = np.empty(L.GUIDE_NPARAMS, np.float32)
        L.check(L.lib.pt_guide_get_params(self.h, L.ptr(p)), "pt_guide_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (L.GUIDE_NPARAMS,)
        L.check(L.lib.pt_guide_set_params(self.h, L.ptr(p)), "pt_guide_set_params")

    def forward(self, x):
        """Raw outputs (n, 41) and processed mixtures (n, 48) of points x (n, 6)."""
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        raw = np.empty((n, L.GUIDE_OUTPUTS), np.float32)
        mix = np.empty((n, L.GUIDE_MIX), np.float32)
        L.check(L.lib.pt_guide_forward(self.h, L.ptr(x), n, L.ptr(raw), L.ptr(mix)), "pt_guide_forward")
        return raw, mix

    def gradient(self, samples):
        samples = np.ascontiguousarray(samples, np.float32)
        g = np.empty(L.GUIDE_NPARAMS, np.float32)
        loss = ctypes.c_float()
        L.check(L.lib.pt_guide_gradient(self.h, L.ptr(samples), samples.shape[0], L.ptr(g), ctypes.byref(loss)),
                "pt_guide_gradient")
        return loss.value, g

    def train_step(self, samples, lr=1e-3) -> float:
        samples = np.ascontiguousarray(samples, np.float32)
        loss = ctypes.c_float()
        L.check(L.lib.pt_guide_train_step(self.h, L.ptr(samples), samples.shape[0], lr, ctypes.byref(loss)),
                "pt_guide_train_step")
        return loss.value

    def set_select_range(self, c_min: float, c_max: float):
        L.check(L.lib.pt_guide_set_select_range(self.h, c_min, c_max), "pt_guide_set_select_range")

    def samples(self, cap: int = 1 << 21) -> np.ndarray:
        """The training samples of the last training pass (n, 11)."""
        out = np.empty((cap, 11), np.float32)
        n = ctypes.c_size_t()
        L.check(L.lib.pt_guide_samples(self.h, L.ptr(out), cap, ctypes.byref(n)), "pt_guide_samples")
        return out[:min(n.value, cap)].copy()

    def render(self, cfg: L.RenderCfg, gcfg: L.GuideCfg | None = None):
        """pt_render_guided: (image, stats dict, guide stats dict)."""
        gcfg = gcfg or L.GuideCfg()
        st, gs = L.Stats(), L.GuideStats()
        L.check(L.lib.pt_render_guided(self.scene.h, ctypes.byref(cfg), self.h, ctypes.byref(gcfg),
                                       ctypes.byref(st), ctypes.byref(gs)), "pt_render_guided")
        img = np.empty((cfg.height, cfg.width, 3), np.float32)
        L.check(L.lib.pt_film_read(self.scene.h, L.ptr(img)), "pt_film_read")
        return img, st.as_dict(), gs.as_dict()

    def close(self):
        if self.h:
            L.lib.pt_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def guide_eval(mix, dirs=None, u=None):
    """Guide pdf at dirs and/or samples (dir, pdf) from uniforms u, on the device."""
    mix = np.ascontiguousarray(mix, np.float32)
    n = mix.shape[0]
    pdf = sdir = spdf = None
    d = None if dirs is None else np.ascontiguousarray(dirs, np.float32)
    uu = None if u is None else np.ascontiguousarray(u, np.float32)
    if d is not None:
        pdf = np.empty(n, np.float32)
    if uu is not None:
        sdir = np.empty((n, 3), np.float32)
        spdf = np.empty(n, np.float32)
    L.check(L.lib.pt_guide_eval(L.ptr(mix), L.ptr(d), L.ptr(uu), n, L.ptr(pdf), L.ptr(sdir), L.ptr(spdf)),
            "pt_guide_eval")
    return pdf, sdir, spdf
