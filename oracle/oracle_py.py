"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_P = ctypes.c_void_p


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    u32, u64, i32, sz, f32, f64 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_size_t,
                                   ctypes.c_float, ctypes.c_double)
    sig = {
        "or_mix_bits": (u64, [u64]),
        "or_permutation_element": (u32, [u32, u32, u32]),
        "or_murmur64a": (u64, [ctypes.c_char_p, sz, u64]),
        "or_hash_base_digit_seed": (u64, [i32, i32, u32]),
        "or_radical_inverse": (f32, [ctypes.c_int, u64]),
        "or_owen_radical_inverse": (f64, [ctypes.c_int, u64, u32]),
        "or_scrambled_radical_inverse": (f64, [ctypes.c_int, u64, u32]),
        "or_halton_dim": (f32, [u32, u32, u32, u64]),
        "or_sobol_matrix_value": (u32, [u32, u32]),
        "or_sobol_owen": (f32, [u32, u32, u32]),
        "or_stratified": (f32, [u32, u32, u32, u32, u32]),
        "or_rng": (f32, [u32, u32, u32]),
        "or_pixel_seed": (u32, [u32, u32]),
        "or_sample_batch": (ctypes.c_int, [u32, u32, u32, _P, _P, _P, sz, _P]),
        "or_bsdf_eval_batch": (ctypes.c_int, [_P, _P, _P, _P, _P, sz, _P, _P]),
        "or_bsdf_sample_batch": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, sz, _P, _P, _P, _P, _P]),
        "or_scene_create": (_P, []),
        "or_scene_destroy": (None, [_P]),
        "or_scene_add_material": (ctypes.c_int, [_P, _P]),
        "or_scene_add_mesh": (ctypes.c_int, [_P, _P, u32, _P, u32, u32, u32]),
        "or_scene_add_mesh_normals": (ctypes.c_int, [_P, _P, u32, _P, u32, _P, u32, _P, u32, u32]),
        "or_scene_add_sphere": (ctypes.c_int, [_P, _P, f32, u32, u32]),
        "or_scene_add_rect": (ctypes.c_int, [_P, ctypes.c_int, f32, f32, f32, f32, f32, u32, u32]),
        "or_scene_set_env": (ctypes.c_int, [_P, _P, u32, u32, f32]),
        "or_scene_set_camera": (ctypes.c_int, [_P, _P]),
        "or_scene_set_medium": (ctypes.c_int, [_P, _P, _P, f32, f32]),
        "or_scene_add_curve": (ctypes.c_int, [_P, _P, f32, f32, u32, u32]),
        "or_scene_add_texture_image": (ctypes.c_int, [_P, _P, u32, u32]),
        "or_scene_add_texture_checker": (ctypes.c_int, [_P, _P, _P, f32]),
        "or_scene_set_mesh_uv": (ctypes.c_int, [_P, u32, _P, u32, _P, u32]),
        "or_texture_batch": (ctypes.c_int, [_P, u32, _P, sz, _P]),
        "or_hit_uv_batch": (ctypes.c_int, [_P, _P, sz, _P]),
        "or_scene_commit": (ctypes.c_int, [_P]),
        "or_scene_num_lights": (ctypes.c_int, [_P]),
        "or_intersect_batch": (ctypes.c_int, [_P, _P, sz, ctypes.c_int, _P, _P]),
        "or_brute_intersect_batch": (ctypes.c_int, [_P, _P, sz, _P, _P]),
        "or_surface_batch": (ctypes.c_int, [_P, _P, sz, _P, _P]),
        "or_light_sample_batch": (ctypes.c_int, [_P, u32, _P, _P, sz, _P, _P, _P, _P]),
        "or_env_sample_batch": (ctypes.c_int, [_P, _P, sz, _P, _P]),
        "or_env_pdf": (f32, [_P, _P]),
        "or_camera_rays": (ctypes.c_int, [_P, _P, _P, _P, sz, _P]),
        "or_render": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, ctypes.c_int]),
        "or_render_range": (f64, [_P, _P, u64, u64, ctypes.c_int, _P]),
        "or_scene_set_guide": (ctypes.c_int, [_P, _P, u32, u32]),
        "or_guide_eval_batch": (ctypes.c_int, [_P, _P, _P, sz, _P, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def sref(struct):
    return ctypes.cast(ctypes.byref(struct), ctypes.c_void_p)


class OracleScene:
    """A SceneSpec loaded into the CPU oracle."""

    def __init__(self, spec):
        # structs of the public header, shared with the product binding
        from pathtracer_amd import _lib as L
        from pathtracer_amd.scenes import camera_desc
        self.L = L
        self.spec = spec
        self.h = lib.or_scene_create()
        for mi, (kind, params) in enumerate(spec.materials):
            m = L.material(kind, params, getattr(spec, "material_flags", {}).get(mi, 0))
            lib.or_scene_add_material(self.h, sref(m))
        for tex in getattr(spec, "textures", []):
            if tex[0] == "image":
                rgb = np.ascontiguousarray(tex[1], np.float32)
                if lib.or_scene_add_texture_image(self.h, ptr(rgb), rgb.shape[1], rgb.shape[0]) < 0:
                    raise ValueError("bad texture")
            else:
                c1, c2 = (np.asarray(c, np.float32) for c in tex[1:3])
                lib.or_scene_add_texture_checker(self.h, ptr(c1), ptr(c2), tex[3])
        n_mesh = 0
        for kind, i in spec.order:
            if kind == "mesh":
                pos, idx, mat, flags = spec.meshes[i]
                if i in spec.mesh_normals:
                    nrm, nidx = spec.mesh_normals[i]
                    rc = lib.or_scene_add_mesh_normals(self.h, ptr(pos), pos.shape[0], ptr(idx), idx.shape[0],
                                                       ptr(nrm), nrm.shape[0], ptr(nidx), mat, flags)
                else:
                    rc = lib.or_scene_add_mesh(self.h, ptr(pos), pos.shape[0], ptr(idx), idx.shape[0], mat, flags)
                if rc < 0:
                    raise ValueError("bad mesh")
                if i in getattr(spec, "mesh_uvs", {}):
                    uv, uvidx = spec.mesh_uvs[i]
                    if lib.or_scene_set_mesh_uv(self.h, n_mesh, ptr(uv), uv.shape[0], ptr(uvidx), uvidx.shape[0]) < 0:
                        raise ValueError("bad mesh uvs")
                n_mesh += 1
            elif kind == "sphere":
                c, r, mat, flags = spec.spheres[i]
                ca = np.asarray(c, np.float32)
                lib.or_scene_add_sphere(self.h, ptr(ca), r, mat, flags)
            elif kind == "curve":
                cps, w0, w1, _ctype, mat, flags = spec.curves[i]
                lib.or_scene_add_curve(self.h, ptr(np.ascontiguousarray(cps, np.float32)), w0, w1, mat, flags)
            else:
                axis, k, a0, a1, b0, b1, mat, flags = spec.rects[i]
                lib.or_scene_add_rect(self.h, axis, k, a0, a1, b0, b1, mat, flags)
        if spec.env is not None:
            rgb, scale = spec.env
            rgb = np.ascontiguousarray(rgb, np.float32)
            lib.or_scene_set_env(self.h, ptr(rgb), rgb.shape[1], rgb.shape[0], scale)
        self.cam = camera_desc(**spec.camera)
        lib.or_scene_set_camera(self.h, sref(self.cam))
        if getattr(spec, "medium", None) is not None:
            sa, ss, scale, _le, _les, gph = spec.medium
            lib.or_scene_set_medium(self.h, ptr(np.asarray(sa, np.float32)), ptr(np.asarray(ss, np.float32)),
                                    scale, gph)
        lib.or_scene_commit(self.h)

    def __del__(self):
        try:
            if self.h:
                lib.or_scene_destroy(self.h)
        except Exception:
            pass

    def cfg(self, **overrides):
        from pathtracer_amd.scenes import render_cfg
        kw = dict(self.spec.render)
        kw.update(overrides)
        return render_cfg(**kw)

    def render(self, cfg, want_var=False, want_samples=False, threads=0):
        W, H = cfg.width, cfg.height
        spp = cfg.spp_x * cfg.spp_y
        sppc = cfg.spp_count or (spp - cfg.spp_begin)
        img = np.empty((H, W, 3), np.float32)
        var = np.empty((H, W, 3), np.float32) if want_var else None
        smp = np.empty((H * W, sppc, 3), np.float32) if want_samples else None
        cnt = np.zeros(8, np.uint64)
        rc = lib.or_render(self.h, sref(cfg), ptr(img), ptr(var), ptr(smp), ptr(cnt), threads)
        if rc != 0:
            raise RuntimeError("or_render failed")
        return img, var, smp, cnt

    def render_range(self, cfg, first, count, threads=0):
        cnt = np.zeros(8, np.uint64)
        lum = lib.or_render_range(self.h, sref(cfg), first, count, threads, ptr(cnt))
        return lum, cnt

    def intersect(self, rays, any_hit=False, brute=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        if brute:
            lib.or_brute_intersect_batch(self.h, ptr(rays), n, ptr(prim), ptr(t))
        else:
            lib.or_intersect_batch(self.h, ptr(rays), n, int(any_hit), ptr(prim), ptr(t))
        return prim, t

    def surface(self, rays):
        """Shading normal (against the ray) and is_inside (-1: miss) of the closest hits."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        nrm = np.empty((n, 3), np.float32)
        inside = np.empty(n, np.int32)
        lib.or_surface_batch(self.h, ptr(rays), n, ptr(nrm), ptr(inside))
        return nrm, inside

    def texture(self, tex_id, uv):
        """Texture tex_id at uv (n, 2): (n, 3) colours."""
        uv = np.ascontiguousarray(uv, np.float32)
        out = np.empty((uv.shape[0], 3), np.float32)
        if lib.or_texture_batch(self.h, tex_id, ptr(uv), uv.shape[0], ptr(out)):
            raise ValueError("bad texture id")
        return out

    def hit_uv(self, rays):
        """HitRecord::texCoords of the closest hits ((-1, -1) on a miss)."""
        rays = np.ascontiguousarray(rays, np.float32)
        uv = np.empty((rays.shape[0], 2), np.float32)
        lib.or_hit_uv_batch(self.h, ptr(rays), rays.shape[0], ptr(uv))
        return uv

    def camera_rays(self, cfg, pixel, k):
        pixel = np.ascontiguousarray(pixel, np.uint32)
        k = np.ascontiguousarray(k, np.uint32)
        out = np.empty((pixel.shape[0], 8), np.float32)
        lib.or_camera_rays(self.h, sref(cfg), ptr(pixel), ptr(k), pixel.shape[0], ptr(out))
        return out

    def light_sample(self, light, u2, ref):
        u2 = np.ascontiguousarray(u2, np.float32)
        ref = np.ascontiguousarray(ref, np.float32)
        n = u2.shape[0]
        p, nrm, le = (np.empty((n, 3), np.float32) for _ in range(3))
        pdf = np.empty(n, np.float32)
        if lib.or_light_sample_batch(self.h, light, ptr(u2), ptr(ref), n, ptr(p), ptr(nrm), ptr(pdf), ptr(le)):
            raise ValueError("bad light index")
        return p, nrm, pdf, le

    def env_sample(self, u2):
        u2 = np.ascontiguousarray(u2, np.float32)
        n = u2.shape[0]
        d = np.empty((n, 3), np.float32)
        pdf = np.empty(n, np.float32)
        lib.or_env_sample_batch(self.h, ptr(u2), n, ptr(d), ptr(pdf))
        return d, pdf

    def set_guide(self, mix, kinds, verts):
        """The guided estimator (cfg estimator OR_EST_GUIDED = 6) with one fixed
        processed mixture (48 floats) at every guided point."""
        self._gmix = np.ascontiguousarray(mix, np.float32).reshape(48)
        lib.or_scene_set_guide(self.h, ptr(self._gmix), kinds, verts)

    def env_pdf(self, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32)
        return np.array([lib.or_env_pdf(self.h, ptr(dirs[i])) for i in range(dirs.shape[0])], np.float32)


def sample(sampler, seed, dims, index, pixel=None, spp=1):
    dims = np.ascontiguousarray(dims, np.uint32)
    index = np.ascontiguousarray(index, np.uint64)
    pix = None if pixel is None else np.ascontiguousarray(pixel, np.uint32)
    out = np.empty(dims.shape[0], np.float32)
    lib.or_sample_batch(sampler, seed, spp, ptr(dims), ptr(index), ptr(pix), dims.shape[0], ptr(out))
    return out


def bsdf_eval(kind, params, wo, wi, nrm, inside=None):
    from pathtracer_amd import _lib as L
    m = L.material(kind, params)
    wo, wi, nrm = (np.ascontiguousarray(a, np.float32) for a in (wo, wi, nrm))
    n = wo.shape[0]
    ins = None if inside is None else np.ascontiguousarray(inside, np.uint8)
    f = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    lib.or_bsdf_eval_batch(sref(m), ptr(wo), ptr(wi), ptr(nrm), ptr(ins), n, ptr(f), ptr(pdf))
    return f, pdf


def bsdf_sample(kind, params, wo, nrm, uc, u2, aux=None, inside=None):
    from pathtracer_amd import _lib as L
    m = L.material(kind, params)
    wo, nrm, uc, u2 = (np.ascontiguousarray(a, np.float32) for a in (wo, nrm, uc, u2))
    n = wo.shape[0]
    ins = None if inside is None else np.ascontiguousarray(inside, np.uint8)
    ax = None if aux is None else np.ascontiguousarray(aux, np.float32)
    wi, f = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    flags = np.empty(n, np.uint32)
    ok = np.empty(n, np.uint8)
    lib.or_bsdf_sample_batch(sref(m), ptr(wo), ptr(nrm), ptr(ins), ptr(uc), ptr(u2), ptr(ax), n, ptr(wi), ptr(f),
                             ptr(pdf), ptr(flags), ptr(ok))
    return wi, f, pdf, flags, ok


EST_GUIDED = 6


def guide_eval(mix, dirs=None, u=None):
    """Per row of mix (n, 48): the mixture pdf at dirs (n, 3) and/or a sample
    direction from uniforms u (n, 3)."""
    mix = np.ascontiguousarray(mix, np.float32)
    n = mix.shape[0]
    d = None if dirs is None else np.ascontiguousarray(dirs, np.float32)
    uu = None if u is None else np.ascontiguousarray(u, np.float32)
    pdf = np.empty(n, np.float32) if d is not None else None
    smp = np.empty((n, 3), np.float32) if uu is not None else None
    lib.or_guide_eval_batch(ptr(mix), ptr(d), ptr(uu), n, ptr(pdf), ptr(smp))
    return pdf, smp


def processed_mixture(raw):
    """The 48-float processed mixture of raw outputs (n, 41), from the numpy
    guide oracle's pi, kappa, mu and c: per lobe (kappa mu, log(pi kappa /
    (2 pi (1 - e^{-2 kappa}))) - kappa), then the 8 weights, then c."""
    from oracle import guide_oracle as G
    pi, kappa, mu, c = G.mixture(raw)
    n = pi.shape[0]
    out = np.zeros((n, 48), np.float64)
    for i in range(8):
        out[:, 4 * i:4 * i + 3] = kappa[:, i:i + 1] * mu[:, i]
        out[:, 4 * i + 3] = np.log(pi[:, i]) + G.log_norm(kappa[:, i]) - kappa[:, i]
    out[:, 32:40] = pi
    out[:, 40] = c
    return out.astype(np.float32)
