"""ctypes binding of libpathtracer_amd.so (the C ABI of include/pathtracer_amd.h).

The library is loaded from this package directory only; there is no CPU
fallback.  If the HIP build is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PATHTRACER_AMD_LIB selects another build of the same library (e.g. one
# compiled with different occupancy targets).
LIB_PATH = os.environ.get("PATHTRACER_AMD_LIB") or os.path.join(_HERE, "libpathtracer_amd.so")

# pt_mat_kind
MAT_LAMBERTIAN, MAT_EMITTER, MAT_CONDUCTOR, MAT_MFDIELECTRIC = 1, 2, 3, 4
MAT_THINDIELECTRIC, MAT_PHONG, MAT_DISNEY, MAT_HAIR, MAT_KAJIYA_KAY = 5, 6, 7, 8, 9
# pt_sampler_kind
SAMPLER_HALTON_OWEN, SAMPLER_HALTON_PERM, SAMPLER_HALTON_NONE = 0, 1, 2
SAMPLER_SOBOL, SAMPLER_STRATIFIED = 3, 4
# pt_estimator
EST_REFERENCE_NEE, EST_NEE_MIS, EST_NAIVE, EST_ROULETTE, EST_RESERVOIR, EST_VOLUME = 0, 1, 2, 3, 4, 5
# render flags
RENDER_COUNTERS, RENDER_SORT_HITS, RENDER_ACCUMULATE = 1, 2, 4
RENDER_KEEP_SAMPLES, RENDER_TIMING = 8, 16
BATCH_NO_GRAPH = 256
TUNE_WAVEFRONT, TUNE_GRIDS, TUNE_LEAF, TUNE_SORT = 1, 2, 4, 8
BUILD_DEVICE_LBVH, BUILD_NO_LDS_NODES = 1, 2
BUILD_ANY_HIT_QNODES, BUILD_ANY_HIT_FULL_NODES = 4, 8
INFO_ANY_HIT_QNODES = 1
GUIDE_NO_TRAIN = 1
GUIDE_PATH_QUERIES = 2
GUIDE_INPUTS, GUIDE_OUTPUTS, GUIDE_NPARAMS, GUIDE_MIX = 6, 41, 27497, 48
PRIM_FLIP_FACE, PRIM_FORCE_INSIDE, PRIM_NOT_LIGHT = 1, 2, 4
NPARAMS = 30


class MaterialDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("p", ctypes.c_float * NPARAMS)]


class CameraDesc(ctypes.Structure):
    _fields_ = [("fov", ctypes.c_float), ("aspect", ctypes.c_float),
                ("focal_dist", ctypes.c_float), ("aperture", ctypes.c_float),
                ("up", ctypes.c_float * 3), ("front", ctypes.c_float * 3),
                ("pos", ctypes.c_float * 3), ("pos1", ctypes.c_float * 3),
                ("shutter", ctypes.c_float), ("normalize_dir", ctypes.c_uint32)]


class RenderCfg(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("spp_x", ctypes.c_uint32), ("spp_y", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("sampler", ctypes.c_uint32),
                ("estimator", ctypes.c_uint32), ("rr_depth", ctypes.c_uint32),
                ("light_samples", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("bg", ctypes.c_float * 3), ("spp_begin", ctypes.c_uint32),
                ("spp_count", ctypes.c_uint32), ("wavefront", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "paths", "segments", "shadow_rays", "node_visits", "prim_tests",
        "shadow_node_visits", "shadow_prim_tests", "nan_samples", "extend_launches")] + \
        [(n, ctypes.c_double) for n in (
            "ms_total", "ms_generate", "ms_extend", "ms_shade", "ms_shadow", "ms_film")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BuildCfg(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_uint32), ("sah_bins", ctypes.c_uint32),
                ("lds_levels", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class SceneInfo(ctypes.Structure):
    _fields_ = [("n_materials", ctypes.c_uint32), ("n_triangles", ctypes.c_uint32),
                ("n_spheres", ctypes.c_uint32), ("n_rects", ctypes.c_uint32),
                ("n_lights", ctypes.c_uint32), ("bvh_nodes", ctypes.c_uint32),
                ("bvh_depth", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64),
                ("build_ms", ctypes.c_double), ("bvh_ms", ctypes.c_double), ("flags", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("wavefront", ctypes.c_uint32), ("trav_blocks_per_cu", ctypes.c_uint32),
                ("shade_blocks_per_cu", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32),
                ("sort_bits", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AutotuneCfg(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("spp_div", ctypes.c_uint32), ("repeats", ctypes.c_uint32)]


class BatchCfg(ctypes.Structure):
    _fields_ = [("wavefront", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class GuideCfg(ctypes.Structure):
    _fields_ = [("train_spp", ctypes.c_uint32), ("first_pass_spp", ctypes.c_uint32), ("batch", ctypes.c_uint32),
                ("max_steps", ctypes.c_uint32), ("record_stride", ctypes.c_uint32),
                ("record_verts", ctypes.c_uint32), ("kinds", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("flags", ctypes.c_uint32), ("guide_verts", ctypes.c_uint32)]


class GuideStats(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_uint32), ("steps", ctypes.c_uint32), ("samples", ctypes.c_uint64),
                ("last_loss", ctypes.c_float), ("ms_train", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# every symbol include/pathtracer_amd.h declares
EXPORTED = [
    "pt_scene_create", "pt_scene_destroy", "pt_scene_add_material", "pt_scene_add_mesh", "pt_scene_add_mesh_normals",
    "pt_scene_add_sphere", "pt_scene_add_rect", "pt_scene_add_curve", "pt_scene_set_env", "pt_scene_set_camera",
    "pt_scene_set_medium", "pt_scene_update_mesh", "pt_scene_add_texture_image", "pt_scene_add_texture_checker",
    "pt_scene_set_mesh_uv",
    "pt_scene_commit", "pt_scene_get_info", "pt_render", "pt_render_noalloc", "pt_film_read", "pt_film_read_samples",
    "pt_scene_set_tuning", "pt_autotune",
    "pt_batch_create", "pt_batch_destroy", "pt_batch_add", "pt_batch_set_cfg", "pt_batch_render",
    "pt_batch_film_read", "pt_batch_set_tuning", "pt_batch_autotune",
    "pt_guide_create", "pt_guide_destroy", "pt_guide_get_params", "pt_guide_set_params", "pt_guide_forward",
    "pt_guide_eval", "pt_guide_gradient", "pt_guide_train_step", "pt_guide_samples", "pt_render_guided",
    "pt_guide_set_select_range",
    "pt_sample", "pt_intersect", "pt_bsdf_eval", "pt_bsdf_sample", "pt_light_sample", "pt_env_sample",
    "pt_device_synchronize", "pt_last_error", "pt_version",
]

_P = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()'). There is no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    i32, u32, sz = ctypes.c_int32, ctypes.c_uint32, ctypes.c_size_t
    sig = {
        "pt_scene_create": (i32, [ctypes.POINTER(_P)]),
        "pt_scene_destroy": (None, [_P]),
        "pt_scene_add_material": (i32, [_P, ctypes.POINTER(MaterialDesc), ctypes.POINTER(u32)]),
        "pt_scene_add_mesh": (i32, [_P, _P, u32, _P, u32, u32, u32]),
        "pt_scene_add_mesh_normals": (i32, [_P, _P, u32, _P, u32, _P, u32, _P, u32, u32]),
        "pt_scene_add_sphere": (i32, [_P, _P, ctypes.c_float, u32, u32]),
        "pt_scene_update_mesh": (i32, [_P, u32, _P, u32]),
        "pt_scene_add_texture_image": (i32, [_P, _P, u32, u32, ctypes.POINTER(u32)]),
        "pt_scene_add_texture_checker": (i32, [_P, _P, _P, ctypes.c_float, ctypes.POINTER(u32)]),
        "pt_scene_set_mesh_uv": (i32, [_P, u32, _P, u32, _P, u32]),
        "pt_scene_add_rect": (i32, [_P, ctypes.c_int] + [ctypes.c_float] * 5 + [u32, u32]),
        "pt_scene_add_curve": (i32, [_P, _P, ctypes.c_float, ctypes.c_float, u32, u32, u32]),
        "pt_scene_set_env": (i32, [_P, _P, u32, u32, ctypes.c_float]),
        "pt_scene_set_camera": (i32, [_P, ctypes.POINTER(CameraDesc)]),
        "pt_scene_set_medium": (i32, [_P, _P, _P, ctypes.c_float, _P, ctypes.c_float, ctypes.c_float]),
        "pt_scene_commit": (i32, [_P, ctypes.POINTER(BuildCfg)]),
        "pt_scene_get_info": (i32, [_P, ctypes.POINTER(SceneInfo)]),
        "pt_render": (i32, [_P, ctypes.POINTER(RenderCfg), ctypes.POINTER(Stats)]),
        "pt_render_noalloc": (i32, [_P, ctypes.POINTER(RenderCfg), _P, ctypes.POINTER(Stats)]),
        "pt_film_read": (i32, [_P, _P]),
        "pt_film_read_samples": (i32, [_P, _P, sz]),
        "pt_scene_set_tuning": (i32, [_P, ctypes.POINTER(Tuning)]),
        "pt_autotune": (i32, [_P, ctypes.POINTER(RenderCfg), ctypes.POINTER(AutotuneCfg), ctypes.POINTER(Tuning)]),
        "pt_batch_create": (i32, [ctypes.POINTER(_P)]),
        "pt_batch_destroy": (None, [_P]),
        "pt_batch_add": (i32, [_P, _P, ctypes.POINTER(RenderCfg), ctypes.POINTER(u32)]),
        "pt_batch_set_cfg": (i32, [_P, u32, ctypes.POINTER(RenderCfg)]),
        "pt_batch_render": (i32, [_P, ctypes.POINTER(BatchCfg), ctypes.POINTER(Stats)]),
        "pt_batch_film_read": (i32, [_P, u32, _P]),
        "pt_batch_set_tuning": (i32, [_P, ctypes.POINTER(Tuning)]),
        "pt_batch_autotune": (i32, [_P, ctypes.POINTER(AutotuneCfg), ctypes.POINTER(Tuning)]),
        "pt_guide_create": (i32, [_P, u32, ctypes.POINTER(_P)]),
        "pt_guide_destroy": (None, [_P]),
        "pt_guide_get_params": (i32, [_P, _P]),
        "pt_guide_set_params": (i32, [_P, _P]),
        "pt_guide_forward": (i32, [_P, _P, sz, _P, _P]),
        "pt_guide_eval": (i32, [_P, _P, _P, sz, _P, _P, _P]),
        "pt_guide_gradient": (i32, [_P, _P, sz, _P, ctypes.POINTER(ctypes.c_float)]),
        "pt_guide_train_step": (i32, [_P, _P, sz, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
        "pt_guide_samples": (i32, [_P, _P, sz, ctypes.POINTER(sz)]),
        "pt_guide_set_select_range": (i32, [_P, ctypes.c_float, ctypes.c_float]),
        "pt_render_guided": (i32, [_P, ctypes.POINTER(RenderCfg), _P, ctypes.POINTER(GuideCfg),
                                   ctypes.POINTER(Stats), ctypes.POINTER(GuideStats)]),
        "pt_sample": (i32, [u32, u32, u32, _P, _P, _P, sz, _P]),
        "pt_intersect": (i32, [_P, _P, sz, ctypes.c_int, _P, _P]),
        "pt_bsdf_eval": (i32, [ctypes.POINTER(MaterialDesc), _P, _P, _P, _P, sz, _P, _P]),
        "pt_bsdf_sample": (i32, [ctypes.POINTER(MaterialDesc), _P, _P, _P, _P, _P, _P, sz,
                                 _P, _P, _P, _P, _P]),
        "pt_light_sample": (i32, [_P, u32, _P, _P, sz, _P, _P, _P, _P]),
        "pt_env_sample": (i32, [_P, _P, sz, _P, _P]),
        "pt_device_synchronize": (i32, []),
        "pt_last_error": (ctypes.c_char_p, []),
        "pt_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class PTError(RuntimeError):
    pass


def check(status: int, what: str = ""):
    if status != 0:
        msg = lib.pt_last_error().decode()
        raise PTError(f"{what}: status {status}: {msg}")


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def mat_texture(tex_id: int) -> int:
    """PT_MAT_TEXTURE(id): the material's colour comes from texture tex_id."""
    return 0x1 | (int(tex_id) << 8)


def material(kind: int, params, flags: int = 0) -> MaterialDesc:
    m = MaterialDesc()
    m.kind = kind
    m.flags = flags
    for i, v in enumerate(params):
        m.p[i] = float(v)
    return m
