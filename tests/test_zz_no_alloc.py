"""pt_render_noalloc: the render entry point that allocates no device memory
(its wavefront lives in a static 64 MiB device array) against pt_render of
the same frame, through the C ABI.  Two calls in a row on one device, the
second with a smaller frame, so the second runs over scratch the first left
behind.  The film is the caller's device buffer (allocated here with
hipMalloc through the HIP runtime the library links)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library links."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def to_host(self, p, shape):
        out = np.empty(shape, np.float32)
        assert self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def _noalloc(L, scene, cfg, hip):
    film = hip.alloc(cfg.width * cfg.height * 3 * 4)
    st = L.Stats()
    try:
        L.check(L.lib.pt_render_noalloc(scene.h, ctypes.byref(cfg), film, ctypes.byref(st)), "pt_render_noalloc")
        img = hip.to_host(film, (cfg.height, cfg.width, 3))
    finally:
        hip.free(film)
    return img, st.as_dict()


def _frames():
    from pathtracer_amd import scenes
    # 96x96 x 16 spp = 147,456 paths: two chunks of the 131,072 slots
    big = scenes.cornell(width=96, height=96, spp_x=4, spp_y=4)
    small = scenes.glass_hdri(width=40, height=40, spp_x=2, spp_y=2, subdiv=4, env_w=128, env_h=64, max_depth=8)
    return big, small


@pytest.mark.parametrize("estimator", [1, 0, 4])
def test_noalloc_matches_render(gpu, estimator):
    from pathtracer_amd import render as R, _lib as L
    hip = _Hip()
    for spec in _frames():           # the second call is the smaller frame
        g = R.Scene(spec)
        cfg = g.cfg(estimator=estimator, flags=L.RENDER_COUNTERS)
        ref, rst = g.render(cfg)     # pt_render + pt_film_read
        img, st = _noalloc(L, g, cfg, hip)
        assert img.shape == ref.shape
        if estimator == 1:
            # one shadow ray per vertex: the same adds in the same order
            assert np.array_equal(img, ref), np.abs(img - ref).max()
        else:
            # several shadow rays of one vertex add to one slot with float
            # atomics, in any order
            np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-7)
        # node visits and primitive tests depend on which rays share a wave
        # (speculative traversal), so only the path counts must be equal
        for k in ("paths", "segments", "shadow_rays", "nan_samples"):
            assert st[k] == rst[k], k
        print(f"{spec.name} est {estimator}: {st['paths']} paths in {st['ms_total']:.1f} ms, "
              f"{st['extend_launches']} extension launches", flush=True)
        g.close()


def test_noalloc_limits(gpu):
    from pathtracer_amd import render as R, scenes, _lib as L
    hip = _Hip()
    g = R.Scene(scenes.cornell(width=16, height=16, spp_x=2, spp_y=2))
    film = hip.alloc(16 * 16 * 3 * 4)
    try:
        for kw, code in ((dict(wavefront=(1 << 17) + 1), 1), (dict(flags=L.RENDER_KEEP_SAMPLES), 4),
                         (dict(flags=L.RENDER_ACCUMULATE), 4), (dict(flags=L.RENDER_TIMING), 4)):
            cfg = g.cfg(**kw)
            assert L.lib.pt_render_noalloc(g.h, ctypes.byref(cfg), film, None) == code, kw
        assert L.lib.pt_render_noalloc(g.h, ctypes.byref(g.cfg()), None, None) == 1
        # a wavefront below the maximum works and gives the same film
        ref, _ = g.render(g.cfg(estimator=1))
        cfg = g.cfg(wavefront=300, estimator=1)
        assert L.lib.pt_render_noalloc(g.h, ctypes.byref(cfg), film, None) == 0
        assert np.array_equal(hip.to_host(film, (16, 16, 3)), ref)
    finally:
        hip.free(film)
