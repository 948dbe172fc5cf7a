"""Textured materials on the GPU against the oracle: ImageTexture and
CheckerTexture on Lambertian and the Disney base colour, over sphere, rect,
interpolated mesh uvs and the planar fallback (the oracle's textures and
texture coordinates are pinned in tests/test_oracle_texture.py).  Per-sample
bars as in tests/helpers.py (check_samples)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import check_samples, dump_to_spec, parse_dump

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "scene_api_check")


@pytest.mark.parametrize("estimator", [0, 1])
def test_textured_room_per_sample(gpu, estimator):
    from pathtracer_amd import render as R, scenes, _lib as L
    spec = scenes.textured_room(width=64, height=64, spp_x=2, spp_y=2, estimator=estimator)
    g = R.Scene(spec)
    cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES)
    _, st = g.render(cfg)
    o = O.OracleScene(spec)
    _, _, os_, _ = o.render(o.cfg(), want_samples=True)
    check_samples(g.samples(cfg), os_, f"textured room est {estimator}")
    assert st["nan_samples"] == 0


def test_one_colour_texture_renders_as_its_colour(gpu):
    """A checker whose two colours are the albedo, on every white surface of
    the Cornell box, against the untextured box: the same samples."""
    from pathtracer_amd import render as R, scenes, _lib as L
    plain = scenes.cornell(width=48, height=48, spp_x=2, spp_y=2)
    tex = scenes.cornell(width=48, height=48, spp_x=2, spp_y=2)
    t = tex.add_texture_checker((0.73, 0.73, 0.73), (0.73, 0.73, 0.73), 10.0)
    tex.material_flags[0] = L.mat_texture(t)
    out = []
    for spec in (plain, tex):
        g = R.Scene(spec)
        cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES, estimator=1)
        g.render(cfg)
        out.append(g.samples(cfg))
    check_samples(out[1], out[0], "one-colour checker vs plain albedo")


def _write_inputs(tmp_path):
    rng = np.random.default_rng(12)
    ppm8 = rng.integers(0, 256, size=(6, 8, 3), dtype=np.uint8)
    a = tmp_path / "a.ppm"
    with open(a, "wb") as f:
        f.write(b"P6\n8 6\n255\n" + ppm8.tobytes())
    pfm = rng.uniform(0.1, 1.0, size=(5, 4, 3)).astype("<f4")
    b = tmp_path / "b.pfm"
    with open(b, "wb") as f:
        f.write(b"PF\n4 5\n-1.0\n" + pfm.tobytes())
    return str(a), str(b)


def test_scene_api_textured_render(gpu, tmp_path):
    """The C++ scene API's textured scene (image and checker textures,
    uv meshes, rotated rect and mesh) through SceneInfo::render against the
    same lowered calls through the Python binding, and per sample against
    the oracle."""
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/scene_api_check is not built (run __graft_entry__.build())")
    from pathtracer_amd import _lib as L, film_io, render as R
    a, b = _write_inputs(tmp_path)
    dump = tmp_path / "tdump.txt"
    subprocess.run([EXE, "tdump", a, b, str(dump)], check=True, timeout=60)
    pfm = tmp_path / "scene.pfm"
    print("scene_api_check trender", flush=True)
    subprocess.run([EXE, "trender", a, b, str(pfm)], check=True, timeout=120)
    cxx = film_io.read_pfm(str(pfm))
    spec = dump_to_spec(parse_dump(str(dump)), width=48, height=48, spp_x=2, spp_y=2, max_depth=5)
    spec.render["bg"] = (0.05, 0.05, 0.05)
    g = R.Scene(spec)
    cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES)
    img, st = g.render(cfg)
    assert cxx.shape == img.shape
    np.testing.assert_allclose(cxx, img, rtol=1e-4, atol=1e-5)
    assert st["nan_samples"] == 0
    o = O.OracleScene(spec)
    _, _, os_, _ = o.render(o.cfg(), want_samples=True)
    check_samples(g.samples(cfg), os_, "textured scene API")
