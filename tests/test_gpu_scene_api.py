"""The C++ scene API end to end on the GPU: the reference-style scene of
tests/cpp/scene_api_check.cpp rendered through SceneInfo::render must equal
the same lowered calls rendered through the Python binding (same library,
same sample streams; only the order of the floating-point additions of
shadow-ray contributions may differ), and agree per sample with the oracle.
The lowering itself is pinned on the CPU in tests/test_scene_api.py."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import check_samples, TETRA_OBJ, dump_to_spec, parse_dump

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "scene_api_check")


def test_scene_api_render_matches_binding_and_oracle(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/scene_api_check is not built (run __graft_entry__.build())")
    from pathtracer_amd import _lib as L, film_io, render as R
    obj = tmp_path / "tetra.obj"
    obj.write_text(TETRA_OBJ)
    dump = tmp_path / "dump.txt"
    subprocess.run([EXE, "dump", str(obj), str(dump)], check=True, timeout=60)
    pfm = tmp_path / "scene.pfm"
    print("scene_api_check render", flush=True)
    subprocess.run([EXE, "render", str(obj), str(pfm)], check=True, timeout=120)
    cxx = film_io.read_pfm(str(pfm))

    spec = dump_to_spec(parse_dump(str(dump)))
    g = R.Scene(spec)
    cfg = g.cfg(flags=L.RENDER_KEEP_SAMPLES)
    img, st = g.render(cfg)
    assert cxx.shape == img.shape
    np.testing.assert_allclose(cxx, img, rtol=1e-4, atol=1e-5)
    assert st["nan_samples"] == 0

    gs = g.samples(cfg)
    o = O.OracleScene(spec)
    _, _, os_, _ = o.render(o.cfg(), want_samples=True)
    check_samples(gs, os_, "scene API")
