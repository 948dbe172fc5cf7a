"""The C ABI library loads and exports every symbol include/pathtracer_amd.h
declares (no compute calls: this runs without a GPU)."""
import os
import re

from pathtracer_amd import _lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pathtracer_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert "pt_render" in names and "pt_scene_commit" in names and "pt_sample" in names
    assert set(names) == set(L.EXPORTED)


def test_library_exports_every_declared_symbol():
    for name in declared_functions():
        assert hasattr(L.lib, name), name


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """ctypes mirrors of the public structs have the sizes gcc gives the header's."""
    import ctypes
    import subprocess
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "pathtracer_amd.h"\nint main(void){printf("%zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(pt_material_desc), sizeof(pt_camera_desc), sizeof(pt_render_cfg), sizeof(pt_stats),'
                   'sizeof(pt_build_cfg), sizeof(pt_scene_info)); printf("%zu\\n", sizeof(pt_tuning)); return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [ctypes.sizeof(c) for c in (L.MaterialDesc, L.CameraDesc, L.RenderCfg, L.Stats, L.BuildCfg,
                                              L.SceneInfo, L.Tuning)]
    assert got[0] == 128


def test_errors_are_reported_not_thrown():
    import ctypes
    h = ctypes.c_void_p()
    assert L.lib.pt_scene_create(ctypes.byref(h)) == 0
    bad = L.material(99, [])
    assert L.lib.pt_scene_add_material(h, ctypes.byref(bad), None) == 1   # PT_ERR_INVALID
    assert b"unknown material kind" in L.lib.pt_last_error()
    assert L.lib.pt_scene_add_rect(h, 5, 0, 0, 1, 0, 1, 0, 0) == 1        # bad axis
    assert L.lib.pt_scene_commit(h, None) == 1                            # empty scene
    assert b"no objects" in L.lib.pt_last_error()
    L.lib.pt_scene_destroy(h)


def test_update_mesh_checks_its_arguments():
    import ctypes
    import numpy as np
    h = ctypes.c_void_p()
    assert L.lib.pt_scene_create(ctypes.byref(h)) == 0
    mid = ctypes.c_uint32()
    assert L.lib.pt_scene_add_material(h, ctypes.byref(L.material(1, [0.5, 0.5, 0.5])), ctypes.byref(mid)) == 0
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    assert L.lib.pt_scene_add_mesh(h, L.ptr(pos), 4, L.ptr(idx), 2, 0, 0) == 0
    assert L.lib.pt_scene_update_mesh(h, 0, L.ptr(pos + 1), 4) == 0
    assert L.lib.pt_scene_update_mesh(h, 1, L.ptr(pos), 4) == 1
    assert b"mesh index out of range" in L.lib.pt_last_error()
    assert L.lib.pt_scene_update_mesh(h, 0, L.ptr(pos[:3].copy()), 3) == 1
    assert b"vertex count" in L.lib.pt_last_error()
    bad = pos.copy()
    bad[2, 1] = np.nan
    assert L.lib.pt_scene_update_mesh(h, 0, L.ptr(bad), 4) == 1
    assert b"non-finite" in L.lib.pt_last_error()
    L.lib.pt_scene_destroy(h)
