"""The source-compatible C++ scene API (include/pathtracer_amd/scene.hpp), on
the CPU: a scene written in the reference's style (tests/cpp/scene_api_check.cpp)
lowers to the C ABI calls its wrappers imply.

* Translate / Rotate<Axis> / Scale flatten to world-space geometry: the
  expected positions are recomputed here from the reference's ray transforms
  (objectTransform.hpp:43-146, objectTransform.cpp:93-213) by inverting them.
* is_inside flags fold the wrapper chain (FlipFace toggles, transforms force
  true); emitters outside the top level are kept out of light discovery.
* loadOBJMesh (meshLoader.cpp:7-58): positive, negative and v/vt/vn index
  forms; non-triangle faces are an error.
* The PNG writer (framebuffer.cpp:27-58 semantics) decodes with zlib and
  passes its CRCs; the Radiance HDR reader matches the RGBE decode of a file
  written here with run-length-encoded scanlines.
"""
import math
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from tests.helpers import TETRA_OBJ, parse_dump

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "scene_api_check.cpp")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("sac")
    exe = str(d / "scene_api_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", exe, SRC, "-L", os.path.join(ROOT, "pathtracer_amd"), "-lpathtracer_amd",
                    "-Wl,-rpath," + os.path.join(ROOT, "pathtracer_amd")], check=True)
    obj = d / "tetra.obj"
    obj.write_text(TETRA_OBJ)
    return exe, str(obj), d


@pytest.fixture(scope="module")
def calls(driver):
    exe, obj, d = driver
    out = str(d / "dump.txt")
    subprocess.run([exe, "dump", obj, out], check=True)
    return parse_dump(out)


def rot_geom(axis, ang):
    """World-from-object map of Rotate<axis>(angle): the transpose of the ray rotation."""
    c, s = math.cos(ang), math.sin(ang)
    i, j = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
    ray = np.eye(3)
    ray[i, i], ray[i, j], ray[j, i], ray[j, j] = c, -s, s, c  # x' = c x - s z (Y), y' = c y - s z (X) ...
    return ray.T


def test_materials_lowered_once_in_first_use_order(calls):
    mats = [c for c in calls if c["kind"] == "material"]
    kinds = [m["mat_kind"] for m in mats]
    # red, white, emitter, gold conductor, Disney, rough glass (white shared by 4 rects and the box)
    assert kinds == [1, 1, 2, 3, 7, 4]
    assert mats[2]["p"][:4] == [1.0, 1.0, 1.0, 15.0]
    np.testing.assert_allclose(mats[3]["p"][:8], [0.14, 0.43, 1.38, 4.54, 2.455, 1.914, 0.2, 0.2], rtol=1e-7)
    # Disney parameters in constructor order
    np.testing.assert_allclose(mats[4]["p"][:15], [0.8, 0.3, 0.2, 0, 0.5, 0, 0.5, 0.4, 0, 0, 0, 0.5, 0, 0, 1.5],
                               rtol=1e-7)


def test_top_level_rects_and_flags(calls):
    rects = [c for c in calls if c["kind"] == "rect"]
    assert [(r["axis"], r["flags"]) for r in rects[:6]] == [(0, 0), (0, 1), (1, 0), (1, 0), (2, 0), (1, 0)]
    assert rects[5]["v"] == pytest.approx([1.99, 0.75, 1.25, 0.75, 1.25])
    # the emitter inside a nested ObjectList: hit, never sampled as a light
    nested = rects[6]
    assert nested["axis"] == 2 and nested["flags"] == 4 and nested["mat"] == 2


def test_rotated_translated_box(calls):
    meshes = [c for c in calls if c["kind"] == "mesh"]
    box = meshes[:6]
    R = rot_geom(1, math.radians(15.0))
    t = np.array([1.1, 0.0, 0.9])
    lo, hi = np.zeros(3), np.array([0.5, 0.9, 0.5])
    faces = [(0, hi[0]), (0, lo[0]), (1, hi[1]), (1, lo[1]), (2, hi[2]), (2, lo[2])]
    for m, (a, k) in zip(box, faces):
        assert m["flags"] == 2 | 4 and m["mat"] == 1 and m["nrm"] is None
        b0, b1 = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
        corners = []
        for j in range(2):
            for i in range(2):
                p = np.zeros(3)
                p[a] = k
                p[b0] = (lo[b0], hi[b0])[i]
                p[b1] = (lo[b1], hi[b1])[j]
                corners.append(R @ p + t)
        np.testing.assert_allclose(m["pos"], np.array(corners), atol=2e-6)
        assert m["idx"].tolist() == [[0, 1, 3], [0, 3, 2]]
    # rotation preserves the box: every face corner lies at the rotated corner set
    allpos = np.concatenate([m["pos"] for m in box])
    assert np.allclose(np.linalg.norm(allpos - t, axis=1).max(), np.linalg.norm(hi), atol=1e-5)


def test_scaled_and_flipped_spheres(calls):
    sph = [c for c in calls if c["kind"] == "sphere"]
    assert sph[0]["c"] == pytest.approx([0.5, 0.3, 0.6]) and sph[0]["r"] == pytest.approx(0.3)
    assert sph[0]["flags"] == 2            # Translate(Scale(.)) forces is_inside
    assert sph[1]["c"] == pytest.approx([1.5, 1.2, 0.4]) and sph[1]["r"] == pytest.approx(0.15)
    assert sph[1]["flags"] == 2 | 1        # FlipFace outside a Translate: is_inside false


def test_obj_mesh_with_normals(calls):
    m = [c for c in calls if c["kind"] == "mesh"][-1]
    assert m["mat"] == 5 and m["flags"] == 2 | 4
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    R = rot_geom(0, math.radians(30.0))
    want = (R @ (0.25 * v).T).T + np.array([0.6, 1.2, 1.2])
    np.testing.assert_allclose(m["pos"], want, atol=2e-6)
    # f -4//-4 -1//-1 -2//-2 resolves to vertices and normals (0, 3, 2)
    assert m["idx"].tolist() == [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]
    assert m["nidx"].tolist() == m["idx"].tolist()
    n = np.array([[-0.577, -0.577, -0.577], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    got = m["nrm"] / np.linalg.norm(m["nrm"], axis=1, keepdims=True)
    exp = (R @ n.T).T
    exp /= np.linalg.norm(exp, axis=1, keepdims=True)
    np.testing.assert_allclose(got, exp, atol=2e-6)


def test_camera_passes_through(calls):
    cam = calls[-1]
    assert cam["kind"] == "camera"
    assert cam["fov"] == pytest.approx(math.radians(40.0), rel=1e-6)
    assert cam["pos"] == pytest.approx([1.0, 1.0, 4.2])


def test_non_triangle_face_is_an_error(driver, tmp_path):
    exe, _, _ = driver
    bad = tmp_path / "quad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    r = subprocess.run([exe, "dump", str(bad), str(tmp_path / "x.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "non-triangle face" in r.stderr


def _png_rows(path):
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, ihdr = 8, b"", None
    while p < len(b):
        n = struct.unpack(">I", b[p:p + 4])[0]
        tag, data = b[p + 4:p + 8], b[p + 8:p + 8 + n]
        assert zlib.crc32(tag + data) & 0xFFFFFFFF == struct.unpack(">I", b[p + 8 + n:p + 12 + n])[0]
        if tag == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", data)
        if tag == b"IDAT":
            idat += data
        p += 12 + n
    w, h = ihdr[0], ihdr[1]
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_png_writer(driver, tmp_path):
    exe, _, _ = driver
    out = str(tmp_path / "f.png")
    subprocess.run([exe, "png", out], check=True)
    img = _png_rows(out)
    film = (np.arange(36, dtype=np.float32) / np.float32(35)).reshape(3, 4, 3)
    film.reshape(-1)[5] = -1.0
    film.reshape(-1)[7] = np.nan
    film.reshape(-1)[8] = 7.0
    f = np.nan_to_num(film, nan=0.0)
    with np.errstate(invalid="ignore"):
        g = np.power(f, np.float32(0.5))
    g = np.nan_to_num(g, nan=0.0)  # pow of a negative is NaN, which clamps to 0 like the reference's max
    want = (np.clip(g, 0, 1) * 255).astype(np.uint8)[::-1]  # vertical flip
    assert (img == want).all()


def _write_hdr_rle(path, rgbe):
    """A Radiance file with new-style run-length scanlines (runs and literals)."""
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            ch = rgbe[y, :, c]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and ch[x + run] == ch[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, ch[x]])
                    x += run
                else:
                    n = min(128, w - x)
                    lit = [ch[x]]
                    k = x + 1
                    while k < w and len(lit) < n and not (k + 2 < w and ch[k] == ch[k + 1] == ch[k + 2]):
                        lit.append(ch[k])
                        k += 1
                    out += bytes([len(lit)] + lit)
                    x += len(lit)
    with open(path, "wb") as f:
        f.write(out)


def test_hdr_reader(driver, tmp_path):
    exe, _, _ = driver
    rng = np.random.default_rng(5)
    h, w = 5, 40
    rgbe = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    rgbe[:, 10:20, :] = rgbe[:, 10:11, :]   # runs
    rgbe[1, :5, 3] = 0                      # zero exponent -> black
    src = str(tmp_path / "env.hdr")
    _write_hdr_rle(src, rgbe)
    out = str(tmp_path / "env.pfm")
    subprocess.run([exe, "hdr", src, out], check=True)
    from pathtracer_amd.film_io import read_pfm
    got = read_pfm(out)
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136)).astype(np.float32)
    want = (rgbe[..., :3].astype(np.float32) * f[..., None])[::-1]  # row 0 = bottom of the image
    assert got.shape == want.shape and (got == want).all()


# ---- textures (texture/texture.hpp:20-94) ----

def _write_ppm(path, rgb8):
    """P6 with the file's first row the image's top row."""
    h, w, _ = rgb8.shape
    with open(path, "wb") as f:
        f.write(b"P6\n# test image\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(rgb8, np.uint8).tobytes())


def _write_pfm(path, rgb):
    """Little-endian PFM, the file's first row the image's bottom row."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(rgb, "<f4").tobytes())


@pytest.fixture(scope="module")
def tex_files(driver):
    exe, _, d = driver
    rng = np.random.default_rng(11)
    ppm8 = rng.integers(0, 256, size=(3, 4, 3), dtype=np.uint8)      # top row first
    pfm = rng.uniform(0.0, 3.0, size=(2, 3, 3)).astype(np.float32)   # bottom row first
    a, b = str(d / "a.ppm"), str(d / "b.pfm")
    _write_ppm(a, ppm8)
    _write_pfm(b, pfm)
    out = str(d / "tdump.txt")
    subprocess.run([exe, "tdump", a, b, out], check=True)
    return a, b, ppm8, pfm, parse_dump(out)


def _texread(exe, tmp_path, src, flip, hdr):
    from pathtracer_amd.film_io import read_pfm
    out = str(tmp_path / "t.pfm")
    subprocess.run([exe, "texread", src, str(int(flip)), str(int(hdr)), out], check=True)
    return read_pfm(out)


def test_image_texture_reading(driver, tex_files, tmp_path):
    """ImageTexture(file, flip_y, isHDR): row 0 the image's bottom row unless
    flip_y; an 8-bit file read as HDR is (v/255)^2.2 and a float file read
    as LDR is quantised to (uint8)(x^(1/2.2) 255 + 0.5) / 255, clamped (stb_image's
    conversions at their default gamma and scale)."""
    exe = driver[0]
    a, b, ppm8, pfm, _ = tex_files
    ldr = ppm8.astype(np.float32) / np.float32(255.0)
    assert np.array_equal(_texread(exe, tmp_path, a, False, False), ldr[::-1])
    assert np.array_equal(_texread(exe, tmp_path, a, True, False), ldr)
    np.testing.assert_allclose(_texread(exe, tmp_path, a, False, True), (ldr.astype(np.float64) ** 2.2)[::-1],
                               rtol=1e-6)
    assert np.array_equal(_texread(exe, tmp_path, b, False, True), pfm)
    assert np.array_equal(_texread(exe, tmp_path, b, True, True), pfm[::-1])
    q = np.clip(np.floor(pfm.astype(np.float64) ** (1 / 2.2) * 255 + 0.5), 0, 255) / 255
    np.testing.assert_allclose(_texread(exe, tmp_path, b, False, False), q, atol=1e-6)
    assert q.max() == 1.0                        # values above 1 clamp to 255
    r = subprocess.run([exe, "texread", str(tmp_path / "x.png"), "0", "0", str(tmp_path / "o.pfm")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and ".hdr, .pfm or .ppm" in r.stderr


def test_texture_lowering(tex_files):
    """Textures are lowered once each before the first material that uses
    them; a textured material carries PT_MAT_TEXTURE(id) and a zero colour;
    a ConstantTexture lowers to the plain colour."""
    _, _, ppm8, pfm, calls = tex_files
    tex = [c for c in calls if c["kind"].startswith("texture")]
    mats = [c for c in calls if c["kind"] == "material"]
    assert [t["kind"] for t in tex] == ["texture_checker", "texture_image", "texture_image"]
    assert tex[0]["scale"] == 10.0
    np.testing.assert_allclose(tex[0]["c1"], [0.2, 0.3, 0.1], rtol=1e-7)
    np.testing.assert_allclose(tex[0]["c2"], [0.9, 0.9, 0.9], rtol=1e-7)
    assert np.array_equal(tex[1]["rgb"], (ppm8.astype(np.float32) / np.float32(255.0))[::-1])
    assert np.array_equal(tex[2]["rgb"], pfm[::-1])          # flip_y = true, isHDR = true
    # checker Lambertian, image Lambertian, emitter, constant-texture Lambertian, textured Disney
    assert [(m["mat_kind"], m["flags"]) for m in mats] == [(1, 1), (1, 1 | 1 << 8), (2, 0), (1, 0), (7, 1 | 2 << 8)]
    assert mats[0]["p"][:3] == [0.0, 0.0, 0.0]
    np.testing.assert_allclose(mats[3]["p"][:3], [0.5, 0.6, 0.7], rtol=1e-7)
    np.testing.assert_allclose(mats[4]["p"][3:15], [0, 0.3, 0, 0.5, 0.4, 0, 0, 0, 0.5, 0, 0, 1.5], rtol=1e-7)


def test_mesh_uv_lowering(tex_files):
    """MeshData uvs pass through; a rotated textured rect (two triangles)
    takes the rect's corner uvs; a rotated mesh without uvs takes its planar
    fallback in object space (triangleMesh.cpp:10-18, 38-51)."""
    calls = tex_files[4]
    seq = [c for c in calls if c["kind"] in ("mesh", "mesh_uv")]
    kinds = [c["kind"] for c in seq]
    assert kinds == ["mesh", "mesh_uv", "mesh", "mesh", "mesh_uv", "mesh", "mesh_uv"]
    quad_uv = seq[1]
    assert quad_uv["uv"].tolist() == [[0, 0], [1, 0], [1, 1], [0, 1]]
    assert quad_uv["uvidx"].tolist() == [[0, 1, 2], [0, 2, 3]]
    # rotated tetra: object-space positions, box padded by 1e-3
    P = np.array([[0, 0, 0], [0.4, 0, 0.1], [0.2, 0, 0.5], [0.2, 0.45, 0.2]], np.float64)
    F = [[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]]
    lo, hi = P.min(0) - 1e-3, P.max(0) + 1e-3
    uv = seq[4]["uv"]
    assert seq[4]["uvidx"].tolist() == [[0, 1, 2], [3, 4, 5], [6, 7, 8], [9, 10, 11]]
    for k, f in enumerate(F):
        n = np.abs(np.cross(P[f[1]] - P[f[0]], P[f[2]] - P[f[0]]))
        au = 1 if (n[0] > n[1] and n[0] > n[2]) else 0
        av = 2 if au == 1 else (2 if n[1] > n[2] else 1)
        for c in range(3):
            q = P[f[c]]
            np.testing.assert_allclose(uv[3 * k + c], [(q[au] - lo[au]) / (hi[au] - lo[au]),
                                                       (q[av] - lo[av]) / (hi[av] - lo[av])], atol=1e-6)
    # the rotated rect's corners (d0, d1) = (0,0), (1,0), (0,1), (1,1) of its extents
    assert seq[6]["uv"].tolist() == [[0, 0], [1, 0], [0, 1], [1, 1]]
    assert seq[6]["uvidx"].tolist() == seq[5]["idx"].tolist()


def test_unsupported_textures_are_errors(driver, tmp_path):
    exe, obj, _ = driver
    r = subprocess.run([exe, "noise", obj, str(tmp_path / "x.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "NoiseTexture has no GPU lowering" in r.stderr
    r = subprocess.run([exe, "temit", "x"], capture_output=True, text=True)
    assert r.returncode == 1 and "constant colours only" in r.stderr
