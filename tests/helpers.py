"""Shared helpers of the test suite."""
import numpy as np


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


def random_dirs(rng, n):
    return unit(rng.normal(size=(n, 3)))


def hemi_dirs(rng, n, nrm):
    d = random_dirs(rng, n)
    s = np.sign(np.sum(d * nrm, axis=1, keepdims=True))
    s[s == 0] = 1
    return (d * s).astype(np.float32)


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)


# ---- the C++ scene-API driver (tests/cpp/scene_api_check.cpp) ----

TETRA_OBJ = """# tetrahedron with corner normals, mixed index forms
v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
vn -0.577 -0.577 -0.577
vn 1 0 0
vn 0 1 0
vn 0 0 1
vt 0 0
f 1//1 3//3 2//2
f 1/1/1 2/1/2 4/1/4
f -4//-4 -1//-1 -2//-2
f 2//2 3//3 4//4
"""


def parse_dump(path):
    """The ABI calls the driver's scene lowered to, as dicts in call order."""
    out = []
    with open(path) as f:
        for line in f:
            t = line.split()
            kind = t[0]
            if kind == "material":
                out.append(dict(kind="material", mat_kind=int(t[1]), p=[float(x) for x in t[2:]]))
            elif kind == "rect":
                out.append(dict(kind="rect", mat=int(t[1]), flags=int(t[2]), axis=int(t[3]),
                                v=[float(x) for x in t[4:9]]))
            elif kind == "sphere":
                out.append(dict(kind="sphere", mat=int(t[1]), flags=int(t[2]), c=[float(x) for x in t[3:6]],
                                r=float(t[6])))
            elif kind == "mesh":
                mat, flags, nv, nt, nn = (int(x) for x in t[1:6])
                vals = t[6:]
                pos = np.array(vals[:3 * nv], np.float32).reshape(nv, 3)
                vals = vals[3 * nv:]
                idx = np.array(vals[:3 * nt], np.int32).reshape(nt, 3)
                vals = vals[3 * nt:]
                nrm = nidx = None
                if nn:
                    nrm = np.array(vals[:3 * nn], np.float32).reshape(nn, 3)
                    nidx = np.array(vals[3 * nn:3 * nn + 3 * nt], np.int32).reshape(nt, 3)
                out.append(dict(kind="mesh", mat=mat, flags=flags, pos=pos, idx=idx, nrm=nrm, nidx=nidx))
            elif kind == "camera":
                v = [float(x) for x in t[1:]]
                out.append(dict(kind="camera", fov=v[0], aspect=v[1], focal=v[2], aperture=v[3], up=v[4:7],
                                front=v[7:10], pos=v[10:13]))
    return out


def dump_to_spec(calls, width=64, height=64, spp_x=2, spp_y=2, max_depth=6):
    """A SceneSpec holding exactly the lowered calls (same order)."""
    import math
    from pathtracer_amd import scenes
    s = scenes.SceneSpec(name="scene_api")
    for c in calls:
        if c["kind"] == "material":
            s.add_material(c["mat_kind"], c["p"])
        elif c["kind"] == "rect":
            s.add_rect(c["axis"], *c["v"], c["mat"], c["flags"])
        elif c["kind"] == "sphere":
            s.add_sphere(c["c"], c["r"], c["mat"], c["flags"])
        elif c["kind"] == "mesh":
            s.add_mesh(c["pos"], c["idx"], c["mat"], c["flags"], normals=c["nrm"], nidx=c["nidx"])
        elif c["kind"] == "camera":
            s.camera = dict(fov_deg=math.degrees(c["fov"]), fov_rad=c["fov"], aspect=c["aspect"], focal_dist=c["focal"],
                            aperture=c["aperture"], up=c["up"], front=c["front"], pos=c["pos"])
    s.render = dict(width=width, height=height, spp_x=spp_x, spp_y=spp_y, max_depth=max_depth, sampler=0,
                    estimator=0, rr_depth=0, light_samples=4, seed=0, bg=(0, 0, 0))
    return s
