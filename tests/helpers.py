"""Shared helpers of the test suite."""
import numpy as np

# ---- conditioning of the oracle: what a last-place libm difference can move ----
EPS = 1e-6       # input perturbation that stands for a last-place difference in a libm call
N_PERTURB = 32


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
                out.append(dict(kind="material", mat_kind=int(t[1]), flags=int(t[2]), p=[float(x) for x in t[3:]]))
            elif kind == "texture_image":
                w, h = int(t[1]), int(t[2])
                out.append(dict(kind="texture_image", rgb=np.array(t[3:], np.float32).reshape(h, w, 3)))
            elif kind == "texture_checker":
                v = [float(x) for x in t[1:]]
                out.append(dict(kind="texture_checker", scale=v[0], c1=v[1:4], c2=v[4:7]))
            elif kind == "mesh_uv":
                nuv, nt = int(t[1]), int(t[2])
                uv = np.array(t[3:3 + 2 * nuv], np.float32).reshape(nuv, 2)
                uvidx = np.array(t[3 + 2 * nuv:3 + 2 * nuv + 3 * nt], np.int32).reshape(nt, 3)
                out.append(dict(kind="mesh_uv", uv=uv, uvidx=uvidx))
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
            m = s.add_material(c["mat_kind"], c["p"])
            if c["flags"]:
                s.material_flags[m] = c["flags"]
        elif c["kind"] == "texture_image":
            s.add_texture_image(c["rgb"])
        elif c["kind"] == "texture_checker":
            s.add_texture_checker(c["c1"], c["c2"], c["scale"])
        elif c["kind"] == "mesh_uv":
            s.set_mesh_uv(len(s.meshes) - 1, c["uv"], c["uvidx"])
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


# ---- per-sample parity of whole paths against the oracle ----
#
# GPU and oracle draw the same sample streams, so a path sample's radiance
# agrees to rounding unless a branch decision (a lobe choice, Russian
# roulette, a hit at a silhouette) falls within a last-place libm difference
# of its boundary: that path then continues elsewhere and its radiance may
# differ by any amount.  Such samples must be rare (MIN_AGREE), and the
# radiance they carry between them small against the film's (MAX_ENERGY).
SAMPLE_RTOL, SAMPLE_ATOL = 1e-3, 1e-4
MIN_AGREE = 0.999
MAX_ENERGY = 5e-4


def sample_stats(gs, os_):
    """(fraction of samples within the bar, radiance of the others relative to
    the film's total, relative difference of the totals)."""
    g = np.asarray(gs, np.float64)
    o = np.asarray(os_, np.float64)
    close = np.all(np.abs(g - o) <= SAMPLE_RTOL * np.maximum(np.abs(g), np.abs(o)) + SAMPLE_ATOL, axis=-1)
    total = max(np.abs(o).sum(), 1e-30)
    energy = np.abs(g - o)[~close].sum() / total
    return float(close.mean()), float(energy), float(abs(g.sum() - o.sum()) / total)


def check_samples(gs, os_, label="", self_moved=None):
    """GPU samples gs against the oracle's os_.  self_moved: the oracle's own
    samples of the same scene with the camera moved by 1e-6, for scenes whose
    paths amplify a last-place difference far beyond the default bars (thin
    fibres with a sharp specular cone); the bars then become a quarter of the
    disagreement the oracle shows with itself under that move, where that is
    looser."""
    frac, energy, dmean = sample_stats(gs, os_)
    n = int(np.prod(np.shape(gs)[:-1]))
    min_agree, max_energy = MIN_AGREE, MAX_ENERGY
    if self_moved is not None:
        sf, se, _ = sample_stats(self_moved, os_)
        min_agree = min(min_agree, 1.0 - 0.25 * (1.0 - sf))
        max_energy = max(max_energy, 0.25 * se)
        print(f"{label}: oracle against itself, camera moved 1e-6: {sf:.4f} within 1e-3, energy {se:.2e}")
    print(f"{label}: {n} samples, {frac:.6f} within 1e-3, disagreeing energy {energy:.2e}, "
          f"mean {dmean:.2e}")
    assert frac >= min_agree, f"{label}: only {frac:.5f} of path samples agree"
    assert energy <= max_energy, f"{label}: disagreeing samples carry {energy:.2e} of the radiance"
    return frac, energy


def perturb(rng, arrays, kinds):
    """The inputs moved by EPS: unit vectors turned (and renormalised), uniforms
    shifted inside [0, 1), flags kept."""
    out = []
    for a, k in zip(arrays, kinds):
        if k == "dir":
            d = a.astype(np.float64) + EPS * rng.uniform(-1, 1, a.shape)
            out.append((d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
        elif k == "u":
            out.append(np.clip(a.astype(np.float64) + EPS * rng.uniform(-1, 1, a.shape), 0.0,
                               1.0 - 2.0 ** -24).astype(np.float32))
        else:
            out.append(a)
    return out


def explained_by_conditioning(fn, arrays, kinds, idx, gpu_out, rtols, atols, rng):
    """Tuples idx where the GPU misses the oracle.  Each is accepted only if the
    oracle itself is ill-conditioned there: the GPU's outputs lie within the
    bar plus the spread of the oracle's outputs over N_PERTURB runs with every
    input moved by EPS (the GPU's libm differs from glibc's in the last place,
    which such a move stands for).  Returns a boolean per tuple."""
    sub = [a[idx] for a in arrays]
    base = fn(*sub)
    spread = [np.zeros(np.shape(b), np.float64) for b in base]
    for _ in range(N_PERTURB):
        out = fn(*perturb(rng, sub, kinds))
        for s, o, b in zip(spread, out, base):
            np.maximum(s, np.abs(np.asarray(o, np.float64) - np.asarray(b, np.float64)), out=s)
    ok = np.ones(len(idx), bool)
    for g, b, s, rt, at in zip(gpu_out, base, spread, rtols, atols):
        g = np.asarray(g, np.float64)[idx]
        b = np.asarray(b, np.float64)
        lim = rt * np.maximum(np.abs(g), np.abs(b)) + at + 2.0 * s
        dev = np.abs(g - b) <= lim
        ok &= dev.reshape(len(idx), -1).all(axis=1)
    return ok
