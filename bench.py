"""Benchmark of the per-sample path-tracing hot path on one MI355X.

Workload (one step = one frame): the triangle-heavy single frame -- a
noise-displaced icosphere of 1,310,720 triangles in rough glass (IOR 1.5,
alpha 0.05) on a 2-triangle copper plane (alpha 0.3), lit by a 1024x512
synthetic sky+sun environment sampled through its 2D piecewise-constant CDF;
1024x1024 pixels, 64 samples per pixel (67,108,864 path samples), Sobol
sampler, max depth 16, thin-lens depth of field with lens diameter 0.02,
NEE+MIS with Russian roulette from vertex 3.  The reference's own estimator
(4 uniform light samples per non-specular vertex, no MIS) is timed on the
same frame as a second line item.

Prints one JSON line.  `value` is path samples per second over K timed frames
(scene resident in HBM).  `roofline` is for the kernel with the largest share
of device time, from algorithmic bytes (HIP-event-timed launches, counters of
an identical counting frame).  `quality` is the relMSE of a 64-spp render
against a 4096-spp render of the same camera at 128x128.  `cpu_baseline`
times the CPU oracle (the C restatement of the reference's path, OpenMP) on
the host cores over a bounded slice of the same frame: the reference's own
estimator as `value`, NEE+MIS beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpath-samples/s on 1 MI355X; relMSE vs 4096-spp reference at equal time"
WORKLOADS = {
    "glass_hdri": "glass_hdri: 1.31M-tri displaced rough-glass icosphere + copper plane + 1024x512 sky/sun HDRI, "
                  "thin-lens DoF",
    "cornell": "cornell: 20-tri Cornell box, 1280-tri gold icosphere, Lambertian sphere, 0.5x0.5 quad light; "
               "256x256, 16 spp Halton, depth 8",
    "disney_hair": "disney_hair: 8x8 Disney spheres with seeded parameters + 8 Marschner hair spheres, 4x4 quad "
                   "light; 1024x1024, 128 spp stratified, depth 12",
}
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit of work, the hot path's cost model: per extension
# segment a ray record (32) read and a hit record (16) written by traversal,
# two path-state records (40 each, read and written) and a material record
# (32) in shading; per shadow ray a 48-byte record written by shading and read
# by any-hit traversal, plus 4 bytes of visibility; 32 bytes per BVH node visit
# and 36 per triangle test (three float3); 12 bytes of film per path sample.
B_RAY, B_HIT, B_STATE, B_MAT = 32, 16, 40, 32
B_SHADOW, B_VIS = 48, 4
B_NODE, B_TRI, B_FILM = 32, 36, 12
# What this build's layout fetches per unit (node and primitive bytes mostly
# served by LDS, L2 and the Infinity Cache): reported as the cache-inclusive
# rate, never as the roofline's `achieved`.
L_EXT_RAY = 24 + 8            # ray (o, d) read + hit (t, prim) written
L_SHADOW_RAY = 44             # shadow-queue entry (o, d, tmax, contribution, slot)
L_NODE = 112                  # BVH4 node: four child boxes (SoA) + four child codes
L_QNODE = 64                  # quantised BVH4 node
L_PRIM = 48                   # triangle record (v0, e1, e2)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_glass_hdri.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


OR_REFERENCE_TRAVERSALS = 1 << 24   # the oracle repeats the reference's second continuation traversal


def _time_oracle_leg(o, cfg, start, total, threads, target_s):
    """Path samples per second of the oracle over consecutive samples from
    `start`, grown until the timed slice takes about target_s."""
    n, t = 4096, 0.0
    while True:
        t1 = time.perf_counter()
        o.render_range(cfg, start, n, threads)
        t = time.perf_counter() - t1
        if t > 1.0 or n >= 1 << 24:
            break
        n *= 4
    n = min(max(n, int(n * target_s / max(t, 1e-3))), total - start)
    t1 = time.perf_counter()
    o.render_range(cfg, start, n, threads)
    return n, time.perf_counter() - t1


def cpu_baseline(spec, cfg, target_s=8.0):
    """The CPU oracle (the C restatement of the reference's path, OpenMP) on
    the host cores, over a bounded slice from the middle of the same frame.
    `value` is the reference's own estimator (maxdepth_nee.hpp:9-256: 4 light
    samples per non-specular vertex, each continuation ray traversed twice as
    the reference does); `nee_mis` times the estimator the GPU runs."""
    from oracle import oracle_py as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    o = O.OracleScene(spec)
    build_s = time.perf_counter() - t0
    spp = cfg.spp_x * cfg.spp_y
    total = cfg.width * cfg.height * spp
    start = (total // 2) // spp * spp
    ref_cfg = o.cfg(estimator=0, flags=OR_REFERENCE_TRAVERSALS)
    n, t = _time_oracle_leg(o, ref_cfg, start, total, threads, target_s)
    # one thread on a sample of about 4 s
    n1 = max(256, min(n, int(n / max(t, 1e-3) / max(threads, 1) * 4.0)))
    t1 = time.perf_counter()
    o.render_range(ref_cfg, start, n1, 1)
    t_one = time.perf_counter() - t1
    m, tm = _time_oracle_leg(o, o.cfg(estimator=1), start, total, threads, target_s)
    return {"value": n / t / 1e6, "unit": "Mpath-samples/s", "cores": threads, "kind": "port",
            "estimator": "reference_nee",
            "sample": f"{n} consecutive path samples (pixel-major) from the middle of the same frame under the "
                      f"reference's estimator, {t:.1f} s on {threads} OpenMP threads; oracle BVH/scene build "
                      f"{build_s:.1f} s untimed",
            "single_thread": {"value": n1 / t_one / 1e6, "cores": 1,
                              "sample": f"the first {n1} of those path samples, {t_one:.1f} s"},
            "nee_mis": {"value": m / tm / 1e6, "cores": threads,
                        "sample": f"{m} path samples from the same place under NEE+MIS, {tm:.1f} s"}}


def read_traffic(kernel, estimator):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary committed beside the profiles (None when absent, or measured for
    another estimator)."""
    try:
        kv = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        if int(kv.get("estimator", -1)) != estimator:
            return None
        return int(float(kv[f"{kernel}.hbm_bytes_per_launch"]))
    except (OSError, KeyError, ValueError):
        return None


def quality(scene, estimator, res=128, spp_side=8, ref_side=64):
    """relMSE of a 64-spp frame against a 4096-spp frame (another seed) of the
    same camera at res x res: mean over pixels and channels of
    (x - r)^2 / (r^2 + 1e-2); relmse_trim leaves out the 0.1% of pixels with
    the largest error."""
    import numpy as np
    t0 = time.perf_counter()
    ref, _ = scene.render(scene.cfg(estimator=estimator, width=res, height=res, spp_x=ref_side, spp_y=ref_side,
                                    seed=101))
    t_ref = time.perf_counter() - t0
    t0 = time.perf_counter()
    img, _ = scene.render(scene.cfg(estimator=estimator, width=res, height=res, spp_x=spp_side, spp_y=spp_side,
                                    seed=7))
    t_img = time.perf_counter() - t0
    e = ((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 1e-2)).mean(-1).ravel()
    k = int(0.999 * e.size)
    return {"relmse": float(e.mean()), "relmse_trim": float(np.partition(e, k)[:k].mean()),
            "spp": spp_side * spp_side, "ref_spp": ref_side * ref_side,
            "resolution": f"{res}x{res}", "ms_render": round(t_img * 1e3, 1), "ms_ref": round(t_ref * 1e3, 1)}


def any_hit_qnodes(scs):
    """Whether any-hit traversal reads the 64-byte quantised nodes: the
    library picks them for a tree whose 128-byte nodes exceed 4 MiB (a batch:
    when every one of its scenes does)."""
    return all(sc.info.bvh_nodes * 128 > (4 << 20) for sc in scs)


def kernel_roofline(cst, ms, launches, read_traffic_for=None, qnodes=True, paths=0):
    """Per-kernel algorithmic bytes (the cost model above, times the frame's
    counters) over HIP-event time; the roofline object is for the kernel with
    the most device time.  `cache_inclusive_gbs` is the same kernel's rate in
    the bytes its layout fetches (most of them from caches)."""
    seg, shr = cst["segments"], cst["shadow_rays"]
    kernels = {
        "k_extend": ("closest-hit traversal",
                     seg * (B_RAY + B_HIT) + cst["node_visits"] * B_NODE + cst["prim_tests"] * B_TRI,
                     seg * L_EXT_RAY + cst["node_visits"] * L_NODE + cst["prim_tests"] * L_PRIM, ms["ms_extend"]),
        "k_shadow": ("any-hit traversal",
                     shr * (B_SHADOW + B_VIS) + cst["shadow_node_visits"] * B_NODE + cst["shadow_prim_tests"] * B_TRI,
                     shr * L_SHADOW_RAY + cst["shadow_node_visits"] * (L_QNODE if qnodes else L_NODE) +
                     cst["shadow_prim_tests"] * L_PRIM, ms["ms_shadow"]),
        "k_shade": ("materials, lights, NEE, queue compaction", seg * (2 * B_STATE + B_MAT) + shr * B_SHADOW,
                    seg * (68 + 24 + 32 + 68) + shr * L_SHADOW_RAY, ms["ms_shade"]),
        "k_film": ("per-pixel sums", paths * B_FILM, paths * 12, ms["ms_film"]),
    }
    gbs = lambda b, t: round(b / (t * 1e-3) / 1e9, 1) if t > 0 else 0.0
    per_kernel = {k: {"what": w, "gb_per_frame": round(b / 1e9, 3), "ms_per_frame": round(t, 2),
                      "achieved_gbs": gbs(b, t), "frac": round(gbs(b, t) / HBM_PEAK_GBS, 4),
                      "cache_inclusive_gbs": gbs(lb, t)}
                  for k, (w, b, lb, t) in kernels.items()}
    kname = max(kernels, key=lambda k: kernels[k][3])
    _, kbytes, lbytes, kms = kernels[kname]
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": (read_traffic(kname, read_traffic_for)
                                                             if read_traffic_for is not None else None),
                "kernel": kname, "launches_per_frame": launches, "bytes_per_launch": int(kbytes / max(launches, 1)),
                "ms_per_launch": round(kms / max(launches, 1), 4),
                "cache_inclusive_gbs": round(lbytes / (kms * 1e-3) / 1e9, 1) if kms > 0 else 0.0}
    path_bytes = sum(v[1] for v in kernels.values())
    path_ms = sum(v[3] for v in kernels.values())
    roofline["path"] = {"bytes_per_frame": int(path_bytes), "bytes_per_path": round(path_bytes / max(paths, 1), 1),
                        "kernel_ms": round(path_ms, 2), "achieved": gbs(path_bytes, path_ms),
                        "frac": round(gbs(path_bytes, path_ms) / HBM_PEAK_GBS, 4)}
    return roofline, per_kernel


def build_flags(args):
    from pathtracer_amd import _lib as L
    return L.BUILD_DEVICE_LBVH if args.build == "lbvh" else 0


def bvh_info(args, scs, seconds):
    """The BVH build of the scenes: kind, build time (host wall clock of the
    build alone, device kernels waited for), whole commit, nodes, depth."""
    return {"build": {"sah": "host_sah", "lbvh": "device_lbvh"}[args.build],
            "bvh_ms": round(sum(sc.info.bvh_ms for sc in scs), 2),
            "commit_ms": round(sum(sc.info.build_ms for sc in scs), 2),
            "scene_s": round(seconds, 2),
            "nodes": sum(sc.info.bvh_nodes for sc in scs),
            "max_depth": max(sc.info.bvh_depth for sc in scs)}


def main_batch(args):
    """64 independent renders (seed 4): displaced geodesic spheres of 1e4-5e5
    triangles (log-uniform), 1-4 quad lights, a material from {Lambertian,
    conductor, rough dielectric, Disney}, 128^2-1024^2 pixels, 4-256 spp
    (log-uniform), depth 8-16.  One step renders all 64 in one wavefront
    (captured chunk graph replayed per chunk); the same 64 renders launched
    one after another with pt_render are timed beside it."""
    from pathtracer_amd import _lib as L, render as R, scenes
    t0 = time.perf_counter()
    specs = scenes.batch64(seed=4, n=8 if args.small else 64)
    if args.small:
        for sp in specs:
            sp.render.update(width=64, height=64)
    log(f"{len(specs)} scenes generated in {time.perf_counter() - t0:.1f} s, "
        f"{sum(sp.n_triangles for sp in specs)} triangles")
    t0 = time.perf_counter()
    scs = [R.Scene(sp, build_flags=build_flags(args)) for sp in specs]
    build = bvh_info(args, scs, time.perf_counter() - t0)
    log(f"BVHs built and uploaded: {build}")
    cfgs = [sc.cfg(estimator=args.estimator) for sc in scs]
    total = sum(c.width * c.height * c.spp_x * c.spp_y for c in cfgs)
    b = R.Batch()
    for sc, c in zip(scs, cfgs):
        b.add(sc, c)
    tuning = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tuning = b.autotune(spp_div=8)
        log(f"autotuned in {time.perf_counter() - t0:.1f} s: {tuning}")
    # counting + per-stage timing pass (eager launches, untimed)
    cst = b.render(flags=L.RENDER_COUNTERS | L.RENDER_TIMING | L.BATCH_NO_GRAPH)
    for i in range(args.warmup):
        b.render()
        log(f"warmup {i} done")
    L.check(L.lib.pt_device_synchronize(), "sync")
    t1 = time.perf_counter()
    for i in range(args.steps):
        st = b.render()
        log(f"step {i}: {st['ms_total']:.1f} ms")
    L.check(L.lib.pt_device_synchronize(), "sync")
    dt = time.perf_counter() - t1
    value = total * args.steps / dt / 1e6
    # the same renders one after another (plain pt_render, default launch parameters)
    for sc, c in zip(scs, cfgs):
        sc.render_device(c)
    L.check(L.lib.pt_device_synchronize(), "sync")
    t2 = time.perf_counter()
    for sc, c in zip(scs, cfgs):
        sc.render_device(c)
    L.check(L.lib.pt_device_synchronize(), "sync")
    sdt = time.perf_counter() - t2
    serial = {"value": round(total / sdt / 1e6, 2), "ms_per_step": round(sdt * 1e3, 2),
              "what": "the same renders launched one after another with pt_render"}
    log(f"serial: {serial}")
    ms = {k: cst[k] for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_generate", "ms_film")}
    roofline, per_kernel = kernel_roofline(cst, ms, cst["extend_launches"], qnodes=any_hit_qnodes(scs),
                                           paths=cst["paths"])
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpath-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch64: 64 seeded renders (displaced geodesic spheres 1e4-5e5 tris, 1-4 quad "
                               "lights, 128^2-1024^2, 4-256 spp, depth 8-16) in one wavefront",
                   "renders": len(specs), "path_samples_per_step": total,
                   "triangles": sum(sp.n_triangles for sp in specs),
                   "estimator": ["reference_nee", "nee_mis"][args.estimator], "tuning": tuning},
        "serial": serial, "speedup_vs_serial": round(sdt / (dt / args.steps), 3),
        "mrays_per_s": round((cst["segments"] + cst["shadow_rays"]) * args.steps / dt / 1e6, 1),
        "kernel_ms_per_step_eager": {k: round(v, 2) for k, v in ms.items()},
        "roofline": roofline, "kernels": per_kernel, "bvh": build, "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(specs[0], cfgs[0])
    print(json.dumps(res), flush=True)


def relmse(img, ref, keep=1.0):
    """relMSE of the films scaled to the reference's unit mean (the room is
    dim); keep < 1 averages only the pixels below that quantile of the error
    (keep = 0.999: the usual 0.1% outlier rejection)."""
    import numpy as np
    mu = float(np.mean(ref))
    x, r = img.astype(np.float64) / mu, ref.astype(np.float64) / mu
    e = ((x - r) ** 2 / (r ** 2 + 1e-2)).mean(-1).ravel()
    if keep < 1.0:
        k = int(keep * e.size)
        e = np.partition(e, k)[:k]
    return float(e.mean())


def time_to_reach(curve, target):
    """Time at which a log-log interpolated (ms, error) curve reaches target
    (extrapolated at slope -1 beyond its end)."""
    import numpy as np
    gm = [(np.log(e[0]), np.log(e[1])) for e in curve]
    for (ta, ea), (tb, eb) in zip(gm, gm[1:]):
        if (ea - np.log(target)) * (eb - np.log(target)) <= 0 and ea != eb:
            return float(np.exp(ta + (np.log(target) - ea) * (tb - ta) / (eb - ea)))
    if gm[-1][1] < np.log(target):
        return float(np.exp(gm[-1][0] + (np.log(target) - gm[-1][1]) * -1.0))
    return None


def guided_quality(scenes, R, L, q, scene=None, seeds=2):
    """relMSE against a plain reference at q x q (8192 spp: four 2048-spp
    films, the frame size a render takes) for the plain and the guided method
    at 16, 64, 256 spp, in full and with the 0.1% largest pixel errors
    rejected, and the guided time to reach the plain method's 256-spp error
    (log-log interpolation) by each.  Measured at the workload's own
    resolution: the guide's training costs a fixed number of Adam steps per
    frame, which a smaller film would over-weight.  `seeds` films of each
    method per spp are averaged."""
    import numpy as np
    qs = scene or R.Scene(scenes.guiding_room(width=q, height=q))
    ref = np.zeros((q, q, 3))
    nref = 4
    for k in range(nref):
        img, _ = qs.render(qs.cfg(spp_x=64, spp_y=32, seed=900 + k))
        ref += img / nref
    curve = {"plain": [], "guided": []}
    for side in (4, 8, 16):
        pe, pt, ge, gt, pe_t, ge_t = [], [], [], [], [], []
        for k in range(seeds):
            c = qs.cfg(spp_x=side, spp_y=side, seed=11 + k)
            L.check(L.lib.pt_device_synchronize(), "sync")
            t = time.perf_counter()
            img, _ = qs.render(c)
            pt.append(time.perf_counter() - t)
            pe.append(relmse(img, ref))
            pe_t.append(relmse(img, ref, 0.999))
            g = R.Guide(qs, seed=5 + k)
            L.check(L.lib.pt_device_synchronize(), "sync")
            t = time.perf_counter()
            img, _, _ = g.render(c, L.GuideCfg())
            gt.append(time.perf_counter() - t)
            ge.append(relmse(img, ref))
            ge_t.append(relmse(img, ref, 0.999))
            g.close()
        curve["plain"].append({"spp": side * side, "ms": round(float(np.mean(pt)) * 1e3, 2),
                               "relmse": float(np.mean(pe)), "relmse_trim": float(np.mean(pe_t))})
        curve["guided"].append({"spp": side * side, "ms": round(float(np.mean(gt)) * 1e3, 2),
                                "relmse": float(np.mean(ge)), "relmse_trim": float(np.mean(ge_t))})
    log(f"quality: {curve}")
    out = {}
    for key in ("relmse", "relmse_trim"):
        target = curve["plain"][-1][key]
        t_need = time_to_reach([(e["ms"], e[key]) for e in curve["guided"]], target)
        out[key] = {"plain_256spp": target, "guided_ms_to_plain": t_need,
                    "time_ratio_guided_vs_plain": (t_need / curve["plain"][-1]["ms"]) if t_need else None}
    return curve, out


def main_churn(args):
    """Geometry that moves every frame: the glass_hdri frame (1.31M triangles,
    1024x1024) at 16 spp, whose icosphere turns about the vertical axis
    between frames.  One step = pt_scene_update_mesh + pt_scene_commit (BVH
    rebuilt) + the render.  Both builds are timed on the same frames; `value`
    is the on-device build's path samples per second, rebuild included."""
    import numpy as np
    from pathtracer_amd import _lib as L, render as R, scenes
    spec = scenes.glass_hdri(spp_x=4, spp_y=4)
    if args.small:
        spec = scenes.glass_hdri(width=256, height=256, spp_x=2, spp_y=2, subdiv=6)
    pos0 = spec.meshes[0][0]
    n = args.warmup + args.steps
    frames = []
    for f in range(n):
        a = 0.05 * (f + 1)
        rot = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32)
        frames.append(np.ascontiguousarray(pos0 @ rot.T, np.float32))
    out = {}
    for build in ("lbvh", "sah"):
        flags = L.BUILD_DEVICE_LBVH if build == "lbvh" else 0
        sc = R.Scene(spec, build_flags=flags)
        cfg = sc.cfg(estimator=args.estimator)
        total = cfg.width * cfg.height * cfg.spp_x * cfg.spp_y
        parts = {"update_ms": 0.0, "commit_ms": 0.0, "bvh_ms": 0.0, "render_ms": 0.0}
        t_all = 0.0
        for f in range(n):
            L.check(L.lib.pt_device_synchronize(), "sync")
            t0 = time.perf_counter()
            sc.update_mesh(0, frames[f])
            t1 = time.perf_counter()
            sc.commit(build_flags=flags)
            t2 = time.perf_counter()
            sc.render_device(cfg)
            L.check(L.lib.pt_device_synchronize(), "sync")
            t3 = time.perf_counter()
            if f >= args.warmup:
                parts["update_ms"] += (t1 - t0) * 1e3
                parts["commit_ms"] += (t2 - t1) * 1e3
                parts["bvh_ms"] += sc.info.bvh_ms
                parts["render_ms"] += (t3 - t2) * 1e3
                t_all += t3 - t0
        out[build] = {"value": round(total * args.steps / t_all / 1e6, 2),
                      "ms_per_step": round(t_all / args.steps * 1e3, 2),
                      **{k: round(v / args.steps, 2) for k, v in parts.items()}}
        log(f"{build}: {out[build]}")
        sc.close()
    res = {
        "metric": METRIC, "value": out["lbvh"]["value"], "unit": "Mpath-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": out["lbvh"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "churn: glass_hdri with the 1.31M-tri icosphere turned every frame, 1024x1024, "
                               "16 spp; update + commit (BVH rebuild) + render per step",
                   "width": cfg.width, "height": cfg.height, "spp": cfg.spp_x * cfg.spp_y,
                   "triangles": spec.n_triangles, "estimator": ["reference_nee", "nee_mis"][args.estimator]},
        "device_lbvh": out["lbvh"], "host_sah": out["sah"],
        "speedup_vs_host_build": round(out["sah"]["ms_per_step"] / out["lbvh"]["ms_per_step"], 3),
    }
    print(json.dumps(res), flush=True)


def main_guided(args):
    """The slit-lit room (4x3x3 Lambertian room, Disney floor, 12.5k-triangle
    displaced conductor statue, a quad light of radiance 200 seen through a
    0.05-wide slit), 1024x1024, 256 spp, depth 16.  One step renders the frame
    with the learned guide, trained online from a fresh network inside the
    step (training passes of 4, 8 and 4 spp, 96 Adam steps after each, then
    the remaining 240 spp with the trained network; the first vertex of a
    path is guided).  The plain NEE+MIS
    tracer renders the same frame beside it.  Quality: relMSE against an
    8192-spp plain reference at the frame's resolution for both methods at
    16, 64 and 256 spp, and the time the guided method needs to reach the
    plain method's relMSE at 256 spp.  relMSE is taken on films scaled to the
    reference's unit mean."""
    import numpy as np
    from pathtracer_amd import _lib as L, render as R, scenes
    t0 = time.perf_counter()
    res_px = 256 if args.small else 1024
    spec = scenes.guiding_room(width=res_px, height=res_px)
    scene = R.Scene(spec)
    log(f"scene built in {time.perf_counter() - t0:.1f} s: {spec.n_triangles} triangles")
    cfg = scene.cfg()
    total = cfg.width * cfg.height * cfg.spp_x * cfg.spp_y
    gcfg = L.GuideCfg()  # the defaults: train on an eighth of the samples

    def guided_frame(c, seed):
        g = R.Guide(scene, seed=seed)
        t = time.perf_counter()
        _, st, gs = g.render(c, gcfg)
        dt = time.perf_counter() - t
        g.close()
        return dt, st, gs

    for i in range(args.warmup):
        scene.render_device(cfg)
        guided_frame(cfg, 100 + i)
        log(f"warmup {i} done")
    gst = []
    t1 = time.perf_counter()
    for i in range(args.steps):
        dt, st, gs = guided_frame(cfg, 1 + i)
        gst.append((dt, st, gs))
        log(f"guided step {i}: {dt * 1e3:.1f} ms, {gs['steps']} Adam steps on {gs['samples']} samples, "
            f"{gs['ms_train']:.1f} ms training")
    L.check(L.lib.pt_device_synchronize(), "sync")
    dt_g = time.perf_counter() - t1
    t2 = time.perf_counter()
    pst = []
    for i in range(args.steps):
        pst.append(scene.render_device(cfg))
    L.check(L.lib.pt_device_synchronize(), "sync")
    dt_p = time.perf_counter() - t2
    value = total * args.steps / dt_g / 1e6
    plain_value = total * args.steps / dt_p / 1e6
    log(f"guided {value:.1f} Mpath/s, plain {plain_value:.1f} Mpath/s")

    # quality at the frame's own resolution against an 8192-spp plain reference
    q = res_px
    curve, reach = None, None
    if not args.no_quality:
        curve, reach = guided_quality(scenes, R, L, q, scene=scene)
    cst = gst[0][1]
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpath-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt_g / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (network bf16/f32)",
        "data": "synthetic",
        "config": {"workload": "guided_room: slit-lit 4x3x3 room, Disney floor, 12.5k-tri conductor statue; "
                               "learned guide trained online inside the step",
                   "width": cfg.width, "height": cfg.height, "spp": cfg.spp_x * cfg.spp_y, "max_depth": cfg.max_depth,
                   "train_spp": cfg.spp_x * cfg.spp_y // 16, "guided_vertices": 1,
                   "triangles": spec.n_triangles},
        "plain": {"value": round(plain_value, 2), "ms_per_step": round(dt_p / args.steps * 1e3, 2)},
        "guide": {k: gst[0][2][k] for k in ("passes", "steps", "samples", "last_loss", "ms_train")},
        "quality": None if curve is None else {
            "resolution": f"{q}x{q}", "ref_spp": 8192, "curve": curve, "reach": reach},
        "segments": cst["segments"], "shadow_rays": cst["shadow_rays"],
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(spec, scene.cfg())
    print(json.dumps(res), flush=True)


def copy_probe(mib):
    """A device-to-device copy of a known size (hipMemcpy through torch's
    copy_, the runtime's copy kernel): what the traffic counters read for it
    scales their per-kernel figures to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_each_way": n * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--estimator", type=int, default=1, help="0: reference NEE, 1: NEE+MIS (default)")
    ap.add_argument("--no-quality", action="store_true", help="skip the relMSE measurement")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-other", action="store_true", help="skip the other estimator's frame (profiling runs)")
    ap.add_argument("--small", action="store_true", help="256x256 quick check (not a bench line)")
    ap.add_argument("--save", default="", help="write the film as PFM/PNG with this path prefix")
    ap.add_argument("--workload", default="glass_hdri",
                    choices=("glass_hdri", "cornell", "disney_hair", "batch64", "guided_room", "churn"),
                    help="glass_hdri: the single triangle-heavy frame (default); cornell: the Cornell box; "
                         "disney_hair: 64 Disney + 8 hair spheres (material divergence); batch64: 64 uneven renders "
                         "in one wavefront against 64 serial renders; guided_room: the slit-lit room with the "
                         "learned guide against the plain tracer; churn: glass_hdri with the mesh moved and the "
                         "BVH rebuilt every frame, device build against host build")
    ap.add_argument("--no-tune", action="store_true", help="batch64: skip the autotuner")
    ap.add_argument("--sort", action="store_true", help="order hits by material kind before shading "
                                                         "(PT_RENDER_SORT_HITS)")
    ap.add_argument("--build", default="sah", choices=("sah", "lbvh"),
                    help="BVH build: sah (host binned SAH, default) or lbvh (on the device)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device three times (the plain copy whose "
                         "FETCH_SIZE / WRITE_SIZE calibrate the traffic counters), then exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by pixel tiles with one film reduction; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        return main_batch(args)
    if args.workload == "guided_room":
        return main_guided(args)
    if args.workload == "churn":
        return main_churn(args)
    from pathtracer_amd import _lib as L, render as R, scenes

    t0 = time.perf_counter()
    small = dict(width=256, height=256, spp_x=4, spp_y=4) if args.small else {}
    spec = {"glass_hdri": scenes.glass_hdri, "cornell": scenes.cornell,
            "disney_hair": scenes.disney_hair_grid}[args.workload](**small)
    log(f"scene generated in {time.perf_counter() - t0:.1f} s: {spec.n_triangles} triangles")
    t0 = time.perf_counter()
    scene = R.Scene(spec, build_flags=build_flags(args))
    build = bvh_info(args, [scene], time.perf_counter() - t0)
    log(f"BVH built and uploaded: {build}")
    cfg = scene.cfg(estimator=args.estimator)
    total = cfg.width * cfg.height * cfg.spp_x * cfg.spp_y

    # counting frame (untimed): traversal counters for the roofline
    sort_flag = L.RENDER_SORT_HITS if args.sort else 0
    ccfg = scene.cfg(estimator=args.estimator, flags=L.RENDER_COUNTERS | sort_flag)
    cst = scene.render_device(ccfg)
    log(f"counting frame: {cst['segments']} extension rays, {cst['shadow_rays']} shadow rays")

    tcfg = scene.cfg(estimator=args.estimator, flags=L.RENDER_TIMING | sort_flag)
    for i in range(args.warmup):
        scene.render_device(tcfg)
        log(f"warmup {i} done")
    L.check(L.lib.pt_device_synchronize(), "sync")
    t1 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(scene.render_device(tcfg))
        log(f"step {i}: {stats[-1]['ms_total']:.1f} ms")
    L.check(L.lib.pt_device_synchronize(), "sync")
    dt = time.perf_counter() - t1
    value = total * args.steps / dt / 1e6

    # the other estimator on the same frame (one warm-up, one timed frame)
    other_res = None
    if not args.no_other:
        other = 1 - args.estimator
        ocfg = scene.cfg(estimator=other, flags=L.RENDER_TIMING | sort_flag)
        scene.render_device(ocfg)
        L.check(L.lib.pt_device_synchronize(), "sync")
        t2 = time.perf_counter()
        scene.render_device(ocfg)
        L.check(L.lib.pt_device_synchronize(), "sync")
        odt = time.perf_counter() - t2
        other_res = {"estimator": ["reference_nee", "nee_mis"][other], "value": round(total / odt / 1e6, 2),
                     "ms_per_step": round(odt * 1e3, 2)}
        log(f"other estimator: {other_res}")

    # roofline of the dominant kernel (HIP-event times of the timed frames)
    ms = {k: sum(s[k] for s in stats) / args.steps for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_generate",
                                                              "ms_film")}
    launches = stats[0]["extend_launches"]
    roofline, per_kernel = kernel_roofline(cst, ms, launches,
                                           read_traffic_for=args.estimator if args.workload == "glass_hdri" else None,
                                           qnodes=any_hit_qnodes([scene]), paths=total)

    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpath-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload] + (" (small)" if args.small else ""),
                   "width": cfg.width, "height": cfg.height, "spp": cfg.spp_x * cfg.spp_y,
                   "path_samples_per_step": total, "max_depth": cfg.max_depth,
                   "sampler": ["halton_owen", "halton_perm", "halton", "sobol", "stratified"][cfg.sampler],
                   "material_sort": bool(args.sort),
                   "estimator": ["reference_nee", "nee_mis"][args.estimator], "rr_depth": cfg.rr_depth,
                   "triangles": spec.n_triangles},
        "mrays_per_s": round((cst["segments"] + cst["shadow_rays"]) * args.steps / dt / 1e6, 1),
        "kernel_ms_per_step": {k: round(v, 2) for k, v in ms.items()},
        "counters_per_frame": {k: cst[k] for k in ("segments", "shadow_rays", "node_visits", "prim_tests",
                                                    "shadow_node_visits", "shadow_prim_tests", "nan_samples")},
        "roofline": roofline,
        "kernels": per_kernel,
        "other_estimator": other_res,
        "bvh": build,
        "quality": None,
        "cpu_baseline": None,
    }
    if not args.no_quality and not args.small and args.workload == "glass_hdri":
        res["quality"] = quality(scene, args.estimator)
        log(f"quality: {res['quality']}")
    if args.save:
        from pathtracer_amd import film_io
        import numpy as np
        img = np.empty((cfg.height, cfg.width, 3), np.float32)
        L.check(L.lib.pt_film_read(scene.h, L.ptr(img)), "pt_film_read")
        film_io.write_pfm(args.save + ".pfm", img)
        film_io.write_png(args.save + ".png", img)
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(spec, cfg)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
