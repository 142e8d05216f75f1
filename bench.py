#!/usr/bin/env python
"""Headline benchmark: plain GPU PathTracer + MIS on the ~2M-triangle sphere-field scene.

Workload: ground quad + 400 sphere meshes (5,120 triangles each) + 4 area-light
quads (8 light triangles), materials drawn from diffuse / GGX / anisotropic GGX /
glass / rough glass, 1920x1080, Halton, depth 8 (64 spp = 64 steps for the full
frame). One step = one Halton sample pass over all 2,073,600 pixels
(generate -> 8 x (extend, shade, occlude, mis, resolve) -> accumulate), with the
scene, BVH and film resident in HBM.

Prints ONE JSON line (metric/value in Msamples/s) with:
  roofline      dominant kernel: algorithmic bytes per launch / average launch
                time (HIP events on the renderer's stream), vs the 8 TB/s HBM peak;
                `traffic` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary.
  cpu_baseline  the CPU oracle (C++ restatement of the reference path) on a
                bounded sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# A renderer's two lanes and their side streams overlap only on different hardware queues.
# HIP spreads a process's streams over GPU_MAX_HW_QUEUES queues (4 by default) in an order
# that depends on every stream created before, and this benchmark creates many renderers:
# with 4 queues the same renderer, created five times in one process, ran at 405, 367, 367,
# 407 and 408 Msamples/s; with 16, at 407-408 every time. The runtime reads the setting when
# it starts, so it is set before the library is loaded (a value already set is kept).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "sphere-field path tracing (400 spheres x 5120 tris + ground + 4 area quads), 1920x1080, depth 8"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(n_spheres=400, level=4, seed=2024):
    """The sphere-field scene in memory (same draws as crystal2_amd.scenes.spheres_scene)."""
    from crystal2_amd import scenes
    return scenes.build_spheres_scene(n_spheres, level, seed)


def scene_bounds(scene):
    """Axis-aligned bounds of all triangles and spheres of a Scene (for the cache's encoding)."""
    d = scene.desc()
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    if d.num_triangles:
        t = np.ctypeslib.as_array(d.triangles, shape=(d.num_triangles * 24,)).reshape(-1, 8)[:, :3]
        lo, hi = np.minimum(lo, t.min(0)), np.maximum(hi, t.max(0))
    if d.num_spheres:
        s = np.ctypeslib.as_array(d.spheres, shape=(d.num_spheres * 22,)).reshape(-1, 22)
        lo = np.minimum(lo, (s[:, :3] - s[:, 3:4]).min(0))
        hi = np.maximum(hi, (s[:, :3] + s[:, 3:4]).max(0))
    return lo.tolist(), hi.tolist()


def per_ray_bytes(nodes, tris):
    # 24 B ray in + 8 B hit out + 32 B per binary-BVH node visit + 36 B per triangle test
    return 24 + 8 + nodes * 32 + tris * 36


def load_traffic(kernel, rays_per_launch=None):
    """HBM bytes per launch of `kernel` from the committed counter summary, scaled from the
    profiled launches' mean ray count to `rays_per_launch` when both are known."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    v = t.get(kernel)
    prof = t.get("_rays_per_launch", {}).get(kernel)
    if v is not None and prof and rays_per_launch:
        v = round(v * rays_per_launch / prof)
    return v


def copy_calibration(mib, reps=3):
    """A plain streaming kernel of known size (the library's k_stream_copy):
    each launch reads exactly `mib` MiB and writes `mib` MiB."""
    import crystal2_amd as c2
    ms = c2.stream_copy(mib << 20, reps)
    log(f"copy calibration: {reps} x k_stream_copy over {mib} MiB ({mib << 20} B read + written per launch), "
        f"{ms:.3f} ms avg, {2 * (mib << 20) / ms / 1e6:.0f} GB/s")


def cpu_baseline(scene, width, height, depth, seconds_budget=20.0, integrator="path"):
    """Oracle (C++ port of the reference path) on a band of rows at sample 0."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    threads, avail, _ = host_cores()
    t0 = time.time()
    orc = pyoracle.OracleScene(scene, variant="portable", integrator=integrator)
    log(f"cpu baseline: oracle scene built in {time.time() - t0:.1f}s, {threads} threads")
    # calibrate on 8 rows, then size the band to about seconds_budget
    mid = height // 2
    _, s = orc.render(width, height, 1, depth, threads=threads, y_range=(mid - 4, mid + 4))
    rate = 8 * width / max(s, 1e-6)
    rows = int(max(8, min(height, seconds_budget * rate / width)))
    y0 = max(0, mid - rows // 2)
    _, s = orc.render(width, height, 1, depth, threads=threads, y_range=(y0, y0 + rows))
    samples = rows * width
    return {"value": samples / s / 1e6, "unit": "Msamples/s", "cores": threads, "host_cpus": avail, "kind": "port",
            "sample": f"rows {y0}..{y0 + rows} of {width}x{height}, Halton sample 0 ({samples} samples, {s:.1f}s)"}


def cpu_reference_baselines(width, height, depth, threads):
    """The reference renderer itself (oracle/_ref/crystal2_ref, compiled from the reference's
    sources) on the same sphere field written in its JSON + OBJ formats, for the path tracer
    and the direct-lighting integrator: Halton sample 0 of the full frame, Integrator::Render
    timed by the driver (scene parsing excluded). Returns {integrator: baseline} or None."""
    import shutil
    import subprocess
    import tempfile
    from crystal2_amd import scenes
    ref = os.path.join(HERE, "oracle", "_ref", "crystal2_ref")
    if not os.path.exists(ref):
        return None
    d = tempfile.mkdtemp(prefix="c2ref_")
    out = {}
    try:
        t0 = time.time()
        scenes.spheres_scene(d, width, height, spp=1, depth=depth, threads=threads)
        scenes.write_config(os.path.join(d, "config_dl.json"), width, height, 1, depth, threads=threads,
                            integrator="DirectLight")
        log(f"cpu reference: scene files written in {time.time() - t0:.1f}s")
        for key, cfg in (("path", "config.json"), ("direct", "config_dl.json")):
            res = subprocess.run([ref, "render", cfg, "scene.json", "film.f32"], cwd=d, check=True,
                                 capture_output=True, text=True, timeout=600).stdout
            rec = json.loads(res.strip().splitlines()[-1])
            out[key] = {"value": rec["samples"] / rec["render_s"] / 1e6, "unit": "Msamples/s",
                        "cores": rec["threads"], "host_cpus": len(os.sched_getaffinity(0)), "kind": "reference",
                        "sample": f"Halton sample 0 of the full {width}x{height} frame, depth {depth} "
                                  f"({rec['samples']} samples, {rec['render_s']:.1f}s in Integrator::Render)"}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def extras_bench(W, H, depth, steps=16, threads=16, cpu=True):
    """Throughput of the rows built beside the path tracer, each with the reference
    renderer timed on the same scene and frame (Halton sample 0, `threads` cores):
    the KD-tree accelerator (and the BVH on the same scene) on a 100-sphere field
    (512,010 entities), image textures with the cubemap skybox, Disk shapes, and the
    Oren-Nayar / Plastic / Beckmann materials. GPU: `steps` passes timed after 2 warm-up
    passes, two lanes and four samples per launch as in the main measurement."""
    import shutil
    import subprocess
    import tempfile
    import crystal2_amd as c2
    from crystal2_amd import scenes
    ref = os.path.join(HERE, "oracle", "_ref", "crystal2_ref")
    golden = os.path.join(HERE, "tests", "golden", "scenes")
    tmp = tempfile.mkdtemp(prefix="c2extras_")
    out = {}
    try:
        field_dir = os.path.join(tmp, "field100")
        scenes.spheres_scene(field_dir, W, H, spp=1, depth=depth, n_spheres=100, threads=threads)
        field = scenes.build_spheres_scene(100, 4, 2024)
        cases = [
            ("kdtree", field, "kdtree", field_dir, "KDTree"),
            ("bvh_same_scene", field, "bvh", field_dir, "BVH"),
            ("textures_skybox", c2.Scene.load(os.path.join(golden, "textured", "scene.json"),
                                              os.path.join(golden, "textured")), "bvh",
             os.path.join(golden, "textured"), "BVH"),
            ("disks", c2.Scene.load(os.path.join(golden, "disks", "scene.json"), os.path.join(golden, "disks")),
             "bvh", os.path.join(golden, "disks"), "BVH"),
            ("orennayar_plastic_beckmann", c2.Scene.load(os.path.join(golden, "zoo_extra", "scene.json"),
                                                         os.path.join(golden, "zoo_extra")), "bvh",
             os.path.join(golden, "zoo_extra"), "BVH"),
        ]
        for name, sc, accel, sdir, refaccel in cases:
            t0 = time.perf_counter()
            pt = c2.WavefrontPathIntegrator(sc, W, H, spp=steps + 2, max_depth=depth, arg_order="rtl", film="f32",
                                            accel=accel)
            build_s = time.perf_counter() - t0
            pt.render(0, 2)
            pt.sync()
            pt.reset()
            a = time.perf_counter()
            pt.render(0, steps)
            pt.sync()
            el = time.perf_counter() - a
            pt.close()
            rec = {"msamples_s": round(W * H * steps / el / 1e6, 2), "entities": sc.num_entities,
                   "accel": accel, "renderer_build_s": round(build_s, 2)}
            if cpu and os.path.exists(ref) and name != "bvh_same_scene":
                cfg = os.path.join(tmp, f"config_{name}.json")
                scenes.write_config(cfg, W, H, 1, depth, threads=threads, accel=refaccel)
                res = subprocess.run([ref, "render", cfg, "scene.json", os.path.join(tmp, "film.f32")], cwd=sdir,
                                     check=True, capture_output=True, text=True, timeout=600).stdout
                r = json.loads(res.strip().splitlines()[-1])
                v = r["samples"] / r["render_s"] / 1e6
                rec["cpu_reference"] = {"value": round(v, 3), "unit": "Msamples/s", "cores": r["threads"],
                                        "kind": "reference", "accel": refaccel,
                                        "sample": f"Halton sample 0 of the full {W}x{H} frame, depth {depth}"}
                rec["vs_cpu_reference"] = round(rec["msamples_s"] / v, 1)
            out[name] = rec
            log(f"extras {name}: {rec}")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def cpu_model():
    """The host CPU as /proc/cpuinfo names it: model name with family / model / stepping
    (some hosts report only "AMD EPYC"), and the logical CPU count."""
    info = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                    info[k] = v.strip()
                if not line.strip() and info:
                    break
    except OSError:
        return None
    name = info.get("model name")
    if name is None:
        return None
    return (f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}; {os.cpu_count()} logical CPUs)")


def host_cores():
    """Threads used for CPU baselines (at most 16: the reference's SamplerIntegrator asserts
    threads <= 16, SamplerIntegrator.cpp:13) and what the host offers."""
    avail = len(os.sched_getaffinity(0))
    return max(1, min(16, avail)), avail, os.cpu_count()


def cornell_bench(reps=5):
    """The Cornell-box workload: 5 wall quads, a 0.5x0.5 light quad (Lemit 15), diffuse and
    mirror spheres r = 0.3, 256x256, 16 spp, depth 5, FOV 40 (1,048,576 samples). GPU: the
    full 16-spp render timed `reps` times after a warm-up (median). The reference renderer
    (GCC build: the right-to-left argument order the GPU renders) renders the same scene
    files on the host cores; its film is the parity check (relMSE, north star <= 1e-4)."""
    import shutil
    import subprocess
    import tempfile
    import crystal2_amd as c2
    from crystal2_amd import scenes
    W = H = 256
    spp, depth = 16, 5
    threads, avail, ncpu = host_cores()
    d = tempfile.mkdtemp(prefix="c2cornell_")
    try:
        scenes.cornell(d, W, H, spp, depth, threads=threads)
        sc = c2.Scene.load(os.path.join(d, "scene.json"), d)
        pt = c2.WavefrontPathIntegrator(sc, W, H, spp, depth, arg_order="rtl")
        pt.render()
        pt.sync()
        times = []
        for _ in range(reps):
            pt.reset()
            a = time.perf_counter()
            pt.render()
            pt.sync()
            times.append(time.perf_counter() - a)
        film = pt.film()
        pt.close()
        el = float(np.median(times))
        out = {"workload": "Cornell box, 256x256, 16 spp, depth 5", "samples": W * H * spp,
               "msamples_s": round(W * H * spp / el / 1e6, 2), "render_ms": round(el * 1e3, 3)}
        ref = os.path.join(HERE, "oracle", "_ref", "crystal2_ref_gcc")
        if os.path.exists(ref):
            res = subprocess.run([ref, "render", "config.json", "scene.json", "film.f32"], cwd=d, check=True,
                                 capture_output=True, text=True, timeout=600).stdout
            r = json.loads(res.strip().splitlines()[-1])
            rf = np.fromfile(os.path.join(d, "film.f32"), np.float32).reshape(H, W, 3)
            v = r["samples"] / r["render_s"] / 1e6
            out["cpu_reference"] = {"value": round(v, 3), "unit": "Msamples/s", "cores": r["threads"],
                                    "host_cpus": avail, "kind": "reference",
                                    "sample": f"the full render ({r['samples']} samples, {r['render_s']:.3f}s)"}
            out["vs_cpu_reference"] = round(out["msamples_s"] / v, 1)
            out["relmse_vs_reference"] = float(_relmse(film, rf))
            out["identical_pixels_vs_reference"] = round(float((film == rf).all(-1).mean()), 5)
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def raycast_bench(reps=5, cpu_rays=1 << 21):
    """The ray-cast workload: a displaced icosphere of 1,310,720 triangles (level 8,
    radius 1 + 0.1 value noise, seed 7) and 2^24 rays (origins uniform in 1.5x its box,
    directions uniform on S^2, tmax = inf), resident in HBM. Closest hit and any hit through
    c2_accel_intersect[_test]_device, each timed `reps` times with HIP events on the launch
    stream. Achieved GB/s uses the algorithmic bytes per ray (24 + 8 + 32 N_node + 36 N_tri,
    counted on the binary SAH tree over 65,536 of the rays). The reference's own BVH casts
    the first `cpu_rays` rays on the host cores (its raycast command); its ids are the
    agreement check (north star: >= 99.99% identical, t within 1e-5 relative)."""
    import shutil
    import subprocess
    import tempfile
    import torch
    import crystal2_amd as c2
    from crystal2_amd import scenes
    t0 = time.time()
    pos, nrm, faces = scenes.displaced_icosphere(8, seed=7)
    sc = c2.Scene.create()
    m = sc.add_material(0, [0.5, 0.5, 0.5])
    sc.add_mesh(pos, nrm, faces, m)
    n = 1 << 24
    rays = scenes.icosphere_rays(n, pos, seed=11)
    bvh = c2.GpuBVH(sc)
    log(f"raycast: {sc.num_entities} triangles, {n} rays, BVH in {time.time() - t0:.1f}s")
    sub = rays[np.random.default_rng(0).choice(n, 1 << 16, replace=False)]
    cn, ct = bvh.count_visits(sub, any_hit=False)
    an, at = bvh.count_visits(sub, any_hit=True)
    dr = torch.from_numpy(rays).cuda()
    dp = torch.empty(n, dtype=torch.int32, device="cuda")
    dt = torch.empty(n, dtype=torch.float32, device="cuda")
    dh = torch.empty(n, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        return float(np.median(ms))

    cms = timed(lambda: bvh.intersect_device(dr.data_ptr(), n, dp.data_ptr(), dt.data_ptr(), sp))
    ams = timed(lambda: bvh.intersect_test_device(dr.data_ptr(), 0, n, dh.data_ptr(), sp))
    ids, tt, hits = dp.cpu().numpy(), dt.cpu().numpy(), dh.cpu().numpy().astype(bool)
    out = {"workload": "ray cast: displaced icosphere 1,310,720 tris, 2^24 rays", "rays": n,
           "closest": {"mrays_s": round(n / cms * 1e-3, 1), "ms": round(cms, 3), "nodes": round(cn, 2),
                       "tris": round(ct, 2), "bytes_per_ray": round(per_ray_bytes(cn, ct), 1),
                       "achieved_gbs": round(n * per_ray_bytes(cn, ct) / (cms * 1e-3) / 1e9, 1),
                       "hit_fraction": round(float((ids >= 0).mean()), 4)},
           "any": {"mrays_s": round(n / ams * 1e-3, 1), "ms": round(ams, 3), "nodes": round(an, 2),
                   "tris": round(at, 2), "bytes_per_ray": round(per_ray_bytes(an, at), 1),
                   "achieved_gbs": round(n * per_ray_bytes(an, at) / (ams * 1e-3) / 1e9, 1)}}
    for k in ("closest", "any"):
        out[k]["frac_hbm"] = round(out[k]["achieved_gbs"] / HBM_PEAK_GBS, 4)
    del dr, dp, dt, dh
    ref = os.path.join(HERE, "oracle", "_ref", "crystal2_ref")
    if cpu_rays and os.path.exists(ref):
        threads, avail, _ = host_cores()
        d = tempfile.mkdtemp(prefix="c2raycast_")
        try:
            scenes.write_obj(os.path.join(d, "mesh.obj"), pos, nrm, faces)
            with open(os.path.join(d, "scene.json"), "w") as fh:
                json.dump({"Camera": {"Type": "Normal", "Position": [0, 0, 4], "LookAt": [0, 0, 0], "Up": [0, 1, 0],
                                      "FOV": 40},
                           "Scene": {"Textures": {}, "Objects": [{"Type": "Geometry", "Shape": {
                               "Type": "TriangleMesh", "ObjFile": "mesh.obj", "Translation": [0, 0, 0],
                               "Scale": [1, 1, 1], "Rotation": [0, 0, 0]},
                               "Material": {"Type": "Diffuse", "Color": [0.5, 0.5, 0.5]}}]}}, fh)
            scenes.write_config(os.path.join(d, "config.json"), 64, 64, 1, 1, threads=threads)
            m_ = min(cpu_rays, n)
            with open(os.path.join(d, "rays.bin"), "wb") as fh:
                np.array([m_], np.int64).tofile(fh)
                rays[:m_].tofile(fh)
            res = subprocess.run([ref, "raycast", "config.json", "scene.json", "rays.bin", str(threads), "out.bin"],
                                 cwd=d, check=True, capture_output=True, text=True, timeout=900).stdout
            r = json.loads(res.strip().splitlines()[-1])
            raw = open(os.path.join(d, "out.bin"), "rb").read()
            rid = np.frombuffer(raw, np.int32, m_, 0)
            rt = np.frombuffer(raw, np.float32, m_, 4 * m_)
            ra = np.frombuffer(raw, np.uint8, m_, 8 * m_).astype(bool)
            agree = ids[:m_] == rid
            both = agree & (rid >= 0)
            rel = np.abs(tt[:m_][both].astype(np.float64) - rt[both]) / np.maximum(rt[both], 1e-30)
            out["cpu_reference"] = {
                "closest_mrays_s": round(m_ / r["closest_s"] / 1e6, 3), "any_mrays_s": round(m_ / r["any_s"] / 1e6, 3),
                "cores": r["threads"], "host_cpus": avail, "kind": "reference",
                "sample": f"the first {m_} rays on the reference's median-split BVH"}
            out["vs_cpu_reference"] = {"closest": round(out["closest"]["mrays_s"] / (m_ / r["closest_s"] / 1e6), 1),
                                       "any": round(out["any"]["mrays_s"] / (m_ / r["any_s"] / 1e6), 1)}
            out["agreement_vs_reference"] = {"ids": round(float(agree.mean()), 6),
                                             "max_rel_dt": float(rel.max()) if rel.size else 0.0,
                                             "any_hit": round(float((hits[:m_] == ra).mean()), 6)}
        finally:
            shutil.rmtree(d, ignore_errors=True)
    bvh.close()
    return out


def _crop(film, W, H, win):
    """Window (x0, y0, w, h) in film coordinates (y up) of a top-down film."""
    x0, y0, w, h = win
    return film[H - y0 - h:H - y0, x0:x0 + w]


def _relmse(x, r):
    return float(np.mean((x.astype(np.float64) - r) ** 2 / (r.astype(np.float64) ** 2 + 1e-2)))


def equal_time(scene, W, H, depth, seconds=0.0, truth_spp=4096, crop=256, cache_bounce=2, train_ratio=48,
               warmup_passes=8, budget_passes=64):
    """relMSE at equal wall time: plain PT vs radiance-cache PT, both measured against a
    plain render of `truth_spp` samples with an FP64 film on a centre crop. The ground truth uses
    Halton samples [truth_offset, truth_offset + truth_spp), the runs samples from 0.
    The budget is fixed by the workload, not by the headline's step count: the wall time of
    `budget_passes` plain passes rendered in one call (or `seconds` when given). The cache
    run's `warmup_passes` training passes (samples outside both runs) count in its time.
    train_ratio 48 trains about 2^16 path vertices per pass at 1920x1080."""
    import crystal2_amd as c2
    win = ((W - crop) // 2, (H - crop) // 2, crop, crop)
    truth_offset = 8192
    t0 = time.perf_counter()
    truth = c2.WavefrontPathIntegrator(scene, W, H, truth_spp, depth, film="f64", window=win)
    truth.render(truth_offset, truth_spp)
    truth.sync()
    ref = _crop(truth.film(), W, H, win)
    truth.close()
    log(f"ground truth: {truth_spp} spp FP64 on {crop}x{crop} in {time.perf_counter() - t0:.1f}s")

    def run(pt, budget, checkpoints, start=0):
        # passes rendered in one call up to the next checkpoint (a renderer overlaps the
        # passes of one call), fewer when the time left would not hold them
        used, k, curve, per = 0.0, start, [], None
        while used < budget:
            done = k - start
            n = min((c for c in checkpoints if c > done), default=done * 2) - done
            if per:
                n = max(1, min(n, int((budget - used) / per)))
            a = time.perf_counter()
            pt.render(k, n)
            pt.sync()
            dt = time.perf_counter() - a
            used += dt
            per = dt / n
            k += n
            if (k - start) in checkpoints or used >= budget:
                curve.append((round(used, 4), k - start, _relmse(_crop(pt.film(), W, H, win), ref)))
        return curve

    cps = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 40, 48, 56, 64, 80, 96, 112, 128, 160, 192, 224, 256, 320, 384, 448,
           512}
    plain = c2.WavefrontPathIntegrator(scene, W, H, 4096, depth, film="f32")
    if seconds <= 0:  # the fixed budget: budget_passes plain passes in one call (after a warm-up)
        plain.render(0, 4)
        plain.sync()
        plain.reset()
        a = time.perf_counter()
        plain.render(0, budget_passes)
        plain.sync()
        seconds = time.perf_counter() - a
        plain.reset()
    plain_curve = run(plain, seconds, cps)
    plain.close()
    lo, hi = scene_bounds(scene)
    cache = c2.RadianceCache(lo, hi, learning_rate=1e-2, batch_capacity=1 << 16, seed=1)
    pt = c2.WavefrontPathIntegrator(scene, W, H, 4096, depth, film="kahan")
    pt.set_cache(cache, cache_bounce=cache_bounce, train_ratio=train_ratio)
    a = time.perf_counter()
    pt.render(4096, warmup_passes)  # training passes on samples outside both runs
    pt.sync()
    warm = time.perf_counter() - a
    pt.reset()
    pt.cache_times()  # clear
    st0 = cache.stats()  # the warm-up's queries and rows are not in the timed kernels
    cache_curve = run(pt, seconds - warm, cps)
    cache_curve = [(round(t + warm, 4), n, r) for (t, n, r) in cache_curve]
    ct = pt.cache_times()
    st = cache.stats()
    st = dict(st, queries=st["queries"] - st0["queries"], train_rows=st["train_rows"] - st0["train_rows"],
              steps_timed=st["steps"] - st0["steps"])
    pt.close()
    target = plain_curve[-1][2]
    reach = next((t for (t, n, r) in cache_curve if r <= target), None)
    flop_q = st["flop_forward_per_query"]
    npass = cache_curve[-1][1]
    infer_tflops = st["queries"] * flop_q / max(ct["infer_ms"], 1e-9) / 1e9
    # training: forward + backward (dX) + weight gradient = 3x the forward FLOP per row, over
    # the training kernels' time (gathering the rows, which runs no MFMA work, is apart)
    train_tflops = st["train_rows"] * 3 * flop_q / max(ct["train_ms"], 1e-9) / 1e9
    return {
        "seconds": seconds,
        "budget": (f"the wall time of {budget_passes} plain passes in one call" if budget_passes else "given"),
        "ground_truth": {"spp": truth_spp, "film": "f64", "window": list(win),
                         "samples": [truth_offset, truth_offset + truth_spp]},
        "plain": {"passes": plain_curve[-1][1], "relmse": plain_curve[-1][2], "curve": plain_curve},
        "cache": {"passes": npass, "warmup_passes": warmup_passes, "warmup_s": round(warm, 4),
                  "relmse": cache_curve[-1][2], "curve": cache_curve, "cache_bounce": cache_bounce,
                  "train_ratio": train_ratio, "adam_steps": st["steps"]},
        "cache_time_to_plain_relmse_frac": None if reach is None else round(reach / seconds, 4),
        "mfma": {"infer_tflops": round(infer_tflops, 2), "train_tflops": round(train_tflops, 2),
                 "infer_util_vs_fp16_peak": round(infer_tflops / 2500.0, 5),
                 "train_util_vs_fp16_peak": round(train_tflops / 2500.0, 5),
                 "infer_ms_per_pass": round(ct["infer_ms"] / max(1, npass), 4),
                 "train_ms_per_pass": round(ct["train_ms"] / max(1, npass), 4),
                 "collect_ms_per_pass": round(ct["collect_ms"] / max(1, npass), 4),
                 "train_rows_per_pass": round(st["train_rows"] / max(1, npass))},
    }


def direct_lighting_bench(scene, W, H, depth, steps, cpu_seconds=0.0):
    """DirectLightingIntegrator on the same scene and film: GPU Msamples/s over `steps`
    sample passes (scene and film resident), and the oracle port on a bounded CPU sample."""
    import crystal2_amd as c2
    pt = c2.WavefrontPathIntegrator(scene, W, H, spp=max(64, steps + 1), max_depth=depth, arg_order="rtl",
                                    integrator="direct")
    pt.render(0, 1)
    pt.sync()
    pt.reset()
    a = time.perf_counter()
    pt.render(0, steps)
    pt.sync()
    el = time.perf_counter() - a
    st = pt.stats()
    pt.close()
    out = {"msamples_s": round(W * H * steps / el / 1e6, 2), "steps": steps,
           "ms_per_step": round(el / steps * 1e3, 3),
           "rays_per_sample": {"extension": round(st["extension_rays"] / (W * H * (steps + 1)), 3),
                               "shadow": round(st["shadow_rays"] / (W * H * (steps + 1)), 3)}}
    if cpu_seconds > 0:
        try:
            out["cpu_baseline"] = cpu_baseline(scene, W, H, depth, cpu_seconds, integrator="direct")
        except Exception as e:  # the GPU figure stands alone if the checker is absent
            log(f"direct-lighting cpu baseline unavailable: {e}")
    return out


def batch_bench(n=64, seed=6400, depth=8):
    """Batched scenes: n sphere-field scenes (triangles log-uniform in [1k, 200k],
    resolution in {128, 256, 512}^2, spp in {4, 16, 64}) rendered (a) one scene at a
    time with its own launches and (b) as one batch replayed from a captured hipGraph
    with the schedule the autotuner keeps (the wavefront's queue kernels or the one
    persistent-kernel launch; the other schedule is timed beside it). Timed wall-clock
    after a warm-up render; every batch film must equal its single-scene film bit for
    bit, under both schedules."""
    import crystal2_amd as c2
    from crystal2_amd import scenes
    t0 = time.time()
    specs = scenes.batch_specs(n, seed)
    items = scenes.build_batch(specs)
    log(f"batch: {n} scenes ({sum(s['triangles'] for s in specs)} triangles) built in {time.time() - t0:.1f}s")
    samples = sum(w * h * spp for _, w, h, spp in items)
    # (a) per-scene launches: each scene its own renderer, rendered in turn
    pts = [c2.WavefrontPathIntegrator(sc, w, h, spp, depth, arg_order="rtl") for sc, w, h, spp in items]
    for pt in pts:  # warm-up
        pt.render(0, 1)
    for pt in pts:
        pt.sync()
        pt.reset()
    a = time.perf_counter()
    for pt in pts:
        pt.render()
    for pt in pts:
        pt.sync()
    single_s = time.perf_counter() - a
    ref = [pt.film() for pt in pts]
    for pt in pts:
        pt.close()
    # (b) one batch, autotuned, replayed from a captured graph
    b = c2.BatchRenderer(items, depth, arg_order="rtl")
    tune = b.autotune()
    b.render(True)  # capture + first replay (warm-up)
    b.sync()
    a = time.perf_counter()
    b.render(True)
    b.sync()
    batch_s = time.perf_counter() - a
    exact = sum(int(np.array_equal(b.film(i), r)) for i, r in enumerate(ref))
    st = b.stats()
    # the schedule the autotuner did not keep, timed the same way
    other = "wavefront" if tune["schedule"] == "persistent" else "persistent"
    b.configure(other, block_threads=tune["block_threads"], waves_per_simd=tune["waves_per_simd"])
    b.render(True)
    b.sync()
    a = time.perf_counter()
    b.render(True)
    b.sync()
    other_s = time.perf_counter() - a
    exact_other = sum(int(np.array_equal(b.film(i), r)) for i, r in enumerate(ref))
    other_nodes = b.stats()["graph_nodes"]
    b.close()
    out = {"scenes": n, "triangles": int(sum(s["triangles"] for s in specs)), "samples": int(samples),
           "depth": depth, "per_scene_s": round(single_s, 4), "batch_graph_s": round(batch_s, 4),
           "per_scene_msamples_s": round(samples / single_s / 1e6, 2),
           "batch_graph_msamples_s": round(samples / batch_s / 1e6, 2),
           "speedup": round(single_s / batch_s, 3), "films_bit_identical": f"{exact}/{n}",
           "schedule": tune["schedule"], "graph_nodes": st["graph_nodes"],
           "autotune": {"paths_in_flight": tune["paths_in_flight"], "queue_blocks": tune["queue_blocks"],
                        "candidates": tune["candidates"], "passes_ms": round(tune["pass_ms"], 3),
                        "bvh_width": tune["bvh_width"], "passes_ms_width8": round(tune["pass_ms_width8"], 3),
                        "persistent_block_threads": tune["block_threads"],
                        "persistent_waves_per_simd": tune["waves_per_simd"],
                        "persistent_sample0_ms": round(tune["sample0_ms_persistent"], 3),
                        "render_ms_wavefront": round(tune["render_ms_wavefront"], 2),
                        "render_ms_persistent": round(tune["render_ms_persistent"], 2)},
           "other_schedule": {"schedule": other, "graph_s": round(other_s, 4),
                              "msamples_s": round(samples / other_s / 1e6, 2), "graph_nodes": other_nodes,
                              "films_bit_identical": f"{exact_other}/{n}"}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--spheres", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cache", action="store_true", help="skip the equal-time radiance-cache comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched-render comparison")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--no-direct", action="store_true", help="skip the DirectLighting integrator line")
    ap.add_argument("--no-cornell", action="store_true", help="skip the Cornell-box (256x256, 16 spp) line")
    ap.add_argument("--no-raycast", action="store_true", help="skip the ray-cast (2^24 rays) line")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the KD-tree / textures / disk / material-model throughput lines")
    ap.add_argument("--equal-seconds", type=float, default=0.0,
                    help="equal-time budget in seconds (0: the wall time of 64 plain passes)")
    ap.add_argument("--truth-spp", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain fp32 negate kernel over MIB MiB (reads MIB, writes MIB): "
                         "the known-size reference that scales rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(several GPUs would each render a Halton sample range and sum the films)", file=sys.stderr)
        sys.exit(2)

    # torch holds the ray-cast line's device buffers; it must find the GPU before the
    # library's HIP runtime claims it in this process
    import torch
    if not torch.cuda.is_available():
        log("torch sees no GPU: the ray-cast line will be skipped")
        args.no_raycast = True
    import crystal2_amd as c2
    W, H, K = args.width, args.height, args.steps
    t0 = time.time()
    scene = build_scene(args.spheres)
    log(f"scene: {scene.num_entities} entities built in {time.time() - t0:.1f}s")
    t0 = time.time()
    pt = c2.WavefrontPathIntegrator(scene, W, H, spp=max(64, args.warmup + K), max_depth=args.depth,
                                    arg_order="rtl", film="f32")
    log(f"renderer + BVH ready in {time.time() - t0:.1f}s")

    # algorithmic bytes per ray from the rays the renderer actually traces: every bounce of
    # sample 0, each bounce's queue captured and subsampled uniformly, the node and triangle
    # counts averaged with the bounce's ray count as its weight
    bvh_for_counts = c2.GpuBVH(scene)
    counts = {}
    rng = np.random.default_rng(1)
    for kind in ("extension", "shadow"):
        tot = wn = wt = 0.0
        per_bounce = []
        for bounce in range(args.depth):
            rays, _ = pt.capture_rays(0, bounce, "shadow" if kind == "shadow" else "extension", n=1 << 23)
            if len(rays) == 0:
                continue
            sub = rays[rng.choice(len(rays), min(len(rays), 8192), replace=False)]
            nodes, tris = bvh_for_counts.count_visits(sub, any_hit=(kind == "shadow"))
            tot += len(rays)
            wn += nodes * len(rays)
            wt += tris * len(rays)
            per_bounce.append((bounce, len(rays), round(nodes, 1)))
        nodes, tris = wn / max(tot, 1), wt / max(tot, 1)
        counts[kind] = (nodes, tris, per_ray_bytes(nodes, tris))
        log(f"{kind} rays (all bounces of sample 0, (bounce, rays, nodes)): {per_bounce} -> {nodes:.1f} node "
            f"visits, {tris:.1f} tri tests -> {counts[kind][2]:.0f} B/ray")
    bvh_for_counts.close()
    pt.reset()

    if args.warmup:
        pt.render(0, args.warmup)
        pt.sync()
    # per-kernel breakdown from a separate, fully instrumented render of 4 passes (not
    # timed); it also names the dominant tracing kernel, the one the roofline is quoted on
    pt.reset()
    pt.set_profiling(True)
    sb = pt.stats()
    pt.render(0, 4)
    pt.sync()
    sa = pt.stats()
    kti = pt.kernel_times()
    for name, v in kti.items():
        if v["launches"]:
            log(f"  {name:10s} {v['ms'] / 4:9.2f} ms/pass  {v['launches'] / 4:6.2f} launches/pass  "
                f"{v['ms'] / v['launches']:.3f} ms avg")
    pt.set_profiling(False)
    dom = max(("extend", "occlude"), key=lambda k: kti[k]["ms"])
    pt.reset()
    st0 = pt.stats()
    # HIP events in the timed region only around the dominant kernel's launches: events
    # around every launch slow it by ~4%, around both tracing kernels by ~2%
    pt.set_profiling(True, kernels=[dom])
    pt.sync()
    t0 = time.perf_counter()
    pt.render(0, K)
    pt.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kt = pt.kernel_times()
    st = pt.stats()
    pt.set_profiling(False)
    samples = W * H * K
    value = samples / elapsed / 1e6
    ext = st["extension_rays"] - st0["extension_rays"]
    sh = st["shadow_rays"] - st0["shadow_rays"]
    mis = st["mis_rays"] - st0["mis_rays"]
    log(f"timed {K} passes in {elapsed:.3f}s: {value:.2f} Msamples/s; rays/sample ext {ext / samples:.2f} "
        f"shadow {sh / samples:.2f} mis {mis / samples:.3f}")
    # rays per second of each tracing stage over its own launches (wall time of those
    # launches; with two lanes another lane's kernels share the GPU meanwhile)
    stages = {}
    for stage, key, kern in (("extend", "extension_rays", "extend"), ("occlude", "shadow_rays", "occlude"),
                             ("mis", "mis_rays", "mis")):
        rays = sa[key] - sb[key]
        ms = kti[kern]["ms"]
        stages[stage] = {"rays": int(rays), "ms": round(ms, 3),
                         "mrays_s": round(rays / (ms / 1e3) / 1e6, 1) if ms > 0 else None}

    rays_dom = ext if dom == "extend" else sh
    bpr = counts["extension" if dom == "extend" else "shadow"][2]
    avg_s = kt[dom]["ms"] / 1e3 / max(1, kt[dom]["launches"])
    bytes_per_launch = rays_dom * bpr / max(1, kt[dom]["launches"])
    achieved = bytes_per_launch / avg_s / 1e9
    kname = {"extend": "k_extend", "occlude": "k_occlude"}[dom]
    traffic = load_traffic(kname, rays_dom / max(1, kt[dom]["launches"]))
    # The path has no dense contraction: HBM is the roofline that applies. What the kernel
    # actually moves through HBM (the committed FETCH_SIZE + WRITE_SIZE summary) over its
    # launch time tells whether it is at that roof or limited below it by memory latency
    # and address issue (the upper tree levels stay in L2 / MALL).
    measured = traffic / avg_s / 1e9 if traffic else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                "bytes_per_launch": round(bytes_per_launch), "avg_launch_ms": round(avg_s * 1e3, 4),
                "measured_hbm_gbs": round(measured, 1) if measured else None,
                "measured_hbm_frac": round(measured / HBM_PEAK_GBS, 4) if measured else None,
                "limiter": (None if measured is None else "hbm bandwidth" if measured >= 0.6 * HBM_PEAK_GBS
                            else "memory latency and address issue (HBM traffic below 60% of peak)")}

    pt.close()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    cornell = ray_cast = None
    if not args.no_cornell:
        cornell = cornell_bench()
        log(f"cornell: {cornell}")
    if not args.no_raycast:
        ray_cast = raycast_bench(cpu_rays=0 if args.no_cpu_baseline else 1 << 21)
        log(f"raycast: {ray_cast}")
    eq = None
    if not args.no_cache:
        eq = equal_time(scene, W, H, args.depth, seconds=args.equal_seconds, truth_spp=args.truth_spp)
        log(f"equal time {eq['seconds']:.2f}s: plain relMSE {eq['plain']['relmse']:.4e} ({eq['plain']['passes']} "
            f"passes), cache relMSE {eq['cache']['relmse']:.4e} ({eq['cache']['passes']} passes); "
            f"cache reaches plain relMSE at {eq['cache_time_to_plain_relmse_frac']} of the time; "
            f"mfma {eq['mfma']}")

    refs = None
    if not args.no_cpu_baseline:
        try:
            refs = cpu_reference_baselines(W, H, args.depth, host_cores()[0])
        except Exception as e:
            log(f"reference renderer unavailable: {e}")

    dl = None
    if not args.no_direct:
        dl = direct_lighting_bench(scene, W, H, args.depth, K,
                                   cpu_seconds=0.0 if args.no_cpu_baseline else min(args.cpu_seconds, 10.0))
        if refs:
            dl["cpu_reference"] = refs["direct"]
        log(f"direct lighting: {dl}")

    extras = None
    if not args.no_extras:
        try:
            extras = extras_bench(W, H, args.depth, threads=host_cores()[0],
                                  cpu=not args.no_cpu_baseline)
        except Exception as e:
            log(f"extras unavailable: {e}")

    bat = None
    if not args.no_batch:
        bat = batch_bench(args.batch_scenes, depth=args.depth)
        log(f"batch: {bat}")

    cpu = cpu_port = None
    if not args.no_cpu_baseline:
        if refs:
            cpu = refs["path"]
            log(f"cpu baseline (reference renderer): {cpu}")
        try:
            cpu_port = cpu_baseline(scene, W, H, args.depth, min(args.cpu_seconds, 10.0))
            log(f"cpu baseline (oracle port): {cpu_port}")
        except Exception as e:  # the GPU figure stands alone if the checker is absent
            log(f"cpu port unavailable: {e}")
        if cpu is None:
            cpu = cpu_port
        if cpu is not None:
            cpu["cpu_model"] = cpu_model()

    out = {
        "metric": "Msamples/s + relMSE at equal time: plain GPU PT vs radiance-cache PT, 1 MI355X",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": WORKLOAD, "width": W, "height": H, "depth": args.depth,
                   "triangles": scene.num_entities, "samples_per_step": W * H,
                   "rays_per_sample": {"extension": round(ext / samples, 3), "shadow": round(sh / samples, 3),
                                       "mis": round(mis / samples, 4)}},
        "roofline": roofline,
        "stages": stages,
        "cpu_baseline": cpu,
        "cpu_port": cpu_port,
        "relmse_equal_time": eq,
        "batch": bat,
        "cornell": cornell,
        "raycast": ray_cast,
        "direct_lighting": dl,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
