#!/usr/bin/env python3
"""Headline benchmark: Whitted frames on one MI355X, Mrays/s.

Workload (the headline frame): displaced UV sphere, 4,194,304 triangles,
3840x2160, 4 point lights, 16 materials, reflection depth 3, FP32.  A "step"
is one full frame (primary rays, one shadow ray per hit and light, reflection
rays up to depth 3, shading, quantisation) with the scene and its LBVH
resident in HBM and the RGBA8 output written to device memory.

value = (primary + shadow + secondary rays per frame) * steps / wall time, the
steps rendered by eos_render_frames (two frames in flight: a frame's primary
rays trace while the previous frame's shadow and reflection launches finish);
single_frame reports one eos_render call per frame (latency).

Also reported:
  roofline      dominant kernel (the primary-ray closest-hit launch): algorithmic
                bytes per launch = 128 B per 4-wide node visit (64 B per binary
                node visit) + 36 B per triangle test + 16 B hit record per ray, from the counting variant of the same
                traversal (deterministic, so equal to the timed launches), less the
                root visit and hit record of each primary ray the tile test resolved
                (the timed launch does neither for them), divided by the launch's
                mean duration (tile test included) from HIP events on its stream.
  cpu_baseline  the timing-faithful port of the reference's OpenMP CPU path on a
                stratified pixel sample of the same frame.
  aa            the config C frame with 2x2 anti-aliasing samples per pixel
                (renderWithAA), Mrays/s.
  glass         ray trees: transparent, transparent-and-reflective and mirror
                spheres (3 x 81,920 triangles) over a plane, 3840x2160, depth 3.
  instanced     config F: one 65,536-triangle mesh placed 64 times through
                per-triangle transformations (4,194,304 instanced triangles),
                3840x2160, depth 2, against the same triangles moved to world
                space beforehand (the cost of local-space tests).
  dense         the config C mesh scaled by 4 and seen from close by, so that
                most primary rays hit (at unit scale the reference's absolute
                determinant test leaves 89% of the triangles unhittable): Mrays/s
                and the primary launch's roofline.
  driver_scene  the reference driver's own scene and benchmark (src/main.cpp:
                330-510): objmodels/bunny.obj translated by (2, -2, -2), two
                lights, 1280x720; GPU frames and the reference's CPU path (port)
                on a stratified sample, the ratio the reference prints as its
                "GPU Speedup".
  host          the CPU the baselines ran on (lscpu fields, affinity and cgroup
                quota, threads used).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = {2: 64, 4: 128}   # binary node / 4-wide node
TRI_BYTES, HIT_BYTES = 36, 16


def _profile(workload: str) -> dict:
    """Per-launch figures measured by rocprofv3 --pmc runs of this bench
    (profiles/traffic_<workload>.json, committed with the summaries they come
    from): HBM bytes per launch and the L2 hit rate, keyed by kernel."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def host_cpu() -> dict:
    """lscpu's model, sockets, cores and SMT, the affinity mask and the cgroup
    CPU quota of this process, and the OpenMP threads the baselines use (the
    quota's share: OMP_NUM_THREADS when set, else the smaller of the affinity
    count and the quota)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = env or max(1, min(affinity, int(quota) if quota else affinity))
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "logical_cpus": info.get("CPU(s)"), "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
            "omp_threads_used": threads,
            "threads_source": "OMP_NUM_THREADS" if env else "min(affinity, cgroup quota)"}


def cpu_baseline(scene, stride: int, threads: int):
    """The reference's CPU path (port) on every stride-th pixel in x and y."""
    import numpy as np
    from oracle import pyoracle as po
    rp = po.RefPort(scene)
    W, H = scene.width, scene.height
    ys = np.arange(stride // 2, H, stride)
    xs = np.arange(stride // 2, W, stride)
    px = (ys[:, None] * W + xs[None, :]).ravel().astype(np.int32)
    _, c = rp.render(px, threads=threads)
    rays = c["primary"] + c["shadow"] + c["secondary"]
    return {
        "value": rays / c["seconds"] / 1e6, "unit": "Mrays/s", "cores": int(c["threads"]), "kind": "port",
        "sample": f"stratified 1/{stride} x 1/{stride} pixel grid: {len(px)} of {W * H} primary rays plus "
                  f"their {c['shadow']} shadow and {c['secondary']} reflection rays; render loop only "
                  f"(reference BVH build {rp.build_seconds:.1f} s not timed)",
        "seconds": c["seconds"],
    }


def cpu_same_algorithm(scene, stride: int, fast: bool = False):
    """The oracle's restatement of the GPU traversal (same LBVH, 4-wide walk,
    OpenMP over rows); fast: built with the reference's -O3 -ffast-math."""
    from oracle import pyoracle as po
    b = po.lbvh_build(scene.verts, 4)
    t = time.perf_counter()
    r = po.render_lbvh(scene, b, row_start=stride // 2, row_step=stride, width=4, fast=fast)
    dt = time.perf_counter() - t
    c = r["counts"]
    rays = c["primary"] + c["shadow"] + c["secondary"]
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "sample": f"every {stride}th row", "seconds": dt}


def hbm_copy_gbs(torch, mib: int = 2048, reps: int = 5) -> float:
    """Device-to-device copy bandwidth (bytes read + written per second, median
    of reps): the box's attainable HBM rate beside the published 8 TB/s."""
    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    rates = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        rates.append(2 * n / (a.elapsed_time(b) * 1e-3) / 1e9)
    del src, dst
    torch.cuda.empty_cache()
    return float(sorted(rates)[len(rates) // 2])


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_batched(E, torch, reps: int):
    """64 soups (config D): sum of sequential eos_render frames vs one batched,
    graph-replayed frame.  Both write RGBA8 to device memory; wall clock with the
    device synchronised around each measurement."""
    from eos_amd import autotune, scenes
    t0 = time.perf_counter()
    data = scenes.config_d()
    sc = [E.Scene(d, build=False) for d in data]
    _progress(f"batch: {len(sc)} scenes, {sum(d.n_tri for d in data)} triangles uploaded in "
              f"{time.perf_counter() - t0:.1f} s")
    # LBVHs: 64 eos_build_lbvh calls (each synchronous) vs one eos_build_lbvh_batch;
    # the second of each is timed (both rebuild the same trees in place)
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        single_gpu = sum(s.build(4)["build_ms"] for s in sc)
        t_single = time.perf_counter() - t
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        bst = E.build_batch(sc, leaf_max=4)
        t_batch = time.perf_counter() - t
    build = {"scenes": len(sc), "triangles": int(sum(d.n_tri for d in data)),
             "sequential_wall_ms": round(t_single * 1e3, 3), "sequential_gpu_ms": round(single_gpu, 3),
             "batched_wall_ms": round(t_batch * 1e3, 3), "batched_gpu_ms": round(bst[0]["build_ms"], 3),
             "speedup_wall": round(t_single / t_batch, 2)}
    _progress(f"batch LBVH build: {build['batched_wall_ms']} ms vs {build['sequential_wall_ms']} ms for 64 builds")
    tuned = autotune.load()
    source = "eos_amd/tuned_launch.json"
    if tuned is None:
        res = autotune.tune(sc, depth=2, reps=2, path=None, workload="config D",
                            grid={"workgroup": (256,), "stack_lds": (16,), "waves_per_cu": (16, 20, 32),
                                  "refill": (16,), "batch": (64, 128), "primary_packets": (0, 1)})
        tuned, source = res["config"], "tuned in this run"
    for s in sc:
        s.set_launch_config(**tuned)
    outs = [torch.empty(d.width * d.height * 4, dtype=torch.uint8, device="cuda:0") for d in data]
    ptrs = [o.data_ptr() for o in outs]

    def sequential():
        rays = 0
        for s, d, p in zip(sc, data, ptrs):
            st = s.render_device(d.camera, 2, p)
            rays += st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
        return rays

    rays = sequential()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        sequential()
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t) / reps
    _, st = E.render_batch(sc, depth=2, use_graph=True, device_ptrs=ptrs)
    assert st["rays"] == rays, (st["rays"], rays)
    torch.cuda.synchronize()
    t = time.perf_counter()
    gpu = []
    for _ in range(reps):
        gpu.append(E.render_batch(sc, depth=2, use_graph=True, device_ptrs=ptrs)[1]["render_ms"])
    torch.cuda.synchronize()
    t_bat = (time.perf_counter() - t) / reps
    # the batched frame walks every scene's tree with the same per-ray steps as
    # that scene's own frame, so its visits and tests are the sum of the scenes'
    # counting frames: the whole frame's algorithmic bytes (128 B per node
    # visit, 36 B per triangle test, 72 B per hit, 4 B per pixel) over the
    # batched frame's GPU time
    visits = tests = hits = pix = 0
    for s, d, p in zip(sc, data, ptrs):
        c = s.render_device(d.camera, 2, p, count_visits=True)
        visits += c["node_visits"]
        tests += c["tri_tests"]
        hits += c["hits"]
        pix += d.width * d.height
    alg = visits * NODE_BYTES[4] + tests * TRI_BYTES + hits * 72 + pix * 4
    gms = float(sum(gpu) / len(gpu))
    return {"scenes": len(sc), "rays_per_frame": int(rays), "sequential_ms": round(t_seq * 1e3, 3),
            "batched_ms": round(t_bat * 1e3, 3), "sequential_mrays": round(rays / t_seq / 1e6, 2),
            "batched_mrays": round(rays / t_bat / 1e6, 2), "speedup": round(t_seq / t_bat, 3),
            "launch_config": tuned, "launch_config_source": source, "hipgraph": True, "lbvh_build": build,
            "roofline": {"bound": "hbm", "scope": "whole batched frame", "alg_bytes": int(alg),
                         "node_visits": int(visits), "tri_tests": int(tests), "hits": int(hits),
                         "gpu_ms": round(gms, 4), "achieved": round(alg / (gms * 1e-3) / 1e9, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def _world_space(scenes, s):
    """The scene with its transformations applied to the vertices and normals
    (float64, rounded once): the same frame without per-triangle transforms."""
    import numpy as np
    v = s.verts.astype(np.float64)
    n = s.normals.astype(np.float64)
    for k, m in enumerate(s.xforms):
        M = m.reshape(4, 4).T.astype(np.float64)
        sel = s.xform_index == k
        v[sel] = v[sel] @ M[:3, :3].T + M[:3, 3]
        n[sel] = n[sel] @ np.linalg.inv(M[:3, :3])
    return scenes.SceneData(v.astype(np.float32), n.astype(np.float32), s.mat, s.materials, s.lights, s.camera,
                            s.depth, s.name + " (world space)")


def bench_instanced(E, torch, reps: int):
    """Config F frames with per-triangle transformations vs the same geometry in
    world space; frames written to device memory, GPU time from HIP events."""
    import numpy as np
    from eos_amd import scenes
    t0 = time.perf_counter()
    s = scenes.config_f()
    w = _world_space(scenes, s)
    _progress(f"instanced: {s.n_tri} triangles ({s.info['instances']} x {s.info['per_instance']} + plane) "
              f"generated in {time.perf_counter() - t0:.1f} s")
    out = torch.empty(s.width * s.height * 4, dtype=torch.uint8, device="cuda:0")
    res = {"workload": "config F: 64 instances x 65,536-tri displaced sphere via per-triangle transforms "
                       "+ 2,048-tri plane, 3840x2160, 2 lights, depth 2",
           "triangles": s.n_tri, "instances": s.info["instances"]}
    for key, sc in (("transformed", s), ("world_space", w)):
        g = E.Scene(sc, leaf_max=4)
        build_ms = g.build_stats["build_ms"]
        g.render_device(s.camera, s.depth, out.data_ptr())
        counted = g.render_device(s.camera, s.depth, out.data_ptr(), count_visits=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        per = [g.render_device(s.camera, s.depth, out.data_ptr()) for _ in range(reps)]
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / reps
        st = per[-1]
        rays = st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
        res[key] = {"mrays_per_s": round(rays / wall / 1e6, 2), "ms_per_frame": round(wall * 1e3, 3),
                    "gpu_render_ms": round(float(np.mean([p["render_ms"] for p in per])), 4),
                    "rays_per_frame": int(rays), "hits": int(st["hits"]), "lbvh_build_ms": round(build_ms, 3),
                    "node_visits": int(counted["node_visits"]), "tri_tests": int(counted["tri_tests"])}
        del g
    res["transform_cost"] = round(res["transformed"]["ms_per_frame"] / res["world_space"]["ms_per_frame"], 3)
    return res


def _frames(g, s, out, reps, **kw):
    """Mean wall time and stats of `reps` frames after one warm-up frame."""
    import torch
    g.render_device(s.camera, s.depth, out.data_ptr(), **kw)
    torch.cuda.synchronize()
    t = time.perf_counter()
    per = [g.render_device(s.camera, s.depth, out.data_ptr(), **kw) for _ in range(reps)]
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps, per[-1]


def bench_aa(E, torch, g, scene, reps: int):
    """The headline frame with 2x2 samples per pixel (four sample frames)."""
    out = torch.empty(scene.width * scene.height * 4, dtype=torch.uint8, device="cuda:0")
    wall, st = _frames(g, scene, out, reps, aa_samples=2)
    rays = st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
    return {"workload": "config C, 2x2 samples per pixel", "samples": 4, "ms_per_frame": round(wall * 1e3, 3),
            "mrays_per_s": round(rays / wall / 1e6, 2), "rays_per_frame": int(rays)}


def bench_glass(E, torch, reps: int):
    """Ray trees at 4K: three level-6 icospheres (glass, water-and-mirror, mirror)."""
    from eos_amd import scenes
    s = scenes.glass(3840, 2160, level=6, depth=3)
    g = E.Scene(s, leaf_max=4)
    out = torch.empty(s.width * s.height * 4, dtype=torch.uint8, device="cuda:0")
    wall, st = _frames(g, s, out, reps)
    rays = st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
    return {"workload": "glass spheres: 3 x 81,920 tris + plane, 3840x2160, 2 lights, depth 3, ray trees",
            "triangles": s.n_tri, "ms_per_frame": round(wall * 1e3, 3), "mrays_per_s": round(rays / wall / 1e6, 2),
            "rays_per_frame": int(rays), "secondary": int(st["secondary_rays"]), "hits": int(st["hits"]),
            "gpu_render_ms": round(st["render_ms"], 4)}


def _primary_roofline(counted, per, width_bytes, n_ids):
    """Algorithmic bytes of the primary-ray launch over its mean dispatch time."""
    import numpy as np
    t = float(np.mean([p["trace_first_ms"] for p in per]))
    culled = int(per[0]["primary_culled"])
    alg = (counted["closest_visits_first"] * width_bytes + counted["closest_tests_first"] * TRI_BYTES + n_ids * HIT_BYTES
           - culled * (width_bytes + HIT_BYTES))
    return {"bound": "hbm", "kernel": "k_trace<MODE_PRIMARY>", "launch_ms": round(t, 4), "alg_bytes_per_launch": int(alg),
            "primary_culled": culled,
            "node_visits": int(counted["closest_visits_first"]), "tri_tests": int(counted["closest_tests_first"]),
            "achieved": round(alg / (t * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def _with_traffic(roof: dict, prefix: str) -> dict:
    """Adds the measured HBM bytes and L2 hit rate of the launch (rocprofv3
    --pmc, profiles/traffic_config_c.json) to a roofline block."""
    prof = _profile("config_c")
    traffic = prof.get(prefix + roof["kernel"])
    t = roof["launch_ms"]
    roof.update({"traffic": traffic, "l2_hit_rate": prof.get(prefix + roof["kernel"] + "_l2_hit_rate"),
                 "traffic_source": prof.get("source"),
                 "hbm_achieved": round(traffic / (t * 1e-3) / 1e9, 2) if traffic else None,
                 "hbm_frac": round(traffic / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None})
    return roof


def bench_dense(E, torch, reps: int):
    """The config C mesh scaled by 4, camera close: a 4M-triangle frame of hits."""
    import numpy as np
    from eos_amd import scenes
    s = scenes.config_c_dense()
    g = E.Scene(s, leaf_max=4)
    W, H = s.width, s.height
    out = torch.empty(W * H * 4, dtype=torch.uint8, device="cuda:0")
    hit = torch.empty(W * H, dtype=torch.int32, device="cuda:0")
    counted = g.render_device(s.camera, s.depth, out.data_ptr(), hit_dev_ptr=hit.data_ptr(), count_visits=True)
    primary_hit_frac = float((hit >= 0).float().mean().item())
    wall, st = _frames(g, s, out, reps)
    per = [g.render_device(s.camera, s.depth, out.data_ptr(), launch_times=True) for _ in range(reps)]
    cams = np.stack([np.array(s.camera, copy=True)] * reps)
    g.render_frames(cams[:1], [out.data_ptr()], depth=s.depth)
    torch.cuda.synchronize()
    t = time.perf_counter()
    g.render_frames(cams, [out.data_ptr()] * reps, depth=s.depth)
    torch.cuda.synchronize()
    wall_seq = (time.perf_counter() - t) / reps
    seq = g.render_frames(cams, [out.data_ptr()] * reps, depth=s.depth, launch_times=True)
    rays = st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
    n_ids = ((W + 7) // 8) * ((H + 7) // 8) * 64
    hits0 = int(counted["hits"])
    return {"workload": "config C mesh scaled x4 (4,194,304 tris), close camera, 3840x2160, 4 lights, depth 3",
            "rays_per_frame": {"primary": st["primary_rays"], "shadow": st["shadow_rays"],
                               "secondary": st["secondary_rays"], "hits": st["hits"]},
            "ms_per_frame_in_flight": round(wall_seq * 1e3, 3), "mrays_per_s": round(rays / wall_seq / 1e6, 2),
            "ms_per_frame_single": round(wall * 1e3, 3), "mrays_per_s_single": round(rays / wall / 1e6, 2),
            "gpu_ms": {"render": round(float(np.mean([p["render_ms"] for p in per])), 4),
                       "trace_closest": round(float(np.mean([p["trace_closest_ms"] for p in per])), 4),
                       "trace_shadow": round(float(np.mean([p["trace_shadow_ms"] for p in per])), 4)},
            "hits_all_levels": hits0, "primary_hit_fraction": round(primary_hit_frac, 4),
            "roofline": _with_traffic(_primary_roofline(counted, per, NODE_BYTES[4], n_ids), "dense:"),
            "roofline_in_flight_launch_ms": round(float(np.mean([p["trace_first_ms"] for p in seq])), 4)}


def driver_scene_data():
    """The reference driver's scene from the committed bunny mesh, read back
    through the library's OBJ reader as TriangleMesh::loadFromFile reads it."""
    import tempfile
    import numpy as np
    from eos_amd import _lib as E
    from eos_amd import scenes
    m = np.load(os.path.join(ROOT, "tests", "golden", "bunny_mesh.npz"), allow_pickle=False)
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "bunny.obj")
        scenes.write_indexed_obj(path, m["positions"], m["faces"])
        v, n = E.load_obj(path)
    return scenes.driver_scene(v, n)


def bench_driver(E, torch, reps: int, cpu_threads: int, cpu_stride: int):
    """The reference driver's scene at 1280x720: GPU frames, and the reference's
    CPU path (port) on every cpu_stride-th pixel in x and y."""
    import numpy as np
    s = driver_scene_data()
    g = E.Scene(s, leaf_max=4)
    W, H = s.width, s.height
    out = torch.empty(W * H * 4, dtype=torch.uint8, device="cuda:0")
    wall, st = _frames(g, s, out, reps)
    cams = np.stack([np.array(s.camera, copy=True)] * reps)
    g.render_frames(cams[:1], [out.data_ptr()], depth=s.depth)
    torch.cuda.synchronize()
    t = time.perf_counter()
    g.render_frames(cams, [out.data_ptr()] * reps, depth=s.depth)
    torch.cuda.synchronize()
    wall_seq = (time.perf_counter() - t) / reps
    rays = st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
    res = {"workload": "reference driver scene: objmodels/bunny.obj (69,666 tris) translated (2,-2,-2), "
                       "2 lights, 1280x720, depth 4 (src/main.cpp:330-411)",
           "rays_per_frame": {"primary": st["primary_rays"], "shadow": st["shadow_rays"],
                              "secondary": st["secondary_rays"], "hits": st["hits"]},
           "ms_per_frame_single": round(wall * 1e3, 4), "mrays_per_s_single": round(rays / wall / 1e6, 2),
           "ms_per_frame_in_flight": round(wall_seq * 1e3, 4), "mrays_per_s": round(rays / wall_seq / 1e6, 2),
           "gpu_render_ms": round(st["render_ms"], 4)}
    if cpu_threads > 0:
        from oracle import pyoracle as po
        rp = po.RefPort(s)
        ys = np.arange(cpu_stride // 2, H, cpu_stride)
        xs = np.arange(cpu_stride // 2, W, cpu_stride)
        px = (ys[:, None] * W + xs[None, :]).ravel().astype(np.int32)
        _, c = rp.render(px, threads=cpu_threads)
        crays = c["primary"] + c["shadow"] + c["secondary"]
        cpu = crays / c["seconds"] / 1e6
        res["cpu_baseline"] = {"value": round(cpu, 4), "unit": "Mrays/s", "cores": int(c["threads"]), "kind": "port",
                               "sample": f"stratified 1/{cpu_stride} x 1/{cpu_stride} pixel grid: {len(px)} of "
                                         f"{W * H} primary rays; render loop only (reference BVH build "
                                         f"{rp.build_seconds:.2f} s not timed)",
                               "frame_seconds_extrapolated": round(c["seconds"] * W * H / len(px), 3)}
        # the reference's own figure: CPU frame time over GPU frame time (src/main.cpp:510)
        res["gpu_speedup_single_frame"] = round(res["cpu_baseline"]["frame_seconds_extrapolated"] / wall, 1)
    return res


def bench_adjoint(E, torch, reps: int, check: bool):
    """Config E: forward frame vs forward + hand-written adjoint (records replayed)."""
    import numpy as np
    from eos_amd import scenes
    s = scenes.config_e()
    g = E.Scene(s, leaf_max=4)
    W, H, depth = s.width, s.height, s.depth
    tm, tl = scenes.perturb_params(s, 1234)
    g.set_materials(tm)
    g.set_lights(tl)
    tgt = torch.empty((H, W, 3), dtype=torch.float32, device="cuda:0")
    g.render_device(s.camera, depth, None, rgb_dev_ptr=tgt.data_ptr())
    g.set_materials(s.materials)
    g.set_lights(s.lights)
    rgb = torch.empty_like(tgt)
    M, L = len(s.materials), len(s.lights)
    gm = torch.zeros((M, 7), dtype=torch.float64, device="cuda:0")
    gl = torch.zeros((L, 6), dtype=torch.float64, device="cuda:0")

    def fwd():
        return g.render_device(s.camera, depth, None, rgb_dev_ptr=rgb.data_ptr())

    def fwd_adj():
        st = g.render_device(s.camera, depth, None, rgb_dev_ptr=rgb.data_ptr(), keep_records=True)
        dl = 2.0 * (rgb - tgt)
        gm.zero_()
        gl.zero_()
        torch.cuda.synchronize()
        ad = g.adjoint_device(s.camera, depth, dl.data_ptr(), gm.data_ptr(), gl.data_ptr())
        return st, ad

    fwd()
    fwd_adj()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        st = fwd()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        _, ad = fwd_adj()
    torch.cuda.synchronize()
    t_fa = (time.perf_counter() - t) / reps
    # k_adjoint's algorithmic bytes: per hit record its 32-byte surface, one
    # visibility byte per light and the pixel's 12-byte upstream gradient
    rec_bytes = 32 + L + 12
    adj_alg = int(ad["records"]) * rec_bytes
    out = {"triangles": s.n_tri, "pixels": W * H, "materials": M, "lights": L, "depth": depth,
           "params": 7 * M + 6 * L, "forward_ms": round(t_f * 1e3, 3), "forward_adjoint_ms": round(t_fa * 1e3, 3),
           "ratio": round(t_fa / t_f, 3), "adjoint_kernel_ms": round(ad["adjoint_ms"], 4),
           "records": int(ad["records"]),
           "roofline": {"bound": "hbm", "kernel": "k_adjoint", "bytes_per_record": rec_bytes,
                        "alg_bytes": adj_alg, "kernel_ms": round(ad["adjoint_ms"], 4),
                        "achieved": round(adj_alg / (ad["adjoint_ms"] * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(adj_alg / (ad["adjoint_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "note": "integer LDS atomics (exact sums) over 55 parameters per record, not bytes, "
                                "bound this kernel"}}
    if check:
        # gradient over every 16th row vs the FP64 forward-mode oracle on the same rows
        from oracle import pyoracle as po
        step = 16
        mask = torch.zeros((H, 1, 1), device="cuda:0")
        mask[::step] = 1.0
        g.render_device(s.camera, depth, None, rgb_dev_ptr=rgb.data_ptr(), keep_records=True)
        dl = 2.0 * (rgb - tgt) * mask
        gm.zero_()
        gl.zero_()
        torch.cuda.synchronize()
        g.adjoint_device(s.camera, depth, dl.data_ptr(), gm.data_ptr(), gl.data_ptr())
        g_gpu = np.concatenate([gm.cpu().numpy().ravel(), gl.cpu().numpy().ravel()])
        orc = po.Oracle(s)
        rr = orc.render(row_start=0, row_step=step, want_recs=True)
        ref = po.grad_fwd64(rr["recs"], s.camera["position"], M, L, po.params_vector(s.materials, s.lights),
                            tgt.cpu().numpy())
        scale = np.abs(ref).max()
        err = np.abs(g_gpu - ref) / np.maximum(np.abs(ref), 1e-3 * scale)
        out["grad_check"] = {"rows": f"every {step}th row", "reference": "FP64 forward mode (oracle)",
                             "max_rel_err": float(err.max()), "median_rel_err": float(np.median(err))}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--leaf-max", type=int, default=4)
    ap.add_argument("--width", type=int, default=4, choices=(2, 4), help="BVH width the renderer walks")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-stride", type=int, default=32)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-instanced", action="store_true")
    ap.add_argument("--skip-aa", action="store_true")
    ap.add_argument("--skip-glass", action="store_true")
    ap.add_argument("--skip-dense", action="store_true")
    ap.add_argument("--skip-driver", action="store_true")
    ap.add_argument("--profile-copy-mib", type=int, default=0,
                    help="before the frames, run one device-to-device copy of this many MiB "
                         "(a kernel of known traffic for calibrating rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's full frame; --gpus {args.gpus} is not supported "
              "(multi-GPU frame sharding is a design note, not built)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from eos_amd import scenes
    from eos_amd import _lib as E

    t0 = time.perf_counter()
    scene = scenes.config_c()
    gen_s = time.perf_counter() - t0
    _progress(f"scene generated in {gen_s:.1f} s ({scene.n_tri} triangles)")
    dev = torch.device("cuda:0")
    torch.cuda.init()
    g = E.Scene(scene, leaf_max=args.leaf_max, width=args.width)
    build = g.build_stats
    W, H, depth = scene.width, scene.height, scene.depth
    out = torch.empty(W * H * 4, dtype=torch.uint8, device=dev)
    cam = scene.camera

    _progress(f"LBVH built in {build['build_ms']:.1f} ms")
    if args.profile_copy_mib > 0:
        n = args.profile_copy_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        _progress(f"calibration copy of {args.profile_copy_mib} MiB done")
    counted = g.render_device(cam, depth, out.data_ptr(), count_visits=True)
    _progress(f"counting frame done: {counted['render_ms']:.1f} ms")
    # one frame per call (latency, per-launch GPU times for the roofline)
    for _ in range(args.warmup):
        g.render_device(cam, depth, out.data_ptr())
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    per = []
    for _ in range(args.steps):
        per.append(g.render_device(cam, depth, out.data_ptr()))
    torch.cuda.synchronize()
    wall_single = time.perf_counter() - t_start
    _progress(f"{args.steps} single frames: {wall_single * 1e3 / args.steps:.2f} ms/frame")
    # the same frames with every launch timed (GPU time per launch kind; the
    # timing events between dependent launches add a few microseconds each, so
    # these frames are not the latency figure)
    per_t = [g.render_device(cam, depth, out.data_ptr(), launch_times=True) for _ in range(max(3, args.steps // 2))]
    torch.cuda.synchronize()
    # the timed steps: the same frames with two in flight (eos_render_frames)
    cams = np.stack([np.array(cam, copy=True)] * args.steps)
    if args.warmup:
        g.render_frames(cams[:args.warmup], [out.data_ptr()] * args.warmup, depth=depth)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    per_seq = g.render_frames(cams, [out.data_ptr()] * args.steps, depth=depth)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    _progress(f"{args.steps} timed frames, two in flight: {wall * 1e3 / args.steps:.2f} ms/frame")
    # the same sequence with each primary-ray launch timed (after the timed run)
    per_seq_t = g.render_frames(cams, [out.data_ptr()] * args.steps, depth=depth, launch_times=True)

    st = per[-1]
    rays = st["primary_rays"] + st["shadow_rays"] + st["secondary_rays"]
    assert all(p["primary_rays"] + p["shadow_rays"] + p["secondary_rays"] == rays for p in per_seq)
    value = rays * args.steps / wall / 1e6
    mean = lambda k: float(np.mean([p[k] for p in per]))  # noqa: E731
    t_first = float(np.mean([p["trace_first_ms"] for p in per_t]))
    n_ids = ((W + 7) // 8) * ((H + 7) // 8) * 64
    # primary rays the tile test resolved: the timed launch neither visits the
    # root for them nor writes their hit records (the counting frame walks them)
    culled = int(per_t[0]["primary_culled"])
    assert all(int(p["primary_culled"]) == culled for p in per_t + per_seq_t)
    alg_bytes = (counted["closest_visits_first"] * NODE_BYTES[build["width"]]
                 + counted["closest_tests_first"] * TRI_BYTES + n_ids * HIT_BYTES
                 - culled * (NODE_BYTES[build["width"]] + HIT_BYTES))
    achieved = alg_bytes / (t_first * 1e-3) / 1e9
    frame_bytes = (counted["node_visits"] * NODE_BYTES[build["width"]] + counted["tri_tests"] * TRI_BYTES
                   + st["hits"] * 72 + W * H * 4 - culled * NODE_BYTES[build["width"]])
    kernel = "k_trace<MODE_PRIMARY>"
    prof = _profile("config_c")
    traffic = prof.get(kernel)
    l2_hit = prof.get(f"{kernel}_l2_hit_rate")
    t_first_flight = float(np.mean([p["trace_first_ms"] for p in per_seq_t]))

    result = {
        "metric": "Mrays/s at 1 MI355X vs host OpenMP; batched and fwd+adjoint vs plain",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "config C: displaced sphere 4,194,304 tris, 3840x2160, 4 lights, depth 3",
                   "triangles": scene.n_tri, "width": W, "height": H, "lights": int(len(scene.lights)),
                   "depth": depth, "leaf_max": args.leaf_max, "bvh_width": build["width"],
                   "parallelism": "single GPU full frame"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel, "launch_ms": round(t_first, 4),
                     "launch_ms_source": "mean dispatch time of the primary launch over single frames with "
                                         "launch timing (HIP events on the scene's stream around that launch)",
                     # the same launch in the timed sequence, where it overlaps the other
                     # frame's shadow and reflection launches
                     "launch_ms_in_flight": round(t_first_flight, 4),
                     "frac_in_flight": round(alg_bytes / (t_first_flight * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "l2_hit_rate": l2_hit,
                     "traffic_source": prof.get("source"),
                     "alg_bytes_per_launch": int(alg_bytes),
                     "primary_culled": culled,
                     "node_visits": int(counted["closest_visits_first"]),
                     "tri_tests": int(counted["closest_tests_first"]),
                     # the measured HBM bytes of the launch at its measured time: node and
                     # triangle records are mostly re-read out of L2 and MALL, so the walk is
                     # bound by its dependent steps, not by HBM bandwidth
                     "hbm_achieved": round(traffic / (t_first * 1e-3) / 1e9, 2) if traffic else None,
                     "hbm_frac": round(traffic / (t_first * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None},
        "rays_per_frame": {"primary": st["primary_rays"], "shadow": st["shadow_rays"],
                           "secondary": st["secondary_rays"], "hits": st["hits"]},
        "gpu_ms": {"render": round(mean("render_ms"), 4),
                   "trace_closest": round(float(np.mean([p["trace_closest_ms"] for p in per_t])), 4),
                   "trace_shadow": round(float(np.mean([p["trace_shadow_ms"] for p in per_t])), 4),
                   "shade": round(float(np.mean([p["shade_ms"] for p in per_t])), 4),
                   "note": "render: frames timed only at their ends (the latency frames); the per-launch "
                           "sums come from frames with every launch timed"},
        "single_frame": {"ms_per_frame": round(wall_single / args.steps * 1e3, 4),
                         "mrays_per_s": round(rays * args.steps / wall_single / 1e6, 3),
                         "note": "one eos_render per frame (frame latency); value keeps two frames in flight"},
        "lbvh_build_ms": round(build["build_ms"], 3),
        "scene_gen_s": round(gen_s, 2),
        "frame_totals": {"node_visits": counted["node_visits"], "tri_tests": counted["tri_tests"]},
        # whole-frame algorithmic bytes: node visits and triangle tests of every
        # ray, 72 B of shading data per hit, the RGBA8 frame
        "frame_roofline": {"alg_bytes": int(frame_bytes), "render_ms": round(mean("render_ms"), 4),
                           "achieved": round(frame_bytes / (mean("render_ms") * 1e-3) / 1e9, 2),
                           "unit": "GB/s", "frac": round(frame_bytes / (mean("render_ms") * 1e-3) / 1e9
                                                         / HBM_PEAK_GBS, 4)},
    }
    try:
        result["device"] = E.device_name()
    except E.EosError:
        pass
    copy = hbm_copy_gbs(torch)
    result["roofline"]["copy_gbs"] = round(copy, 1)
    result["roofline"]["copy_note"] = ("measured device-to-device copy of 2 GiB (read + write bytes per second, "
                                       "torch's copy kernel): the attainable HBM rate beside the published peak")
    _progress(f"HBM copy: {copy:.0f} GB/s")
    if not args.skip_batch:
        result["batched"] = bench_batched(E, torch, reps=3)
        _progress(f"batched: {result['batched']['speedup']}x over sequential")
    if not args.skip_adjoint:
        result["adjoint"] = bench_adjoint(E, torch, reps=3, check=not args.no_cpu)
        _progress(f"adjoint: fwd+adj / fwd = {result['adjoint']['ratio']}")
    if not args.skip_aa:
        result["aa"] = bench_aa(E, torch, g, scene, reps=3)
        _progress(f"aa 2x2: {result['aa']['mrays_per_s']} Mrays/s, {result['aa']['ms_per_frame']} ms")
    if not args.skip_glass:
        result["glass"] = bench_glass(E, torch, reps=3)
        _progress(f"glass: {result['glass']['mrays_per_s']} Mrays/s, {result['glass']['ms_per_frame']} ms")
    if not args.skip_dense:
        result["dense"] = bench_dense(E, torch, reps=3)
        _progress(f"dense: {result['dense']['mrays_per_s']} Mrays/s, primary hits "
                  f"{result['dense']['primary_hit_fraction']:.3f}, roofline {result['dense']['roofline']['frac']}")
    host = host_cpu()
    threads = host["omp_threads_used"]
    if not args.skip_driver:
        result["driver_scene"] = bench_driver(E, torch, reps=40, cpu_threads=0 if args.no_cpu else threads,
                                              cpu_stride=4)
        _progress(f"driver scene: {result['driver_scene']['mrays_per_s']} Mrays/s, GPU speedup "
                  f"{result['driver_scene'].get('gpu_speedup_single_frame')}")
    if not args.skip_instanced:
        result["instanced"] = bench_instanced(E, torch, reps=5)
        _progress(f"instanced: {result['instanced']['transformed']['mrays_per_s']} Mrays/s, "
                  f"{result['instanced']['transform_cost']}x the world-space frame")
    result["host"] = host
    if not args.no_cpu:
        _progress(f"CPU baseline (reference port, {threads} threads)")
        cb = cpu_baseline(scene, args.cpu_stride, threads)
        _progress(f"CPU baseline: {cb['value']:.6f} Mrays/s in {cb['seconds']:.1f} s")
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        result["gpu_vs_cpu_baseline"] = round(value / cb["value"], 1)
        sa = cpu_same_algorithm(scene, 64)
        sf = cpu_same_algorithm(scene, 64, fast=True)
        _progress(f"CPU same-algorithm: {sa['value']:.3f} Mrays/s (-O2 parity build), {sf['value']:.3f} (-O3 -ffast-math)")
        result["cpu_same_algorithm"] = {"value": round(sf["value"], 4), "unit": "Mrays/s",
                                        "sample": sf["sample"], "cores": threads,
                                        "note": "oracle restatement of the GPU's 4-wide LBVH walk and shading, "
                                                "built -O3 -ffast-math (the reference's timing flags), OpenMP",
                                        "parity_build_value": round(sa["value"], 4),
                                        "gpu_ratio": round(value / sf["value"], 1)}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
