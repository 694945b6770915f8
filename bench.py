"""Benchmark: Mrays/s (trace+sort+reorder) of the wavefront bounce loop.

Workload (one GPU): the medium interior scene (263,250 triangles), 1920x1080,
1 spp, 8 bounces, Russian roulette after bounce 3, 60-bit keys (30 origin
bits + 30 octahedral direction bits) sorted as u64 keys with a u32 payload,
8-bit digits.  One untimed frame of the C++ frame loop captures every
bounce's path-ray wavefront (rays + slot-to-pixel LUT) in HBM.  A timed step
is one pass of the hot path over all captured bounces of that frame:
    keys -> sort -> reorder -> closest-hit trace          (Tracer::traceSort)
value = rays through the step / wall time of the timed steps, the step
captured once in a HIP graph and replayed (the eager-launch rate is
reported beside it).

Also reported: per-stage GPU times (hipEvents on the stream the kernels run
on), the roofline of every stage and of the dominant one, the unsorted-trace
rate of the same wavefronts (the reference's sort ablation, Renderer.cpp:
280-283), the reference-style full-frame figures of the C++ frame loop, and a
CPU baseline (the oracle's port of the same stages on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def d2d(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device copy out of renderer-owned buffers (hipMemcpy)."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--env", default="config1.json")
    ap.add_argument("--digit-bits", type=int, default=8)
    ap.add_argument("--frames", type=int, default=3, help="reference-style full frames timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable workload")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-large", action="store_true", help="skip the large-scene reorder ablation")
    ap.add_argument("--no-fuse", action="store_true",
                    help="reorder and trace as separate kernels (wf_reorder + wf_trace) instead of "
                         "the fused wf_reorder_trace")
    ap.add_argument("--no-ablations", action="store_true",
                    help="skip the unsorted-trace and key-layout ablations of the main workload "
                         "(profiling runs: every trace launch is then one of the measured steps)")
    ap.add_argument("--batch-views", type=int, default=64)
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="wf_set_tuning knob applied before the scene is built (repeatable)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (scales HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the wavefront loop runs on a single device and has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from vulkanwavefrontpathtracer_amd import stages
    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    for kv in args.tune:
        k, v = kv.split("=")
        stages.call("wf_set_tuning", k.encode(), int(v))
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nb} bytes read + {nb} bytes written")
        del src, dst
    env = env_path(args.env)
    cfg = json.loads(env.read_text())
    log(f"building scene for {args.env}")
    r = Renderer(env, {"Renderer.digitBits": args.digit_bits})
    r.build_scene()
    scene_h = r.scene_handle()
    r.set_capture(True)
    r.render()                       # untimed: captures the bounce wavefronts
    frame_stats = r.stats()
    mode = KEY_MODES[cfg["Renderer"]["keyMode"]]
    kbytes = 4 if mode < 2 else 8
    kbits = stages.key_bits(mode)
    smin, smax = (C.c_float * 3)(), (C.c_float * 3)()
    stages.call("wf_scene_bounds", scene_h, smin, smax)

    # Resident per-bounce buffers (captured rays are copied so the renderer can be reused).
    bounces = []
    b = 0
    while True:
        rp, sp, n = r.captured(b)
        if n == 0:
            break
        rays = torch.empty((n, 8), dtype=torch.float32, device=dev)
        s2i = torch.empty(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        d2d(rays.data_ptr(), rp, 32 * n)
        d2d(s2i.data_ptr(), sp, 4 * n)
        bounces.append(dict(
            n=n, rays=rays, s2i=s2i,
            keys=torch.empty(n, dtype=torch.int64 if kbytes == 8 else torch.int32, device=dev),
            idx=torch.empty(n, dtype=torch.int32, device=dev),
            out_rays=torch.empty_like(rays), out_s2i=torch.empty_like(s2i),
            out_i2s=torch.empty(r.size[0] * r.size[1], dtype=torch.int32, device=dev),
            hits=torch.empty((n, 4), dtype=torch.int32, device=dev),
            hits_u=torch.empty((n, 4), dtype=torch.int32, device=dev),
            sorter=stages.Sorter()))
        b += 1
    total_rays = sum(x["n"] for x in bounces)
    log(f"captured {len(bounces)} bounces, {total_rays} path rays: {[x['n'] for x in bounces]}")

    class SceneRef:
        handle = scene_h

    sc = SceneRef()
    stream = torch.cuda.current_stream()
    s_ptr = C.c_void_p(stream.cuda_stream)
    lib = stages._capi.load()
    mn3, mx3 = (C.c_float * 3)(*smin), (C.c_float * 3)(*smax)
    # pre-size the sort scratch
    for x in bounces:
        need = lib.wf_sort_scratch_bytes(x["n"], kbytes, kbits, args.digit_bits)
        x["sorter"].scratch = torch.empty(need, dtype=torch.uint8, device=dev)

    names = ["keys", "sort", "reorder", "trace"]
    ev = {k: [] for k in names}

    fused = not args.no_fuse

    def step(record=False, sp=None, key_mode=None):
        s_ptr = sp if sp is not None else C.c_void_p(stream.cuda_stream)
        km = mode if key_mode is None else key_mode
        for x in bounces:
            n = x["n"]
            marks = []
            if record:
                marks.append(torch.cuda.Event(enable_timing=True))
                marks[-1].record(stream)
            stages._capi.check("wf_morton_keys", lib.wf_morton_keys(
                C.c_void_p(x["rays"].data_ptr()), C.c_void_p(x["keys"].data_ptr()),
                C.c_void_p(x["idx"].data_ptr()), n, km, mn3, mx3, s_ptr, None))
            if record:
                marks.append(torch.cuda.Event(enable_timing=True))
                marks[-1].record(stream)
            stages._capi.check("wf_sort_pairs", lib.wf_sort_pairs(
                C.c_void_p(x["keys"].data_ptr()), C.c_void_p(x["idx"].data_ptr()), n, kbytes, kbits,
                args.digit_bits, C.c_void_p(x["sorter"].scratch.data_ptr()),
                x["sorter"].scratch.numel(), s_ptr, None))
            if record:
                marks.append(torch.cuda.Event(enable_timing=True))
                marks[-1].record(stream)
            if not fused:
                stages._capi.check("wf_reorder", lib.wf_reorder(
                    C.c_void_p(x["idx"].data_ptr()), C.c_void_p(x["rays"].data_ptr()),
                    C.c_void_p(x["out_rays"].data_ptr()), C.c_void_p(x["s2i"].data_ptr()),
                    C.c_void_p(x["out_s2i"].data_ptr()), C.c_void_p(x["out_i2s"].data_ptr()), n,
                    s_ptr, None))
            if record:
                marks.append(torch.cuda.Event(enable_timing=True))
                marks[-1].record(stream)
            if fused:
                # reorder + trace in one kernel: the same outputs as the two calls
                stages._capi.check("wf_reorder_trace", lib.wf_reorder_trace(
                    sc.handle, C.c_void_p(x["idx"].data_ptr()), C.c_void_p(x["rays"].data_ptr()),
                    C.c_void_p(x["out_rays"].data_ptr()), C.c_void_p(x["s2i"].data_ptr()),
                    C.c_void_p(x["out_s2i"].data_ptr()), C.c_void_p(x["out_i2s"].data_ptr()),
                    C.c_void_p(x["hits"].data_ptr()), n, 1, s_ptr, None))
            else:
                stages._capi.check("wf_trace", lib.wf_trace(
                    sc.handle, C.c_void_p(x["out_rays"].data_ptr()), None,
                    C.c_void_p(x["hits"].data_ptr()), n, 1, s_ptr, None))
            if record:
                marks.append(torch.cuda.Event(enable_timing=True))
                marks[-1].record(stream)
                for k, name in enumerate(names):
                    ev[name].append((marks[k], marks[k + 1], n))

    log(f"warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} eager steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall_eager = time.perf_counter() - t0
    value_eager = total_rays * args.steps / wall_eager / 1e6
    log(f"eager: {value_eager:.1f} Mrays/s ({wall_eager * 1e3 / args.steps:.3f} ms/step)")

    # The same step captured once in a HIP graph (every stage call is
    # stream-ordered with no host synchronisation) and replayed.
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step(sp=C.c_void_p(torch.cuda.current_stream().cuda_stream))
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    log(f"timing {args.steps} graph-replayed steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = total_rays * args.steps / wall / 1e6
    log(f"graph: {value:.1f} Mrays/s ({ms_per_step:.3f} ms/step)")
    # the graph's last replay must equal an eager step's hits
    graph_hits = [x["hits"].clone() for x in bounces]
    step()
    torch.cuda.synchronize()
    graph_equal = all(torch.equal(h, x["hits"]) for h, x in zip(graph_hits, bounces))
    del graph

    # Per-stage times on the kernels' stream (separate, event-bracketed steps).
    for _ in range(3):
        step(record=True)
    torch.cuda.synchronize()
    stage_ms = {k: sum(a.elapsed_time(b_) for a, b_, _ in ev[k]) / 3 for k in names}
    # Algorithmic bytes per element: keys 32+K+4, sort P*2*(K+4)+K,
    # reorder 80, trace 32 read + 16 write.
    passes = -(-kbits // args.digit_bits)
    # Fused reorder + trace: 4 (index) + 32 (ray gathered once) + 32 (reordered
    # ray) + 4 + 4 + 4 (LUTs) + 16 (hit) = 96 B, all under "trace".
    bpe = {"keys": 32 + kbytes + 4, "sort": passes * 2 * (kbytes + 4) + kbytes, "reorder": 80,
           "trace": 96 if fused else 48}
    stages_roof = {}
    for k in names:
        t = stage_ms[k] / 1e3
        if fused and k == "reorder":
            stages_roof[k] = {"ms_per_frame": 0.0, "fused_into": "trace"}
            continue
        gbs = total_rays * bpe[k] / t / 1e9 if t > 0 else 0.0
        stages_roof[k] = {"ms_per_frame": round(stage_ms[k], 4), "bytes_per_ray": bpe[k],
                          "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if fused:
        stages_roof["trace"]["fused"] = "reorder + trace (wf_reorder_trace)"
    dominant = max(names, key=lambda k: stage_ms[k])
    dom = stages_roof[dominant]
    # per-launch average duration of the dominant kernel at the largest bounce
    big = max(range(len(bounces)), key=lambda i: bounces[i]["n"])
    recs = [e for e in ev[dominant]]
    launch_ms = [a.elapsed_time(b_) for (a, b_, n) in recs if n == bounces[big]["n"]]
    avg_launch = sum(launch_ms) / len(launch_ms)
    launch_bytes = bounces[big]["n"] * bpe[dominant]
    achieved = launch_bytes / (avg_launch / 1e3) / 1e9

    traffic = measured_traffic(dominant, bounces[big]["n"])

    # Unsorted trace of the same wavefronts (sort ablation).
    ts = []
    for _ in range(0 if args.no_ablations else 3):
        for x in bounces:
            a = torch.cuda.Event(enable_timing=True)
            b_ = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            stages._capi.check("wf_trace", lib.wf_trace(sc.handle, C.c_void_p(x["rays"].data_ptr()), None,
                                                  C.c_void_p(x["hits_u"].data_ptr()), x["n"], 1,
                                                  s_ptr, None))
            b_.record(stream)
            ts.append((a, b_))
    torch.cuda.synchronize()
    unsorted_ms = sum(a.elapsed_time(b_) for a, b_ in ts) / 3 if ts else None
    unsorted_mrays = total_rays / (unsorted_ms / 1e3) / 1e6 if ts else None
    sorted_ms = sum(stage_ms.values())
    sorted_mrays = total_rays / (sorted_ms / 1e3) / 1e6
    trace_only_sorted = total_rays / (stage_ms["trace"] / 1e3) / 1e6
    # parity of the sorted trace with the unsorted one (same hits per ray)
    same = all(torch.equal(x["hits"][torch.argsort(x["out_s2i"])], x["hits_u"][torch.argsort(x["s2i"])])
               for x in bounces) if ts else None

    # Key layout ablation: the same 60 origin + direction bits concatenated
    # (origin30 << 30 | oct30) or interleaved as a 5-D Morton code.
    key_layout = None
    if mode in (2, 4) and not args.no_ablations:
        key_layout = {}
        for km, name in ((2, "origin30_oct30"), (4, "origin_dir_morton60")):
            for k in names:
                ev[k].clear()
            for _ in range(3):
                step(record=True, key_mode=km)
            torch.cuda.synchronize()
            kms = {k: sum(a.elapsed_time(b_) for a, b_, _ in ev[k]) / 3 for k in names}
            key_layout[name] = {
                "stage_ms": {k: round(v, 4) for k, v in kms.items()},
                "mrays_s": round(total_rays / (sum(kms.values()) / 1e3) / 1e6, 1),
                "trace_speedup_vs_unsorted": round(unsorted_ms / kms["trace"], 3)}

    # Reference-style full frames (Benchmark::run): path Mrays/s includes shade.
    frame_walls, fstats = [], []
    for _ in range(args.frames):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r.render()
        frame_walls.append(time.perf_counter() - t1)
        fstats.append(r.stats())
    ref_style = None
    if fstats:
        fs = fstats[-1]
        path_rays = fs["path_rays"]
        ref_style = {
            "path_rays": path_rays,
            "path_mrays_s": round(path_rays / (fs["path_ms"] / 1e3) / 1e6, 1) if fs["path_ms"] > 0 else None,
            "frame_ms_wall": round(1e3 * sum(frame_walls) / len(frame_walls), 2),
            "rays_per_frame": fs["primary_rays"] + fs["shadow_rays"] + path_rays,
            "stage_ms": {k: round(v, 3) for k, v in fs["stage_ms"].items()},
        }

    large = None if args.no_large else large_section(dev)
    adjoint = None if args.no_adjoint else adjoint_section(dev)
    batched = None if args.no_batch else batch_section(dev, args.batch_views)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(bounces[0], scene_h, smin, smax, mode, kbits, args.cpu_sample, r)

    out = {
        "metric": "Mrays/s (trace+sort+reorder), plain vs batched & adjoint, at RMSE≤1e-3 vs FP64",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural interior scene, captured bounce wavefronts)",
        "config": {
            "workload": "medium plain: interior 263,250 tris, 1920x1080, 1 spp, 8 bounces, RR after "
                        "bounce 3, 60-bit keys (30 origin + 30 octahedral bits, " + cfg["Renderer"]["keyMode"]
                        + "), u64 key sort + u32 payload",
            "rays_per_step": total_rays, "bounces": len(bounces),
            "rays_per_bounce": [x["n"] for x in bounces], "digit_bits": args.digit_bits,
            "parallelism": "1 GPU",
        },
        "tuning": tuning_state(),
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "launch_ms": round(avg_launch, 4),
                     "launch_elems": bounces[big]["n"], "bytes_per_elem": bpe[dominant]},
        "stages": stages_roof,
        "value_eager_launches": round(value_eager, 2),
        "graph_replay_equals_eager": bool(graph_equal),
        "sorted_pipeline_mrays_s": round(sorted_mrays, 1),
        "trace_only_sorted_mrays_s": round(trace_only_sorted, 1),
        "unsorted_trace_mrays_s": round(unsorted_mrays, 1) if ts else None,
        "sort_speedup_trace": round(unsorted_ms / stage_ms["trace"], 3) if ts else None,
        "sorted_vs_unsorted_hits_equal": same,
        "key_layout_ablation": key_layout,
        "reference_style_frame": ref_style,
        "untimed_capture_frame": {"path_rays": frame_stats["path_rays"]},
        "large_ablation": large,
        "adjoint": adjoint,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def adjoint_section(dev, reps=3):
    """Differentiable workload: the interior scene with 512 material groups,
    512x512, 16 spp, 6 bounces, L2 loss against a target rendered from
    parameters perturbed by N(0, 0.05).  Forward = the C++ frame loop (frame
    1); adjoint = the path-replay kernel.  Reports both GPU times, their
    ratio and the adjoint's replayed rays per second."""
    import numpy as np
    import torch

    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    log("adjoint workload: building the 512-material interior scene")
    r = Renderer(env_path("config4.json"))
    r.build_scene()
    a0, e0 = r.materials()
    rng = np.random.default_rng(5)
    at = np.clip(a0 + rng.normal(0, 0.05, a0.shape), 0, 1).astype(np.float32)
    et = np.clip(e0 + rng.normal(0, 0.05, e0.shape), 0, None).astype(np.float32)
    r.set_materials(at, et)
    r.reset_frame_index()
    r.render()
    target = r.pixels_tensor()
    r.set_materials(a0, e0)
    fwd_ms, fwd_wall, adj_ms, adj_wall = [], [], [], []
    for _ in range(reps):
        r.reset_frame_index()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd_ms.append(r.render())
        torch.cuda.synchronize()
        fwd_wall.append(time.perf_counter() - t0)
        img = r.pixels_tensor()
        d = 2.0 * (img - target)
        d[..., 3] = 0
        t0 = time.perf_counter()
        ga, ge = r.adjoint(d, frame=1)
        torch.cuda.synchronize()
        adj_wall.append(time.perf_counter() - t0)
        adj_ms.append(r.last_adjoint_ms)
    st = r.stats()
    rays = st["primary_rays"] + st["shadow_rays"] + st["path_rays"]
    loss = float(((img[..., :3] - target[..., :3]).double() ** 2).sum())
    f_ms = min(fwd_wall) * 1e3
    a_ms = min(adj_ms)
    return {"workload": "interior 263,250 tris, 512 material groups, 512x512, 16 spp, 6 bounces, "
                        "L2 loss vs perturbed-parameter target",
            "forward_ms_wall": round(f_ms, 3), "forward_stage_gpu_ms": round(min(fwd_ms), 3),
            "adjoint_ms": round(a_ms, 3), "adjoint_ms_wall": round(min(adj_wall) * 1e3, 3),
            "adjoint_over_forward": round(a_ms / f_ms, 3),
            "rays_per_frame": rays, "forward_mrays_s": round(rays / (f_ms / 1e3) / 1e6, 1),
            "adjoint_mrays_s": round(rays / (a_ms / 1e3) / 1e6, 1), "loss": loss,
            "grad_norm": float(torch.sqrt((ga.double() ** 2).sum() + (ge.double() ** 2).sum()))}


def large_section(dev, reps=3):
    """Large scene reorder ablation: the city (~2.8M triangles), 1920x1080,
    4 spp, 8 bounces.  For each key (30-bit origin keys in u32, 60-bit
    origin + octahedral keys in u64, concatenated or interleaved as a 5-D
    Morton code) and sort period (every bounce, every 2nd,
    never) a frame of the C++ frame loop with those settings captures every
    pass's path wavefronts; the 4 passes of each bounce are concatenated into
    one wavefront (8.3M rays, all samples in flight) and the timed step runs
    keys -> sort -> reorder + closest-hit trace (fused, wf_reorder_trace) on
    the bounces the period sorts and the trace alone on the others."""
    import torch

    from vulkanwavefrontpathtracer_amd import stages
    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    env = env_path("config2.json")
    cfg = json.loads(env.read_text())
    depth, spp = cfg["Renderer"]["recursionDepth"], cfg["Renderer"]["numberOfPrimarySamples"]
    log("large ablation: building the city scene")
    base = Renderer(env)
    base.build_scene()

    class SceneRef:
        handle = base.scene_handle()

    smin, smax = (C.c_float * 3)(), (C.c_float * 3)()
    stages.call("wf_scene_bounds", SceneRef.handle, smin, smax)
    lib = stages._capi.load()
    stream = torch.cuda.current_stream()
    s_ptr = C.c_void_p(stream.cuda_stream)
    modes = [("32-bit", "origin30", 1, 4, 30), ("60-bit", "origin30_oct30", 2, 8, 60),
             ("60-bit 5-D", "origin_dir_morton60", 4, 8, 60)]
    rows, sort_rates = [], {}
    for name, key_mode, mode, kbytes, kbits in modes:
        for period in (1, 2, 0):
            r = Renderer(env, {"Renderer.keyMode": key_mode, "Renderer.reorderPeriod": period,
                               "Renderer.sortPathRays": period != 0,
                               "Renderer.reorderPathRays": period != 0})
            r.set_scene(SceneRef, list(smin), list(smax))
            r.set_capture(True)
            r.render()
            waves = []
            for b in range(depth):
                parts = [r.captured(p * depth + b) for p in range(spp)]
                n = sum(x[2] for x in parts)
                if n == 0:
                    break
                rays = torch.empty((n, 8), dtype=torch.float32, device=dev)
                s2i = torch.empty(n, dtype=torch.int32, device=dev)
                o = 0
                torch.cuda.synchronize()
                for rp, sp, k in parts:
                    if k:
                        d2d(rays[o:].data_ptr(), rp, 32 * k)
                        d2d(s2i[o:].data_ptr(), sp, 4 * k)
                        o += k
                waves.append(dict(n=n, rays=rays, s2i=s2i))
            del r
            nmax = max(w["n"] for w in waves)
            keys = torch.empty(nmax, dtype=torch.int64 if kbytes == 8 else torch.int32, device=dev)
            idx = torch.empty(nmax, dtype=torch.int32, device=dev)
            out_rays = torch.empty((nmax, 8), dtype=torch.float32, device=dev)
            out_s2i = torch.empty(nmax, dtype=torch.int32, device=dev)
            out_i2s = torch.empty(1920 * 1080, dtype=torch.int32, device=dev)
            hits = torch.empty((nmax, 4), dtype=torch.int32, device=dev)
            best = None
            for digit_bits in ((11, 8) if period == 1 else (sort_rates[name]["digit_bits"],)):
                need = lib.wf_sort_scratch_bytes(nmax, kbytes, kbits, digit_bits)
                scratch = torch.empty(need, dtype=torch.uint8, device=dev)
                ms = {"keys": 0.0, "sort": 0.0, "reorder": 0.0, "trace": 0.0}
                sort_big = []
                for rep in range(reps + 1):
                    evs = []
                    for b, w in enumerate(waves):
                        n = w["n"]
                        srt = period > 0 and b % period == 0
                        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
                        ev[0].record(stream)
                        if srt:
                            stages._capi.check("wf_morton_keys", lib.wf_morton_keys(
                                C.c_void_p(w["rays"].data_ptr()), C.c_void_p(keys.data_ptr()),
                                C.c_void_p(idx.data_ptr()), n, mode, smin, smax, s_ptr, None))
                        ev[1].record(stream)
                        if srt:
                            stages._capi.check("wf_sort_pairs", lib.wf_sort_pairs(
                                C.c_void_p(keys.data_ptr()), C.c_void_p(idx.data_ptr()), n, kbytes,
                                kbits, digit_bits, C.c_void_p(scratch.data_ptr()), need, s_ptr, None))
                        ev[2].record(stream)
                        ev[3].record(stream)
                        if srt:
                            # reorder gather fused into the trace (its time counts as trace)
                            stages._capi.check("wf_reorder_trace", lib.wf_reorder_trace(
                                SceneRef.handle, C.c_void_p(idx.data_ptr()),
                                C.c_void_p(w["rays"].data_ptr()), C.c_void_p(out_rays.data_ptr()),
                                C.c_void_p(w["s2i"].data_ptr()), C.c_void_p(out_s2i.data_ptr()),
                                C.c_void_p(out_i2s.data_ptr()), C.c_void_p(hits.data_ptr()), n, 1,
                                s_ptr, None))
                        else:
                            stages._capi.check("wf_trace", lib.wf_trace(
                                SceneRef.handle, C.c_void_p(w["rays"].data_ptr()), None,
                                C.c_void_p(hits.data_ptr()), n, 1, s_ptr, None))
                        ev[4].record(stream)
                        evs.append((ev, n, srt))
                    torch.cuda.synchronize()
                    if rep == 0:
                        continue                     # warm-up
                    for ev, n, srt in evs:
                        for k, nm in enumerate(["keys", "sort", "reorder", "trace"]):
                            ms[nm] += ev[k].elapsed_time(ev[k + 1]) / reps
                        if srt and n == nmax:
                            sort_big.append(ev[1].elapsed_time(ev[2]))
                total_rays = sum(w["n"] for w in waves)
                t = sum(ms.values())
                res = {"keys": name, "period": period, "digit_bits": digit_bits,
                       "mrays_s": round(total_rays / (t / 1e3) / 1e6, 1),
                       "ms_per_frame": {k: round(v, 3) for k, v in ms.items()},
                       "rays": total_rays, "rays_per_bounce": [w["n"] for w in waves]}
                if sort_big:
                    passes = -(-kbits // digit_bits)
                    bpe = passes * 2 * (kbytes + 4) + kbytes
                    gbs = nmax * bpe / (min(sort_big) / 1e3) / 1e9
                    res["sort_GBs"] = round(gbs, 1)
                    res["sort_frac_hbm"] = round(gbs / HBM_PEAK_GBS, 4)
                if best is None or res["mrays_s"] > best["mrays_s"]:
                    best = res
                del scratch
            if period == 1:
                sort_rates[name] = best
            rows.append(best)
            log(f"large ablation {name} period {period}: {best['mrays_s']} Mrays/s")
    out = {"workload": "city ~2.8M tris, 1920x1080, 4 spp in flight (8.3M rays/bounce), 8 bounces",
           "rows": rows}
    for name, *_ in modes:
        every = next(x for x in rows if x["keys"] == name and x["period"] == 1)
        never = next(x for x in rows if x["keys"] == name and x["period"] == 0)
        out[f"speedup_every_vs_never_{name}"] = round(every["mrays_s"] / never["mrays_s"], 3)
    return out


def batch_section(dev, num_views=64, reps=2):
    """Batched workload: num_views scenes from the boxes generator (triangles
    log-uniform in [1k, 500k]), one view each with resolution in
    {128, 256, 512, 1024}^2 and spp in 1..16, 6 bounces, 2 shadow samples.
    The batch (one queue, scene-prefixed keys, graph-replayed passes,
    autotuned) against the same views rendered one after another by the
    single-view frame loop (one renderer per view, buffers allocated by an
    untimed warm-up frame)."""
    import torch

    from vulkanwavefrontpathtracer_amd import stages
    from vulkanwavefrontpathtracer_amd.batch import Batch, boxes_env_view, workload
    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    boxes, seeds, vspec = workload(num_views=num_views)
    log(f"batched workload: building {num_views} scenes ({sum(boxes) * 192:,} triangles)")
    hosts = [stages.HostScene(0, sd, num_boxes=nb) for nb, sd in zip(boxes, seeds)]
    scenes = [stages.Scene.from_host(h, dev) for h in hosts]
    env = env_path("config0.json")
    views = [boxes_env_view(env, v["scene"], v["width"], v["height"], v["spp"]) for v in vspec]
    depth, S = 6, 2
    b = Batch(scenes, views, depth=depth, shadow_samples=S)
    log("batched workload: autotuning")
    t0 = time.perf_counter()
    best_pass_ms = b.autotune()
    tune_s = time.perf_counter() - t0
    tuning = b.tuning()
    b.render()                                   # graph capture + warm-up
    walls = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gpu_ms = b.render()
        walls.append(time.perf_counter() - t0)
    st = b.stats()
    rays = st["primary_rays"] + st["shadow_rays"] + st["path_rays"]
    batch_s = min(walls)

    log("batched workload: sequential single-view frames")
    cfg = json.loads(env.read_text())
    rs = []
    for v in views:
        c = json.loads(json.dumps(cfg))
        c["Resolution"] = {"width": v.width, "height": v.height}
        c["Renderer"].update({"numberOfPrimarySamples": v.spp, "numberOfShadowSamples": S,
                              "recursionDepth": depth, "rayType": "path", "russianRoulette": False,
                              "sortPathRays": True, "reorderPathRays": True,
                              "keyMode": "origin30_oct30"})
        r = Renderer(c)
        r.set_scene(scenes[v.scene], hosts[v.scene].bmin, hosts[v.scene].bmax)
        r.render()                               # allocation + warm-up
        rs.append(r)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in rs:
        r.reset_frame_index()
        r.render()
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    # spot parity: the largest and the smallest view against the batch
    import numpy as np
    sizes = [v.width * v.height * v.spp for v in views]
    check = sorted({int(np.argmax(sizes)), int(np.argmin(sizes))})
    equal = all(np.array_equal(b.pixels(k)[:, :3], rs[k].pixels()[:, :3]) for k in check)
    return {"workload": f"{num_views} boxes scenes (tris log-uniform 1k-500k), res in "
                        "{128,256,512,1024}^2, spp U{1..16}, 6 bounces, 2 shadow samples",
            "triangles": sum(boxes) * 192 + 4 * num_views, "pixels": st["pixels"],
            "rays_per_frame": rays, "batched_s": round(batch_s, 4),
            "batched_gpu_ms": round(gpu_ms, 2), "batched_mrays_s": round(rays / batch_s / 1e6, 1),
            "sequential_s": round(seq_s, 4), "sequential_mrays_s": round(rays / seq_s / 1e6, 1),
            "speedup_vs_sequential": round(seq_s / batch_s, 3), "tuning": tuning,
            "autotune_s": round(tune_s, 2), "autotune_best_pass_ms": round(best_pass_ms, 3),
            "views_checked_equal": equal}


def tuning_state():
    from vulkanwavefrontpathtracer_amd import stages

    out = {}
    for k in ("bvh.presplit", "bvh.split_depth", "bvh.leaf_max", "bvh.builder", "bvh.collapse",
              "trace.lds_stack", "trace.block", "trace.persistent", "trace.refill", "trace.width",
              "trace.loop", "trace.group64", "trace.group16", "trace.group4", "sort.onesweep_tiles"):
        v = C.c_int()
        stages.call("wf_get_tuning", k.encode(), C.byref(v))
        out[k] = v.value
    return out


KEY_MODES = {"aila32": 0, "origin30": 1, "origin30_oct30": 2, "origin_dir_morton60": 4}
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_interior_1080p.txt"
STAGE_KERNEL = {"trace": "trace_kernel", "sort": "sort_bin_kernel", "reorder": "reorder_kernel",
                "keys": "keys_kernel"}


def measured_traffic(stage, elems):
    """HBM bytes per launch of the stage's kernel at this launch size, as
    measured by the two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, scaled
    by a plain 1 GiB copy) summarised in profiles/; None when not measured."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        f = line.split()
        if len(f) == 4 and f[0] == "traffic" and f[1] == STAGE_KERNEL.get(stage) and int(f[2]) == elems:
            return int(f[3])
    return None


def cpu_baseline(b0, scene_h, smin, smax, mode, kbits, sample, renderer):
    """The oracle's port of keys -> sort -> reorder -> trace on the host cores,
    over the first `sample` rays of the largest bounce (bounded CPU work)."""
    import numpy as np

    from oracle import oracle
    from vulkanwavefrontpathtracer_amd import stages

    n = min(sample, b0["n"])
    rays = b0["rays"][:n].cpu().numpy()
    s2i = b0["s2i"][:n].cpu().numpy()
    env = json.loads((ROOT / "vulkanwavefrontpathtracer_amd" / "env" / "config1.json").read_text())
    hs = stages.HostScene(1, env["Scene"]["seed"])
    tv = hs.tri_verts()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"cpu baseline: {n} rays, {threads} threads")
    t0 = time.perf_counter()
    keys = oracle.keys_oct60(rays, list(smin), list(smax)) if mode == 2 else \
        oracle.keys_aila32(rays, list(smin), list(smax))
    t1 = time.perf_counter()
    k, v = oracle.sort_pairs(keys, np.arange(n, dtype=np.uint32), kbits)
    t2 = time.perf_counter()
    out_r, out_s, out_i = oracle.reorder(v, rays, s2i, int(s2i.max()) + 1)
    t3 = time.perf_counter()
    t_, tri = oracle.trace_bvh_f64(out_r, tv, closest=True)
    t4 = time.perf_counter()
    total = t4 - t0
    return {"value": round(n / total / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {n} path rays of bounce 0 of the same frame; oracle keys/sort/reorder "
                      f"single-threaded ({t1 - t0:.2f}+{t2 - t1:.2f}+{t3 - t2:.2f} s), fp64 BVH "
                      f"trace on {threads} OpenMP threads ({t4 - t3:.2f} s, BVH build included)"}


if __name__ == "__main__":
    sys.exit(main())
