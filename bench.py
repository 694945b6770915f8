"""Benchmark: Mrays/s (trace+sort+reorder) of the wavefront bounce loop.

Headline workload (one GPU): the large plain scene, a procedural city of
2,821,186 triangles, 1920x1080 x 4 spp with all samples in flight (8.3M
primary rays, 5.5M path rays at the first bounce), 8 bounces, the env file's
ray key (the reference's 32-bit Aila key sorted as u32 keys with a u32
payload, 8-bit digits), reorder every bounce.  One untimed frame of the C++ frame loop
captures every pass's path wavefronts in HBM; the 4 passes of each bounce are
concatenated into one wavefront.  A timed step is one pass of the hot path
over all captured bounces of that frame:
    keys -> sort -> reorder + closest-hit trace    (Tracer::traceSort)
with the reorder gather fused into the trace kernel (wf_reorder_trace).
value = rays through the step / wall time of the timed steps, the step
captured once in a HIP graph and replayed (the eager-launch rate is
reported beside it).

Also reported: per-stage GPU times (hipEvents on the stream the kernels run
on); the reference's per-bounce sort log (SortLog: rays, key, sort, sorted
trace and unsorted trace times of each bounce, Renderer.cpp:280-289); the
sort ablation (the same wavefronts traced unsorted, sorted every 2nd bounce,
with the other key layouts and with 11-bit digits); the dominant launch's
roofline, with hardware counters read in this run by rocprofv3 child passes;
the medium interior workload by the same method; the reference-style
full-frame figures of the C++ frame loop; the batched and differentiable
workloads; and a CPU baseline (the oracle's port of the same stages, in the
headline's key mode, over the headline step's wavefronts).
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
KEY_MODES = {"aila32": 0, "origin30": 1, "origin30_oct30": 2, "origin_dir_morton60": 4,
             "origin_dir_morton30": 5}
KEY_DESC = {0: "32-bit Aila keys (7 origin + 4 direction bits per axis), u32 key sort + u32 payload",
            1: "30-bit origin Morton keys (10 bits per axis), u32 key sort + u32 payload",
            2: "60-bit keys (30 origin bits then 30 octahedral direction bits), u64 key sort + u32 payload",
            4: "60-bit 5-D Morton keys (30 origin + 30 octahedral direction bits interleaved), "
               "u64 key sort + u32 payload",
            5: "30-bit 5-D Morton keys (the top 30 bits of the 60-bit one), u32 key sort + u32 payload"}
PROFILE_DIR = ROOT / "profiles"


_T0 = time.perf_counter()


def log(msg):
    print(f"[bench +{time.perf_counter() - _T0:.1f}s] {msg}", file=sys.stderr, flush=True)


def d2d(dst: int, src: int, nbytes: int) -> None:
    from vulkanwavefrontpathtracer_amd._capi import memcpy_d2d

    memcpy_d2d(dst, src, nbytes)


# --------------------------------------------------------------- wavefronts
def capture_waves(dev, env_name, overrides=None):
    """One untimed frame of the C++ frame loop with capture on; per bounce,
    the passes' pre-sort wavefronts concatenated into one resident wavefront
    (slot-to-pixel values offset by pass * pixels so that every slot has its
    own inverse-LUT entry).  Returns (renderer, scene handle, bounds, waves)."""
    import torch

    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    env = env_path(env_name)
    cfg = json.loads(env.read_text())
    r = Renderer(env, overrides or {})
    log(f"{env_name}: building the scene")
    r.build_scene()
    r.set_capture(True)
    r.render()
    depth = cfg["Renderer"]["recursionDepth"]
    spp = cfg["Renderer"]["numberOfPrimarySamples"]
    w, h = r.size
    npx = w * h
    waves = []
    for b in range(depth):
        parts = [r.captured(p * depth + b) for p in range(spp)]
        n = sum(x[2] for x in parts)
        if n == 0:
            break
        rays = torch.empty((n, 8), dtype=torch.float32, device=dev)
        s2i = torch.empty(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        o = 0
        for p, (rp, sp, k) in enumerate(parts):
            if k:
                d2d(rays[o:].data_ptr(), rp, 32 * k)
                d2d(s2i[o:].data_ptr(), sp, 4 * k)
                s2i[o:o + k] += p * npx
                o += k
        waves.append(dict(n=n, rays=rays, s2i=s2i))
    smin, smax = (C.c_float * 3)(), (C.c_float * 3)()
    from vulkanwavefrontpathtracer_amd import stages

    stages.call("wf_scene_bounds", r.scene_handle(), smin, smax)
    log(f"{env_name}: {len(waves)} bounces, {sum(x['n'] for x in waves)} path rays "
        f"{[x['n'] for x in waves]}")
    return r, r.scene_handle(), (smin, smax), waves, spp * npx, cfg


class Pipeline:
    """Tracer::traceSort over captured wavefronts: per bounce keys -> sort ->
    reorder + trace (fused), or the trace alone on bounces that are not
    sorted; every stage call is stream-ordered, so a step can be captured in
    a HIP graph."""

    def __init__(self, dev, scene, bounds, waves, index_cap, key_mode, digit_bits, fused=True):
        import torch

        from vulkanwavefrontpathtracer_amd import stages

        self.stages = stages
        self.lib = stages._capi.load()
        self.scene = scene
        self.mn3, self.mx3 = bounds
        self.waves = waves
        self.fused = fused
        self.set_keys(key_mode, digit_bits)
        for x in waves:
            n = x["n"]
            x["keys"] = torch.empty(n, dtype=torch.int64, device=dev)
            x["idx"] = torch.empty(n, dtype=torch.int32, device=dev)
            x["out_rays"] = torch.empty_like(x["rays"])
            x["out_s2i"] = torch.empty_like(x["s2i"])
            x["out_i2s"] = torch.empty(index_cap, dtype=torch.int32, device=dev)
            x["hits"] = torch.empty((n, 4), dtype=torch.int32, device=dev)
            x["hits_u"] = torch.empty((n, 4), dtype=torch.int32, device=dev)
        need = max(self.lib.wf_sort_scratch_bytes(x["n"], 8, 64, d) for x in waves for d in (8, 11))
        self.scratch = torch.empty(need, dtype=torch.uint8, device=dev)
        self.names = ["keys", "sort", "reorder", "trace"]

    def set_keys(self, key_mode, digit_bits):
        self.mode = key_mode
        self.kbytes = self.lib.wf_key_bytes(key_mode)
        self.kbits = self.stages.key_bits(key_mode)
        self.digit_bits = digit_bits

    def bytes_per_ray(self):
        passes = -(-self.kbits // self.digit_bits)
        return {"keys": 32 + self.kbytes + 4,
                "sort": passes * 2 * (self.kbytes + 4) + self.kbytes,
                "reorder": 80, "trace": 96 if self.fused else 48}

    def step(self, stream_ptr, period=1, events=None, stream=None):
        """One pass over all bounces; bounce b is sorted when period > 0 and
        b % period == 0.  events: list receiving (bounce, stage, start, end)."""
        import torch

        st = self.stages._capi
        lib = self.lib
        V = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        for b, x in enumerate(self.waves):
            n = x["n"]
            srt = period > 0 and b % period == 0

            def mark():
                if events is None:
                    return None
                e = torch.cuda.Event(enable_timing=True)
                e.record(stream)
                return e

            m0 = mark()
            if srt:
                st.check("wf_morton_keys", lib.wf_morton_keys(
                    V(x["rays"]), V(x["keys"]), V(x["idx"]), n, self.mode, self.mn3, self.mx3,
                    stream_ptr, None))
            m1 = mark()
            if srt:
                st.check("wf_sort_pairs", lib.wf_sort_pairs(
                    V(x["keys"]), V(x["idx"]), n, self.kbytes, self.kbits, self.digit_bits,
                    V(self.scratch), self.scratch.numel(), stream_ptr, None))
            m2 = mark()
            if srt and not self.fused:
                st.check("wf_reorder", lib.wf_reorder(
                    V(x["idx"]), V(x["rays"]), V(x["out_rays"]), V(x["s2i"]), V(x["out_s2i"]),
                    V(x["out_i2s"]), n, stream_ptr, None))
            m3 = mark()
            if srt and self.fused:
                st.check("wf_reorder_trace", lib.wf_reorder_trace(
                    self.scene, V(x["idx"]), V(x["rays"]), V(x["out_rays"]), V(x["s2i"]),
                    V(x["out_s2i"]), V(x["out_i2s"]), V(x["hits"]), n, 1, stream_ptr, None))
            elif srt:
                st.check("wf_trace", lib.wf_trace(self.scene, V(x["out_rays"]), None, V(x["hits"]),
                                                  n, 1, stream_ptr, None))
            else:
                st.check("wf_trace", lib.wf_trace(self.scene, V(x["rays"]), None, V(x["hits"]), n,
                                                  1, stream_ptr, None))
            m4 = mark()
            if events is not None:
                for k, (a, e) in enumerate(((m0, m1), (m1, m2), (m2, m3), (m3, m4))):
                    events.append((b, self.names[k], a, e))

    def trace_unsorted(self, stream_ptr, events, stream):
        import torch

        for b, x in enumerate(self.waves):
            a = torch.cuda.Event(enable_timing=True)
            e = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            self.stages._capi.check("wf_trace", self.lib.wf_trace(
                self.scene, C.c_void_p(x["rays"].data_ptr()), None, C.c_void_p(x["hits_u"].data_ptr()),
                x["n"], 1, stream_ptr, None))
            e.record(stream)
            events.append((b, "trace_unsorted", a, e))

    def timed(self, stream, reps=3, period=1):
        """Per bounce and stage GPU ms (mean over reps), via hipEvents on the
        kernels' stream."""
        import torch

        evs = []
        for _ in range(reps):
            self.step(C.c_void_p(stream.cuda_stream), period, evs, stream)
        torch.cuda.synchronize()
        out = {}
        for b, k, a, e in evs:
            out.setdefault(b, {}).setdefault(k, 0.0)
            out[b][k] += a.elapsed_time(e) / reps
        return out


def run_workload(dev, args, env_name, ablations=True, overrides=None, key_mode=None):
    import torch

    r, scene, bounds, waves, index_cap, cfg = capture_waves(dev, env_name, overrides)
    mode = KEY_MODES[key_mode or cfg["Renderer"]["keyMode"]]
    p = Pipeline(dev, scene, bounds, waves, index_cap, mode, args.digit_bits, fused=not args.no_fuse)
    total_rays = sum(x["n"] for x in waves)
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)

    log(f"{env_name}: warmup {args.warmup} steps")
    for _ in range(args.warmup):
        p.step(sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p.step(sp)
    torch.cuda.synchronize()
    wall_eager = time.perf_counter() - t0
    value_eager = total_rays * args.steps / wall_eager / 1e6

    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        p.step(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    log(f"{env_name}: timing {args.steps} graph-replayed steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = total_rays * args.steps / wall / 1e6
    log(f"{env_name}: {value:.1f} Mrays/s graph-replayed, {value_eager:.1f} eager")
    graph_hits = [x["hits"].clone() for x in waves]
    p.step(sp)
    torch.cuda.synchronize()
    graph_equal = all(torch.equal(h, x["hits"]) for h, x in zip(graph_hits, waves))
    del graph

    per_b = p.timed(stream)
    unsorted_ev = []
    for _ in range(3):
        p.trace_unsorted(sp, unsorted_ev, stream)
    torch.cuda.synchronize()
    for b, k, a, e in unsorted_ev:
        per_b[b].setdefault(k, 0.0)
        per_b[b][k] += a.elapsed_time(e) / 3
    same = all(torch.equal(x["hits"][torch.argsort(x["out_s2i"])], x["hits_u"][torch.argsort(x["s2i"])])
               for x in waves)

    stage_ms = {k: sum(per_b[b][k] for b in per_b) for k in p.names}
    unsorted_ms = sum(per_b[b]["trace_unsorted"] for b in per_b)
    bpe = p.bytes_per_ray()
    stages_roof = {}
    for k in p.names:
        if p.fused and k == "reorder":
            stages_roof[k] = {"ms_per_step": 0.0, "fused_into": "trace"}
            continue
        t = stage_ms[k] / 1e3
        gbs = total_rays * bpe[k] / t / 1e9 if t > 0 else 0.0
        stages_roof[k] = {"ms_per_step": round(stage_ms[k], 4), "bytes_per_ray": bpe[k],
                          "achieved_GBs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
    # SortLog per bounce (Renderer.h:29-46, Renderer.cpp:280-289, printed by
    # Benchmark.cpp:159-195): the sorted pipeline's stages and the unsorted
    # trace of the same wavefront
    sort_log = []
    for b, x in enumerate(waves):
        d = per_b[b]
        sort_log.append({"depth": b + 1, "ray_count": x["n"], "morton_ms": round(d["keys"], 4),
                         "sort_ms": round(d["sort"], 4), "reorder_ms": round(d["reorder"], 4),
                         "trace_sort_ms": round(d["trace"], 4),
                         "trace_ms": round(d["trace_unsorted"], 4),
                         "sorting_pays": d["keys"] + d["sort"] + d["reorder"] + d["trace"] <
                         d["trace_unsorted"]})
    sorted_ms = sum(stage_ms.values())
    big = max(range(len(waves)), key=lambda i: waves[i]["n"])
    nbig = waves[big]["n"]
    # traversal work of the largest bounce (per-ray node visits and triangle
    # tests do not depend on the order the rays are traced in)
    nv = torch.zeros(nbig, dtype=torch.int32, device=dev)
    nt = torch.zeros_like(nv)
    scratch_hits = torch.empty((nbig, 4), dtype=torch.int32, device=dev)
    p.stages._capi.check("wf_trace_stats", p.lib.wf_trace_stats(
        scene, C.c_void_p(waves[big]["rays"].data_ptr()), C.c_void_p(scratch_hits.data_ptr()), nbig, 1,
        C.c_void_p(nv.data_ptr()), C.c_void_p(nt.data_ptr()), None))
    torch.cuda.synchronize()
    nodes_per_ray, tris_per_ray = nv.double().mean().item(), nt.double().mean().item()
    del nv, nt, scratch_hits
    # a library reference point for the sort of the largest bounce: torch's
    # stable sort of the same keys (rocPRIM's radix sort on this image),
    # event-timed like the stages
    lib_keys, _ = p.stages.morton_keys(waves[big]["rays"], mode, list(p.mn3), list(p.mx3))
    torch.sort(lib_keys, stable=True)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    lib_ms = []
    for _ in range(5):
        ev0.record(stream)
        torch.sort(lib_keys, stable=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        lib_ms.append(ev0.elapsed_time(ev1))
    lib_sort_ms = min(lib_ms)
    del lib_keys
    sort_big_ms = per_b[big]["sort"]
    sort_gbs = nbig * bpe["sort"] / (sort_big_ms / 1e3) / 1e9
    out = {
        "value": value, "value_eager": value_eager, "ms_per_step": ms_per_step,
        "rays": total_rays, "rays_per_bounce": [x["n"] for x in waves],
        "graph_replay_equals_eager": bool(graph_equal), "sorted_vs_unsorted_hits_equal": bool(same),
        "stages": stages_roof, "stage_ms": stage_ms, "sort_log": sort_log,
        "sorted_pipeline_mrays_s": round(total_rays / (sorted_ms / 1e3) / 1e6, 1),
        "unsorted_trace_mrays_s": round(total_rays / (unsorted_ms / 1e3) / 1e6, 1),
        "per_bounce_reorder_vs_unsorted": round(unsorted_ms / sorted_ms, 3),
        "trace_speedup_sorted_vs_unsorted": round(unsorted_ms / stage_ms["trace"], 3),
        "sort_largest_bounce": {"pairs": nbig, "ms": round(sort_big_ms, 4),
                                "bytes_per_pair": bpe["sort"], "achieved_GBs": round(sort_gbs, 1),
                                "frac_hbm": round(sort_gbs / HBM_PEAK_GBS, 4),
                                "key_bits": p.kbits, "digit_bits": p.digit_bits,
                                "library_reference": {
                                    "what": "torch.sort(keys, stable=True) of the same keys (values and "
                                            "indices out; rocPRIM radix sort)",
                                    "ms": round(lib_sort_ms, 4),
                                    "achieved_GBs_same_basis": round(nbig * bpe["sort"] / (lib_sort_ms / 1e3) / 1e9, 1)}},
        "dominant_launch": {"kernel": "trace_kernel<reorder>" if p.fused else "trace_kernel",
                            "elems": nbig, "ms": per_b[big]["trace"], "bytes_per_elem": bpe["trace"],
                            "unsorted_ms": per_b[big]["trace_unsorted"],
                            "nodes_per_ray": nodes_per_ray, "tris_per_ray": tris_per_ray},
        "mode": mode, "cfg": cfg, "renderer": r, "scene": scene, "bounds": bounds, "waves": waves,
    }
    if ablations:
        abl = {}
        variants = ablations if isinstance(ablations, list) else [
            ("every 2nd bounce", mode, args.digit_bits, 2),
            ("32-bit Aila keys", 0, args.digit_bits, 1),
            ("30-bit origin keys", 1, args.digit_bits, 1),
            ("60-bit 5-D Morton keys", 4, args.digit_bits, 1),
            ("30-bit 5-D Morton keys", 5, args.digit_bits, 1),
            ("60-bit concatenated keys", 2, args.digit_bits, 1),
            ("11-bit digits", mode, 11, 1)]
        for name, km, db, period in variants:
            if km == mode and db == args.digit_bits and period == 1:
                continue                          # the headline itself
            p.set_keys(km, db)
            pb = p.timed(stream, period=period)
            ms = sum(sum(d.values()) for d in pb.values())
            abl[name] = {"mrays_s": round(total_rays / (ms / 1e3) / 1e6, 1),
                         "vs_unsorted": round(unsorted_ms / ms, 3),
                         "ms_per_step": {k: round(sum(pb[b].get(k, 0.0) for b in pb), 4)
                                         for k in p.names}}
        p.set_keys(mode, args.digit_bits)
        # sorting only the bounces whose own log shows a gain
        pays = [s["sorting_pays"] for s in sort_log]
        best_ms = sum(sum(per_b[b][k] for k in p.names) if pays[b] else per_b[b]["trace_unsorted"]
                      for b in per_b)
        abl["sort only where the per-bounce log shows a gain"] = {
            "mrays_s": round(total_rays / (best_ms / 1e3) / 1e6, 1),
            "vs_unsorted": round(unsorted_ms / best_ms, 3), "sorted_bounces": pays}
        out["ablations"] = abl
    return out


def reference_frame(r, frames):
    """Reference-style full frames of the C++ frame loop (Benchmark::run):
    path Mrays/s includes shade; the renderer's own SortLog (with
    Benchmark.printSortLogs, an unsorted trace of every sorted wavefront)."""
    import torch

    walls, stats = [], None
    for _ in range(frames):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.render()
        walls.append(time.perf_counter() - t0)
        stats = r.stats()
    if stats is None:
        return None
    path_rays = stats["path_rays"]
    return {
        "path_rays": path_rays,
        "path_mrays_s": round(path_rays / (stats["path_ms"] / 1e3) / 1e6, 1) if stats["path_ms"] > 0 else None,
        "frame_ms_wall": round(1e3 * min(walls), 2),
        "rays_per_frame": stats["primary_rays"] + stats["shadow_rays"] + path_rays,
        "stage_ms": {k: round(v, 3) for k, v in stats["stage_ms"].items()},
        "renderer_sort_log_path": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in x.items()}
                                   for x in r.sort_logs("path")],
    }


# In-run hardware counters.  rocprofv3 cannot attach to this process, so
# each counter set is read by a child process (this script with
# --counter-pass) under `rocprofv3 --pmc`: it rebuilds the headline
# wavefronts, runs one plain 1 GiB device copy (the scale for FETCH_SIZE and
# WRITE_SIZE), then the fused reorder + trace of the largest bounce and the
# unsorted trace of the same bounce, in that order.
COUNTER_SETS = (
    ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "GRBM_GUI_ACTIVE",
     "TCC_HIT_sum", "TCC_MISS_sum"),
    ("FETCH_SIZE",),
    ("WRITE_SIZE",),
)
SIMDS = 1024                  # 256 CUs x 4 SIMDs
ENGINE_GHZ = 2.4
# wave64 VALU instructions per SIMD per clock behind the published 157.3
# TFLOP/s fp32 (1024 SIMDs x 2.4 GHz x 64 lanes x 2 flops x 0.5)
VALU_ISSUE_PEAK = 0.5
FP32_PEAK_TFLOPS = 157.3
# fp32 operations (an FMA counts 2) of one 4-wide node test and one
# triangle test as traverse.h and intersect.h write them: node = 3 scale
# multiplies + 3 origin FMAs + per child (6 plane FMAs, 6 min/max, 1
# compare) + 5 compares of the ordering network; triangle = 9 vertex
# subtractions + 6 shear FMAs + 3 edge functions (multiply + FMA) + 6 sign
# compares + 3 zero tests (the miss path; a hit adds the determinant and t).
FLOPS_PER_NODE = 3 + 6 + 4 * (12 + 6 + 1) + 5
FLOPS_PER_TRI = 9 + 12 + 9 + 6 + 3


def counter_child(args):
    """Child of one counter pass (runs under rocprofv3 --pmc)."""
    import torch

    from vulkanwavefrontpathtracer_amd import stages

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    for kv in args.tune:
        k, v = kv.split("=")
        stages.set_tuning(k, int(v))
    r, scene, bounds, waves, index_cap, cfg = capture_waves(dev, args.env)
    mode = KEY_MODES[args.key_mode or cfg["Renderer"]["keyMode"]]
    p = Pipeline(dev, scene, bounds, waves, index_cap, mode, args.digit_bits)
    big = max(range(len(waves)), key=lambda i: waves[i]["n"])
    p.waves = [waves[big]]
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p.step(sp)                                   # keys + sort (+ a first trace)
    torch.cuda.synchronize()
    nb = 1 << 30
    src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    x = waves[big]
    V = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    stages._capi.check("wf_reorder_trace", p.lib.wf_reorder_trace(
        scene, V(x["idx"]), V(x["rays"]), V(x["out_rays"]), V(x["s2i"]), V(x["out_s2i"]),
        V(x["out_i2s"]), V(x["hits"]), x["n"], 1, sp, None))
    torch.cuda.synchronize()
    stages._capi.check("wf_trace", p.lib.wf_trace(scene, V(x["rays"]), None, V(x["hits_u"]), x["n"], 1,
                                                  sp, None))
    torch.cuda.synchronize()
    log(f"counter pass: {x['n']} rays traced sorted and unsorted")
    return 0


def read_counter_csv(path):
    """Dispatches of one rocprofv3 --pmc CSV in dispatch order: (name,
    grid, {counter: value})."""
    import csv

    per = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            d = per.setdefault(int(row["Dispatch_Id"]), [row["Kernel_Name"], int(row["Grid_Size"]), {}])
            d[2][row["Counter_Name"]] = d[2].get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    return [per[k] for k in sorted(per)]


def in_run_counters(args, timeout_s=240):
    """Counters of the largest bounce's fused reorder + trace launch and of
    the unsorted trace of the same rays, each pass a rocprofv3 child; None
    (with the reason logged) when rocprofv3 is missing or a pass fails."""
    import shutil
    import signal
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        log("counters: rocprofv3 not found")
        return None
    tmp = tempfile.mkdtemp(prefix="wfpt_counters_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    child = [sys.executable, str(ROOT / "bench.py"), "--counter-pass", "--env", args.env,
             "--digit-bits", str(args.digit_bits)]
    if args.key_mode:
        child += ["--key-mode", args.key_mode]
    for kv in args.tune:
        child += ["--tune", kv]
    found = {}
    for i, counters in enumerate(COUNTER_SETS):
        out = os.path.join(tmp, f"p{i}")
        cmd = [prof, "--kernel-include-regex", "trace_kernel|copyBuffer", "--pmc", *counters,
               "-d", out, "-o", "run", "--output-format", "csv", "--", *child]
        log(f"counters: pass {i + 1}/{len(COUNTER_SETS)} ({' '.join(counters)})")
        with open(os.path.join(tmp, f"p{i}.log"), "w") as lf:
            pr = subprocess.Popen(cmd, cwd=tmp, env=env, stdout=lf, stderr=subprocess.STDOUT,
                                  start_new_session=True)
            try:
                rc = pr.wait(timeout=timeout_s)
            except subprocess.TimeoutExpired:
                os.killpg(pr.pid, signal.SIGKILL)
                pr.wait()
                log(f"counters: pass {i + 1} timed out")
                return None
        csv_path = os.path.join(out, "run_counter_collection.csv")
        if not os.path.exists(csv_path):
            log(f"counters: pass {i + 1} wrote no results (exit {rc})")
            return None
        disp = read_counter_csv(csv_path)
        copies = [k for k, d in enumerate(disp) if "copyBuffer" in d[0]]
        if not copies:
            log("counters: no copy dispatch found")
            return None
        cal = disp[copies[-1]]
        traces = [d for d in disp[copies[-1] + 1:] if "trace_kernel" in d[0]]
        if len(traces) < 2:
            log("counters: the two trace launches were not found")
            return None
        found.setdefault("calibration", {}).update(cal[2])
        found.setdefault("sorted", {"kernel": traces[0][0].split("(")[0], "grid": traces[0][1]})
        found["sorted"].update(traces[0][2])
        found.setdefault("unsorted", {"kernel": traces[1][0].split("(")[0], "grid": traces[1][1]})
        found["unsorted"].update(traces[1][2])
    log("counters: done")
    return found


def counter_figures(c, launch_ms):
    """Issue fraction, lane utilisation, L2 hit rate and HBM bytes of one
    launch from its counters and its live duration."""
    if c is None:
        return None
    waves = c["SQ_WAVES"]
    issue = c["SQ_INSTS_VALU"] / SIMDS / (launch_ms * 1e-3 * ENGINE_GHZ * 1e9)
    return {
        "valu_insts_per_wave": round(c["SQ_INSTS_VALU"] / waves, 1),
        "valu_issue_per_simd_clk": round(issue, 4),
        "valu_issue_frac": round(issue / VALU_ISSUE_PEAK, 4),
        "lane_utilization": round(c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU"]), 4),
        "l2_hit_rate": round(c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 4),
    }


def dominant_roofline(dl, counters):
    """Roofline of the dominant launch (the fused reorder + trace of the
    largest bounce).  Traversal is bound by VALU instruction issue, not by
    HBM: achieved = algorithmic fp32 work (node and triangle tests the rays
    make) over the live launch time against the published fp32 peak; beside
    it the hardware's VALU issue rate, lane utilisation and L2 hit rate for
    the sorted launch and for the unsorted trace of the same rays, and the
    HBM bytes of the sorted launch."""
    n, ms = dl["elems"], dl["ms"]
    flops = n * (dl["nodes_per_ray"] * FLOPS_PER_NODE + dl["tris_per_ray"] * FLOPS_PER_TRI)
    achieved = flops / (ms * 1e-3) / 1e12
    out = {
        "bound": "valu", "kernel": dl["kernel"], "achieved": round(achieved, 2),
        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
        "traffic": None, "launch_ms": round(ms, 4), "launch_elems": n,
        "algorithmic": {"nodes_per_ray": round(dl["nodes_per_ray"], 3),
                        "tris_per_ray": round(dl["tris_per_ray"], 3),
                        "flops_per_node": FLOPS_PER_NODE, "flops_per_tri": FLOPS_PER_TRI,
                        "streamed_bytes_per_ray": dl["bytes_per_elem"]},
        "valu_issue_peak_per_simd_clk": VALU_ISSUE_PEAK,
        "counters": None,
    }
    if counters:
        cal = counters["calibration"]
        fetch_scale = (1 << 30) / (cal["FETCH_SIZE"] * 1024.0)
        write_scale = (1 << 30) / (cal["WRITE_SIZE"] * 1024.0)
        srt, uns = counters["sorted"], counters["unsorted"]
        rd = srt["FETCH_SIZE"] * 1024.0 * fetch_scale
        wr = srt["WRITE_SIZE"] * 1024.0 * write_scale
        out["traffic"] = int(rd + wr)
        sf = counter_figures(srt, ms)
        out.update({"valu_issue_frac": sf["valu_issue_frac"], "lane_utilization": sf["lane_utilization"],
                    "l2_hit_rate": sf["l2_hit_rate"]})
        uf = counter_figures(uns, dl["unsorted_ms"])
        uf["traffic"] = int(uns["FETCH_SIZE"] * 1024.0 * fetch_scale + uns["WRITE_SIZE"] * 1024.0 * write_scale)
        sf["traffic"] = out["traffic"]
        out["counters"] = {
            "source": "rocprofv3 --pmc child passes of this run (one counter set each)",
            "fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
            "sorted_launch": dict(sf, kernel=srt["kernel"], grid=srt["grid"], ms=round(ms, 4)),
            "unsorted_launch": dict(uf, kernel=uns["kernel"], grid=uns["grid"],
                                    ms=round(dl["unsorted_ms"], 4)),
            "traffic_over_streamed_bytes": round(out["traffic"] / (n * dl["bytes_per_elem"]), 2),
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--env", default="config2.json")
    ap.add_argument("--digit-bits", type=int, default=8)
    ap.add_argument("--frames", type=int, default=2, help="reference-style full frames timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 30,
                    help="rays per bounce of the CPU baseline (default: all)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable workload")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-interior", action="store_true", help="skip the medium interior workload")
    ap.add_argument("--no-fuse", action="store_true",
                    help="reorder and trace as separate kernels (wf_reorder + wf_trace)")
    ap.add_argument("--no-ablations", action="store_true",
                    help="skip the sort ablations (profiling runs)")
    ap.add_argument("--batch-views", type=int, default=64)
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="wf_set_tuning knob applied before the scene is built (repeatable)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (scales HBM counters)")
    ap.add_argument("--key-mode", choices=sorted(KEY_MODES), default=None,
                    help="ray key of the timed step (default: the env file's Renderer.keyMode)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes of the dominant launch")
    ap.add_argument("--counter-pass", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.counter_pass:
        return counter_child(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the wavefront loop "
              "runs on a single device); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from vulkanwavefrontpathtracer_amd import stages

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    for kv in args.tune:
        k, v = kv.split("=")
        stages.set_tuning(k, int(v))
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nb} bytes read + {nb} bytes written")
        del src, dst

    head = run_workload(dev, args, args.env, ablations=not args.no_ablations, key_mode=args.key_mode)
    cfg = head["cfg"]
    R = cfg["Renderer"]
    dl = head["dominant_launch"]
    counters = None if args.no_counters else in_run_counters(args)
    roofline = dominant_roofline(dl, counters)

    interior = None
    if not args.no_interior:
        it = run_workload(dev, args, "config1.json",
                          ablations=[] if args.no_ablations else
                          [("30-bit 5-D Morton keys", 5, args.digit_bits, 1)])
        interior = {k: it[k] for k in ("value", "ms_per_step", "rays", "rays_per_bounce",
                                       "per_bounce_reorder_vs_unsorted",
                                       "trace_speedup_sorted_vs_unsorted", "sorted_pipeline_mrays_s",
                                       "unsorted_trace_mrays_s", "stages", "sort_largest_bounce")}
        if it.get("ablations"):
            interior["ablations"] = it["ablations"]
        interior["value"] = round(interior["value"], 2)
        interior["workload"] = ("medium plain: interior 263,250 tris, 1920x1080, 1 spp, 8 bounces, "
                                "RR after bounce 3, 60-bit 5-D Morton keys")
        del it

    r = head["renderer"]
    r_sortlogs = None
    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    fr = Renderer(env_path(args.env), {"Benchmark.printSortLogs": True})
    fr.set_scene(type("S", (), {"handle": head["scene"]})(), list(head["bounds"][0]),
                 list(head["bounds"][1]))
    ref_style = reference_frame(fr, args.frames)
    del fr

    adjoint = None if args.no_adjoint else adjoint_section(dev)
    batched = None if args.no_batch else batch_section(dev, args.batch_views)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(head, args.cpu_sample)

    split_depth = cfg.get("Scene", {}).get("splitDepth", -1)
    if split_depth < 0:
        split_depth = stages.get_tuning("bvh.split_depth")
    out = {
        "metric": "Mrays/s (trace+sort+reorder), plain vs batched & adjoint, at RMSE≤1e-3 vs FP64",
        "value": round(head["value"], 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural city scene, captured bounce wavefronts)",
        "config": {
            "workload": "large plain: city 2,821,186 tris, 1920x1080 x 4 spp in flight, "
                        f"{R['recursionDepth']} bounces, {KEY_DESC[head['mode']]}, reorder every bounce",
            "key_mode": {v: k for k, v in KEY_MODES.items()}[head["mode"]],
            "rays_per_step": head["rays"], "bounces": len(head["rays_per_bounce"]),
            "rays_per_bounce": head["rays_per_bounce"], "digit_bits": args.digit_bits,
            "bvh_split_depth": split_depth,
            "bvh_builder": "binned SAH" if stages.get_tuning("bvh.builder") == 1 else "radix tree",
            "parallelism": "1 GPU",
        },
        "roofline": roofline,
        "stages": head["stages"],
        "value_eager_launches": round(head["value_eager"], 2),
        "graph_replay_equals_eager": head["graph_replay_equals_eager"],
        "sorted_vs_unsorted_hits_equal": head["sorted_vs_unsorted_hits_equal"],
        "sorted_pipeline_mrays_s": head["sorted_pipeline_mrays_s"],
        "unsorted_trace_mrays_s": head["unsorted_trace_mrays_s"],
        "per_bounce_reorder_vs_unsorted": head["per_bounce_reorder_vs_unsorted"],
        "trace_speedup_sorted_vs_unsorted": head["trace_speedup_sorted_vs_unsorted"],
        "sort_largest_bounce": head["sort_largest_bounce"],
        "sort_log": head["sort_log"],
        "ablations": head.get("ablations"),
        "reference_style_frame": ref_style,
        "interior": interior,
        "adjoint": adjoint,
        "batched": batched,
        "tuning": tuning_state(),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def adjoint_section(dev, reps=3):
    """Differentiable workload: the interior scene with 512 material groups,
    512x512, 16 spp, 6 bounces, L2 loss against a target rendered from
    parameters perturbed by N(0, 0.05).  Forward = the C++ frame loop (frame
    1); adjoint = the path-replay kernel.  Both are timed on the GPU (the
    forward as the sum of its stage events, the adjoint by its own events)
    and on the wall clock; the ratio is GPU over GPU."""
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
    f_gpu, a_gpu = min(fwd_ms), min(adj_ms)
    f_wall, a_wall = min(fwd_wall) * 1e3, min(adj_wall) * 1e3
    return {"workload": "interior 263,250 tris, 512 material groups, 512x512, 16 spp, 6 bounces, "
                        "L2 loss vs perturbed-parameter target",
            "forward_gpu_ms": round(f_gpu, 3), "forward_wall_ms": round(f_wall, 3),
            "adjoint_gpu_ms": round(a_gpu, 3), "adjoint_wall_ms": round(a_wall, 3),
            "adjoint_over_forward": round(a_gpu / f_gpu, 3),
            "adjoint_over_forward_wall": round(a_wall / f_wall, 3),
            "rays_per_frame": rays, "forward_mrays_s": round(rays / (f_gpu / 1e3) / 1e6, 1),
            "adjoint_mrays_s": round(rays / (a_gpu / 1e3) / 1e6, 1), "loss": loss,
            "grad_norm": float(torch.sqrt((ga.double() ** 2).sum() + (ge.double() ** 2).sum()))}


def batch_section(dev, num_views=64, reps=2):
    """Batched workload: num_views scenes from the boxes generator (triangles
    log-uniform in [1k, 500k]), one view each with resolution in
    {128, 256, 512, 1024}^2 and spp in 1..16, 6 bounces, 2 shadow samples.
    The batch (one queue, scene-prefixed keys, graph-replayed passes,
    autotuned) against the same views rendered one after another by the
    single-view frame loop with the batch's tuned knobs (digit width, trace
    block size, reorder period; one renderer per view, buffers allocated by
    an untimed warm-up frame)."""
    import numpy as np
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

    log("batched workload: sequential single-view frames with the batch's knobs")
    cfg = json.loads(env.read_text())
    saved_block = stages.get_tuning("trace.block")
    stages.set_tuning("trace.block", tuning["block"])
    try:
        rs = []
        for v in views:
            c = json.loads(json.dumps(cfg))
            c["Resolution"] = {"width": v.width, "height": v.height}
            c["Renderer"].update({"numberOfPrimarySamples": v.spp, "numberOfShadowSamples": S,
                                  "recursionDepth": depth, "rayType": "path",
                                  "russianRoulette": False,
                                  "sortPathRays": tuning["reorder_period"] > 0,
                                  "reorderPathRays": tuning["reorder_period"] > 0,
                                  "reorderPeriod": tuning["reorder_period"],
                                  "digitBits": tuning["digit_bits"], "keyMode": "origin30_oct30"})
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
    finally:
        stages.set_tuning("trace.block", saved_block)
    sizes = [v.width * v.height * v.spp for v in views]
    check = sorted({int(np.argmax(sizes)), int(np.argmin(sizes))})
    equal = all(np.array_equal(b.pixels(k)[:, :3], rs[k].pixels()[:, :3]) for k in check)
    return {"workload": f"{num_views} boxes scenes (tris log-uniform 1k-500k), res in "
                        "{128,256,512,1024}^2, spp U{1..16}, 6 bounces, 2 shadow samples",
            "triangles": sum(boxes) * 192 + 4 * num_views, "pixels": st["pixels"],
            "rays_per_frame": rays, "batched_s": round(batch_s, 4),
            "batched_gpu_ms": round(gpu_ms, 2), "batched_mrays_s": round(rays / batch_s / 1e6, 1),
            "sequential_s": round(seq_s, 4), "sequential_mrays_s": round(rays / seq_s / 1e6, 1),
            "sequential_knobs": "the batch's digit width, trace block and reorder period",
            "speedup_vs_sequential": round(seq_s / batch_s, 3), "tuning": tuning,
            "autotune_s": round(tune_s, 2), "autotune_best_pass_ms": round(best_pass_ms, 3),
            "views_checked_equal": equal}


def tuning_state():
    from vulkanwavefrontpathtracer_amd import stages

    out = {}
    for k in ("bvh.builder", "bvh.presplit", "bvh.split_depth", "bvh.leaf_max", "bvh.collapse", "trace.block",
              "trace.xcd", "trace.group64", "trace.group16", "trace.group4", "sort.onesweep_tiles",
              "sort.rank_max"):
        out[k] = stages.get_tuning(k)
    return out


def cpu_baseline(head, sample):
    """The oracle's port of keys -> sort -> reorder -> trace on the host
    cores, in the headline's key mode, over the headline step's bounce
    wavefronts (all of them unless `sample` caps the rays per bounce): keys,
    reorder and trace on OMP_NUM_THREADS threads, the stable sort on one; the
    oracle's BVH is built before the clock starts, and the trace is its fp32
    watertight one (the same records as the GPU's)."""
    import numpy as np

    from oracle import oracle
    from vulkanwavefrontpathtracer_amd import stages

    cfg = head["cfg"]
    hs = stages.HostScene(stages.SCENE_CITY, cfg["Scene"]["seed"])
    mn, mx = list(head["bounds"][0]), list(head["bounds"][1])
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    waves = [(x["rays"][:sample].cpu().numpy(), x["s2i"][:sample].cpu().numpy()) for x in head["waves"]]
    n = sum(len(r) for r, _ in waves)
    log(f"cpu baseline: building the oracle BVH, then {n} rays on {threads} threads")
    bvh = oracle.Bvh(hs.tri_verts())
    t = {"keys": 0.0, "sort": 0.0, "reorder": 0.0, "trace": 0.0}
    hits = 0
    mode = head["mode"]
    keyfn, kbits = {0: (oracle.keys_aila32, 32), 1: (oracle.keys_origin30, 30),
                    2: (oracle.keys_oct60, 60), 4: (oracle.keys_morton5d60, 60)}[mode]
    for rays, s2i in waves:
        t0 = time.perf_counter()
        keys = keyfn(rays, mn, mx)
        t1 = time.perf_counter()
        k, v = oracle.sort_pairs(keys, np.arange(len(rays), dtype=np.uint32), kbits)
        t2 = time.perf_counter()
        out_r, out_s, out_i = oracle.reorder(v, rays, s2i, int(s2i.max()) + 1)
        t3 = time.perf_counter()
        h = bvh.trace_f32(out_r, closest=True)
        t4 = time.perf_counter()
        hits += int((h[:, 0] >= 0).sum())
        for name, a, b in (("keys", t0, t1), ("sort", t1, t2), ("reorder", t2, t3), ("trace", t3, t4)):
            t[name] += b - a
    total = sum(t.values())
    return {"value": round(n / total / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{n} path rays: {'the whole headline step' if n == head['rays'] else f'the first {sample} rays of each bounce of the headline step'}; "
                      f"oracle {KEY_DESC[mode].split(',')[0]} and reorder on {threads} OpenMP threads, "
                      f"stable LSD sort on one thread ({t['keys']:.2f}+{t['sort']:.2f}+{t['reorder']:.2f} s), "
                      f"fp32 watertight BVH trace on {threads} OpenMP threads ({t['trace']:.2f} s, BVH "
                      f"built before timing); {hits} hits"}


if __name__ == "__main__":
    sys.exit(main())
