#!/usr/bin/env python3
"""Benchmark: OIT fragment resolve on one MI355X.

Workload: 1920x1080, 1000 analytic spheres (r in [0.01, 0.05], alpha in
[0.1, 0.5], seed 1), Linked List with K = 16 and a node pool of 1.5 x the
frame's fragments. A step is one frame: clear, fragment store + link
(one append of the whole stream, already resident in HBM), composite.

Prints one JSON line: G fragments/s resolved (F / frame time), the HBM
roofline fraction of the dominant kernel (per-kernel HIP events on the
launch stream), per-frame latency p50/p99, and the single-threaded CPU
oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (scaled by a 1 GiB copy: FETCH x 2, WRITE x 1) on
# this workload (profiles/linkedlist_k16_1920x1080_1000spheres.md).
MEASURED_TRAFFIC = {("linkedlist_k16_1920x1080_1000spheres", "store", 0): 559654502}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU model (first /proc/cpuinfo model name), the logical
    CPUs the OS reports (nproc) and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(tg, d, c, K, cap, budget_s):
    """The oracle (single-threaded C restatement) on the same frame, repeated
    until about `budget_s` seconds of CPU work; median frame time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oit_oracle

    cfg = oit_oracle.make_config(oit_oracle.LINKEDLIST, K, 1, 1920, 1080, ll_capacity=cap)
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        oit_oracle.run(cfg, tg, d, c)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 10:
            break
    times.sort()
    med = times[len(times) // 2]
    return {
        "value": len(tg) / med / 1e9,
        "unit": "G fragments/s",
        "cores": 1,
        **host_cpu(),
        "kind": "port",
        "sample": f"{len(times)} full frames of the same workload (1920x1080, 1000 spheres, LL K={K}), "
                  f"median {med * 1e3:.1f} ms/frame, oracle/liboit_oracle.so single thread",
    }


def oracle_frame_banded(oit_oracle, tg, d, c, K, W, H, bands, ex):
    """One Linked List frame of the oracle split over pixel bands: the
    fragments are bucketed by band (a stable sort of the band index, so each
    band keeps draw order), each band runs with its targets renumbered on the
    executor `ex` (the oracle's C calls release the GIL), and the band images
    are written into the frame's image. Pixels are independent, so the image
    equals the full-frame run bit for bit."""
    import numpy as np

    rows = (H + bands - 1) // bands
    span = np.uint32(W * rows)
    out = np.empty((W * H, 4), np.float32)
    band = (tg // span).astype(np.uint16)
    order = np.argsort(band, kind="stable")
    starts = np.concatenate([[0], np.cumsum(np.bincount(band, minlength=bands))])

    def work(b):
        h = min(rows, H - b * rows)
        if h <= 0:
            return
        sel = order[starts[b]:starts[b + 1]]
        first = b * W * rows
        if len(sel) == 0:
            cfg = oit_oracle.make_config(oit_oracle.LINKEDLIST, K, 1, W, h, ll_capacity=1)
            out[first:first + W * h] = oit_oracle.run(cfg, tg[:0], d[:0], c[:0])["out"]
            return
        cfg = oit_oracle.make_config(oit_oracle.LINKEDLIST, K, 1, W, h, ll_capacity=len(sel) + 1)
        out[first:first + W * h] = oit_oracle.run(cfg, tg[sel] - np.uint32(first), d[sel], c[sel])["out"]

    list(ex.map(work, range(bands)))
    return out


def cpu_baseline_all_cores(tg, d, c, K, budget_s, threads):
    """The same frame on `threads` host cores (oracle_frame_banded with four
    pixel bands per thread; the bucketing is timed with the runs)."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oit_oracle

    oit_oracle.tables()  # tables built once, before the threads
    W, H = 1920, 1080
    bands = threads * 4
    times = []
    t_start = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while True:
            t0 = time.perf_counter()
            oracle_frame_banded(oit_oracle, tg, d, c, K, W, H, bands, ex)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_start > budget_s or len(times) >= 10:
                break
    times.sort()
    med = times[len(times) // 2]
    return {
        "value": len(tg) / med / 1e9,
        "unit": "G fragments/s",
        "cores": threads,
        **host_cpu(),
        "kind": "port",
        "sample": f"{len(times)} full frames of the same workload (1920x1080, 1000 spheres, LL K={K}), "
                  f"median {med * 1e3:.1f} ms/frame, fragments bucketed into {bands} pixel bands, "
                  f"oracle/liboit_oracle.so on {threads} threads",
    }


ALGORITHM_ROWS = ((1, "linkedlist", 8), (1, "linkedlist", 16), (3, "loop64", 8), (3, "loop64", 16),
                  (4, "spinlock", 8), (4, "spinlock", 16), (0, "simple", 8), (0, "simple", 16),
                  (6, "weighted", 1), (2, "loop32", 8), (2, "loop32", 16), (5, "interlock", 8),
                  (5, "interlock", 16))


def bench_algorithms(tg_d, d_d, c_d, F, p, steps, stream, only=None, variant=0, counts=None):
    """Every algorithm on the same 1920x1080 stream (Linked List pool 1.5F,
    tail blending on): device time per frame, G fragments/s and the
    fraction of the HBM roofline by the algorithm's compulsory bytes.
    `only`: a set of "name:K" strings to restrict the rows to. `counts`:
    fragments per target (host), for the slots the composites touch."""
    from vk_order_independent_transparency_amd import DeviceBuffer, Event, FrameConfig, Renderer
    from vk_order_independent_transparency_amd import roofline as rl
    import numpy as np

    P = p.targets
    out = DeviceBuffer((P, 4), "float32")
    rows = []
    for alg, name, K in ALGORITHM_ROWS:
        if only and f"{name}:{K}" not in only:
            continue
        with Renderer(0) as r:
            r.configure(FrameConfig(algorithm=alg, layers=K, tail_blend=alg != 6, width=p.width, height=p.height,
                                    ll_pool_nodes=F + F // 2 if alg == 1 else 0, max_fragments=F,
                                    variant=variant))
            for _ in range(5):
                r.render(tg_d, d_d, c_d, out, stream)
            e0, e1 = Event(), Event()
            stream.synchronize()
            e0.record(stream)
            for _ in range(steps):
                r.render(tg_d, d_d, c_d, out, stream)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / steps
        if alg == 1:
            nbytes = rl.ll_frame_bytes_reference(F, P)
        elif alg == 3:
            occ = None if counts is None else int(np.minimum(counts, K).sum())
            nbytes = sum(rl.loop64_stage_bytes(F, P, K, occupied=occ).values())
        elif alg in (0, 4):
            nbytes = sum(rl.kbuffer_stage_bytes(F, P, K, alg == 4).values())
        elif alg == 2:
            occ = None if counts is None else int(np.minimum(counts, K).sum())
            nbytes = sum(rl.loop32_stage_bytes(F, P, K, occupied=occ).values())
        elif alg == 5:
            nbytes = sum(rl.interlock_stage_bytes(F, P).values())
        else:
            nbytes = sum(rl.wboit_stage_bytes(F, P).values())
        gbs, fr = rl.frac(nbytes, ms / 1e3)
        rows.append({"algorithm": name, "layers": K, "ms_per_frame": round(ms, 4),
                     "G_fragments_per_s": round(F / ms / 1e6, 3), "bytes": nbytes, "frac": round(fr, 4)})
        log(json.dumps(rows[-1]))
    return rows


def bench_4k_4x(steps, variant=0):
    """3840x2160 with 4 samples per pixel (sample shading), 4000 spheres,
    radius [0.005, 0.03], alpha [0.05, 0.4], seed 3: WBOIT (FP32 accumulation
    and revealage) and the Linked List at K = 32 (the exact-image reference,
    pool of count + 1 nodes) on the same device-generated stream. Device time
    per frame, G fragment-samples/s and the fraction of the HBM roofline by
    the compulsory bytes."""
    from vk_order_independent_transparency_amd import (DeviceBuffer, Event, FrameConfig, OIT_LINKEDLIST,
                                                       OIT_WEIGHTED, Renderer, Stream, scene)
    from vk_order_independent_transparency_amd import roofline as rl

    p = scene.SceneParams(seed=3, width=3840, height=2160, samples=4, spheres=4000, radius=(0.005, 0.03),
                          alpha=(0.05, 0.4))
    st = Stream()
    g = scene.DeviceScene(p)
    t, d, c = g.buffers()
    g.emit(t, d, c, st)
    st.synchronize()
    F = g.fragments
    g.close()
    P = p.targets
    out = DeviceBuffer((P, 4), "float32")
    rows = []
    for alg, name, K in ((OIT_WEIGHTED, "weighted", 1), (OIT_LINKEDLIST, "linkedlist", 32)):
        with Renderer(0) as r:
            r.configure(FrameConfig(algorithm=alg, layers=K, tail_blend=True, width=p.width, height=p.height,
                                    samples=p.samples, ll_pool_nodes=F + 1 if alg == OIT_LINKEDLIST else 0,
                                    variant=variant))
            for _ in range(2):
                r.render(t, d, c, out, st, count=F)
            e0, e1 = Event(), Event()
            st.synchronize()
            e0.record(st)
            for _ in range(steps):
                r.render(t, d, c, out, st, count=F)
            e1.record(st)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / steps
        nbytes = (sum(rl.wboit_stage_bytes(F, P).values()) if alg == OIT_WEIGHTED
                  else rl.ll_frame_bytes_reference(F, P))
        gbs, fr = rl.frac(nbytes, ms / 1e3)
        rows.append({"algorithm": name, "layers": K, "ms_per_frame": round(ms, 4),
                     "G_fragment_samples_per_s": round(F / ms / 1e6, 3), "bytes": nbytes, "frac": round(fr, 4)})
        log("4K 4x: " + json.dumps(rows[-1]))
    return {"workload": "wboit_and_linkedlist_k32_3840x2160x4_4000spheres", "fragment_samples": F,
            "sample_targets": P, "rows": rows}


def bench_batch(seed, frames_n, resolutions=(256, 512, 1024, 2048)):
    """The 64-frame batched workload: every frame one at a time (clear, append,
    resolve per frame, one context each) against the batched Linked List
    (one append launch, one resolve launch per layer group on forked streams,
    replayed from a graph, variants tuned). Device time from HIP events; the
    batched images are checked bit for bit against the one-at-a-time ones.
    `resolutions`: the square frame sizes drawn from."""
    import numpy as np

    from vk_order_independent_transparency_amd import (Batch, BatchFrame, DeviceBuffer, Event, FrameConfig,
                                                       OIT_LINKEDLIST, Renderer, Stream, scene)

    work = scene.batch_workload(seed, frames_n, resolutions)
    st = Stream()
    frames, total = [], 0
    for fr in work:
        g = scene.DeviceScene(fr.scene)
        t, d, c = g.buffers()
        g.emit(t, d, c, st)
        out = DeviceBuffer((fr.scene.targets, 4), "float32")
        frames.append(BatchFrame(width=fr.scene.width, height=fr.scene.height, layers=fr.layers, target=t, depth=d,
                                 rgba=c, count=g.fragments, out=out))
        total += g.fragments
        g.close()
    st.synchronize()
    log(f"batch workload: {len(frames)} frames, {total} fragments")
    rends, seq_out = [], []
    for f in frames:
        r = Renderer(0)
        r.configure(FrameConfig(algorithm=OIT_LINKEDLIST, layers=f.layers, tail_blend=True, width=f.width,
                                height=f.height, ll_pool_nodes=f.count + 1))
        rends.append(r)
        seq_out.append(DeviceBuffer((f.width * f.height, 4), "float32"))
    ev = [(Event(), Event()) for _ in frames]

    def seq_pass():
        for i, (r, f) in enumerate(zip(rends, frames)):
            ev[i][0].record(st)
            r.render(f.target, f.depth, f.rgba, seq_out[i], st, count=f.count)
            ev[i][1].record(st)

    seq_pass()
    st.synchronize()
    seq = []
    for _ in range(3):
        seq_pass()
        st.synchronize()
        seq.append(sum(a.elapsed_ms(b) for a, b in ev))
    seq_ms = min(seq)
    want = [o.numpy() for o in seq_out]
    for r in rends:
        r.close()
    del seq_out
    def timed(frame_order):
        b = Batch(0, concurrent=True, graph=True, frame_order=frame_order)
        b.run(frames, st)
        st.synchronize()
        b.tune(frames, st)
        b.run(frames, st)
        st.synchronize()
        e0, e1 = Event(), Event()
        runs = []
        for _ in range(5):
            e0.record(st)
            b.run(frames, st)
            e1.record(st)
            e1.synchronize()
            runs.append(e0.elapsed_ms(e1))
        ok = all(np.array_equal(f.out.numpy().view(np.uint32), w.view(np.uint32)) for f, w in zip(frames, want))
        t = b.tuning
        b.close()
        return min(runs), ok, t

    order_ms, order_ok, _ = timed(True)
    bat_ms, exact, t = timed(False)
    exact = exact and order_ok
    log(f"batch frame after frame: {order_ms:.3f} ms, interleaved: {bat_ms:.3f} ms")
    log(f"batch: sequential {seq_ms:.3f} ms, batched {bat_ms:.3f} ms, {seq_ms / bat_ms:.2f}x, bit-exact {exact}")
    sizes = "" if tuple(resolutions) == (256, 512, 1024, 2048) else "_" + "_".join(str(r) for r in resolutions)
    return {"workload": f"batch_{len(frames)}_frames_seed{seed}{sizes}", "frames": len(frames), "fragments": total,
            "sequential_ms": round(seq_ms, 4), "batched_ms": round(bat_ms, 4),
            "speedup": round(seq_ms / bat_ms, 3), "G_fragments_per_s": round(total / bat_ms / 1e6, 3),
            "frame_order_ms": round(order_ms, 4),
            "bit_exact_vs_sequential": bool(exact), "append_chunks": int(t.append_chunks),
            "bin_density": float(t.bin_density),
            "binned_frames": sum(1 for f in frames if f.count >= t.bin_density * f.width * f.height * f.samples
                                 and f.count > 0)}


def bench_serve(seed, n, rates, slots):
    """Served workload at each arrival rate: p50/p99 request latency (arrival
    to result on the device) and G fragments/s over the busy span."""
    from vk_order_independent_transparency_amd import serve

    out = []
    srv = serve.Server(0, slots=slots)
    # warm pass: every request of the trace once, all at t = 0, so the
    # contexts and stream buffers reach their resident sizes
    warm, _ = serve.served_workload(seed, n, rates[0])
    srv.replay(warm, [0.0] * len(warm))
    for rate in rates:
        reqs, arr = serve.served_workload(seed, n, rate)
        done = srv.replay(reqs, arr)
        stt = srv.stats()
        ok = [c for c in done if c.status == 0]
        frags = sum(c.fragments for c in ok)
        span_us = max(c.t_done_us for c in ok) - min(c.t_submit_us for c in ok)
        lat = serve.latency_summary(done)
        offered_s = float(arr[-1]) / 1e6
        row = {"rate_per_s": rate, "requests": n, "completed": len(ok), "failed": len(done) - len(ok),
               "offered_s": round(offered_s, 4), "span_s": round(span_us / 1e6, 4),
               # kept up: the requests completed at (within 5 % of) the rate they arrived
               "saturated": bool(span_us / 1e6 > 1.05 * offered_s),
               "p50_ms": round(lat["p50_ms"], 3), "p99_ms": round(lat["p99_ms"], 3),
               "G_fragments_per_s": round(frags / span_us / 1e3, 3),
               "peak_held_GB": round(stt.peak_held_bytes / 2 ** 30, 2), "peak_in_flight": int(stt.peak_in_flight)}
        log(json.dumps(row))
        out.append(row)
    # the same requests all at once: the slots against one request at a time
    # (one slot, so each request's launches run alone on the device); both
    # with warm resident contexts
    reqs, _ = serve.served_workload(seed, n, rates[-1])
    burst = serve_burst(serve, srv, reqs, slots)
    log("served burst vs one at a time: " + json.dumps(burst))
    srv.close()
    # the same generator with 256x256 and 512x512 scenes only: requests whose
    # device work is short, so the per-request steps (count, sizing, launches)
    # weigh more
    small, _ = serve.served_workload(seed, n, rates[-1], resolutions=((256, 256), (512, 512)))
    srv = serve.Server(0, slots=slots)
    srv.replay(small, [0.0] * len(small))
    burst_small = serve_burst(serve, srv, small, slots)
    log("served small requests, burst vs one at a time: " + json.dumps(burst_small))
    srv.close()
    return out, burst, burst_small


def serve_burst(serve, srv, reqs, slots):
    """All of `reqs` submitted at once on the warm server `srv` against the
    same requests on a warm one-slot server: fragments per second over each
    busy span."""
    burst = {}
    for name, server in (("slots", srv), ("one_at_a_time", None)):
        one = server is None
        if one:
            server = serve.Server(0, slots=1)
            server.replay(reqs, [0.0] * len(reqs))
        done = server.replay(reqs, [0.0] * len(reqs))
        ok = [c for c in done if c.status == 0]
        span_us = max(c.t_done_us for c in ok) - min(c.t_submit_us for c in ok)
        # the requests' own device time (count, generation, OIT passes; host gaps
        # excluded) over the span: near 1 one at a time means the device is never idle
        # between requests
        device_us = sum(c.device_us for c in ok)
        burst[name] = {"G_fragments_per_s": round(sum(c.fragments for c in ok) / span_us / 1e3, 3),
                       "span_s": round(span_us / 1e6, 4), "completed": len(ok),
                       "device_busy_fraction": round(device_us / span_us, 3)}
        if one:
            server.close()
    burst["slots"]["slots"] = slots
    burst["speedup"] = round(burst["slots"]["G_fragments_per_s"] / burst["one_at_a_time"]["G_fragments_per_s"], 3)

    return burst


def _under_load(served):
    """p50/p99 at the highest arrival rate the GPU keeps up with (else the lowest rate)."""
    if not served:
        return None
    kept = [r for r in served if not r["saturated"]]
    row = kept[-1] if kept else served[0]
    return {"p50": row["p50_ms"], "p99": row["p99_ms"],
            "under_load": f"served workload at {row['rate_per_s']:g} requests/s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--layers", type=int, default=16)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=0, help="OIT_VARIANT_* bits (kernel schedule; results identical)")
    ap.add_argument("--no-algorithms", action="store_true", help="skip the other algorithms on the same stream")
    ap.add_argument("--only-algorithms", type=str, default="",
                    help="comma-separated name:K rows of the algorithms pass (e.g. loop64:16,spinlock:16)")
    ap.add_argument("--algorithm-steps", type=int, default=50, help="timed frames per algorithm row")
    ap.add_argument("--no-4k", action="store_true", help="skip the 3840x2160x4 WBOIT / Linked List K=32 frames")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-frame batched workload")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true", help="skip the served workload")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rates", type=str, default="100,300,600", help="arrival rates, requests/s")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="also run one device-to-device copy of MB megabytes (a known-traffic kernel for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the frame and request workloads "
              f"shard by frame or screen tile with no cross-GPU reduction, and this build measures one GPU.",
              file=sys.stderr)
        sys.exit(2)

    from vk_order_independent_transparency_amd import (DeviceBuffer, Event, FrameConfig, OIT_LINKEDLIST, Renderer,
                                                       Stream, scene)
    from vk_order_independent_transparency_amd import roofline as rl
    import numpy as np

    K = args.layers
    p = scene.HD_1080P
    t0 = time.perf_counter()
    tg, d, c = scene.generate(p)
    F = len(tg)
    P = p.targets
    cap = F + F // 2
    log(f"workload: {F} fragments, {P} pixels, generated in {time.perf_counter() - t0:.2f}s")

    stream = Stream()
    tg_d = DeviceBuffer.from_numpy(tg, stream)
    d_d = DeviceBuffer.from_numpy(d, stream)
    c_d = DeviceBuffer.from_numpy(c, stream)
    out = DeviceBuffer((P, 4), "float32")
    r = Renderer(0)
    r.configure(FrameConfig(algorithm=OIT_LINKEDLIST, layers=K, tail_blend=True, width=p.width,
                            height=p.height, ll_pool_nodes=cap, variant=args.variant))

    for _ in range(args.warmup):
        r.render(tg_d, d_d, c_d, out, stream)
    stream.synchronize()
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src, dst = DeviceBuffer((nbytes,), "uint8"), DeviceBuffer((nbytes,), "uint8")
        from vk_order_independent_transparency_amd import _abi
        _abi.check("oit_memcpy_d2d", _abi.lib().oit_memcpy_d2d(dst.handle, src.handle, nbytes, stream.handle))
        stream.synchronize()
        log(f"calibration copy: {nbytes} bytes device to device")
        del src, dst

    # timed region: exactly `steps` frames between two HIP events on the
    # launch stream, synchronised on both sides
    e_start, e_end = Event(), Event()
    stream.synchronize()
    e_start.record(stream)
    for i in range(args.steps):
        r.render(tg_d, d_d, c_d, out, stream)
    e_end.record(stream)
    stream.synchronize()
    total_ms = e_start.elapsed_ms(e_end)
    ms_per_step = total_ms / args.steps
    value = F / (ms_per_step / 1e3) / 1e9
    log(f"{ms_per_step:.4f} ms/frame -> {value:.2f} G fragments/s")

    # per-frame latency (separate pass: an event after every frame)
    ev = [Event() for _ in range(args.steps + 1)]
    stream.synchronize()
    ev[0].record(stream)
    for i in range(args.steps):
        r.render(tg_d, d_d, c_d, out, stream)
        ev[i + 1].record(stream)
    stream.synchronize()
    per_frame = sorted(ev[i].elapsed_ms(ev[i + 1]) for i in range(args.steps))

    # per-kernel durations on the launch stream (separate pass)
    r.profile_enable(True)
    r.profile_reset()
    for _ in range(args.steps):
        r.render(tg_d, d_d, c_d, out, stream)
    stream.synchronize()
    prof = r.profile_read()
    r.profile_enable(False)
    stage_bytes = rl.ll_stage_bytes(F, P, fused=not (args.variant & 1))
    stages = {}
    for name, (ms, n) in prof.items():
        if n:
            stages[name] = {"avg_ms": ms / n, "launches": int(n)}
            if name in stage_bytes:
                gbs, fr = rl.frac(stage_bytes[name], ms / n / 1e3)
                stages[name].update({"bytes": stage_bytes[name], "GBps": round(gbs, 1), "frac": round(fr, 4)})
    dom = max((s for s in stages if s in stage_bytes), key=lambda s: stages[s]["avg_ms"])
    frame_gbs, frame_frac = rl.frac(rl.ll_frame_bytes_reference(F, P), ms_per_step / 1e3)
    log(json.dumps(stages))

    st = r.stats(stream)
    r.close()
    only = {x.strip() for x in args.only_algorithms.split(",") if x.strip()}
    counts = np.bincount(tg[tg < P].astype(np.int64), minlength=P)  # fragments per target
    algorithms = None if args.no_algorithms else bench_algorithms(tg_d, d_d, c_d, F, p, args.algorithm_steps, stream,
                                                                  only, args.variant, counts)

    frames_4k = None if args.no_4k else bench_4k_4x(10, args.variant)
    batch = None if args.no_batch else bench_batch(1, args.batch_frames)
    # the same generator with only 256x256 and 512x512 frames: per-frame launch
    # and clear costs weigh more than the fragment work
    batch_small = None if args.no_batch else bench_batch(1, args.batch_frames, (256, 512))
    served = served_burst = served_small = None
    if not args.no_serve:
        served, served_burst, served_small = bench_serve(1, args.serve_requests,
                                                         [float(x) for x in args.serve_rates.split(",")],
                                                         args.serve_slots)

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(tg, d, c, K, cap, args.cpu_budget)
        log(f"cpu baseline: {cpu['value']:.4f} G fragments/s")
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        cpu_all = cpu_baseline_all_cores(tg, d, c, K, args.cpu_budget / 2, threads)
        log(f"cpu baseline, {threads} threads: {cpu_all['value']:.4f} G fragments/s")

    result = {
        "metric": "G fragments/s resolved + % HBM roofline; p50/p99 composite latency under load",
        "value": round(value, 3),
        "unit": "G fragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"linkedlist_k{K}_1920x1080_1000spheres",
            "algorithm": "linkedlist",
            "layers": K,
            "width": p.width,
            "height": p.height,
            "spheres": p.spheres,
            "seed": p.seed,
            "fragments": F,
            "pool_nodes": cap,
            "tail_blend": True,
            "variant": args.variant,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(stages[dom]["GBps"], 1),
            "peak": rl.HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(stages[dom]["frac"], 4),
            "traffic": MEASURED_TRAFFIC.get((f"linkedlist_k{K}_1920x1080_1000spheres", dom, args.variant)),
        },
        "frame_roofline": {"bytes": rl.ll_frame_bytes_reference(F, P), "GBps": round(frame_gbs, 1),
                           "frac": round(frame_frac, 4)},
        "frame_latency_ms": {"p50": round(per_frame[len(per_frame) // 2], 4),
                             "p99": round(per_frame[min(len(per_frame) - 1, int(0.99 * len(per_frame)))], 4)},
        "latency_ms": _under_load(served),
        "algorithms": algorithms,
        "frames_4k_4x": frames_4k,
        "batch": batch,
        "batch_small_frames": batch_small,
        "serve": served,
        "serve_vs_one_at_a_time": served_burst,
        "serve_small_requests_vs_one_at_a_time": served_small,
        "stages": stages,
        "ll_overflow": int(st.ll_overflow),
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
