"""Benchmark of the DDM analysis path on one MI355X.

Default workload (the largest single-GPU configuration): 1000 frames of
2048x2048 uint16 (12-bit) synthetic Brownian video, ROI 1024^2, scales
1024/512/256/128/64/32, 64 log-spaced lags 1..999, 40 log-spaced lambdas
2..512 px, mask tolerance 20 %, chunk 30, FP32 path; lag block, FFT batch and
transform families chosen by the engine's autotuner at engine set-up (timed
once, then read from its cache).  One step = one whole analysis of the video
with the frames resident in HBM when the timer starts (engine creation,
autotuning included, outside the timed region).

Prints one JSON line: analysed frames/s; the dominant kernel's roofline
(algorithmic bytes / launch time from HIP events on the engine's stream,
against 8 TB/s); chunk-to-result latency of the same video streamed chunk by
chunk from pinned host memory; the PCIe-inclusive rate; the max / mean
relative I(q, tau) error of tile 0 of every scale against the FP64 oracle;
and a multithreaded FP32 CPU baseline timed on this host.

`--serving` runs the many-stream serving workload instead.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MI355X_CUS, MI355X_CLOCK_HZ = 256, 2.4e9  # compute units, peak engine clock
CHECK_TOL = 1e-4  # largest relative error against the FP64 oracle the metric allows
METRIC = "analysed frames/s (all scales x lags) + p99 chunk latency, at <=1e-4 err vs FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


SPECTRUM = {"masked": 0, "full": 1}  # ddm_desc.spectrum


def pmc_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3
    PMC summary profiles/pmc_traffic_<workload>.md (FETCH_SIZE and WRITE_SIZE
    passes, scaled by a calibration copy): its line
    `traffic_bytes_per_launch <kernel> <bytes>`."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.md")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def distinct_frames(t0, n, lags):
    s = set(range(t0, t0 + n))
    for t in range(t0, t0 + n):
        for tau in lags:
            if t - tau >= 0:
                s.add(t - tau)
    return len(s)


def lag_engine_bytes(E, lags, n_frames, block, compensated=True):
    """Algorithmic HBM bytes of every lag-engine launch of one analysis (the
    engine runs one launch per block of `block` frames, the last one partial):
    each distinct ring frame a launch pairs with read once (8 B per complex
    element) plus the FP32 accumulators of every lag read and written once,
    and with compensated accumulation their FP32 compensation terms too."""
    acc = 16 if compensated else 8
    total, launches = 0, 0
    for t0 in range(0, n_frames, block):
        m = min(block, n_frames - t0)
        total += 8 * E * distinct_frames(t0, m, lags) + acc * E * len(lags)
        launches += 1
    return total, launches


def percentile(a, p):
    return float(np.percentile(np.asarray(a, dtype=np.float64), p)) if len(a) else None


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"model": model, "hardware_concurrency": os.cpu_count(), "affinity": affinity}


def serving(args):
    """Serving workload: `--streams` streams of the serving mix (frame size
    uniform over 256^2 / 512^2 / 1024^2 uint8, D uniform in 0.2..2.0, scales
    from the frame size down to 32, 24 log lags up to 300, 30 log lambdas
    4..256), chunks of 30 frames arriving as a Poisson process at --rate
    chunks/s in aggregate for --seconds.  Streams are admitted by device
    footprint; with --chunks-per-stream 0 every admitted stream stays
    resident, else a stream leaves after that many chunks and a waiting one
    takes its place.  Reports frames/s and chunk-to-result latency
    percentiles."""
    import torch

    import cuda_multiddm_amd as ddm
    from cuda_multiddm_amd.workloads import serving_streams

    streams = serving_streams(args.streams)
    rng = np.random.default_rng(11)
    pool = {}
    for size in sorted({s.size for s in streams}):
        proto = next(s for s in streams if s.size == size)
        f = proto.make_frames(30)
        pinned = torch.empty(f.size, dtype=torch.uint8, pin_memory=True).numpy()  # uploads skip the staging copy
        pinned[:] = f.reshape(-1)
        pool[size] = pinned

    srv = ddm.Server(max_streams=args.streams, batched=args.batched, group=args.group, hold_ms=args.hold_ms)
    waiting = list(range(args.streams))
    active, sent = {}, {}
    admitted, refused = 0, 0
    host_s = {"open": 0.0, "submit": 0.0, "results": 0.0, "close": 0.0}

    def timed(name, fn, *a):
        t = time.perf_counter()
        try:
            return fn(*a)
        finally:
            host_s[name] += time.perf_counter() - t

    def admit():
        nonlocal admitted, refused
        while waiting:
            k = waiting[0]
            try:
                sid = timed("open", lambda: srv.open(spectrum=SPECTRUM[args.spectrum],
                                                     **streams[k].desc(args.serving_lag_block)))
            except ddm.DDMError as e:
                if e.code != ddm.DDM_EBUSY:
                    raise
                refused += 1
                return
            waiting.pop(0)
            active[k] = sid
            sent[k] = 0
            admitted += 1

    admit()
    log(f"serving: {len(active)} streams resident, offered {args.rate} chunks/s")
    t_start = time.perf_counter()
    t_next = t_start
    frames = 0
    completed = 0
    while time.perf_counter() - t_start < args.seconds:
        now = time.perf_counter()
        if now < t_next:
            time.sleep(min(t_next - now, 0.002))
            continue
        t_next += rng.exponential(1.0 / args.rate)
        if not active:
            admit()
            continue
        keys = list(active)
        k = keys[int(rng.integers(len(keys)))]
        timed("submit", srv.submit, active[k], pool[streams[k].size])
        frames += 30
        sent[k] += 1
        if args.chunks_per_stream and sent[k] >= args.chunks_per_stream:
            timed("results", srv.results, active[k])
            timed("close", srv.close_stream, active.pop(k))
            completed += 1
            waiting.append(k)  # the stream comes back later as a new request
            admit()
    srv.drain()
    elapsed = time.perf_counter() - t_start
    lat = srv.latencies()
    st = srv.stats()
    sizes = [streams[k].size for k in active]
    out = {
        "metric": "serving: analysed frames/s + chunk-to-result latency",
        "value": round(frames / elapsed, 1),
        "unit": "frames/s",
        "offered_chunks_per_s": args.rate,
        "seconds": round(elapsed, 2),
        "p50_chunk_ms": percentile(lat, 50),
        "p99_chunk_ms": percentile(lat, 99),
        "max_chunk_ms": float(np.max(lat)) if len(lat) else None,
        "chunks": int(st.chunks),
        "streams_admitted": admitted,
        "streams_completed": completed,
        "admission_refusals": refused,
        "active_streams_at_end": st.streams,
        "active_sizes": {str(s): sizes.count(s) for s in (256, 512, 1024)},
        "ticks": int(st.ticks), "graph_ticks": int(st.graph_ticks),
        "device_budget_GB": round(st.budget_bytes / 1e9, 1),
        "device_used_GB_at_end": round(st.used_bytes / 1e9, 1),
        "host_seconds": {k: round(v, 2) for k, v in host_s.items()},
        "config": {"workload": f"{args.streams} streams, frame sizes {{256,512,1024}}^2 u8, scales to 32, "
                               f"24 log lags <= 300, 30 log lambdas 4-256, chunk 30, lag block "
                               f"{args.serving_lag_block}, Poisson arrivals {args.rate} chunks/s, "
                               + ("streams resident" if not args.chunks_per_stream else
                                  f"streams replaced after {args.chunks_per_stream} chunks")
                               + (", batched launches" if args.batched else ", per-stream launches")
                               + (f", chunks grouped per lag pass (hold {args.hold_ms} ms)" if args.group else "")
                               + f", {args.spectrum} spectrum"},
    }
    srv.close()
    print(json.dumps(out))


def fp64_check(W, raw, res, out):
    """Tile 0 of every scale over all frames against the FP64 oracle."""
    from cuda_multiddm_amd.workloads import host_threads
    from oracle import oracle as o
    cfg = o.Config(W.size, W.size, W.frames, W.scales, W.lags, W.lambdas, np.float32(W.tol).item(), W.bpp,
                   W.pixfmt)
    per = o.split_results(cfg, res)
    errs, mean_errs = {}, {}
    for si, s in enumerate(W.scales):
        ref = o.split_results(cfg, o.run(cfg, raw, "f64", nthreads=host_threads(), only_scale=si, tiles=[0]))[s][0]
        g = per[s][0]
        errs[s] = o.max_rel_err(g, ref)
        nz = ref != 0
        mean_errs[s] = float(np.mean(np.abs(g[nz] - ref[nz]) / np.abs(ref[nz]))) if nz.any() else 0.0
        log(f"FP64 check scale {s} tile 0: max {errs[s]:.3e} mean {mean_errs[s]:.3e}")
    out["max_rel_err_vs_fp64"] = {f"scale{s}_tile0": float(f"{v:.3e}") for s, v in errs.items()}
    out["mean_rel_err_vs_fp64"] = {f"scale{s}_tile0": float(f"{v:.3e}") for s, v in mean_errs.items()}


def cpu_baseline(W, raw, nf, out, budget_s=20.0):
    """The FP32 multithreaded C restatement, built for this host, over nf
    frames: each scale is first timed on its tile 0; a scale whose whole run
    is projected within budget_s seconds (the work of a scale is exactly
    proportional to its tiles) is then run and timed on every tile, the
    others are counted as tile 0's time scaled by their tile count."""
    from cuda_multiddm_amd.workloads import host_threads
    from oracle import oracle as o
    kind = "-O3 -march=native"
    try:
        o.use_library(o.build_native(tempfile.mkdtemp(prefix="ddm_cpu_")))
    except Exception as ex:  # no compiler: the portable build shipped with the tree
        kind = f"-O3 portable (native build failed: {type(ex).__name__})"
        o.use_library(os.path.join(ROOT, "oracle", "_build", "libddm_oracle.so"))
    threads = host_threads()
    lags = [l for l in W.lags]
    cfg = o.Config(W.size, W.size, nf, W.scales, lags, W.lambdas, np.float32(W.tol).item(), W.bpp, W.pixfmt)
    data = raw[:nf * W.size * W.size * W.bpp]
    total, per, how = 0.0, {}, {}
    for si, s in enumerate(W.scales):
        tiles = (W.scales[0] // s) ** 2
        t0 = time.perf_counter()
        o.run(cfg, data, "f32", nthreads=threads, only_scale=si, tiles=[0])
        dt = time.perf_counter() - t0
        if tiles > 1 and dt * tiles <= budget_s:
            t0 = time.perf_counter()
            o.run(cfg, data, "f32", nthreads=threads, only_scale=si)
            dt = time.perf_counter() - t0
            how[s] = "all tiles"
        else:
            how[s] = "all tiles" if tiles == 1 else f"tile 0 x {tiles}"
            dt *= tiles
        log(f"CPU baseline scale {s}: {dt:.2f} s ({how[s]})")
        per[s] = dt
        total += dt
    info = cpu_info()
    out["cpu_baseline"] = {
        "value": round(nf / total, 3),
        "unit": "frames/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 multithreaded C restatement of the pipeline ({kind}, {threads} threads) over {nf} frames, "
                  f"full spectrum, per scale: "
                  + ", ".join(f"{s}: {per[s]:.2f} s ({how[s]})" for s in W.scales),
        "cpu_model": info["model"],
        "hardware_concurrency": info["hardware_concurrency"],
        "cpus_available": info["affinity"],
    }


def streaming_latency(W, raw, out):
    """Chunk-to-result latency: the same video streamed chunk by chunk from
    pinned host memory through a one-stream server (every chunk uploaded,
    transformed and accumulated before it counts as done)."""
    import torch

    import cuda_multiddm_amd as ddm
    h = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True).numpy()
    h[:] = raw
    fb = W.size * W.size * W.bpp
    with ddm.Server(max_streams=1) as srv:
        sid = srv.open(width=W.size, height=W.size, scales=W.scales, lags=W.lags, lambdas=W.lambdas,
                       tolerance=W.tol, chunk=W.chunk, bpp=W.bpp, pixfmt=W.pixfmt, lag_block=32)
        t0 = time.perf_counter()
        for i in range(0, W.frames, W.chunk):
            n = min(W.chunk, W.frames - i)
            srv.submit(sid, h[i * fb:(i + n) * fb].reshape(n, -1))
        srv.drain()
        dt = time.perf_counter() - t0
        lat = srv.latencies(sid)
    out["p50_chunk_ms"] = percentile(lat, 50)
    out["p99_chunk_ms"] = percentile(lat, 99)
    out["latency_run"] = {"frames_per_s": round(W.frames / dt, 1), "chunks": len(lat),
                          "input": f"pinned host memory, {W.chunk}-frame chunks submitted back to back, "
                                   f"lag block 32 (every chunk accumulated before it counts as done)"}
    del h


def pcie_rate(W, raw, make, out):
    """One whole analysis with the frames in pinned host memory (ROI rows
    cross PCIe inside the timed region)."""
    import torch
    h = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True).numpy()
    h[:] = raw
    e = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.submit_host(h, W.frames)
    e.results()
    dt = time.perf_counter() - t0
    e.close()
    out["pcie_inclusive_frames_per_s"] = round(W.frames / dt, 1)
    del h


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="brownian2048_u16",
                    choices=["brownian2048_u16", "brownian512", "uniform128"])
    ap.add_argument("--frames", type=int, default=0, help="frames to analyse (0 = the workload's count)")
    ap.add_argument("--lag-block", type=int, default=0, help="frames per lag-engine pass (0 = engine default)")
    ap.add_argument("--host-input", action="store_true",
                    help="time the steps with frames in pinned host memory (PCIe-inclusive; not the headline)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x2", "fp16x2"],
                    help="64/128 tile transforms: FP32 FFT or split-operand DFT on the matrix cores")
    ap.add_argument("--accumulation", default="compensated", choices=["compensated", "plain"],
                    help="lag sums: compensated FP32 folded into FP64 every 10^4 pairs, or plain FP32")
    ap.add_argument("--spectrum", default="masked", choices=["masked", "full"],
                    help="spectrum elements kept: the band holding every q-ring member, or all of them")
    ap.add_argument("--tiler", default="auto", choices=["auto", "fused", "per-scale"],
                    help="multi-scale tiler: each frame read once for the 8..128 scales and once for the 256..1024 "
                         "row passes (fused), or one transform launch per scale; auto: the autotuner's choice "
                         "(per scale with --no-autotune)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the engine's default schedule instead of the autotuner's choice of lag block, FFT "
                         "batch and transform families (ddm_autotune, run at engine set-up)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=0, help="frames for the CPU baseline sample (0 = all)")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 tile checks")
    ap.add_argument("--no-extras", action="store_true", help="skip the latency and PCIe-inclusive runs")
    ap.add_argument("--profile-steps", action="store_true",
                    help="only run warmup + timed steps (for rocprofv3 runs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="with --profile-steps: finish with a device copy of this many MiB")
    ap.add_argument("--serving", action="store_true", help="run the multi-stream serving workload instead")
    ap.add_argument("--batched", action="store_true", help="serving: batched launches across streams")
    ap.add_argument("--group", action="store_true",
                    help="serving, batched: a tick runs all of a stream's waiting chunks (up to its lag block) as "
                         "one lag pass")
    ap.add_argument("--hold-ms", type=float, default=0.0,
                    help="serving, grouped: how long a stream's chunk may wait for more of its chunks")
    ap.add_argument("--streams", type=int, default=64)
    ap.add_argument("--rate", type=float, default=200.0, help="serving: offered chunks per second")
    ap.add_argument("--seconds", type=float, default=20.0, help="serving: arrival window")
    ap.add_argument("--chunks-per-stream", type=int, default=0,
                    help="serving: chunks after which a stream is replaced (0 = streams stay resident)")
    ap.add_argument("--serving-lag-block", type=int, default=32, choices=[32, 64, 128, 256],
                    help="serving: each stream's lag block (every chunk is flushed through the lag engine)")
    args = ap.parse_args()
    args.autotune = not args.no_autotune
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.serving:
        serving(args)
        return

    import torch

    import cuda_multiddm_amd as ddm
    from cuda_multiddm_amd.workloads import WORKLOADS, host_threads

    W = WORKLOADS[args.workload]
    if args.frames:
        from dataclasses import replace
        W = replace(W, frames=args.frames)
    N, M = W.frames, W.scales[0]
    lags, lams = W.lags, W.lambdas
    log(f"generating {N} frames {W.size}x{W.size} ({W.bits}-bit)")
    frames = W.make_frames(N, host_threads())
    raw = frames.reshape(-1).view(np.uint8)
    log(f"frames ready ({raw.nbytes / 1e9:.2f} GB)")
    torch.cuda.set_device(0)
    if args.host_input:
        h_frames = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True).numpy()
        h_frames[:] = raw
    else:
        d_frames = torch.from_numpy(raw).cuda()
    torch.cuda.synchronize()

    prec = {"fp32": ddm.PREC_FP32, "bf16x2": ddm.PREC_BF16X2, "fp16x2": ddm.PREC_FP16X2}[args.precision]
    tune_cache = os.path.join(tempfile.mkdtemp(prefix="ddm_tune_"), "schedule.txt") if args.autotune else None
    schedule = {}

    def make():
        e = ddm.Engine(*W.engine_args(), lag_block=args.lag_block,
                       accumulation=ddm.ACC_PLAIN if args.accumulation == "plain" else ddm.ACC_COMPENSATED,
                       spectrum=SPECTRUM[args.spectrum])
        e.set_precision(prec)
        if args.autotune:
            # timed on the first engine, read back from the cache by the others
            # (before any frame, outside the timed region)
            schedule["lag_block"], schedule["fft_batch_frames"] = e.autotune(tune_cache)
            schedule["fft_families"] = ["lds" if f == ddm.FFT_LDS else "register" for f in e.fft_families()]
        if args.tiler != "auto":
            e.set_tiler(ddm.TILER_FUSED if args.tiler == "fused" else ddm.TILER_PER_SCALE)
        schedule["tiler"] = "fused" if e.tiler() == ddm.TILER_FUSED else "per-scale"
        # buffers the first submission would allocate (pass buffer, staging)
        # belong to engine creation, outside the timed region
        e.reserve(N)
        if "buf" not in res_host:
            res_host["buf"] = torch.empty(e.result_count, dtype=torch.float64, pin_memory=True).numpy()
        return e

    res_host = {}

    def step(e):
        if args.host_input:
            e.submit_host(h_frames, N)  # ROI rows only cross PCIe
        else:
            e.submit_device(d_frames.data_ptr(), N)
        # results land in pinned host memory, allocated once outside the timed steps
        return e.results(out=res_host["buf"])

    for _ in range(args.warmup):
        with make() as e:
            step(e)
    torch.cuda.synchronize()
    elapsed = 0.0
    res = None
    log("warmup done")
    for _ in range(args.steps):
        e = make()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = step(e)  # ends with a stream synchronisation
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        e.close()
    res = res.copy()  # the pinned buffer is reused by the runs below
    fps = args.steps * N / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    if args.profile_steps:
        if args.calib_copy_mib:
            # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE
            # calibrate the counters of the engine's kernels
            a = torch.ones(args.calib_copy_mib << 20, dtype=torch.uint8, device="cuda")
            b = torch.empty_like(a)
            torch.cuda.synchronize()
            b.copy_(a)
            torch.cuda.synchronize()
        print(json.dumps({"frames_per_s": fps, "ms_per_step": ms_per_step}))
        return
    log(f"timed steps: {ms_per_step:.2f} ms per step, {fps:.0f} frames/s")

    # stage timing with HIP events on the engine's stream (separate run)
    with make() as e:
        e.set_timing(True)
        step(e)
        st = e.stage_stats()
        batch_ms = e.chunk_times()
        info = e.info()
        lag_block = info.lag_block
    E = info.elems_per_frame
    lag_bytes, lag_launches = lag_engine_bytes(E, lags, N, lag_block, args.accumulation == "compensated")
    assert lag_launches == st.lag_launches, (lag_launches, st.lag_launches)
    lag_avg_ms = st.lag_ms / st.lag_launches
    achieved = lag_bytes / (st.lag_ms * 1e-3) / 1e9
    # whole-pipeline algorithmic bytes per frame (ROI read, spectra write,
    # lag engine reads + accumulators), for context
    pipe_bytes = (M * M * W.bpp + 8 * E) * N + lag_bytes
    pipe_gbs = pipe_bytes / (elapsed / args.steps) / 1e9

    kernel = f"lag_kernel<{lag_block}, {'true' if args.accumulation == 'compensated' else 'false'}>"
    traffic, traffic_src = pmc_traffic(args.workload, kernel.replace(" ", ""))
    # the lag engine's compute floor: every valid pair costs four FP32 VALU
    # instructions per element (two subtractions, two fused multiply-adds),
    # and a CU issues at most one wave64 VALU instruction per clock
    # (4 SIMDs x 16 lanes) at the 2.4 GHz peak engine clock
    pairs = sum(max(0, N - tau) for tau in lags) * E
    valu_floor_ms = pairs * 4 / 64 / lag_launches / (MI355X_CUS * MI355X_CLOCK_HZ) * 1e3
    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": W.text.replace(f"{WORKLOADS[args.workload].frames}x", f"{N}x"),
            "name": args.workload,
            "input": "pinned host memory (PCIe-inclusive)" if args.host_input else "device memory",
            "frames": N, "frame_size": W.size, "scales": W.scales, "lags": len(lags), "lambdas": len(lams),
            "chunk": W.chunk, "elems_per_frame": E, "ring_slots": info.ring_slots,
            # device-resident frames are transformed whole chunks at a time, up to
            # ~128 frames per batch (the engine's FFT batch)
            "fft_batch_frames": schedule.get("fft_batch_frames", W.chunk * max(1, 128 // W.chunk)),
            "lag_block": lag_block, "schedule": "autotuned" if args.autotune else "default",
            "fft_families": schedule.get("fft_families", ["register"] * len(W.scales)),
            "precision": args.precision, "accumulation": args.accumulation, "spectrum": args.spectrum, "tiler": schedule.get("tiler"),
        },
        "stage_ms_per_step": {"fft": round(st.fft_ms, 3), "lag": round(st.lag_ms, 3),
                              "azimuthal": round(st.azimuthal_ms, 3)},
        "batch_device_ms": {"p50": percentile(batch_ms, 50), "p99": percentile(batch_ms, 99),
                            "what": "device time of each FFT batch (tile + FFT + the lag launches it completes)"},
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": int(lag_bytes / lag_launches),
            "avg_launch_ms": round(lag_avg_ms, 4),
            "launches": lag_launches,
            "valu_floor": {"pairs_per_launch": int(pairs / lag_launches), "ms": round(valu_floor_ms, 4),
                           "frac": round(valu_floor_ms / lag_avg_ms, 4),
                           "assumes": "4 FP32 VALU instructions per pair and element, one wave64 VALU "
                                      "instruction per CU per clock, 256 CUs at 2.4 GHz"},
        },
        "pipeline_algorithmic_GBps": round(pipe_gbs, 1),
    }
    if prec != ddm.PREC_FP32:
        # matrix-core work of the split-operand DFT: 3 products in the row stage
        # and 6 in the column stage, each S^3 multiply-adds per tile
        out["mfma_flops_per_step"] = sum(18 * s ** 3 * (M // s) ** 2 for s in W.scales if s in (64, 128)) * N
    if not args.no_extras and not args.host_input:
        if not args.frames or args.frames == WORKLOADS[args.workload].frames:
            del d_frames
        streaming_latency(W, raw, out)
        log(f"chunk-to-result latency p50 {out['p50_chunk_ms']:.2f} ms p99 {out['p99_chunk_ms']:.2f} ms")
        pcie_rate(W, raw, make, out)
        log(f"PCIe-inclusive {out['pcie_inclusive_frames_per_s']:.0f} frames/s")
    if not args.no_check:
        fp64_check(W, raw, res, out)
    if not args.no_cpu_baseline:
        cpu_baseline(W, raw, min(args.cpu_frames or N, N), out)
    worst = max(out.get("max_rel_err_vs_fp64", {}).values(), default=0.0)
    if worst > CHECK_TOL:
        # a number whose results miss the FP64 check is not a measurement
        print(json.dumps(out), file=sys.stderr)
        sys.exit(f"[bench] FP64 check failed: max relative error {worst:.3e} > {CHECK_TOL:g}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
