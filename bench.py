"""Benchmark of the DDM analysis path on one MI355X.

Workload (one step = one whole analysis of the video, inputs resident in
HBM when the timer starts): 1024 frames of 512x512 uint8 synthetic
Brownian video (2000 Gaussian spots, sigma 2 px, amplitude 80, D = 0.5
px^2/frame, background 100, noise sigma 5), scales 512/256/128/64, 32
log-spaced lags 1..512, 30 log-spaced lambdas 4..256 px, mask tolerance
20 %, chunk 30 frames, plain FP32 path.

Prints one JSON line: analysed frames/s, p50/p99 chunk latency, the lag
engine's roofline (algorithmic bytes / measured kernel time against 8 TB/s
HBM), the max relative I(q, tau) error of checked tiles against the FP64
oracle, and a multithreaded FP32 CPU baseline timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MI355X_CUS, MI355X_CLOCK_HZ = 256, 2.4e9  # compute units, peak engine clock
CHECK_TOL = 1e-4  # largest relative error against the FP64 oracle the metric allows


def pmc_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3
    PMC summary profiles/pmc_traffic_<workload>.md (FETCH_SIZE and WRITE_SIZE
    passes, scaled by a calibration copy): its line
    `traffic_bytes_per_launch <kernel> <bytes>`."""
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "profiles", f"pmc_traffic_{workload}.md")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2]), os.path.relpath(path, here)
    return None, None

WORKLOADS = {
    # the headline workload: 1024 frames of 512^2 uint8 Brownian video
    "brownian512": dict(
        frames=1024, size=512, bits=8, scales=[512, 256, 128, 64], lag_max=512, lag_count=32,
        lam_lo=4.0, lam_hi=256.0, lam_count=30, tol=1.2, chunk=30, cpu_frames=0,
        spots=2000, sigma=2.0, amplitude=80.0, diffusion=0.5, background=100.0, noise=5.0, seed=2026,
        text="single-stream Brownian 1024x512^2 u8, scales 512/256/128/64, 32 log lags 1-512, "
             "30 log lambdas 4-256 px, tol 20%, chunk 30, plain FP32"),
    # 1000 frames of 2048^2 uint16 with 12-bit values; the analysed ROI is the
    # largest scale (1024^2 at the origin)
    "brownian2048_u16": dict(
        frames=1000, size=2048, bits=12, scales=[1024, 512, 256, 128, 64, 32], lag_max=999, lag_count=64,
        lam_lo=2.0, lam_hi=512.0, lam_count=40, tol=1.2, chunk=30, cpu_frames=60,
        spots=8000, sigma=2.0, amplitude=80.0, diffusion=1.0, background=1000.0, noise=30.0, seed=2027,
        text="single-stream Brownian 1000x2048^2 u16 (12-bit), ROI 1024^2, scales 1024..32, 64 log lags "
             "1-999, 40 log lambdas 2-512 px, tol 20%, chunk 30, plain FP32"),
}
WORKLOAD = WORKLOADS["brownian512"]


def distinct_frames(t0, n, lags):
    s = set(range(t0, t0 + n))
    for t in range(t0, t0 + n):
        for tau in lags:
            if t - tau >= 0:
                s.add(t - tau)
    return len(s)


def lag_engine_bytes(E, lags, n_frames, block):
    """Algorithmic HBM bytes of every lag-engine launch of one analysis (the
    engine runs one launch per block of `block` frames, the last one partial):
    each distinct ring frame a launch pairs with read once (8 B per complex
    element) plus the FP32 accumulators of every lag read and written once."""
    total, launches = 0, 0
    for t0 in range(0, n_frames, block):
        m = min(block, n_frames - t0)
        total += 8 * E * distinct_frames(t0, m, lags) + 8 * E * len(lags)
        launches += 1
    return total, launches


def percentile(a, p):
    return float(np.percentile(np.asarray(a, dtype=np.float64), p)) if len(a) else None


def serving(args):
    """Serving workload: 64 streams of 256^2 / 512^2 / 1024^2 uint8 frames
    (size drawn uniformly), scales from the frame size down to 32, 24 log
    lags up to 300, 30 log lambdas 4..256, chunks of 30 frames arriving as a
    Poisson process at --rate chunks/s in aggregate for --seconds.  Streams
    are admitted by device footprint; a stream leaves after --chunks-per-stream
    chunks and a waiting one takes its place.  Reports frames/s and
    chunk-to-result latency percentiles."""
    import cuda_multiddm_amd as ddm

    rng = np.random.default_rng(7)
    sizes = [int(s) for s in rng.choice([256, 512, 1024], size=args.streams)]
    lags = ddm.log_lags(300, 24)
    lams = ddm.log_lambdas(4, 256, 30)
    import torch
    pool = {}
    for s in (256, 512, 1024):
        f = ddm.synth_brownian(30, s, s, 8, max(1, 2000 * s * s // (512 * 512)), seed=s,
                               nthreads=min(os.cpu_count() or 1, 16))
        pinned = torch.empty(f.size, dtype=torch.uint8, pin_memory=True).numpy()  # uploads skip the staging copy
        pinned[:] = f.reshape(-1)
        pool[s] = pinned

    def desc(s):
        sc = [1 << k for k in range(int(np.log2(s)), 4, -1)]
        return dict(width=s, height=s, scales=sc, lags=lags, lambdas=lams, chunk=30, lag_block=args.serving_lag_block)

    srv = ddm.Server(max_streams=args.streams)
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
                sid = timed("open", lambda: srv.open(**desc(sizes[k])))
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
        k = list(active)[rng.integers(len(active))]
        timed("submit", srv.submit, active[k], pool[sizes[k]])
        frames += 30
        sent[k] += 1
        if sent[k] >= args.chunks_per_stream:
            timed("results", srv.results, active[k])
            timed("close", srv.close_stream, active.pop(k))
            completed += 1
            waiting.append(k)  # the stream comes back later as a new request
            admit()
    srv.drain()
    elapsed = time.perf_counter() - t_start
    lat = srv.latencies()
    st = srv.stats()
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
        "device_budget_GB": round(st.budget_bytes / 1e9, 1),
        "device_used_GB_at_end": round(st.used_bytes / 1e9, 1),
        "host_seconds": {k: round(v, 2) for k, v in host_s.items()},
        "config": {"workload": f"{args.streams} streams, frame sizes {{256,512,1024}}^2 u8, scales to 32, "
                               f"24 log lags <= 300, 30 log lambdas 4-256, chunk 30, lag block {args.serving_lag_block}, "
                               f"Poisson arrivals {args.rate} chunks/s"},
    }
    srv.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="brownian512", choices=sorted(WORKLOADS))
    ap.add_argument("--frames", type=int, default=0, help="frames to analyse (0 = the workload's count)")
    ap.add_argument("--lag-block", type=int, default=0, help="frames per lag-engine pass (0 = engine default)")
    ap.add_argument("--host-input", action="store_true",
                    help="frames start in pinned host memory (PCIe-inclusive rate; not the headline value)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x2", "fp16x2"],
                    help="64/128 tile transforms: FP32 FFT or split-operand DFT on the matrix cores")
    ap.add_argument("--autotune", action="store_true",
                    help="let the engine's autotuner pick the lag block and FFT batch (ddm_autotune)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=0, help="frames for the CPU baseline sample (0 = all)")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 tile checks")
    ap.add_argument("--profile-steps", action="store_true",
                    help="only run warmup + timed steps (for rocprofv3 runs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="with --profile-steps: finish with a device copy of this many MiB")
    ap.add_argument("--serving", action="store_true", help="run the multi-stream serving workload instead")
    ap.add_argument("--streams", type=int, default=64)
    ap.add_argument("--rate", type=float, default=200.0, help="serving: offered chunks per second")
    ap.add_argument("--seconds", type=float, default=20.0, help="serving: arrival window")
    ap.add_argument("--chunks-per-stream", type=int, default=12)
    ap.add_argument("--serving-lag-block", type=int, default=32, choices=[32, 64, 128, 256],
                    help="serving: each stream's lag block (every chunk is flushed through the lag engine)")
    args = ap.parse_args()
    if args.serving and args.gpus == 1:
        serving(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import cuda_multiddm_amd as ddm

    W = dict(WORKLOADS[args.workload])
    if args.frames:
        W["frames"] = args.frames
    N, S = W["frames"], W["size"]
    M = W["scales"][0]  # analysed ROI: M x M at the frame origin
    bpp, pixfmt = (1, ddm.PIX_U8) if W["bits"] <= 8 else (2, ddm.PIX_U16LE)
    lags = ddm.log_lags(W["lag_max"], W["lag_count"])
    lams = ddm.log_lambdas(W["lam_lo"], W["lam_hi"], W["lam_count"])
    print(f"[bench] generating {N} frames {S}x{S}", file=sys.stderr, flush=True)
    frames = ddm.synth_brownian(N, S, S, W["bits"], W["spots"], W["sigma"], W["amplitude"], W["diffusion"],
                                W["background"], W["noise"], W["seed"], min(os.cpu_count() or 1, 16))
    raw = frames.reshape(-1).view(np.uint8)
    print(f"[bench] frames ready ({raw.nbytes / 1e9:.2f} GB)", file=sys.stderr, flush=True)
    torch.cuda.set_device(0)
    if args.host_input:
        h_frames = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True).numpy()
        h_frames[:] = raw
    else:
        d_frames = torch.from_numpy(raw).cuda()
    torch.cuda.synchronize()

    prec = {"fp32": ddm.PREC_FP32, "bf16x2": ddm.PREC_BF16X2, "fp16x2": ddm.PREC_FP16X2}[args.precision]

    tune_cache = None
    if args.autotune:
        import tempfile
        tune_cache = os.path.join(tempfile.mkdtemp(prefix="ddm_tune_"), "schedule.txt")
    schedule = {}

    def make():
        e = ddm.Engine(S, S, W["scales"], lags, lams, W["tol"], W["chunk"], bpp=bpp, pixfmt=pixfmt,
                       lag_block=args.lag_block)
        e.set_precision(prec)
        if args.autotune:
            # timed on the first engine, read back from the cache by the others
            # (before any frame, outside the timed region)
            schedule["lag_block"], schedule["fft_batch_frames"] = e.autotune(tune_cache)
        # buffers the first submission would allocate (pass buffer, staging)
        # belong to engine creation, outside the timed region
        e.reserve(N)
        return e

    def step(e):
        if args.host_input:
            e.submit_host(h_frames, N)  # ROI rows only cross PCIe
        else:
            e.submit_device(d_frames.data_ptr(), N)
        return e.results()

    for _ in range(args.warmup):
        with make() as e:
            step(e)
    torch.cuda.synchronize()
    elapsed = 0.0
    res = None
    print("[bench] warmup done", file=sys.stderr, flush=True)
    for _ in range(args.steps):
        e = make()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = step(e)  # ends with a stream synchronisation
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        e.close()
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

    # stage timing with HIP events on the engine's stream (separate run)
    with make() as e:
        e.set_timing(True)
        step(e)
        st = e.stage_stats()
        chunk_ms = e.chunk_times()
        info = e.info()
        lag_block = info.lag_block
    E = info.elems_per_frame
    lag_bytes, lag_launches = lag_engine_bytes(E, lags, N, lag_block)
    assert lag_launches == st.lag_launches, (lag_launches, st.lag_launches)
    lag_avg_ms = st.lag_ms / st.lag_launches
    achieved = lag_bytes / (st.lag_ms * 1e-3) / 1e9
    # whole-pipeline algorithmic bytes per frame (ROI read, spectra write,
    # lag engine reads + accumulators), for context
    pipe_bytes = (M * M * bpp + 8 * E) * N + lag_bytes
    pipe_gbs = pipe_bytes / (elapsed / args.steps) / 1e9

    traffic, traffic_src = pmc_traffic(args.workload, f"lag_kernel<{lag_block}>")
    # the lag engine's compute floor: every valid pair costs four FP32 VALU
    # instructions per element (two subtractions, two fused multiply-adds),
    # and a CU issues at most one wave64 VALU instruction per clock
    # (4 SIMDs x 16 lanes) at the 2.4 GHz peak engine clock
    pairs = sum(max(0, N - tau) for tau in lags) * E
    valu_floor_ms = pairs * 4 / 64 / lag_launches / (MI355X_CUS * MI355X_CLOCK_HZ) * 1e3
    out = {
        "metric": "analysed frames/s (all scales x lags) + p99 chunk latency, at <=1e-4 err vs FP64",
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
            "workload": W["text"].replace(f"{WORKLOADS[args.workload]['frames']}x", f"{N}x"),
            "name": args.workload,
            "input": "pinned host memory (PCIe-inclusive)" if args.host_input else "device memory",
            "frames": N, "frame_size": S, "scales": W["scales"], "lags": len(lags), "lambdas": len(lams),
            "chunk": W["chunk"], "elems_per_frame": E, "ring_slots": info.ring_slots,
            # device-resident frames are transformed whole chunks at a time, up to
            # ~128 frames per batch (the engine's FFT batch); host input goes chunk
            # by chunk.  Chunk latencies below are per batch.
            "fft_batch_frames": W["chunk"] if args.host_input else
            schedule.get("fft_batch_frames", W["chunk"] * max(1, 128 // W["chunk"])),
            "lag_block": lag_block, "schedule": "autotuned" if args.autotune else "default",
        },
        "p50_chunk_ms": percentile(chunk_ms, 50),
        "p99_chunk_ms": percentile(chunk_ms, 99),
        "stage_ms_per_step": {"fft": round(st.fft_ms, 3), "lag": round(st.lag_ms, 3),
                              "azimuthal": round(st.azimuthal_ms, 3)},
        "roofline": {
            "bound": "hbm",
            "kernel": f"lag_kernel<{lag_block}>",
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
        "precision": args.precision,
    }
    if prec != ddm.PREC_FP32:
        # matrix-core work of the split-operand DFT: 3 products in the row stage
        # and 6 in the column stage, each S^3 multiply-adds per tile
        mf = sum(18 * s ** 3 * (M // s) ** 2 for s in W["scales"] if s in (64, 128)) * N
        out["mfma_flops_per_step"] = mf

    if not args.no_check:
        from oracle import oracle as o
        cfg = o.Config(S, S, N, W["scales"], lags, lams, W["tol"], bpp, pixfmt)
        flat = res
        offs, off = {}, 0
        for s in W["scales"]:
            offs[s] = off
            off += (W["scales"][0] // s) ** 2 * len(lams) * len(lags)
        errs, mean_errs = {}, {}
        for si, s in enumerate(W["scales"]):
            if s > 256:
                continue
            ref = o.run(cfg, raw, "f64", nthreads=min(os.cpu_count() or 1, 16), only_scale=si, max_tiles=1)
            n1 = len(lams) * len(lags)
            g, r = flat[offs[s]:offs[s] + n1], ref[offs[s]:offs[s] + n1]
            errs[s] = o.max_rel_err(g, r)
            nz = r != 0
            mean_errs[s] = float(np.mean(np.abs(g[nz] - r[nz]) / np.abs(r[nz]))) if nz.any() else 0.0
            print(f"[bench] FP64 check scale {s}: {errs[s]:.3e}", file=sys.stderr, flush=True)
        out["max_rel_err_vs_fp64"] = {f"scale{s}_tile0": float(f"{v:.3e}") for s, v in errs.items()}
        out["mean_rel_err_vs_fp64"] = {f"scale{s}_tile0": float(f"{v:.3e}") for s, v in mean_errs.items()}

    if not args.no_cpu_baseline:
        from oracle import oracle as o
        threads = min(os.cpu_count() or 1, 16)
        nf = min(args.cpu_frames or W["cpu_frames"] or N, N)
        cfg = o.Config(S, S, nf, W["scales"], lags, lams, W["tol"], bpp, pixfmt)
        total = 0.0
        per = {}
        for si, s in enumerate(W["scales"]):
            t0 = time.perf_counter()
            o.run(cfg, raw[:nf * S * S * bpp], "f32", nthreads=threads, only_scale=si, max_tiles=1)
            dt = time.perf_counter() - t0
            print(f"[bench] CPU baseline scale {s}: {dt:.2f} s", file=sys.stderr, flush=True)
            tiles = (W["scales"][0] // s) ** 2
            per[s] = dt
            total += dt * tiles
        out["cpu_baseline"] = {
            "value": round(nf / total, 3),
            "unit": "frames/s",
            "cores": threads,
            "kind": "port",
            "sample": f"FP32 multithreaded C restatement, tile 0 of each scale over {nf} frames, "
                      f"time scaled by the tile count of each scale "
                      f"({', '.join(f'{s}: {per[s]:.2f}s' for s in W['scales'])})",
        }
    worst = max(out.get("max_rel_err_vs_fp64", {}).values(), default=0.0)
    if worst > CHECK_TOL:
        # a number whose results miss the FP64 check is not a measurement
        print(json.dumps(out), file=sys.stderr)
        sys.exit(f"[bench] FP64 check failed: max relative error {worst:.3e} > {CHECK_TOL:g}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
