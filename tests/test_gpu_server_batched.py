"""Batched serving (DDM_SERVER_BATCHED): the waiting chunks of all streams run
as ticks of one launch per kernel kind over a work queue, captured as HIP
graphs.  Results are bit-identical to the per-stream server, with and without
graphs, alone and in a mix, and match the FP64 oracle."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from cuda_multiddm_amd.workloads import host_threads, serving_streams
from oracle import oracle as o

pytestmark = pytest.mark.gpu

STREAMS = {
    # 8 .. 64 tiles, lag block 32: every kernel batched
    "a": dict(desc=dict(width=64, height=64, scales=[64, 32, 16, 8], lags=[1, 2, 5, 17, 40], lambdas=[2.0, 4.0, 8.0],
                        chunk=8, lag_block=32), n=60, bits=8, seed=21),
    # 12-bit little endian in a wider frame, ROI offsets, lag block 256: the
    # lag pass of this stream runs as its own launch inside the tick
    "b": dict(desc=dict(width=80, height=72, scales=[64, 16], lags=[1, 3, 9, 40], lambdas=[2.0, 3.0, 6.0, 12.0],
                        chunk=16, bpp=2, pixfmt=ddm.PIX_U16LE, x_off=9, y_off=5), n=70, bits=12, seed=22),
    # 256 tiles (rows + columns through the pass buffer), 128 and 32 tiles
    "c": dict(desc=dict(width=256, height=256, scales=[256, 128, 32], lags=ddm.log_lags(100, 12),
                        lambdas=ddm.log_lambdas(3, 60, 8), chunk=30, lag_block=32), n=150, bits=8, seed=23),
    # 3-byte pixels: tile transforms of this layout are launched on their own
    "d": dict(desc=dict(width=40, height=40, scales=[32, 16], lags=[1, 2, 7], lambdas=[3.0, 5.0], chunk=10, bpp=3,
                        lag_block=32), n=40, bits=8, seed=24),
}


def _frames(name):
    s = STREAMS[name]
    d = s["desc"]
    if d.get("bpp", 1) == 3:
        rng = np.random.default_rng(s["seed"])
        return rng.integers(0, 256, (s["n"], d["height"], d["width"], 3), dtype=np.uint8)
    return ddm.synth_uniform(s["n"], d["width"], d["height"], s["bits"], seed=s["seed"])


def _chunks(name):
    f = _frames(name)
    c = STREAMS[name]["desc"]["chunk"]
    return [f[i:i + c] for i in range(0, len(f), c)]


def _run(names, batched, graph=True, order_seed=0):
    with ddm.Server(batched=batched, graph=graph) as srv:
        ids = {n: srv.open(**STREAMS[n]["desc"]) for n in names}
        left = {n: _chunks(n) for n in names}
        rng = np.random.default_rng(order_seed)
        while any(left.values()):
            n = names[int(rng.choice([i for i, k in enumerate(names) if left[k]]))]
            srv.submit(ids[n], left[n].pop(0))
        out = {n: srv.results(ids[n]) for n in names}
        st = srv.stats()
        lat = srv.latencies()
        assert len(lat) == sum(len(_chunks(n)) for n in names)
    return out, st


def test_batched_equals_per_stream_server(gpu):
    names = list(STREAMS)
    ref, _ = _run(names, batched=False)
    got, st = _run(names, batched=True, order_seed=1)
    print(f"ticks {st.ticks}, graph ticks {st.graph_ticks}, items {st.tick_items}")
    assert st.ticks > 0
    for n in names:
        assert np.array_equal(got[n], ref[n]), n
    nog, st2 = _run(names, batched=True, graph=False, order_seed=2)
    assert st2.graph_ticks == 0
    for n in names:
        assert np.array_equal(nog[n], ref[n]), n


def test_batched_graph_ticks_when_all_kernels_batch(gpu):
    """Streams whose every kernel has a batched form run their ticks as graphs."""
    got, st = _run(["a", "c"], batched=True, order_seed=3)
    assert st.ticks > 0 and st.graph_ticks == st.ticks
    for n in ("a", "c"):
        s = STREAMS[n]
        d = s["desc"]
        cfg = o.Config(d["width"], d["height"], s["n"], d["scales"], d["lags"], d["lambdas"], 1.2)
        assert o.max_rel_err(got[n], o.run(cfg, _frames(n))) <= 1e-5, n


def test_batched_serving_mix_alone_equals_mix(gpu):
    """Streams of the serving mix (256^2 / 512^2 / 1024^2, scales to 32, 24 log
    lags to 300): each alone in a batched server and all together give the
    same bits, and the oracle agrees on the first and last tile of each scale."""
    pool = serving_streams(64)
    picked = [next(s for s in pool if s.size == size) for size in (256, 512, 1024)]
    N = 330
    frames = [s.make_frames(N) for s in picked]
    chunks = [[f[i:i + 30] for i in range(0, N, 30)] for f in frames]
    alone = []
    for k, s in enumerate(picked):
        with ddm.Server(batched=True) as srv:
            sid = srv.open(**s.desc())
            for ch in chunks[k]:
                srv.submit(sid, ch)
            alone.append(srv.results(sid))
    rng = np.random.default_rng(9)
    with ddm.Server(batched=True) as srv:
        ids = [srv.open(**s.desc()) for s in picked]
        left = [list(c) for c in chunks]
        while any(left):
            k = int(rng.choice([i for i in range(len(left)) if left[i]]))
            srv.submit(ids[k], left[k].pop(0))
        mixed = [srv.results(i) for i in ids]
        st = srv.stats()
    print(f"mix: {st.ticks} ticks ({st.graph_ticks} as graphs), {st.tick_items} queue items")
    for k, s in enumerate(picked):
        assert np.array_equal(alone[k], mixed[k]), k
        d = s.desc()
        cfg = o.Config(s.size, s.size, N, d["scales"], d["lags"], d["lambdas"], np.float32(1.2).item())
        ref = np.zeros_like(mixed[k])
        for si, sc in enumerate(d["scales"]):
            T = (s.size // sc) ** 2
            ref += o.run(cfg, frames[k].reshape(-1), "f64", nthreads=host_threads(), only_scale=si, tiles=[0, T - 1])
        g, r = o.split_results(cfg, mixed[k]), o.split_results(cfg, ref)
        for sc in d["scales"]:
            T = (s.size // sc) ** 2
            err = o.max_rel_err(g[sc][[0, T - 1]], r[sc][[0, T - 1]])
            assert err <= 1e-5, (k, sc, err)


def test_batched_results_mid_stream_and_close_while_others_run(gpu):
    with ddm.Server(batched=True) as srv:
        a, c = srv.open(**STREAMS["a"]["desc"]), srv.open(**STREAMS["c"]["desc"])
        ca, cc = _chunks("a"), _chunks("c")
        for ch in ca[:3]:
            srv.submit(a, ch)
        mid = srv.results(a)  # 24 frames so far
        for ch in ca[3:]:
            srv.submit(a, ch)
            srv.submit(c, cc.pop(0))
        full_a = srv.results(a)
        srv.close_stream(a)
        for ch in cc:
            srv.submit(c, ch)
        full_c = srv.results(c)
        srv.drain()
    s = STREAMS["a"]
    d = s["desc"]
    f = _frames("a")
    for got, n in ((mid, 24), (full_a, s["n"])):
        cfg = o.Config(64, 64, n, d["scales"], d["lags"], d["lambdas"], 1.2)
        assert o.max_rel_err(got, o.run(cfg, f[:n])) <= 1e-5
    ref, _ = _run(["c"], batched=False)
    assert np.array_equal(full_c, ref["c"])


def test_grouped_chunks_share_a_lag_pass(gpu):
    """DDM_SERVER_GROUP: a tick takes all of a stream's waiting chunks up to
    its lag block into one lag pass.  Chunks submitted back to back pile up,
    so passes of several chunks happen; results stay within the FP32 bound of
    the oracle, and alone vs mix agree to within that bound."""
    pool = serving_streams(64)
    picked = [next(s for s in pool if s.size == size) for size in (256, 512)]
    N = 330
    frames = [s.make_frames(N) for s in picked]
    chunks = [[f[i:i + 30] for i in range(0, N, 30)] for f in frames]
    res = {}
    for mode in ("alone", "mix"):
        out = []
        with ddm.Server(batched=True, group=True) as srv:
            if mode == "alone":
                for k, s in enumerate(picked):
                    sid = srv.open(**s.desc(lag_block=128))
                    for ch in chunks[k]:
                        srv.submit(sid, ch)
                    out.append(srv.results(sid))
            else:
                ids = [srv.open(**s.desc(lag_block=128)) for s in picked]
                for i in range(len(chunks[0])):
                    for k in range(len(picked)):
                        srv.submit(ids[k], chunks[k][i])
                out = [srv.results(i) for i in ids]
            st = srv.stats()
            print(f"{mode}: {st.chunks} chunks in {st.ticks} ticks")
        res[mode] = out
    for k, s in enumerate(picked):
        d = s.desc()
        cfg = o.Config(s.size, s.size, N, d["scales"], d["lags"], d["lambdas"], np.float32(1.2).item())
        ref = o.run(cfg, frames[k].reshape(-1), "f64", nthreads=host_threads(), tiles=[0])
        g = {m: o.split_results(cfg, res[m][k]) for m in res}
        r = o.split_results(cfg, ref)
        for sc in d["scales"]:
            for m in res:
                assert o.max_rel_err(g[m][sc][0], r[sc][0]) <= 1e-5, (k, sc, m)
        a, b = res["alone"][k], res["mix"][k]
        nz = a != 0
        assert np.all(b[~nz] == 0) and np.max(np.abs(a[nz] - b[nz]) / np.abs(a[nz])) <= 1e-6
