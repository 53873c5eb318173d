"""Serving: streams of different frame sizes, scale sets and lag sets share
one device.  Each stream's result is bit-identical alone and in the mix (and
matches the FP64 oracle), admission is by device footprint, and every chunk
gets a chunk-to-result latency."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o

pytestmark = pytest.mark.gpu

STREAMS = {
    "a": dict(desc=dict(width=64, height=64, scales=[64, 32], lags=[1, 2, 5, 17], lambdas=[2.0, 4.0, 8.0], chunk=8),
              n=60, bits=8, seed=21),
    "b": dict(desc=dict(width=80, height=72, scales=[64, 16], lags=[1, 3, 9, 40], lambdas=[2.0, 3.0, 6.0, 12.0],
                        chunk=16, bpp=2, pixfmt=ddm.PIX_U16LE, x_off=9, y_off=5), n=70, bits=12, seed=22),
    "c": dict(desc=dict(width=128, height=128, scales=[128, 64, 32], lags=ddm.log_lags(100, 12),
                        lambdas=ddm.log_lambdas(3, 60, 8), chunk=30, lag_block=32), n=120, bits=8, seed=23),
}


def _frames(name):
    s = STREAMS[name]
    d = s["desc"]
    return ddm.synth_uniform(s["n"], d["width"], d["height"], s["bits"], seed=s["seed"])


def _chunks(name):
    f = _frames(name)
    c = STREAMS[name]["desc"]["chunk"]
    return [f[i:i + c] for i in range(0, len(f), c)]


def _oracle(name):
    s = STREAMS[name]
    d = dict(s["desc"])
    f = _frames(name)
    cfg = o.Config(d["width"], d["height"], s["n"], d["scales"], d["lags"], d["lambdas"], 1.2,
                   d.get("bpp", 1), d.get("pixfmt", 0), d.get("x_off", 0), d.get("y_off", 0))
    return o.run(cfg, f.view(np.uint8))


def test_alone_equals_mix(gpu):
    alone = {}
    for name, s in STREAMS.items():
        with ddm.Server() as srv:
            sid = srv.open(**s["desc"])
            for ch in _chunks(name):
                srv.submit(sid, ch)
            alone[name] = srv.results(sid)
            assert len(srv.latencies(sid)) == len(_chunks(name))
    with ddm.Server() as srv:
        ids = {name: srv.open(**s["desc"]) for name, s in STREAMS.items()}
        queues = {name: _chunks(name) for name in STREAMS}
        step = 0
        while any(queues.values()):  # interleave the streams' chunks as they "arrive"
            for name in STREAMS:
                if queues[name] and (step % (1 + len(name) % 2) == 0 or len(queues[name]) > 2):
                    srv.submit(ids[name], queues[name].pop(0))
            step += 1
        srv.drain()
        st = srv.stats()
        assert st.streams == 3 and st.chunks == sum(len(_chunks(n)) for n in STREAMS)
        assert st.chunks_in_flight == 0
        lat = srv.latencies()
        assert len(lat) == st.chunks and np.all(lat > 0)
        for name in STREAMS:
            mixed = srv.results(ids[name])
            assert np.array_equal(mixed, alone[name]), name
    for name in STREAMS:
        assert o.max_rel_err(alone[name], _oracle(name)) <= 1e-5, name


def test_admission_by_footprint(gpu):
    fp = {name: ddm.validate(**s["desc"])[2].device_bytes for name, s in STREAMS.items()}
    assert fp["c"] > fp["a"]
    budget = fp["b"] + fp["c"] + 1000
    with ddm.Server(device_budget=budget, max_streams=8) as srv:
        a = srv.open(**STREAMS["a"]["desc"])
        b = srv.open(**STREAMS["b"]["desc"])
        with pytest.raises(ddm.DDMError) as ex:
            srv.open(**STREAMS["c"]["desc"])
        assert ex.value.code == ddm.DDM_EBUSY
        assert srv.stats().used_bytes == fp["a"] + fp["b"]
        srv.close_stream(a)
        c = srv.open(**STREAMS["c"]["desc"])
        assert srv.stats().used_bytes == fp["b"] + fp["c"]
        for ch in _chunks("c"):
            srv.submit(c, ch)
        assert o.max_rel_err(srv.results(c), _oracle("c")) <= 1e-5
        del b
    with ddm.Server(max_streams=2) as srv:
        srv.open(**STREAMS["a"]["desc"])
        srv.open(**STREAMS["a"]["desc"])
        with pytest.raises(ddm.DDMError) as ex:
            srv.open(**STREAMS["a"]["desc"])
        assert ex.value.code == ddm.DDM_EBUSY


def test_admission_refuses_when_device_memory_runs_out(gpu):
    """A budget larger than the device: a stream whose spectrum ring alone
    (2048^2 tiles, lags up to 20000: ~340 GB) cannot be allocated is refused
    with DDM_EBUSY, not an allocation error, and the server keeps serving."""
    huge = dict(width=2048, height=2048, scales=[2048], lags=[1, 20000], lambdas=[8.0], chunk=30)
    assert ddm.validate(**huge)[2].device_bytes > (300 << 30)
    with ddm.Server(device_budget=1 << 40, max_streams=4) as srv:
        with pytest.raises(ddm.DDMError) as ex:
            srv.open(**huge)
        assert ex.value.code == ddm.DDM_EBUSY
        assert srv.stats().used_bytes == 0
        c = srv.open(**STREAMS["c"]["desc"])
        for ch in _chunks("c"):
            srv.submit(c, ch)
        assert o.max_rel_err(srv.results(c), _oracle("c")) <= 1e-5


def test_server_rejects_bad_requests(gpu):
    with ddm.Server() as srv:
        with pytest.raises(ddm.DDMError):
            srv.open(64, 64, [48], [1], [4.0])
        with pytest.raises(ddm.DDMError):
            srv.results(12345)
