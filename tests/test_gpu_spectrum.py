"""Masked spectrum (DDM_SPECTRUM_MASKED, the default): the ring and the lag
sums hold only the part of each tile's half spectrum that reaches the q rings
(for 256 .. 1024 tiles, column after column, the low |ky| rows of each column
down to its last ring member; for the others, row after row, the prefix of
each row up to its last ring member).
The averages read only ring members and sum them in member order, so
I(q, tau) must be bit-identical to the full spectrum (DDM_SPECTRUM_FULL)
for every scale, pixel format, lag block, precision mode and launch path,
and per-element reads must match the full engine on the kept elements and
be NaN elsewhere."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o
from tests.layout_util import kept_mask

pytestmark = pytest.mark.gpu

CASES = {
    # the reference's benchmark-mode parameters: 128 tiles, 20 linear lambdas
    "uniform128": dict(desc=dict(width=128, height=128, scales=[128], lags=list(range(1, 17)),
                                 lambdas=ddm.lin_lambdas(4, 64, 20), chunk=30), n=64, bits=8, gen="uniform"),
    # every tile size 8 .. 256 (column order for 256), log lags, lag block 64
    "multi256": dict(desc=dict(width=256, height=256, scales=[256, 128, 64, 32, 16, 8], lags=ddm.log_lags(200, 16),
                               lambdas=ddm.log_lambdas(4, 128, 12), chunk=30, lag_block=64),
                     n=150, bits=8, gen="brownian"),
    # sparse rings with gaps between them (tolerance 5 %); no member at all for 8 tiles
    "gaps": dict(desc=dict(width=256, height=256, scales=[256, 64, 8], lags=[1, 4, 9],
                           lambdas=[3.0, 11.0, 40.0], tolerance=1.05, chunk=20), n=60, bits=8, gen="uniform"),
    # 12-bit little-endian pixels in a wider frame at an ROI offset, 512 and 32 tiles
    "u16roi": dict(desc=dict(width=530, height=520, scales=[512, 32], lags=[1, 5, 30], lambdas=[6.0, 20.0, 50.0],
                             chunk=16, bpp=2, pixfmt=ddm.PIX_U16LE, x_off=17, y_off=5), n=48, bits=12,
                   gen="uniform"),
    # 1024 tiles and lag block 256 over 300 frames
    "big1024": dict(desc=dict(width=1024, height=1024, scales=[1024, 128], lags=ddm.log_lags(280, 10),
                              lambdas=ddm.log_lambdas(8, 256, 8), chunk=30, lag_block=256),
                    n=300, bits=8, gen="brownian"),
}


def _frames(case):
    c = CASES[case]
    d = c["desc"]
    w, h = d["width"], d["height"]
    if c["gen"] == "uniform":
        f = ddm.synth_uniform(c["n"], w, h, c["bits"], seed=len(case))
    else:
        f = ddm.synth_brownian(c["n"], w, h, c["bits"], max(1, w * h // 200), seed=len(case), nthreads=8)
    return np.ascontiguousarray(f)


def _run(case, frames, spectrum, precision=ddm.PREC_FP32):
    d = dict(CASES[case]["desc"])
    w, h = d.pop("width"), d.pop("height")
    with ddm.Engine(w, h, spectrum=spectrum, **d) as e:
        e.set_precision(precision)
        e.submit_host(frames.reshape(-1).view(np.uint8), len(frames))
        return e.results(), e.info()


@pytest.mark.parametrize("case", list(CASES))
def test_masked_results_bit_identical_to_full(gpu, case):
    frames = _frames(case)
    print(f"[{case}] {frames.shape}", flush=True)
    full, fi = _run(case, frames, ddm.SPECTRUM_FULL)
    masked, mi = _run(case, frames, ddm.SPECTRUM_MASKED)
    print(f"[{case}] elements per frame: masked {mi.elems_per_frame}, full {fi.elems_per_frame}")
    assert mi.elems_per_frame < fi.elems_per_frame
    assert np.array_equal(masked, full)
    # and the oracle on the first and last tile of every scale
    d = CASES[case]["desc"]
    M = max(d["scales"])
    cfg = o.Config(d["width"], d["height"], len(frames), d["scales"], d["lags"], d["lambdas"],
                   float(np.float32(d.get("tolerance", 1.2))), bpp=d.get("bpp", 1), pixfmt=d.get("pixfmt", ddm.PIX_U8),
                   x_off=d.get("x_off", 0), y_off=d.get("y_off", 0))
    ref = np.zeros_like(masked)
    for si, s in enumerate(cfg.sorted_scales()):
        ref += o.run(cfg, frames.reshape(-1), "f64", nthreads=8, only_scale=si, tiles=[0, (M // s) ** 2 - 1])
    g, r = o.split_results(cfg, masked), o.split_results(cfg, ref)
    for s in cfg.sorted_scales():
        T = (M // s) ** 2
        assert o.max_rel_err(g[s][[0, T - 1]], r[s][[0, T - 1]]) <= 1e-5, s


@pytest.mark.parametrize("mode", [ddm.PREC_BF16X2, ddm.PREC_FP16X2])
def test_masked_matrix_core_transforms(gpu, mode):
    """The 64 / 128 tile DFTs on the matrix cores write the band too."""
    frames = _frames("multi256")
    full, _ = _run("multi256", frames, ddm.SPECTRUM_FULL, mode)
    masked, _ = _run("multi256", frames, ddm.SPECTRUM_MASKED, mode)
    assert np.array_equal(masked, full)


def test_masked_spectra_and_accumulators(gpu):
    """Per-element reads: equal to the full engine on the kept elements, NaN
    elsewhere, and the kept elements are the rows' prefixes / columns' bands
    reaching the last q-ring member (so every member is kept)."""
    case = "multi256"
    frames = _frames(case)
    d = dict(CASES[case]["desc"])
    w, h = d.pop("width"), d.pop("height")
    out = {}
    for sp in (ddm.SPECTRUM_FULL, ddm.SPECTRUM_MASKED):
        with ddm.Engine(w, h, spectrum=sp, **d) as e:
            e.submit_host(frames.reshape(-1).view(np.uint8), len(frames))
            out[sp] = ([e.spectra(len(frames) - 1, si) for si in range(len(e.scales))],
                       [e.accumulators(si) for si in range(len(e.scales))], list(e.scales))
    for si, s in enumerate(out[ddm.SPECTRUM_FULL][2]):
        union = np.zeros((s, s // 2 + 1), dtype=bool)
        for lam in d["lambdas"]:
            union |= o.mask(s, lam, 1.2).astype(bool)
        fs, ms = out[ddm.SPECTRUM_FULL][0][si], out[ddm.SPECTRUM_MASKED][0][si]
        fa, ma = out[ddm.SPECTRUM_FULL][1][si], out[ddm.SPECTRUM_MASKED][1][si]
        kept = ~np.isnan(ms.real[0])
        assert np.all(kept[union]), s  # every ring member is held
        assert np.array_equal(kept, kept_mask(s, d["lambdas"], float(np.float32(1.2)))), s
        assert np.all(~np.isnan(ms.real)[:, kept]) and np.all(np.isnan(ms.real)[:, ~kept])
        assert np.array_equal(ms[:, kept], fs[:, kept]), s
        assert np.array_equal(ma[:, :, kept], fa[:, :, kept]), s
        assert np.all(np.isnan(ma[:, :, ~kept]))
        print(f"scale {s}: {int(kept.sum())} of {kept.size} elements kept, {int(union.sum())} ring members")


def test_masked_server_streams(gpu):
    """Server streams (per-stream and batched launches): masked and full
    spectra give the same bits, and agree with an engine over the same
    frames to FP32 rounding (a server runs a lag pass per chunk, an engine
    per lag block)."""
    frames = _frames("multi256")
    d = dict(CASES["multi256"]["desc"])
    d["lag_block"] = 32
    with ddm.Engine(spectrum=ddm.SPECTRUM_FULL, **d) as e:
        e.submit_host(frames.reshape(-1).view(np.uint8), len(frames))
        engine = e.results()
    for batched in (False, True):
        got = {}
        for sp in (ddm.SPECTRUM_FULL, ddm.SPECTRUM_MASKED):
            with ddm.Server(batched=batched) as srv:
                sid = srv.open(spectrum=sp, **d)
                for i in range(0, len(frames), 30):
                    srv.submit(sid, frames[i:i + 30])
                got[sp] = srv.results(sid)
        err = o.max_rel_err(got[ddm.SPECTRUM_MASKED], engine)
        print(f"batched={batched}: max rel err vs the engine {err:.3e}")
        assert np.array_equal(got[ddm.SPECTRUM_MASKED], got[ddm.SPECTRUM_FULL]), batched
        assert err <= 1e-5
