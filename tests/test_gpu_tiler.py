"""Multi-scale tiler (DDM_TILER_FUSED): each frame's pixels are
read once for all 8..128 tile scales (128 x 128 blocks staged in LDS) and
once for the row passes of all 256..1024 scales (row bands), where the
reference cuts every scale from the frame separately
(src/DDM.cu:171-192, src/DDM_kernel.cuh:43-78).

The fused launches run the same transform arithmetic on the same pixel
values as the per-scale launches (DDM_TILER_PER_SCALE), so spectra and
I(q, tau) must be the same bit for bit, for every pixel format and global
load width (16-byte, 4-byte and per-pixel loads, chosen by the alignment of
the region), both spectrum layouts, regions of 128 .. 2048 pixels, and
scales the tiler leaves to per-scale launches (2 / 4 / 2048 tiles, the LDS
transform family, matrix-core transforms).  The oracle checks the fused
results on the first and last tile of every scale."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o

pytestmark = pytest.mark.gpu

CASES = {
    # every small scale from one 128 block per tile; 256 stays per scale
    "small_all_u8": dict(width=256, height=256, scales=[256, 128, 64, 32, 16, 8], bits=8),
    # the 512^2 u8 workload's scale set: two band scales, two block scales
    "u8_512": dict(width=512, height=512, scales=[512, 256, 128, 64], bits=8),
    # 12-bit little-endian pixels at a 16-byte aligned region offset: 16-byte loads
    "u16_a16": dict(width=1088, height=1040, scales=[1024, 512, 256, 128, 64, 32], bits=12, bpp=2,
                    pixfmt=ddm.PIX_U16LE, x_off=64, y_off=8),
    # 4-byte aligned offset: 4-byte loads
    "u16_a4": dict(width=540, height=520, scales=[512, 256, 128, 32], bits=12, bpp=2, pixfmt=ddm.PIX_U16LE,
                   x_off=2, y_off=3),
    # odd offset: pixel-by-pixel loads
    "u8_odd": dict(width=531, height=515, scales=[512, 256, 64, 16, 8], bits=8, x_off=17, y_off=1),
    # big-endian pixels and 3-byte pixels: pixel-by-pixel loads
    "u16be": dict(width=256, height=256, scales=[256, 128, 64, 32], bits=12, bpp=2, pixfmt=ddm.PIX_U16BE),
    "rgb3": dict(width=256, height=256, scales=[128, 64, 32, 8], bits=8, bpp=3),
    # a 2048 region: 2048 per scale, 1024 / 512 / 256 from 8-row bands
    "region2048": dict(width=2048, height=2048, scales=[2048, 1024, 512, 256, 128, 4], bits=8, n=24),
    # a 128 region: one block per frame
    "region128": dict(width=128, height=128, scales=[128, 64, 32, 16, 8, 2], bits=8),
}
LAGS = [1, 2, 3, 5, 8, 13]
LAMBDAS = [2.5, 4.0, 7.0, 12.0, 30.0, 60.0]


def _frames(case):
    c = CASES[case]
    n = c.get("n", 40)
    f = ddm.synth_uniform(n, c["width"], c["height"], c["bits"], seed=len(case) + 3)
    if c.get("bpp", 1) == 3:
        f = np.repeat(f[..., None], 3, axis=-1)  # channel 0 is read; the others equal it
        f[..., 1] ^= 0x5a
    return np.ascontiguousarray(f)


def _engine(case, **kw):
    c = CASES[case]
    return ddm.Engine(c["width"], c["height"], c["scales"], LAGS, LAMBDAS, 1.2, 16, bpp=c.get("bpp", 1),
                      pixfmt=c.get("pixfmt", ddm.PIX_U8), x_off=c.get("x_off", 0), y_off=c.get("y_off", 0),
                      lag_block=32, **kw)


def _run(case, frames, tiler, spectrum=ddm.SPECTRUM_MASKED, setup=None, host=False):
    """Whole frames from device memory (the region at its offset inside
    them, so the load width follows the offset), or from host memory (the
    engine uploads the region's rows only)."""
    import torch
    with _engine(case, spectrum=spectrum) as e:
        e.set_tiler(tiler)
        assert e.tiler() == tiler
        if setup:
            setup(e)
        if host:
            e.submit_host(frames.reshape(-1).view(np.uint8), len(frames))
        else:
            d = torch.from_numpy(frames.reshape(-1).view(np.uint8).copy()).cuda()
            torch.cuda.synchronize()  # the engine's stream does not wait for torch's
            e.submit_device(d.data_ptr(), len(frames))
            e.sync()
            del d
        res = e.results()
        spec = [e.spectra(len(frames) - 1, si) for si in range(len(e.scales))]
        spec0 = [e.spectra(len(frames) - 2, si) for si in range(len(e.scales))]
        return res, spec, spec0, e.stage_stats().fft_launches


def _same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


@pytest.mark.parametrize("spectrum", [ddm.SPECTRUM_MASKED, ddm.SPECTRUM_FULL])
@pytest.mark.parametrize("case", list(CASES))
def test_fused_tiler_bit_identical(gpu, case, spectrum):
    frames = _frames(case)
    print(f"[{case}] spectrum={spectrum} {frames.shape}", flush=True)
    per, sp_per, sp0_per, l_per = _run(case, frames, ddm.TILER_PER_SCALE, spectrum)
    fus, sp_fus, sp0_fus, l_fus = _run(case, frames, ddm.TILER_FUSED, spectrum)
    print(f"[{case}] transform launches: per scale {l_per}, fused {l_fus}")
    c = CASES[case]
    for si, s in enumerate(sorted(c["scales"], reverse=True)):
        a, b = sp_fus[si], sp_per[si]
        ok = ~np.isnan(b.real)
        big = np.abs(b[ok]).max()
        d = np.abs(a[ok] - b[ok])
        print(f"[{case}] scale {s}: spectra max |diff| / max |X| = {d.max() / big:.3e}, "
              f"differing {int((d > 0).sum())} of {d.size}, first at {np.argwhere(np.abs(a - b) > 0)[:1].tolist()}")
    print(f"[{case}] I(q,tau) max rel diff fused vs per-scale {o.max_rel_err(fus, per):.3e}")
    assert l_fus < l_per
    assert np.array_equal(fus, per)
    assert _same(sp_fus, sp_per) and _same(sp0_fus, sp0_per)
    if spectrum == ddm.SPECTRUM_FULL:
        return
    M = max(s for s in c["scales"])
    cfg = o.Config(c["width"], c["height"], len(frames), c["scales"], LAGS, LAMBDAS, float(np.float32(1.2)),
                   bpp=c.get("bpp", 1), pixfmt=c.get("pixfmt", ddm.PIX_U8), x_off=c.get("x_off", 0),
                   y_off=c.get("y_off", 0))
    ref = np.zeros_like(fus)
    for si, s in enumerate(cfg.sorted_scales()):
        ref += o.run(cfg, frames.reshape(-1), "f64", nthreads=8, only_scale=si, tiles=[0, (M // s) ** 2 - 1])
    g, r = o.split_results(cfg, fus), o.split_results(cfg, ref)
    for s in cfg.sorted_scales():
        T = (M // s) ** 2
        err = o.max_rel_err(g[s][[0, T - 1]], r[s][[0, T - 1]])
        assert err <= 1e-5, (s, err)


def test_fused_tiler_with_lds_family_and_matrix_cores(gpu):
    """Scales left to per-scale launches (an LDS-family scale, matrix-core
    64 / 128 transforms) beside fused ones give the same bits as all
    per-scale launches."""
    case = "small_all_u8"
    frames = _frames(case)

    def lds32(e):
        e.set_fft_family(list(e.scales).index(32), ddm.FFT_LDS)

    for setup in (lds32, lambda e: e.set_precision(ddm.PREC_BF16X2)):
        per = _run(case, frames, ddm.TILER_PER_SCALE, setup=setup)
        fus = _run(case, frames, ddm.TILER_FUSED, setup=setup)
        assert np.array_equal(fus[0], per[0])
        assert _same(fus[1], per[1])


@pytest.mark.parametrize("case", ["u16_a16", "u8_odd"])
def test_fused_tiler_host_input(gpu, case):
    """Host frames: the engine uploads the region's rows into staging
    buffers and transforms them a chunk at a time."""
    frames = _frames(case)
    per = _run(case, frames, ddm.TILER_PER_SCALE, host=True)
    fus = _run(case, frames, ddm.TILER_FUSED, host=True)
    dev = _run(case, frames, ddm.TILER_FUSED)
    assert np.array_equal(fus[0], per[0]) and np.array_equal(fus[0], dev[0])
    assert _same(fus[1], per[1]) and _same(fus[1], dev[1])


def test_tiler_mode_fixed_after_first_frame(gpu):
    frames = _frames("region128")
    with _engine("region128") as e:
        e.set_tiler(ddm.TILER_FUSED)
        e.submit_host(frames[:4].reshape(-1).view(np.uint8), 4)
        with pytest.raises(ddm.DDMError):
            e.set_tiler(ddm.TILER_PER_SCALE)


def test_tiler_default_by_region(gpu):
    """Regions up to 256 x 256 default to the fused tiler (their per-scale
    transforms are short launches), larger ones to per-scale launches."""
    for size, want in ((128, ddm.TILER_FUSED), (256, ddm.TILER_FUSED), (512, ddm.TILER_PER_SCALE),
                       (1024, ddm.TILER_PER_SCALE)):
        scales = [s for s in (1024, 512, 256, 128, 64, 32) if s <= size]
        with ddm.Engine(size, size, scales, LAGS, LAMBDAS, 1.2, 16) as e:
            assert e.tiler() == want, size

