"""The whole HIP path (tile -> FFT -> lag engine -> azimuthal average)
through the C ABI, against the FP64 CPU oracle on the same seeded inputs.
Tolerance: max relative I(q, tau) error vs FP64 <= 1e-5 for the plain FP32
path (entries that are exactly 0 in FP64 -- empty rings, lags without
pairs -- must be exactly 0)."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _both(frames, scales, lags, lambdas, tol=1.2, chunk=30, bpp=1, pixfmt=ddm.PIX_U8, x_off=0, y_off=0,
          norm=ddm.NORM_FRAMES, device_input=False, lag_block=0):
    n, h, w = frames.shape[0], frames.shape[1], frames.shape[2]
    raw = np.ascontiguousarray(frames).view(np.uint8)
    with ddm.Engine(w, h, scales, lags, lambdas, tol, chunk, bpp, pixfmt, x_off, y_off, norm, lag_block) as e:
        print(f"engine {w}x{h} n={n} scales={e.scales} lags={len(e.lags)} chunk={chunk}", flush=True)
        if device_input:
            import torch
            t = torch.from_numpy(raw.reshape(-1).copy()).cuda()
            torch.cuda.synchronize()  # the engine's stream does not wait for torch's
            e.submit_device(t.data_ptr(), n)
            e.sync()
        else:
            e.submit_host(raw)
        got = e.results()
    cfg = o.Config(w, h, n, scales, lags, lambdas, np.float32(tol).item(), bpp, pixfmt, x_off, y_off, norm)
    ref = o.run(cfg, raw)
    return got, ref


def test_uniform_128_whole_pipeline(gpu):
    """64 frames 128^2 uint8 uniform, scale 128, tau 1..16, 20 lambdas 4..64, tol 20%, chunk 30."""
    frames = ddm.synth_uniform(64, 128, 128, 8, seed=1)
    got, ref = _both(frames, [128], list(range(1, 17)), ddm.lin_lambdas(4, 64, 20), 1.2, 30)
    assert got.shape == ref.shape
    err = o.max_rel_err(got, ref)
    print("uniform 128^2 max rel err", err)
    assert err <= TOL


def test_multiscale_u16_lags_beyond_chunk(gpu):
    """12-bit values in uint16 (little endian), ROI at an offset, four scales,
    lags larger than the chunk, lambdas giving empty rings, a ragged last chunk."""
    rng = np.random.default_rng(2)
    frames = rng.integers(0, 4096, (70, 75, 70)).astype("<u2")
    lags = [1, 2, 3, 7, 16, 17, 30, 45, 69, 80]
    lams = ddm.log_lambdas(2, 90, 12)
    got, ref = _both(frames, [8, 64, 32, 16], lags, lams, 1.2, 16, bpp=2, pixfmt=ddm.PIX_U16LE, x_off=5,
                     y_off=7)
    err = o.max_rel_err(got, ref)
    print("multiscale max rel err", err)
    assert err <= TOL
    # lag 80 > N - 1: no pair, exactly zero
    assert np.all(got.reshape(-1, len(lags))[:, -1] == 0)


def test_three_byte_pixels_channel0_big_endian(gpu):
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, (40, 40, 36, 3), dtype=np.uint8)
    got, ref = _both(rgb, [32, 16], [1, 3, 9, 33], [2.0, 3.0, 5.0, 8.0], 1.2, 11, bpp=3, x_off=4, y_off=2)
    assert o.max_rel_err(got, ref) <= TOL
    be = rng.integers(0, 65535, (33, 32, 32)).astype(">u2")
    got, ref = _both(be, [32], [1, 2, 4, 8], [2.0, 4.0, 6.0], 1.2, 10, bpp=2, pixfmt=ddm.PIX_U16BE)
    assert o.max_rel_err(got, ref) <= TOL


def test_pair_mean_normalisation(gpu):
    frames = ddm.synth_uniform(50, 64, 64, 8, seed=4)
    got, ref = _both(frames, [64, 32], [1, 5, 20, 49], [3.0, 6.0, 12.0], 1.2, 20, norm=ddm.NORM_PAIRS)
    assert o.max_rel_err(got, ref) <= TOL


def test_host_and_device_submission_identical(gpu):
    frames = ddm.synth_uniform(45, 64, 64, 8, seed=5)
    a, _ = _both(frames, [64, 16], [1, 2, 7, 31], [2.0, 5.0], 1.2, 30)
    b, ref = _both(frames, [64, 16], [1, 2, 7, 31], [2.0, 5.0], 1.2, 30, device_input=True)
    assert np.array_equal(a, b)
    c, _ = _both(frames, [64, 16], [1, 2, 7, 31], [2.0, 5.0], 1.2, 30, device_input=True)
    assert np.array_equal(b, c)  # deterministic
    for chunk, lb in [(7, 32), (30, 128), (45, 64), (9, 256)]:  # other chunk / lag-block splits
        d, _ = _both(frames, [64, 16], [1, 2, 7, 31], [2.0, 5.0], 1.2, chunk, lag_block=lb)
        assert o.max_rel_err(d, ref) <= TOL


@pytest.mark.parametrize("chunk,lb,lags", [
    (7, 32, [1, 2, 3, 5, 9, 20, 60]),       # batch of 126 limited by the ring room (39 - waiting)
    (30, 256, [1, 4, 17, 90, 200]),         # 120-frame batches, lag blocks straddling them
    (13, 64, [1, 2, 3, 4, 5, 6, 7, 8, 40]),  # batches of 117, room 77 - waiting
])
def test_device_batches_match_chunked_host_submission(gpu, chunk, lb, lags):
    """Device-resident frames are transformed several chunks per batch, as
    many as the ring's free slots allow; the results equal a chunk-by-chunk
    host submission bit for bit, and the oracle."""
    frames = ddm.synth_uniform(300, 64, 64, 8, seed=chunk + lb)
    a, ref = _both(frames, [64, 32], lags, [2.0, 5.0, 9.0], 1.2, chunk, lag_block=lb)
    b, _ = _both(frames, [64, 32], lags, [2.0, 5.0, 9.0], 1.2, chunk, lag_block=lb, device_input=True)
    assert np.array_equal(a, b)
    assert o.max_rel_err(b, ref) <= TOL


def test_reserve_allocates_up_front_and_changes_nothing(gpu):
    """ddm_reserve allocates the pass buffer (256 tiles) and batch-sized
    staging before the first frame; the results of device and bulk host
    submissions are bit-identical to an engine without it."""
    import torch
    frames = ddm.synth_uniform(200, 256, 256, 8, seed=5)
    args = (256, 256, [256, 64], [1, 2, 3, 7, 40], [3.0, 9.0, 30.0], 1.2, 30)
    d = torch.from_numpy(frames.reshape(-1).copy()).cuda()
    torch.cuda.synchronize()
    outs = []
    for reserve in (False, True):
        for src in ("device", "host"):
            with ddm.Engine(*args) as e:
                before = e.info().device_bytes
                if reserve:
                    e.reserve(len(frames))
                    assert e.info().device_bytes > before
                    before = e.info().device_bytes
                if src == "device":
                    e.submit_device(d.data_ptr(), len(frames))
                else:
                    e.submit_host(frames, len(frames))
                if reserve:
                    assert e.info().device_bytes == before  # nothing allocated on the way
                outs.append(e.results())
    for o_ in outs[1:]:
        assert np.array_equal(outs[0], o_)


@pytest.mark.parametrize("w,h,x_off,y_off,bpp,fmt", [
    (96, 64, 32, 0, 1, ddm.PIX_U8),      # ROI rows of consecutive frames evenly spaced: one 2-D copy
    (80, 90, 8, 20, 2, ddm.PIX_U16LE),   # general ROI: one 2-D copy per frame
    (64, 64, 0, 0, 3, ddm.PIX_U8),       # 3-byte pixels, ROI = frame
])
def test_pinned_host_frames_roi_upload(gpu, w, h, x_off, y_off, bpp, fmt):
    """Frames in pinned host memory are uploaded by the copy engine (ROI rows
    only); the results equal the pageable-host and device submissions."""
    import torch
    rng = np.random.default_rng(w + h + bpp)
    n = 40
    raw = rng.integers(0, 256, (n, h, w * bpp), dtype=np.uint8)
    pinned = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True).numpy()
    pinned[:] = raw.reshape(-1)
    scales, lags, lams = [64, 32], [1, 3, 8, 33], [3.0, 6.0]
    outs = []
    for src in ("pinned", "pageable", "device"):
        with ddm.Engine(w, h, scales, lags, lams, 1.2, 16, bpp, fmt, x_off, y_off) as e:
            print(f"roi upload {src} {w}x{h} off=({x_off},{y_off}) bpp={bpp}", flush=True)
            if src == "pinned":
                e.submit_host(pinned, n)
            elif src == "pageable":
                e.submit_host(raw.reshape(-1).copy(), n)
            else:
                t = torch.from_numpy(raw.reshape(-1).copy()).cuda()
                torch.cuda.synchronize()
                e.submit_device(t.data_ptr(), n)
                e.sync()
            outs.append(e.results())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])
    cfg = o.Config(w, h, n, scales, lags, lams, np.float32(1.2).item(), bpp, fmt, x_off, y_off)
    assert o.max_rel_err(outs[0], o.run(cfg, raw)) <= TOL


def test_results_mid_stream(gpu):
    """Reading results flushes a partial lag block; the stream then goes on
    and the final answer still equals a single pass."""
    frames = ddm.synth_uniform(100, 64, 64, 8, seed=12)
    lags, lams = [1, 2, 5, 17, 60, 90], [2.0, 4.0, 9.0]
    with ddm.Engine(64, 64, [64, 32], lags, lams, 1.2, 30, lag_block=64) as e:
        e.submit_host(frames[:37])
        mid = e.results()
        e.submit_host(frames[37:])
        end = e.results()
    assert o.max_rel_err(mid, o.run(o.Config(64, 64, 37, [64, 32], lags, lams), frames[:37])) <= TOL
    assert o.max_rel_err(end, o.run(o.Config(64, 64, 100, [64, 32], lags, lams), frames)) <= TOL


def test_spectra_and_accumulators(gpu):
    frames = ddm.synth_uniform(100, 64, 64, 8, seed=6)
    lags = [1, 3, 8]
    with ddm.Engine(64, 64, [64, 32], lags, [4.0, 8.0], 1.2, 8, lag_block=32, spectrum=ddm.SPECTRUM_FULL) as e:
        e.submit_host(frames)
        assert e.info().ring_slots == 8 + 32 + 8
        for fr in (99, 60):
            spec = e.spectra(fr, 1)  # scale 32
            for tile in range(4):
                ty, tx = divmod(tile, 2)
                ref = np.fft.rfft2(frames[fr, ty * 32:(ty + 1) * 32, tx * 32:(tx + 1) * 32].astype(float))
                assert np.max(np.abs(spec[tile] - ref)) / np.max(np.abs(ref)) < 4e-6
        with pytest.raises(ddm.DDMError):
            e.spectra(40, 0)  # no longer held: the ring keeps chunk + lag block + max lag frames
        with pytest.raises(ddm.DDMError):
            e.spectra(100, 0)  # not submitted yet
        acc = e.accumulators(0)
    cfg = o.Config(64, 64, 100, [64, 32], lags, [4.0, 8.0])
    _, ref_acc = o.run(cfg, frames, acc_scale=0)
    assert acc.shape == ref_acc.shape
    assert np.max(np.abs(acc - ref_acc)) / np.max(np.abs(ref_acc)) < 1e-5


def test_reset_window(gpu):
    """After a reset the engine reports pairs whose later frame arrived since
    the reset, normalised by the frames since the reset:
    I_window = (b I_b - a I_a) / (b - a) with I_a, I_b the oracle over [0, a), [0, b)."""
    frames = ddm.synth_uniform(60, 32, 32, 8, seed=7)
    lags, lams = [1, 2, 5, 12], [2.0, 3.0, 6.0]
    with ddm.Engine(32, 32, [32, 16], lags, lams, 1.2, 10) as e:
        e.submit_host(frames[:30])
        first = e.results()
        e.reset()
        e.submit_host(frames[30:])
        win = e.results()
        assert e.info().frames_since_reset == 30
    Ia = o.run(o.Config(32, 32, 30, [32, 16], lags, lams), frames[:30])
    Ib = o.run(o.Config(32, 32, 60, [32, 16], lags, lams), frames)
    assert o.max_rel_err(first, Ia) <= TOL
    want = (60 * Ib - 30 * Ia) / 30
    assert np.max(np.abs(win - want) / np.maximum(np.abs(want), 1e-30)) <= 1e-4


def test_brownian_256_multiscale(gpu):
    """Reduced single-stream Brownian case: 128 frames 256^2, scales 256..32,
    log-spaced lags to 100 (beyond the 30-frame chunk), log lambdas 4..128."""
    frames = ddm.synth_brownian(128, 256, 256, 8, n_spots=500, sigma=2.0, amplitude=80.0, diffusion=0.5,
                                background=100.0, noise=5.0, seed=8)
    got, ref = _both(frames, [256, 128, 64, 32], ddm.log_lags(100, 20), ddm.log_lambdas(4, 128, 16), 1.2, 30)
    err = o.max_rel_err(got, ref)
    print("brownian 256 max rel err", err)
    assert err <= TOL


def test_create_rejects_invalid_descriptors(gpu):
    with pytest.raises(ddm.DDMError) as ex:
        ddm.Engine(64, 64, [48], [1], [4.0])
    assert "powers of two" in str(ex.value)
    with pytest.raises(ddm.DDMError):
        ddm.Engine(64, 64, [64], [1], [4.0], x_off=1)
    with pytest.raises(ddm.DDMError):
        ddm.Engine(64, 64, [64], [2, 2], [4.0])


def test_autotune_schedule_and_cache(gpu, tmp_path):
    """The autotuner picks a lag block and an FFT batch from its candidates,
    results still match the oracle, a second engine of the same geometry
    reads the choice from the cache, and tuning after the first frame is
    refused."""
    import torch
    frames = ddm.synth_uniform(200, 128, 128, 8, seed=21)
    scales, lags, lams = [128, 64], [1, 2, 3, 5, 9, 17, 40], [4.0, 8.0, 16.0]
    cache = str(tmp_path / "schedule.txt")
    raw = np.ascontiguousarray(frames).view(np.uint8).reshape(-1)
    t = torch.from_numpy(raw.copy()).cuda()
    torch.cuda.synchronize()
    with ddm.Engine(128, 128, scales, lags, lams, 1.2, 20) as e:
        j, b = e.autotune(cache)
        fam, tiler = e.fft_families(), e.tiler()
        print(f"autotune: lag block {j}, FFT batch {b}, transform families {fam}, tiler {tiler}", flush=True)
        assert j in (64, 128, 256) and b in (20, 40, 80, 160)
        assert all(f in (ddm.FFT_REGISTER, ddm.FFT_LDS) for f in fam)
        assert tiler in (ddm.TILER_FUSED, ddm.TILER_PER_SCALE)
        assert e.info().lag_block == j
        e.submit_device(t.data_ptr(), frames.shape[0])
        e.sync()
        got = e.results()
        with pytest.raises(ddm.DDMError):
            e.autotune(cache)
    ref = o.run(o.Config(128, 128, frames.shape[0], scales, lags, lams, np.float32(1.2).item()), frames)
    assert o.max_rel_err(got, ref) <= TOL
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].endswith(f"\t{j} {b} {tiler} " + " ".join(str(f) for f in fam))
    with ddm.Engine(128, 128, scales, lags, lams, 1.2, 20) as e2:
        assert e2.autotune(cache) == (j, b) and e2.fft_families() == fam and e2.tiler() == tiler
    assert len(open(cache).read().splitlines()) == 1  # a hit appends nothing


@pytest.mark.parametrize("spectrum", [ddm.SPECTRUM_MASKED, ddm.SPECTRUM_FULL])
def test_lds_transform_family(gpu, spectrum):
    """The radix-4/2 LDS transforms of the 8 .. 128 tiles (the autotuner's
    alternative to the register FFTs) give the same I(q, tau) to FP32
    rounding and match the oracle; scales without the family and changes
    after the first frame are refused."""
    frames = ddm.synth_uniform(90, 256, 256, 12, seed=29)
    scales, lags, lams = [256, 128, 64, 32, 16, 8], [1, 2, 5, 11, 30], ddm.log_lambdas(3, 64, 9)
    args = (256, 256, scales, lags, lams, 1.2, 30, 2, ddm.PIX_U16LE)
    out = {}
    for fam in (ddm.FFT_REGISTER, ddm.FFT_LDS):
        with ddm.Engine(*args, spectrum=spectrum) as e:
            for si, s in enumerate(e.scales):
                if 8 <= s <= 128:
                    e.set_fft_family(si, fam)
                elif fam == ddm.FFT_LDS:
                    with pytest.raises(ddm.DDMError):
                        e.set_fft_family(si, fam)
            assert e.fft_families() == [fam if 8 <= s <= 128 else ddm.FFT_REGISTER for s in e.scales]
            e.submit_host(frames.reshape(-1).view(np.uint8), len(frames))
            out[fam] = e.results()
            with pytest.raises(ddm.DDMError):
                e.set_fft_family(1, ddm.FFT_REGISTER)
    ref = o.run(o.Config(256, 256, 90, scales, lags, lams, np.float32(1.2).item(), 2, ddm.PIX_U16LE), frames)
    for fam in out:
        err = o.max_rel_err(out[fam], ref)
        print(f"spectrum {spectrum} family {fam}: max rel err {err:.2e}")
        assert err <= TOL
    assert o.max_rel_err(out[ddm.FFT_LDS], out[ddm.FFT_REGISTER]) <= 2e-6
