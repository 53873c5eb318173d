"""Stage kernels through the C ABI on caller-owned device buffers:
tile cut + FFT against numpy (FP64) with the DC bin bit-exact, and the lag
engine bit-exact on integer-valued spectra injected into the ring."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm

pytestmark = pytest.mark.gpu

# FP32 FFT: error relative to the largest bin, bound ~ log2(s^2) * 2^-24
FFT_TOL = 4e-6


G = 64  # ring group: ring[((e // G) * R + slot) * G + e % G]


def pad64(e):
    return (e + G - 1) // G * G


def ring_to_frames(ring_flat, R, E_pad):
    """Blocked ring layout -> [slot][elem] complex."""
    r = np.asarray(ring_flat).reshape(E_pad // G, R, G, 2).transpose(1, 0, 2, 3).reshape(R, E_pad, 2)
    return r[..., 0].astype(np.float64) + 1j * r[..., 1]


def frames_to_ring(re, im):
    """[slot][elem] integer parts (elem count a multiple of 64) -> blocked flat float32 ring."""
    R, E = re.shape
    a = np.stack([re, im], -1).astype(np.float32).reshape(R, E // G, G, 2).transpose(1, 0, 2, 3)
    return np.ascontiguousarray(a).reshape(-1)


def _run_tile_fft(gpu, frames, bpp, fmt, M, s, x_off=0, y_off=0, ring_slots=None, first_frame=0, width=None):
    import torch
    n, h = frames.shape[0], frames.shape[1]
    w = width or frames.shape[2]
    raw = np.ascontiguousarray(frames).view(np.uint8).reshape(n, -1)
    frame_bytes = raw.shape[1]
    T = (M // s) ** 2
    E = T * s * (s // 2 + 1)
    Ep = pad64(E)
    R = ring_slots or n
    d_frames = torch.from_numpy(raw.copy()).to(gpu)
    ring = torch.zeros(R * Ep * 2, dtype=torch.float32, device=gpu)
    print(f"tile_fft M={M} s={s} n={n} fmt={fmt} bpp={bpp}", flush=True)
    ddm.k_tile_fft(d_frames, n, frame_bytes, w, h, bpp, fmt, x_off, y_off, M, s, ring, R, Ep, 0, first_frame)
    torch.cuda.synchronize()
    return ring_to_frames(ring.cpu().numpy(), R, Ep)[:, :E].reshape(R, T, s, s // 2 + 1)


def _tile(img, M, s, tile, x_off, y_off):
    tps = M // s
    ty, tx = divmod(tile, tps)
    return img[y_off + ty * s:y_off + (ty + 1) * s, x_off + tx * s:x_off + (tx + 1) * s].astype(np.float64)


@pytest.mark.parametrize("s", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048])
def test_tile_fft_single_tile(gpu, s):
    rng = np.random.default_rng(s)
    n = 2 if s >= 1024 else 3
    frames = rng.integers(0, 256, (n, s, s), dtype=np.uint8)
    spec = _run_tile_fft(gpu, frames, 1, ddm.PIX_U8, s, s)
    for f in range(n):
        ref = np.fft.rfft2(frames[f].astype(np.float64))
        err = np.max(np.abs(spec[f, 0] - ref)) / np.max(np.abs(ref))
        assert err < FFT_TOL, (s, f, err)
        total = frames[f].astype(np.int64).sum()
        if total < 2 ** 24:  # DC is a sum of integers: exact in FP32 below 2^24
            assert spec[f, 0, 0, 0] == total


@pytest.mark.parametrize("s", [512, 1024, 2048])
def test_tile_fft_dc_exact_large(gpu, s):
    """Values 0..3 keep every partial sum of a tile below 2^24: DC bit-exact."""
    frames = np.random.default_rng(s + 1).integers(0, 4, (1, s, s), dtype=np.uint8)
    spec = _run_tile_fft(gpu, frames, 1, ddm.PIX_U8, s, s)
    assert spec[0, 0, 0, 0] == frames[0].astype(np.int64).sum()


@pytest.mark.parametrize("M,s", [(64, 8), (64, 16), (128, 32), (256, 64), (256, 128), (512, 256), (1024, 512)])
def test_tile_fft_tiles_offsets_and_ring_wrap(gpu, M, s):
    rng = np.random.default_rng(M + s)
    n, x_off, y_off = 3, 5, 3
    frames = rng.integers(0, 256, (n, M + 9, M + 7), dtype=np.uint8)
    R, first = 4, 2  # frames land in slots 2, 3, 0
    spec = _run_tile_fft(gpu, frames, 1, ddm.PIX_U8, M, s, x_off, y_off, R, first)
    for f in range(n):
        slot = (first + f) % R
        for tile in range((M // s) ** 2):
            t = _tile(frames[f], M, s, tile, x_off, y_off)
            ref = np.fft.rfft2(t)
            got = spec[slot, tile]
            if t.sum() < 2 ** 24:
                assert got[0, 0] == t.sum(), (f, tile)  # tile membership, bit-exact
            assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < FFT_TOL


@pytest.mark.parametrize("fmt,bpp", [(ddm.PIX_U16LE, 2), (ddm.PIX_U16BE, 2), (ddm.PIX_U8, 3), (ddm.PIX_U16LE, 4)])
def test_tile_fft_pixel_formats(gpu, fmt, bpp):
    rng = np.random.default_rng(fmt * 10 + bpp)
    M, s, n = 64, 32, 2
    # 12-bit values: a 32x32 tile sums below 2^24, so DC is exact
    vals = rng.integers(0, 4096 if fmt != ddm.PIX_U8 else 256, (n, M, M + 4)).astype(np.uint32)
    raw = np.zeros((n, M, M + 4, bpp), dtype=np.uint8)
    if fmt == ddm.PIX_U8:
        raw[..., 0] = vals
        raw[..., 1:] = rng.integers(0, 256, raw[..., 1:].shape)  # other channels ignored
    elif fmt == ddm.PIX_U16LE:
        raw[..., 0], raw[..., 1] = vals & 0xFF, vals >> 8
    else:
        raw[..., 0], raw[..., 1] = vals >> 8, vals & 0xFF
    if bpp > 2 and fmt != ddm.PIX_U8:
        raw[..., 2:] = 77
    spec = _run_tile_fft(gpu, raw.reshape(n, M, -1), bpp, fmt, M, s, x_off=4, width=M + 4)
    spec = spec.reshape(n, 4, s, s // 2 + 1)
    for f in range(n):
        for tile in range(4):
            t = _tile(vals[f], M, s, tile, 4, 0)
            assert spec[f, tile, 0, 0] == t.sum()
            ref = np.fft.rfft2(t)
            assert np.max(np.abs(spec[f, tile] - ref)) / np.max(np.abs(ref)) < FFT_TOL


@pytest.mark.parametrize("s", [64, 128, 256, 512, 1024])
@pytest.mark.parametrize("fmt,bpp", [(ddm.PIX_U16LE, 2), (ddm.PIX_U16BE, 2), (ddm.PIX_U16LE, 4), (ddm.PIX_U8, 1)])
def test_tile_fft_pixel_formats_large_tiles(gpu, s, fmt, bpp):
    """The 64..1024 tile kernels read 1-byte and 2-byte little-endian pixels
    with their own loads; other layouts take the general pixel reader."""
    rng = np.random.default_rng(s + fmt * 10 + bpp)
    M, n, xo, yo = 2 * s, 2, 3, 1
    hi = 256 if fmt == ddm.PIX_U8 else 4096
    vals = rng.integers(0, hi, (n, M + yo, M + xo + 2)).astype(np.uint32)
    raw = np.zeros(vals.shape + (bpp,), dtype=np.uint8)
    if fmt == ddm.PIX_U8:
        raw[..., 0] = vals
    elif fmt == ddm.PIX_U16LE:
        raw[..., 0], raw[..., 1] = vals & 0xFF, vals >> 8
    else:
        raw[..., 0], raw[..., 1] = vals >> 8, vals & 0xFF
    if bpp > 2:
        raw[..., 2:] = 77
    spec = _run_tile_fft(gpu, raw.reshape(n, M + yo, -1), bpp, fmt, M, s, x_off=xo, y_off=yo, width=M + xo + 2)
    spec = spec.reshape(n, 4, s, s // 2 + 1)
    for f in range(n):
        for tile in range(4):
            t = _tile(vals[f], M, s, tile, xo, yo)
            if t.sum() < 2 ** 24:
                assert spec[f, tile, 0, 0] == t.sum()
            ref = np.fft.rfft2(t)
            assert np.max(np.abs(spec[f, tile] - ref)) / np.max(np.abs(ref)) < FFT_TOL


def _lag_expected(ring, R, t0, n_new, lags):
    E = ring.shape[1]
    acc = np.zeros((len(lags), E), dtype=np.int64)
    for j in range(n_new):
        t = t0 + j
        for li, tau in enumerate(lags):
            if t - tau < 0:
                continue
            d = ring[t % R] - ring[(t - tau) % R]
            acc[li] += (d.real.astype(np.int64) ** 2 + d.imag.astype(np.int64) ** 2)
    return acc


@pytest.mark.parametrize("J", [32, 64, 128, 256])
@pytest.mark.parametrize("t0,n_new", [(0, None), (3, 7), (100, None), (57, 1), (200, 29)])
@pytest.mark.parametrize("nlags", [10, 40, 70])
def test_lag_engine_bit_exact(gpu, J, t0, n_new, nlags):
    """Integer spectra: every |dF|^2 and every partial sum is an integer below
    2^24, so FP32 gives the exact answer whatever the summation order.
    With 256-frame blocks, 10 lags take the history-major path and 40 or 70
    the per-lag units."""
    import torch
    n_new = J if n_new is None else n_new
    base = [0, 1, 2, 5, 15, 16, 17, 31, 33, 40, 63, 64, 65, 127, 128, 129, 150]
    lags = sorted(set(base + list(range(3, 3 + 2 * nlags, 2))))[:nlags]
    R = max(lags) + J + 3
    E = 15 * G
    rng = np.random.default_rng(t0 * 7 + J + n_new + nlags)
    re = rng.integers(-8, 9, (R, E))
    im = rng.integers(-8, 9, (R, E))
    ring_np = re + 1j * im
    ring = torch.from_numpy(frames_to_ring(re, im)).to(gpu)
    acc0 = rng.integers(0, 1000, (len(lags), E))
    acc = torch.from_numpy(acc0.astype(np.float32).reshape(-1)).to(gpu)
    d_lags = torch.tensor(lags, dtype=torch.int32, device=gpu)
    print(f"lag J={J} t0={t0} n_new={n_new} L={len(lags)}", flush=True)
    ddm.k_lag_accumulate(ring, R, E, t0, n_new, d_lags, max(lags), acc, J)
    torch.cuda.synchronize()
    got = acc.cpu().numpy().reshape(len(lags), E).astype(np.int64)
    want = acc0 + _lag_expected(ring_np, R, t0, n_new, lags)
    assert np.array_equal(got, want)


DENSE_LAG_SETS = {
    # leading runs of consecutive lags take the register-window path
    "log32": list(range(1, 15)) + [17, 20, 25, 31, 37, 46, 56, 68, 84, 102, 125, 153, 187, 229],
    "run70": list(range(1, 71)) + [90, 300],  # run capped at 64 (32 for J = 256), rest single
    "from5": list(range(5, 30)) + [40, 41],   # run not starting at 1
    "zero4": [0, 1, 2, 3, 9],                 # shortest run that is dense, lag 0 included
    "short3": [1, 2, 3, 50],                  # run too short: all single
    "run16": list(range(2, 18)),               # exactly one full chunk
    # 33-run: 32 window lags, the 33rd joins the 31 log lags, 32 history-major lags in all
    "log64": list(range(1, 34)) + [37, 42, 46, 52, 58, 64, 72, 80, 90, 100, 112, 124, 139, 155, 173, 193, 215,
                                   240, 268, 299, 334, 372, 416, 464, 517, 577, 644, 719, 802, 895, 999],
    "sparse33": [3 * k + 1 for k in range(33)],  # no run, one lag too many for the history-major path
    # more than 16 history-major lags beside a window run that is not a
    # multiple of 16: the partial sums meet in two rounds and the second
    # round's lags are not aligned with the wave index (the 512^2 workload's
    # 32 log lags up to 512, and a 9-run with 23 lags behind it)
    "log32_512": list(range(1, 15)) + [17, 20, 25, 31, 37, 46, 56, 68, 84, 102, 125, 153, 187, 229, 280, 342,
                                       419, 512],
    "run9_hist23": list(range(1, 10)) + [12 + 9 * k for k in range(23)],
}


@pytest.mark.parametrize("J", [32, 64, 128, 256])
@pytest.mark.parametrize("t0,n_new", [(0, None), (3, 7), (20, None), (57, 1), (300, 29), (500, None)])
@pytest.mark.parametrize("name", sorted(DENSE_LAG_SETS))
def test_lag_engine_dense_runs_bit_exact(gpu, J, t0, n_new, name):
    """Integer spectra, lag sets with runs of consecutive lags; the ring is
    only max lag + n_new slots when that is legal, so stale slots sit right
    behind the valid history."""
    import torch
    lags = DENSE_LAG_SETS[name]
    n_new = J if n_new is None else n_new
    R = max(lags) + n_new + (0 if t0 % 2 else 5)
    E = 3 * G
    rng = np.random.default_rng(t0 * 11 + J + n_new + len(lags))
    re = rng.integers(-8, 9, (R, E))
    im = rng.integers(-8, 9, (R, E))
    ring_np = re + 1j * im
    ring = torch.from_numpy(frames_to_ring(re, im)).to(gpu)
    acc0 = rng.integers(0, 1000, (len(lags), E))
    acc = torch.from_numpy(acc0.astype(np.float32).reshape(-1)).to(gpu)
    d_lags = torch.tensor(lags, dtype=torch.int32, device=gpu)
    print(f"dense lag J={J} t0={t0} n_new={n_new} set={name} R={R}", flush=True)
    ddm.k_lag_accumulate(ring, R, E, t0, n_new, d_lags, max(lags), acc, J)
    torch.cuda.synchronize()
    got = acc.cpu().numpy().reshape(len(lags), E).astype(np.int64)
    want = acc0 + _lag_expected(ring_np, R, t0, n_new, lags)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("t0", [0, 300])
@pytest.mark.parametrize("nhist", [1, 5, 15, 16, 17, 18, 24, 31, 32, 33])
@pytest.mark.parametrize("run", [0, 8, 9, 14, 16, 23, 32])
def test_lag_engine_run_by_history_grid(gpu, run, nhist, t0):
    """256-frame blocks, a window run of `run` consecutive lags followed by
    `nhist` spread lags: the window / history split, the per-wave history
    block lists and the rounds in which partial sums meet (16 or 32 lags per
    round, lag counts on either side of each round) against exact integer
    sums, with nonzero accumulators so that a lag given another lag's
    accumulator shows."""
    _check_run_history(gpu, run, nhist, 7, t0)


@pytest.mark.parametrize("t0", [0, 4500])
@pytest.mark.parametrize("run,nhist", [(0, 18), (14, 18), (9, 32), (16, 31)])
def test_lag_engine_run_by_history_wide(gpu, run, nhist, t0):
    """As above with history lags 130 frames apart (largest 2 213 … 4 042):
    with 31 or 32 of them the history window holds more 16-frame blocks
    than the plan lists (261, 269 > 252), so the history waves take blocks
    round-robin; with 18 the plan lists ~155 blocks."""
    _check_run_history(gpu, run, nhist, 130, t0)


def _check_run_history(gpu, run, nhist, step, t0):
    import torch
    J = 256
    lags = list(range(1, run + 1)) + [run + 3 + step * k for k in range(nhist)]
    n_new = J
    R = max(lags) + n_new + 5
    E = 2 * G
    rng = np.random.default_rng(1000 * run + 10 * nhist + t0 + step)
    re = rng.integers(-8, 9, (R, E))
    im = rng.integers(-8, 9, (R, E))
    ring_np = re + 1j * im
    ring = torch.from_numpy(frames_to_ring(re, im)).to(gpu)
    acc0 = rng.integers(0, 1 << 16, (len(lags), E))
    acc = torch.from_numpy(acc0.astype(np.float32).reshape(-1)).to(gpu)
    d_lags = torch.tensor(lags, dtype=torch.int32, device=gpu)
    print(f"lag grid run={run} nhist={nhist} step={step} t0={t0} R={R}", flush=True)
    ddm.k_lag_accumulate(ring, R, E, t0, n_new, d_lags, max(lags), acc, J)
    torch.cuda.synchronize()
    got = acc.cpu().numpy().reshape(len(lags), E).astype(np.int64)
    want = acc0 + _lag_expected(ring_np, R, t0, n_new, lags)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", range(24))
def test_lag_engine_random_lists_bit_exact(gpu, seed):
    """Random lag lists (a run of consecutive lags of random length and start,
    then spread lags), block sizes, block starts and ring sizes, so that the
    staged frames before the block, the 16-byte staging of one run of ring
    slots and the per-frame staging across the ring's wrap, the register
    window and the history-major blocks meet in combinations the fixed cases
    do not pick; exact integer sums."""
    import torch
    rng = np.random.default_rng(7000 + seed)
    J = int(rng.choice([32, 64, 128, 256, 256, 256]))
    run0 = int(rng.integers(0, 3))
    run = int(rng.integers(0, 40))
    lags = list(range(run0, run0 + run))
    nspread = int(rng.integers(1, 34))
    last = lags[-1] if lags else run0
    for _ in range(nspread):
        last += int(rng.integers(2, 60))
        lags.append(last)
    lags = sorted(set(lags))
    n_new = int(rng.integers(1, J + 1)) if rng.random() < 0.4 else J
    R = max(lags) + n_new + int(rng.integers(0, 50))
    t0 = int(rng.integers(0, 3 * R))
    E = int(rng.integers(1, 5)) * G
    re = rng.integers(-8, 9, (R, E))
    im = rng.integers(-8, 9, (R, E))
    ring_np = re + 1j * im
    ring = torch.from_numpy(frames_to_ring(re, im)).to(gpu)
    acc0 = rng.integers(0, 1 << 16, (len(lags), E))
    acc = torch.from_numpy(acc0.astype(np.float32).reshape(-1)).to(gpu)
    d_lags = torch.tensor(lags, dtype=torch.int32, device=gpu)
    print(f"random lag seed={seed} J={J} t0={t0} n_new={n_new} R={R} L={len(lags)} run={run}@{run0}", flush=True)
    ddm.k_lag_accumulate(ring, R, E, t0, n_new, d_lags, max(lags), acc, J)
    torch.cuda.synchronize()
    got = acc.cpu().numpy().reshape(len(lags), E).astype(np.int64)
    want = acc0 + _lag_expected(ring_np, R, t0, n_new, lags)
    assert np.array_equal(got, want)


def test_stage_entry_points_reject_bad_arguments(gpu):
    import torch
    ring = torch.zeros(16, dtype=torch.float32, device=gpu)
    frames = torch.zeros(64, dtype=torch.uint8, device=gpu)
    lags = torch.tensor([1], dtype=torch.int32, device=gpu)
    with pytest.raises(ddm.DDMError):  # scale not a power of two
        ddm.k_tile_fft(frames, 1, 64, 8, 8, 1, 0, 0, 0, 8, 6, ring, 1, 8, 0, 0)
    with pytest.raises(ddm.DDMError):  # ROI outside the frame
        ddm.k_tile_fft(frames, 1, 64, 8, 8, 1, 0, 1, 0, 8, 8, ring, 1, 40, 0, 0)
    with pytest.raises(ddm.DDMError):  # spectra would overrun the ring element range
        ddm.k_tile_fft(frames, 1, 64, 8, 8, 1, 0, 0, 0, 8, 8, ring, 1, 8, 0, 0)
    big = torch.zeros(64 * 40 * 2, dtype=torch.float32, device=gpu)
    with pytest.raises(ddm.DDMError):  # n_new above the lag block
        ddm.k_lag_accumulate(big, 40, 64, 0, 33, lags, 1, big, 32)
    with pytest.raises(ddm.DDMError):  # ring too short for max lag
        ddm.k_lag_accumulate(big, 40, 64, 0, 32, lags, 9, big, 32)
    with pytest.raises(ddm.DDMError):  # element count not a whole number of ring groups
        ddm.k_lag_accumulate(big, 40, 60, 0, 4, lags, 1, big, 32)
    with pytest.raises(ddm.DDMError):  # unsupported lag block
        ddm.k_lag_accumulate(big, 40, 64, 0, 4, lags, 1, big, 16)
    for bad in ([2, 1], [1, 1], [1, 5]):  # not strictly ascending, or above max_lag
        with pytest.raises(ddm.DDMError):
            ddm.k_lag_accumulate(big, 40, 64, 0, 4, torch.tensor(bad, dtype=torch.int32, device=gpu), 4, big, 32)
