"""Pins the CPU oracle (oracle/ddm_oracle.c) before it is trusted as the
checker: golden vectors from numpy, closed forms of the DFT, brute force,
closed forms of the lag sums and the average, and exact arithmetic for the
q-ring mask.  Each check is chosen so that a plausible slip (a dropped
term, a wrong sign or index, a transposed operand) fails it."""
import os

import numpy as np
import pytest

from oracle import oracle as o

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("s", [2, 4, 8, 16, 32, 64])
def test_rfft2_matches_numpy_golden(s):
    g = np.load(os.path.join(GOLD, "rfft2_golden.npz"))
    got = o.rfft2(g[f"tile_{s}"].astype(np.float64))
    ref = g[f"rfft2_{s}"]
    assert got.shape == (s, s // 2 + 1)
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("s", [2, 4, 8, 16])
def test_rfft2_matches_brute_force_dft(s):
    t = np.random.default_rng(s).integers(0, 256, (s, s)).astype(np.float64)
    assert np.allclose(o.rfft2(t), o.dft2_brute(t), rtol=0, atol=1e-9)


def test_rfft2_closed_forms():
    s = 32
    # constant tile: only DC = s^2 c
    F = o.rfft2(np.full((s, s), 7.0))
    assert F[0, 0] == pytest.approx(7.0 * s * s)
    F[0, 0] = 0
    assert np.max(np.abs(F)) < 1e-9
    # delta at (y0, x0): F = exp(-2 pi i (ky y0 + kx x0) / s)
    y0, x0 = 5, 11
    d = np.zeros((s, s))
    d[y0, x0] = 1.0
    ky, kx = np.meshgrid(np.arange(s), np.arange(s // 2 + 1), indexing="ij")
    assert np.allclose(o.rfft2(d), np.exp(-2j * np.pi * (ky * y0 + kx * x0) / s), atol=1e-12)
    # plane wave cos(2 pi (a x + b y) / s): s^2/2 at (b, a) and (s-b, ...) only
    a, b = 3, 7
    yy, xx = np.mgrid[0:s, 0:s]
    F = o.rfft2(np.cos(2 * np.pi * (a * xx + b * yy) / s))
    assert F[b, a] == pytest.approx(s * s / 2, abs=1e-9)
    F[b, a] = 0
    assert np.max(np.abs(F)) < 1e-9
    # transposed operand would put the peak at (a, b)
    assert np.abs(o.rfft2(np.cos(2 * np.pi * (a * xx + b * yy) / s))[a, b]) < 1e-9


def test_rfft2_parseval_and_linearity():
    s = 64
    rng = np.random.default_rng(3)
    x, y = rng.random((s, s)), rng.random((s, s))
    Fx, Fy = o.rfft2(x), o.rfft2(y)
    assert np.allclose(o.rfft2(2 * x - 3 * y), 2 * Fx - 3 * Fy, atol=1e-9)
    # Parseval over the half spectrum: columns 1..s/2-1 count twice
    w = np.ones(s // 2 + 1)
    w[1:s // 2] = 2
    assert (np.abs(Fx) ** 2 * w).sum() == pytest.approx(s * s * (x * x).sum(), rel=1e-12)


def test_rfft2_fp32_baseline_close_to_fp64():
    s = 128
    t = np.random.default_rng(1).integers(0, 256, (s, s)).astype(np.float64)
    ref = o.rfft2(t)
    assert np.max(np.abs(o.rfft2(t, "f32") - ref)) < 1e-6 * np.max(np.abs(ref))


# --------------------------------------------------------------------------- mask
@pytest.mark.parametrize("s", [8, 64, 512])
def test_mask_matches_exact_radius_away_from_rounding(s):
    """Membership equals 1 <= r^2/q^2 <= tol^2 in exact arithmetic wherever the
    exact ratio is not within FP32 rounding of a bound."""
    tol = np.float32(1.2)
    for lam in [2.0, 3.7, 4.0, 10.0, 33.3, s * 0.5]:
        m = o.mask(s, lam, tol)
        q = float(np.float32(s) / np.float32(lam))
        x = np.arange(s // 2 + 1)
        y = np.arange(s)
        xs = (x + s // 2) % s - s // 2
        ys = (y + s // 2) % s - s // 2
        r2 = (ys[:, None] ** 2 + xs[None, :] ** 2).astype(np.float64)
        ratio = r2 / (q * q)
        t2 = float(tol) ** 2
        clear = (np.abs(ratio - 1.0) > 1e-5) & (np.abs(ratio - t2) > 1e-5)
        exact = (ratio >= 1.0) & (ratio <= t2)
        assert np.array_equal(m.astype(bool)[clear], exact[clear])
        # the frequency shift: kx = s/2 stands for -s/2, ky >= s/2 for ky - s
        assert m.shape == (s, s // 2 + 1)


def test_mask_special_cases():
    # lambda > tol * s: q < 1/tol, every r >= 1 is outside, DC (r = 0) too
    assert o.mask(64, 64 * 1.25, np.float32(1.2)).sum() == 0
    # lambda = s: q = 1, ring holds r^2 in {1} (r^2 = 2 > 1.44): the 4 axis
    # neighbours of DC, of which the half spectrum stores 3 (kx = 0 twice, kx = 1)
    m = o.mask(64, 64.0, np.float32(1.2))
    assert m.sum() == 3 and m[1, 0] and m[63, 0] and m[0, 1]


def test_mask_matches_numpy_fp32_restatement():
    from tests.golden.make_golden import ring_mask_np
    for s in (16, 128):
        for lam in (1.5, 2.0, 4.2, 7.0, 12.0):
            assert np.array_equal(o.mask(s, lam, np.float32(1.2)).astype(bool),
                                  ring_mask_np(s, lam, np.float32(1.2)))


# --------------------------------------------------------------------------- pipeline
def test_pipeline_matches_numpy_golden():
    g = np.load(os.path.join(GOLD, "isf_golden.npz"))
    cfg = o.Config(32, 32, 12, list(g["scales"]), list(g["lags"]), list(g["lambdas"]), float(g["tol"]))
    got = o.run(cfg, g["frames"])
    assert o.max_rel_err(got, g["iqt"]) < 1e-12


def _cfg(frames, scales, lags, lambdas, **kw):
    n, h, w = frames.shape[:3]
    return o.Config(w, h, n, scales, lags, lambdas, **kw)


def test_lag_sums_identical_frames_are_zero():
    f = np.repeat(np.random.default_rng(0).integers(0, 256, (1, 32, 32), dtype=np.uint8), 9, axis=0)
    _, acc = o.run(_cfg(f, [32], [1, 2, 5], [4.0]), f, acc_scale=0)
    assert np.all(acc == 0)


def test_lag_sums_alternating_frames():
    """a, b, a, b, ...: odd lags see |F_a - F_b|^2 in every pair, even lags 0."""
    rng = np.random.default_rng(1)
    a, b = rng.integers(0, 256, (2, 16, 16), dtype=np.uint8)
    N = 11
    f = np.stack([a if t % 2 == 0 else b for t in range(N)])
    lags = [1, 2, 3, 4]
    _, acc = o.run(_cfg(f, [16], lags, [4.0]), f, acc_scale=0)
    D = np.abs(np.fft.rfft2(a.astype(float) - b.astype(float))) ** 2 / 16 ** 4
    for li, tau in enumerate(lags):
        want = (N - tau) * D if tau % 2 else np.zeros_like(D)
        assert np.allclose(acc[li, 0], want, rtol=1e-12, atol=1e-12)


def test_lag_sums_linear_ramp():
    """Frame t = base + t: F_t - F_{t+tau} = -tau s^2 at DC only;
    the DC sum is (N - tau) tau^2 (per-term 1/s^2 factor included)."""
    base = np.random.default_rng(2).integers(0, 100, (16, 16)).astype(np.uint8)
    N = 20
    f = np.stack([base + t for t in range(N)]).astype(np.uint8)
    lags = [1, 3, 7]
    _, acc = o.run(_cfg(f, [16], lags, [4.0]), f, acc_scale=0)
    for li, tau in enumerate(lags):
        assert acc[li, 0, 0, 0] == pytest.approx((N - tau) * tau * tau, rel=1e-12)
        rest = acc[li, 0].copy()
        rest[0, 0] = 0
        assert np.max(np.abs(rest)) < 1e-18


def test_average_closed_form_alternating():
    """For alternating frames I(q, tau odd) = 2 (N - tau) / N * mean_ring |F_a - F_b|^2 / s^4."""
    rng = np.random.default_rng(5)
    a, b = rng.integers(0, 256, (2, 32, 32), dtype=np.uint8)
    N, s = 10, 32
    f = np.stack([a if t % 2 == 0 else b for t in range(N)])
    lams = [2.0, 5.0, 9.0]
    res = o.split_results(_cfg(f, [s], [1, 3], lams), o.run(_cfg(f, [s], [1, 3], lams), f))[s][0]
    D = np.abs(np.fft.rfft2(a.astype(float) - b.astype(float))) ** 2 / s ** 4
    for qi, lam in enumerate(lams):
        m = o.mask(s, lam, np.float32(1.2)).astype(bool)
        for li, tau in enumerate([1, 3]):
            assert res[qi, li] == pytest.approx(2 * (N - tau) / N * D[m].mean(), rel=1e-12)


def test_pair_mean_normalisation():
    f = np.random.default_rng(6).integers(0, 256, (9, 16, 16), dtype=np.uint8)
    lags = [1, 4]
    c1 = _cfg(f, [16], lags, [3.0, 5.0])
    c2 = _cfg(f, [16], lags, [3.0, 5.0], norm_mode=o.NORM_PAIRS)
    r1, r2 = o.run(c1, f).reshape(2, 2), o.run(c2, f).reshape(2, 2)
    for li, tau in enumerate(lags):
        assert np.allclose(r2[:, li], r1[:, li] * 9 / (9 - tau), rtol=1e-13)


def test_tiling_reference_order():
    """tile j = ty * (M/s) + tx at ROI offset; 16-bit and multi-byte pixels."""
    rng = np.random.default_rng(7)
    h, w = 20, 24
    f16 = rng.integers(0, 4096, (2, h, w)).astype("<u2")
    cfg = o.Config(w, h, 2, [16, 8], [1], [4.0], bpp=2, pixfmt=o.PIX_U16LE, x_off=3, y_off=2)
    t = o.tile(cfg, f16.view(np.uint8), 8, 3, 1)  # tile 3 = (ty 1, tx 1)
    assert np.array_equal(t, f16[1, 2 + 8:2 + 16, 3 + 8:3 + 16].astype(float))
    be = f16.astype(">u2")
    cfgb = o.Config(w, h, 2, [16, 8], [1], [4.0], bpp=2, pixfmt=o.PIX_U16BE, x_off=3, y_off=2)
    assert np.array_equal(o.tile(cfgb, be.view(np.uint8), 8, 3, 1), t)
    rgb = rng.integers(0, 256, (1, h, w, 3), dtype=np.uint8)
    cfg_rgb = o.Config(w, h, 1, [8], [1], [4.0], bpp=3, x_off=5, y_off=4)
    assert np.array_equal(o.tile(cfg_rgb, rgb, 8, 0, 0), rgb[0, 4:12, 5:13, 0].astype(float))


def test_fp32_baseline_close_to_oracle():
    f = np.random.default_rng(8).integers(0, 256, (40, 64, 64), dtype=np.uint8)
    cfg = _cfg(f, [64, 32], [1, 2, 5, 11], [4.0, 6.0, 9.0, 14.0, 20.0])
    r64 = o.run(cfg, f, "f64")
    r32 = o.run(cfg, f, "f32", nthreads=4)
    assert o.max_rel_err(r32, r64) < 1e-4


def test_oracle_rejects_bad_parameters():
    f = np.zeros((4, 16, 16), np.uint8)
    with pytest.raises(ValueError):
        o.run(_cfg(f, [12], [1], [4.0]), f)          # scale not a power of two
    with pytest.raises(ValueError):
        o.run(_cfg(f, [16], [1, 1], [4.0]), f)       # repeated lag
    with pytest.raises(ValueError):
        o.run(o.Config(16, 16, 4, [16], [1], [4.0], x_off=1), f)  # ROI outside
