"""Writes the golden fixtures under tests/golden/ with numpy alone (no code
of this repository and none of the reference is run):

  rfft2_golden.npz  seeded uint8 tiles, s = 2..64, and numpy.fft.rfft2 of
                    each in float64 -- pins the oracle's FFT.
  isf_golden.npz    a tiny end-to-end case (12 frames of 32x32 uint8,
                    scales 32/16/8, lags 1,2,3,5,13, lambdas 1.5..40) with
                    I(q, tau) from numpy: rfft2 per tile, direct lag loops
                    over valid pairs, the reference's FP32 ring test
                    (src/azimuthal_average.cu:235-268) in numpy float32, and
                    I = 2 * sum / count / N (src/azimuthal_average.cu:334-336).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def ring_mask_np(s, lam, tol):
    q = np.float32(s) / np.float32(lam)
    q2 = np.float32(q * q)
    tol2 = np.float32(np.float32(tol) * np.float32(tol))
    x = np.arange(s // 2 + 1)
    y = np.arange(s)
    xs = (x + s // 2) % s - s // 2
    ys = (y + s // 2) % s - s // 2
    r2 = (ys[:, None] ** 2 + xs[None, :] ** 2).astype(np.float32)
    ratio = (r2 / q2).astype(np.float32)
    return (ratio >= np.float32(1.0)) & (ratio <= tol2)


def isf_np(frames, scales, lags, lambdas, tol):
    N = frames.shape[0]
    M = max(scales)
    out = []
    for s in sorted(scales, reverse=True):
        tps = M // s
        masks = [ring_mask_np(s, lam, tol) for lam in sorted(lambdas)]
        for tile in range(tps * tps):
            ty, tx = divmod(tile, tps)
            t = frames[:, ty * s:(ty + 1) * s, tx * s:(tx + 1) * s].astype(np.float64)
            F = np.fft.rfft2(t) / (s * s)
            res = np.zeros((len(lambdas), len(lags)))
            for li, tau in enumerate(sorted(lags)):
                acc = np.zeros(F.shape[1:])
                for a in range(0, N - tau):
                    acc += np.abs(F[a] - F[a + tau]) ** 2
                for qi, m in enumerate(masks):
                    c = int(m.sum())
                    res[qi, li] = 2.0 * acc[m].sum() / c / N if c else 0.0
            out.append(res)
    return np.concatenate([r.ravel() for r in out])


def main():
    rng = np.random.default_rng(20261203)
    tiles, specs = {}, {}
    for s in (2, 4, 8, 16, 32, 64):
        t = rng.integers(0, 256, size=(s, s), dtype=np.uint8)
        tiles[f"tile_{s}"] = t
        specs[f"rfft2_{s}"] = np.fft.rfft2(t.astype(np.float64))
    np.savez_compressed(os.path.join(HERE, "rfft2_golden.npz"), **tiles, **specs)

    frames = rng.integers(0, 256, size=(12, 32, 32), dtype=np.uint8)
    scales = [8, 32, 16]          # unsorted on purpose: the engine sorts descending
    lags = [5, 1, 2, 3, 13]       # 13 > frames - 1: no valid pair, I = 0
    lambdas = [4.0, 1.5, 2.0, 3.0, 6.5, 40.0]  # 40 > 1.2 * 32: empty ring
    tol = np.float32(1.0 + 20 / 100.0)
    iqt = isf_np(frames, scales, lags, lambdas, tol)
    np.savez_compressed(os.path.join(HERE, "isf_golden.npz"), frames=frames, scales=np.array(scales),
                        lags=np.array(lags), lambdas=np.array(lambdas, dtype=np.float32),
                        tol=np.float32(tol), iqt=iqt)


if __name__ == "__main__":
    main()
