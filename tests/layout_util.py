"""The spectrum elements a masked engine keeps per tile, computed from the
oracle's q-ring masks: for 8 .. 128 tiles the prefix of every row up to its
last ring member, for 256 .. 1024 tiles the low |ky| rows of every column
down to its last ring member (both signs of ky); the whole half spectrum when
that is 90 % of it or more, or for 2 / 4 / 2048 tiles.  A scale without any
ring member keeps its DC element."""
import numpy as np

from oracle import oracle as o


def kept_mask(s, lambdas, tol):
    H = s // 2
    full = np.ones((s, H + 1), dtype=bool)
    if s < 8 or s > 1024:
        return full
    union = np.zeros((s, H + 1), dtype=bool)
    for lam in lambdas:
        union |= o.mask(s, lam, tol).astype(bool)
    if not union.any():
        union[0, 0] = True
    keep = np.zeros_like(union)
    if s >= 256:
        for kx in range(H + 1):
            ky = np.nonzero(union[:, kx])[0]
            lo = int(ky[ky < H].max()) + 1 if np.any(ky < H) else 0
            hi = int(s - ky[ky >= H].min()) if np.any(ky >= H) else 0
            keep[:lo, kx] = True
            if hi:
                keep[s - hi:, kx] = True
    else:
        for ky in range(s):
            kx = np.nonzero(union[ky])[0]
            if len(kx):
                keep[ky, :kx.max() + 1] = True
    return full if keep.sum() * 10 >= full.size * 9 else keep
