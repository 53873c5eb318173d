"""ctypes front-end of the CPU oracle (ddm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by the ``cpu_baseline`` leg of ``bench.py`` (the FP32
multithreaded build).  The product (``cuda_multiddm_amd``) never imports it.

The restated reference functions and their citations are listed in the
header of ``ddm_oracle.c``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libddm_oracle.so")
_lib = None

PIX_U8, PIX_U16LE, PIX_U16BE = 0, 1, 2
NORM_FRAMES, NORM_PAIRS = 0, 1


class _Params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("bpp", ctypes.c_int), ("pixfmt", ctypes.c_int),
        ("x_off", ctypes.c_int), ("y_off", ctypes.c_int),
        ("n_frames", ctypes.c_int),
        ("scales", ctypes.POINTER(ctypes.c_int)), ("scale_count", ctypes.c_int),
        ("lags", ctypes.POINTER(ctypes.c_int)), ("lag_count", ctypes.c_int),
        ("lambdas", ctypes.POINTER(ctypes.c_float)), ("lambda_count", ctypes.c_int),
        ("tolerance", ctypes.c_float),
        ("norm_mode", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("only_scale", ctypes.c_int),
        ("max_tiles", ctypes.c_int),
        ("tile_list", ctypes.POINTER(ctypes.c_int)), ("tile_list_count", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc) and return the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def build_native(out_dir: str) -> str:
    """Compile the oracle for the host it runs on (-O3 -march=native) into
    out_dir and return the library path (the CPU baseline's build)."""
    path = os.path.join(os.path.abspath(out_dir), "libddm_oracle_native.so")
    subprocess.run(["make", "-s", "-B", f"OUT={path}", "EXTRA=-march=native"], cwd=_HERE, check=True,
                   capture_output=True)
    return path


def use_library(path: str) -> None:
    """Load the oracle from `path` from now on (e.g. a build_native library)."""
    global _lib, _LIB_PATH
    _lib = None
    _LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ddmo_run_f64.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p, dp, ctypes.c_int, dp]
        L.ddmo_run_f32.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p, dp, ctypes.c_int, dp]
        L.ddmo_rfft2_f64.argtypes = [dp, ctypes.c_int, dp]
        L.ddmo_rfft2_f32.argtypes = [dp, ctypes.c_int, dp]
        L.ddmo_dft2_brute.argtypes = [dp, ctypes.c_int, dp]
        L.ddmo_mask.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
        L.ddmo_mask.restype = ctypes.c_int
        L.ddmo_tile.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Config:
    """Parameters of one analysis, in the reference's terms (runDDM args)."""
    width: int
    height: int
    n_frames: int
    scales: Sequence[int]
    lags: Sequence[int]
    lambdas: Sequence[float]
    tolerance: float = 1.2
    bpp: int = 1
    pixfmt: int = PIX_U8
    x_off: int = 0
    y_off: int = 0
    norm_mode: int = NORM_FRAMES
    _keep: list = field(default_factory=list, repr=False)

    def sorted_scales(self):
        return sorted(self.scales, reverse=True)

    def sorted_lags(self):
        return sorted(self.lags)

    def sorted_lambdas(self):
        return sorted(np.float32(x) for x in self.lambdas)

    def tiles(self):
        M = max(self.scales)
        return [(M // s) ** 2 for s in self.sorted_scales()]

    def result_count(self):
        return sum(self.tiles()) * len(self.lambdas) * len(self.lags)

    def _params(self, nthreads=1, only_scale=-1, max_tiles=0, tiles=None) -> _Params:
        sc = (ctypes.c_int * len(self.scales))(*self.scales)
        lg = (ctypes.c_int * len(self.lags))(*self.lags)
        lm = (ctypes.c_float * len(self.lambdas))(*[float(x) for x in self.lambdas])
        tl = (ctypes.c_int * max(1, len(tiles or [])))(*(tiles or [0]))
        self._keep = [sc, lg, lm, tl]
        return _Params(self.width, self.height, self.bpp, self.pixfmt, self.x_off, self.y_off,
                       self.n_frames, sc, len(self.scales), lg, len(self.lags), lm,
                       len(self.lambdas), ctypes.c_float(self.tolerance).value, self.norm_mode,
                       nthreads, only_scale, max_tiles, tl, len(tiles or []))


def split_results(cfg: Config, flat: np.ndarray):
    """Split the flat [scale][tile][q][lag] array into {scale: array[T, Q, L]}."""
    out, off = {}, 0
    Q, L = len(cfg.lambdas), len(cfg.lags)
    for s, T in zip(cfg.sorted_scales(), cfg.tiles()):
        out[s] = flat[off:off + T * Q * L].reshape(T, Q, L)
        off += T * Q * L
    return out


def run(cfg: Config, frames: np.ndarray, precision: str = "f64", nthreads: int = 0,
        acc_scale: int = -1, only_scale: int = -1, max_tiles: int = 0, tiles=None):
    """Run the whole pipeline on the CPU.

    ``only_scale`` (index into the descending scale list), ``max_tiles`` (the
    first tiles of each scale) and ``tiles`` (a list of tile indices, applied
    to every scale) restrict the work; results of skipped tiles stay 0.
    Returns the flat I(q, tau) array [scale][tile][q][lag] (float64) and, when
    ``acc_scale >= 0``, the normalised lag sums of that scale [lag][tile][ky][kx].
    """
    frames = np.ascontiguousarray(frames)
    need = cfg.n_frames * cfg.width * cfg.height * cfg.bpp
    if frames.nbytes < need:
        raise ValueError(f"frames buffer holds {frames.nbytes} bytes, need {need}")
    if nthreads <= 0:
        nthreads = os.cpu_count() or 1
    iqt = np.zeros(cfg.result_count(), dtype=np.float64)
    acc = None
    if acc_scale >= 0:
        s = cfg.sorted_scales()[acc_scale]
        T = cfg.tiles()[acc_scale]
        acc = np.zeros((len(cfg.lags), T, s, s // 2 + 1), dtype=np.float64)
    p = cfg._params(nthreads, only_scale, max_tiles, tiles)
    fn = lib().ddmo_run_f64 if precision == "f64" else lib().ddmo_run_f32
    rc = fn(ctypes.byref(p), frames.ctypes.data_as(ctypes.c_void_p), _dptr(iqt), acc_scale,
            _dptr(acc) if acc is not None else None)
    if rc != 0:
        raise ValueError(f"oracle rejected the configuration (code {rc})")
    return (iqt, acc) if acc_scale >= 0 else iqt


def rfft2(tile: np.ndarray, precision: str = "f64") -> np.ndarray:
    """Unnormalised forward 2-D R2C transform of one s x s tile -> [s, s/2+1] complex."""
    tile = np.ascontiguousarray(tile, dtype=np.float64)
    s = tile.shape[0]
    out = np.zeros((s, s // 2 + 1, 2), dtype=np.float64)
    fn = lib().ddmo_rfft2_f64 if precision == "f64" else lib().ddmo_rfft2_f32
    if fn(_dptr(tile), s, _dptr(out)) != 0:
        raise ValueError("bad tile size")
    return out[..., 0] + 1j * out[..., 1]


def dft2_brute(tile: np.ndarray) -> np.ndarray:
    tile = np.ascontiguousarray(tile, dtype=np.float64)
    s = tile.shape[0]
    out = np.zeros((s, s // 2 + 1, 2), dtype=np.float64)
    lib().ddmo_dft2_brute(_dptr(tile), s, _dptr(out))
    return out[..., 0] + 1j * out[..., 1]


def mask(s: int, lam: float, tol: float) -> np.ndarray:
    """q-ring mask of scale s for wavelength lam, [s, s/2+1] uint8."""
    m = np.zeros((s, s // 2 + 1), dtype=np.uint8)
    lib().ddmo_mask(s, ctypes.c_float(lam), ctypes.c_float(tol), m.ctypes.data_as(ctypes.c_void_p))
    return m


def tile(cfg: Config, frames: np.ndarray, s: int, tile_idx: int, frame: int) -> np.ndarray:
    out = np.zeros((s, s), dtype=np.float64)
    p = cfg._params()
    lib().ddmo_tile(ctypes.byref(p), np.ascontiguousarray(frames).ctypes.data_as(ctypes.c_void_p),
                    max(cfg.scales), s, tile_idx, frame, _dptr(out))
    return out


def max_rel_err(got: np.ndarray, ref: np.ndarray) -> float:
    """Max relative error over entries with ref != 0; entries with ref == 0 must be 0."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    nz = ref != 0
    if np.any(got[~nz] != 0):
        return float("inf")
    if not np.any(nz):
        return 0.0
    return float(np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])))
