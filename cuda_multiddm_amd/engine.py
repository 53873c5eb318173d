"""Python front end of the MI355X DDM engine (ctypes over the C ABI in
include/ddm_amd.h, built as cuda_multiddm_amd/_build/libddm_amd.so).

``run_ddm`` mirrors the reference driver runDDM (src/DDM.cu:268-292,
declared src/main.cpp:40-64): same parameters, same sorting and checks, a
whole video in one call.  ``Engine`` exposes the streaming interface
(create / submit frames / results / reset) the C ABI offers.  There is no
CPU fallback: every call runs the HIP kernels of the library, and loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libddm_amd.so")
CLI_PATH = os.path.join(_HERE, "_build", "ddm")

DDM_OK = 0
STATUS = {0: "DDM_OK", -1: "DDM_EINVAL", -2: "DDM_ENOMEM", -3: "DDM_EDEVICE", -4: "DDM_EROI",
          -5: "DDM_ESCALE", -6: "DDM_ELAG", -7: "DDM_ELAMBDA", -8: "DDM_ESTATE", -9: "DDM_ERANGE",
          -10: "DDM_EBUSY"}
DDM_EBUSY = -10
PIX_U8, PIX_U16LE, PIX_U16BE = 0, 1, 2
NORM_FRAMES, NORM_PAIRS = 0, 1
PREC_FP32, PREC_BF16X2, PREC_FP16X2 = 0, 1, 2
ACC_COMPENSATED, ACC_PLAIN = 0, 1
SPECTRUM_MASKED, SPECTRUM_FULL = 0, 1
FFT_REGISTER, FFT_LDS = 0, 1
TILER_FUSED, TILER_PER_SCALE = 0, 1
SERVER_BATCHED, SERVER_NO_GRAPH, SERVER_GROUP = 1, 2, 4

# Every symbol include/ddm_amd.h declares.
EXPORTED = [
    "ddm_create", "ddm_validate", "ddm_destroy", "ddm_submit_host", "ddm_submit_device", "ddm_sync",
    "ddm_results", "ddm_reset", "ddm_get_info", "ddm_get_params", "ddm_get_accumulators",
    "ddm_get_spectra", "ddm_set_timing", "ddm_get_stage_stats", "ddm_get_chunk_times",
    "ddm_last_error", "ddm_status_string", "ddm_k_tile_fft", "ddm_k_tile_fft_noalloc", "ddm_k_lag_accumulate",
    "ddm_ring_mask",
    "ddm_synth_uniform", "ddm_synth_brownian", "ddm_write_iqt_file", "ddm_flush",
    "ddm_server_create", "ddm_server_destroy", "ddm_server_open", "ddm_server_submit", "ddm_server_results",
    "ddm_server_engine", "ddm_server_drain", "ddm_server_close", "ddm_server_latencies", "ddm_server_stats",
    "ddm_server_last_error", "ddm_server_create2", "ddm_server_set_hold", "ddm_fit_results", "ddm_k_fit", "ddm_set_precision", "ddm_autotune", "ddm_reserve",
    "ddm_set_fft_family", "ddm_get_fft_families", "ddm_set_tiler", "ddm_get_tiler",
]


class DDMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("bytes_per_pixel", ctypes.c_int),
        ("pixel_format", ctypes.c_int), ("x_off", ctypes.c_int), ("y_off", ctypes.c_int),
        ("scales", ctypes.POINTER(ctypes.c_int)), ("scale_count", ctypes.c_int),
        ("lags", ctypes.POINTER(ctypes.c_int)), ("lag_count", ctypes.c_int),
        ("lambdas", ctypes.POINTER(ctypes.c_float)), ("lambda_count", ctypes.c_int),
        ("mask_tolerance", ctypes.c_float), ("chunk_frames", ctypes.c_int),
        ("norm_mode", ctypes.c_int), ("lag_block", ctypes.c_int), ("accumulation", ctypes.c_int),
        ("spectrum", ctypes.c_int),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("largest_scale", ctypes.c_int), ("scale_count", ctypes.c_int), ("lag_count", ctypes.c_int),
        ("lambda_count", ctypes.c_int), ("ring_slots", ctypes.c_int), ("lag_block", ctypes.c_int),
        ("elems_per_frame", ctypes.c_longlong), ("elems_per_frame_padded", ctypes.c_longlong),
        ("result_count", ctypes.c_longlong),
        ("frames_submitted", ctypes.c_longlong), ("frames_since_reset", ctypes.c_longlong),
        ("device_bytes", ctypes.c_size_t), ("pinned_bytes", ctypes.c_size_t),
    ]


class StageStats(ctypes.Structure):
    _fields_ = [
        ("fft_ms", ctypes.c_double), ("lag_ms", ctypes.c_double), ("azimuthal_ms", ctypes.c_double),
        ("fft_launches", ctypes.c_longlong), ("lag_launches", ctypes.c_longlong),
        ("azimuthal_launches", ctypes.c_longlong), ("chunks", ctypes.c_longlong),
    ]


class ServerInfo(ctypes.Structure):
    _fields_ = [
        ("budget_bytes", ctypes.c_size_t), ("used_bytes", ctypes.c_size_t), ("streams", ctypes.c_int),
        ("max_streams", ctypes.c_int), ("frames", ctypes.c_longlong), ("chunks", ctypes.c_longlong),
        ("chunks_in_flight", ctypes.c_longlong), ("ticks", ctypes.c_longlong), ("graph_ticks", ctypes.c_longlong),
        ("tick_items", ctypes.c_longlong),
    ]


_lib = None


def lib():
    """Loads the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"DDM engine library not built: {LIB_PATH} missing "
                           "(run `python -m cuda_multiddm_amd.build`)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, fp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), \
        ctypes.POINTER(ctypes.c_double)
    ll = ctypes.c_longlong
    L.ddm_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(vp)]
    L.ddm_validate.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(Info), ctypes.c_char_p, ctypes.c_size_t]
    L.ddm_destroy.argtypes = [vp]
    L.ddm_destroy.restype = None
    L.ddm_submit_host.argtypes = [vp, vp, ctypes.c_int]
    L.ddm_submit_device.argtypes = [vp, vp, ctypes.c_int]
    L.ddm_sync.argtypes = [vp]
    L.ddm_results.argtypes = [vp, dp]
    L.ddm_fit_results.argtypes = [vp, ctypes.c_float, dp, dp]
    L.ddm_k_fit.argtypes = [vp, vp, ctypes.c_int, ll, ctypes.c_int, vp, vp]
    L.ddm_autotune.argtypes = [vp, ctypes.c_char_p, ip, ip]
    L.ddm_set_fft_family.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.ddm_get_fft_families.argtypes = [vp, ip]
    L.ddm_set_tiler.argtypes = [vp, ctypes.c_int]
    L.ddm_get_tiler.argtypes = [vp]
    L.ddm_reserve.argtypes = [vp, ctypes.c_int]
    L.ddm_reset.argtypes = [vp]
    L.ddm_get_info.argtypes = [vp, ctypes.POINTER(Info)]
    L.ddm_get_params.argtypes = [vp, ip, ip, fp]
    L.ddm_get_accumulators.argtypes = [vp, ctypes.c_int, fp]
    L.ddm_get_spectra.argtypes = [vp, ll, ctypes.c_int, fp]
    L.ddm_set_timing.argtypes = [vp, ctypes.c_int]
    L.ddm_set_precision.argtypes = [vp, ctypes.c_int]
    L.ddm_get_stage_stats.argtypes = [vp, ctypes.POINTER(StageStats)]
    L.ddm_get_chunk_times.argtypes = [vp, fp, ctypes.c_int]
    L.ddm_last_error.argtypes = [vp]
    L.ddm_last_error.restype = ctypes.c_char_p
    L.ddm_status_string.argtypes = [ctypes.c_int]
    L.ddm_status_string.restype = ctypes.c_char_p
    L.ddm_k_tile_fft.argtypes = [vp, ctypes.c_int, ll, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ll,
                                 ll, ll, vp]
    L.ddm_k_tile_fft_noalloc.argtypes = L.ddm_k_tile_fft.argtypes
    L.ddm_k_lag_accumulate.argtypes = [vp, ctypes.c_int, ll, ll, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                       vp, ctypes.c_int, vp]
    L.ddm_ring_mask.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, vp]
    L.ddm_synth_uniform.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint64, ll]
    L.ddm_synth_brownian.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_uint64, ctypes.c_int]
    L.ddm_write_iqt_file.argtypes = [ctypes.c_char_p, dp, fp, ctypes.c_int, ip, ctypes.c_int, ctypes.c_float]
    L.ddm_flush.argtypes = [vp]
    L.ddm_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(vp)]
    L.ddm_server_create2.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.ddm_server_set_hold.argtypes = [vp, ctypes.c_double]
    L.ddm_server_destroy.argtypes = [vp]
    L.ddm_server_destroy.restype = None
    L.ddm_server_open.argtypes = [vp, ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_int)]
    L.ddm_server_submit.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
    L.ddm_server_results.argtypes = [vp, ctypes.c_int, dp]
    L.ddm_server_engine.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.ddm_server_drain.argtypes = [vp]
    L.ddm_server_close.argtypes = [vp, ctypes.c_int]
    L.ddm_server_latencies.argtypes = [vp, ctypes.c_int, fp, ctypes.c_int]
    L.ddm_server_stats.argtypes = [vp, ctypes.POINTER(ServerInfo)]
    L.ddm_server_last_error.argtypes = [vp]
    L.ddm_server_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _arr(ctype, values):
    return (ctype * len(values))(*values)


def make_desc(width, height, scales, lags, lambdas, tolerance=1.2, chunk=30, bpp=1, pixfmt=PIX_U8,
              x_off=0, y_off=0, norm=NORM_FRAMES, lag_block=0, accumulation=ACC_COMPENSATED,
              spectrum=SPECTRUM_MASKED):
    """Builds a ddm_desc; the returned tuple keeps the parameter arrays alive."""
    sc = _arr(ctypes.c_int, [int(s) for s in scales])
    lg = _arr(ctypes.c_int, [int(t) for t in lags])
    lm = _arr(ctypes.c_float, [float(x) for x in lambdas])
    d = Desc(int(width), int(height), int(bpp), int(pixfmt), int(x_off), int(y_off), sc, len(scales), lg,
             len(lags), lm, len(lambdas), ctypes.c_float(tolerance).value, int(chunk), int(norm), int(lag_block),
             int(accumulation), int(spectrum))
    return d, (sc, lg, lm)


def validate(**kw):
    """(code, message, Info) for a descriptor, without touching the device."""
    d, keep = make_desc(**kw)
    info = Info()
    buf = ctypes.create_string_buffer(512)
    rc = lib().ddm_validate(ctypes.byref(d), ctypes.byref(info), buf, 512)
    return rc, buf.value.decode(), info


class Engine:
    """Streaming DDM analysis on the current HIP device."""

    def __init__(self, width, height, scales, lags, lambdas, tolerance=1.2, chunk=30, bpp=1,
                 pixfmt=PIX_U8, x_off=0, y_off=0, norm=NORM_FRAMES, lag_block=0, accumulation=ACC_COMPENSATED,
                 spectrum=SPECTRUM_MASKED):
        self._h = None
        L = lib()
        d, self._keep = make_desc(width, height, scales, lags, lambdas, tolerance, chunk, bpp, pixfmt, x_off,
                                  y_off, norm, lag_block, accumulation, spectrum)
        h = ctypes.c_void_p()
        rc = L.ddm_create(ctypes.byref(d), ctypes.byref(h))
        if rc != DDM_OK:
            raise DDMError(rc, L.ddm_last_error(None).decode())
        self._h = h
        self.width, self.height, self.bpp = int(width), int(height), int(bpp)
        info = self.info()
        self.scales = [0] * info.scale_count
        self.lags = [0] * info.lag_count
        self.lambdas = [0.0] * info.lambda_count
        s = (ctypes.c_int * info.scale_count)()
        t = (ctypes.c_int * info.lag_count)()
        q = (ctypes.c_float * info.lambda_count)()
        self._check(L.ddm_get_params(self._h, s, t, q))
        self.scales, self.lags, self.lambdas = list(s), list(t), list(q)
        self.largest_scale = info.largest_scale
        self.result_count = info.result_count

    def _check(self, rc):
        if rc < 0:
            raise DDMError(rc, lib().ddm_last_error(self._h).decode())
        return rc

    def close(self):
        if self._h is not None:
            lib().ddm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> Info:
        i = Info()
        self._check(lib().ddm_get_info(self._h, ctypes.byref(i)))
        return i

    def submit_host(self, frames: np.ndarray, n_frames: Optional[int] = None):
        frames = np.ascontiguousarray(frames)
        fb = self.width * self.height * self.bpp
        n = frames.nbytes // fb if n_frames is None else int(n_frames)
        if n * fb > frames.nbytes:
            raise ValueError("frame buffer too small")
        self._check(lib().ddm_submit_host(self._h, frames.ctypes.data_as(ctypes.c_void_p), n))

    def submit_device(self, ptr: int, n_frames: int):
        """Frames already in device memory at address ptr (e.g. tensor.data_ptr())."""
        self._check(lib().ddm_submit_device(self._h, ctypes.c_void_p(ptr), int(n_frames)))

    def sync(self):
        self._check(lib().ddm_sync(self._h))

    def results(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """I(q, tau) of every (scale, tile, q, lag), flat.  `out` (optional): a
        contiguous float64 array of result_count values to write into, e.g. in
        pinned host memory, which the copy from the device then reaches
        directly."""
        if out is None:
            out = np.zeros(self.result_count, dtype=np.float64)
        elif out.dtype != np.float64 or out.size != self.result_count or not out.flags.c_contiguous:
            raise ValueError(f"out must be a contiguous float64 array of {self.result_count} values")
        self._check(lib().ddm_results(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def results_by_scale(self) -> Dict[int, np.ndarray]:
        flat, out, off = self.results(), {}, 0
        Q, L = len(self.lambdas), len(self.lags)
        for s in self.scales:
            T = (self.largest_scale // s) ** 2
            out[s] = flat[off:off + T * Q * L].reshape(T, Q, L)
            off += T * Q * L
        return out

    def fit(self, fps: float = 1.0):
        """On-device fit of A (1 - exp(-F tau)) + B to every result curve,
        tau = lag / fps.  Returns (params[scale][tile, q, 4] = F, A, B, rms
        as a dict keyed by scale, D[scale][tile] as a dict of arrays)."""
        Q = len(self.lambdas)
        tiles = [(self.largest_scale // s) ** 2 for s in self.scales]
        prm = np.zeros((sum(tiles) * Q, 4), dtype=np.float64)
        dco = np.zeros(sum(tiles), dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(lib().ddm_fit_results(self._h, float(fps), prm.ctypes.data_as(dp), dco.ctypes.data_as(dp)))
        params, D, off = {}, {}, 0
        for s, T in zip(self.scales, tiles):
            params[s] = prm[off * Q:(off + T) * Q].reshape(T, Q, 4)
            D[s] = dco[off:off + T]
            off += T
        return params, D

    def set_precision(self, mode: int):
        """PREC_FP32 (default), PREC_BF16X2 or PREC_FP16X2 for the 64 / 128 tile transforms."""
        self._check(lib().ddm_set_precision(self._h, int(mode)))

    def autotune(self, cache_path: Optional[str] = None):
        """Times the lag block, the FFT batch and each 8..128 scale's transform
        family on this engine (before the first frame) and keeps the fastest;
        returns (lag_block, fft_batch_frames); fft_families() reads the families."""
        j, b = ctypes.c_int(0), ctypes.c_int(0)
        path = cache_path.encode() if cache_path else None
        self._check(lib().ddm_autotune(self._h, path, ctypes.byref(j), ctypes.byref(b)))
        return j.value, b.value

    def set_fft_family(self, scale_idx: int, family: int):
        """Transform family (FFT_REGISTER or FFT_LDS) of scale scale_idx (sorted
        descending), before the first frame."""
        self._check(lib().ddm_set_fft_family(self._h, int(scale_idx), int(family)))

    def fft_families(self):
        out = (ctypes.c_int * len(self.scales))()
        self._check(lib().ddm_get_fft_families(self._h, out))
        return list(out)

    def set_tiler(self, mode: int):
        """Multi-scale tiler (TILER_FUSED: each frame read once for the 8..128
        scales and once for the 256..1024 row passes; TILER_PER_SCALE: one
        transform launch per scale), before the first frame."""
        self._check(lib().ddm_set_tiler(self._h, int(mode)))

    def tiler(self) -> int:
        return int(lib().ddm_get_tiler(self._h))

    def reserve(self, max_frames: int):
        """Allocates now the transient buffers a submission of up to
        max_frames frames per call would allocate on first use."""
        self._check(lib().ddm_reserve(self._h, int(max_frames)))

    def reset(self):
        self._check(lib().ddm_reset(self._h))

    def accumulators(self, scale_idx: int) -> np.ndarray:
        s = self.scales[scale_idx]
        T = (self.largest_scale // s) ** 2
        out = np.zeros((len(self.lags), T, s, s // 2 + 1), dtype=np.float32)
        self._check(lib().ddm_get_accumulators(self._h, scale_idx,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def spectra(self, frame: int, scale_idx: int) -> np.ndarray:
        s = self.scales[scale_idx]
        T = (self.largest_scale // s) ** 2
        out = np.zeros((T, s, s // 2 + 1, 2), dtype=np.float32)
        self._check(lib().ddm_get_spectra(self._h, frame, scale_idx,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out[..., 0] + 1j * out[..., 1]

    def set_timing(self, on: bool):
        self._check(lib().ddm_set_timing(self._h, 1 if on else 0))

    def stage_stats(self) -> StageStats:
        s = StageStats()
        self._check(lib().ddm_get_stage_stats(self._h, ctypes.byref(s)))
        return s

    def chunk_times(self, max_n: int = 1 << 20) -> np.ndarray:
        buf = np.zeros(max_n, dtype=np.float32)
        n = self._check(lib().ddm_get_chunk_times(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                  max_n))
        return buf[:n].copy()


class Server:
    """Many DDM streams on one device (serving): admission by device
    footprint, chunks of different streams overlapping on the GPU, and
    chunk-to-result latencies."""

    def __init__(self, device_budget: int = 0, max_streams: int = 256, batched: bool = False, graph: bool = True,
                 group: bool = False, hold_ms: float = 0.0):
        """batched: chunks of all streams run as ticks of batched launches
        (DDM_SERVER_BATCHED), as HIP graphs unless graph is False; group: a
        tick takes all of a stream's waiting chunks up to its lag block into
        one lag pass (DDM_SERVER_GROUP)."""
        h = ctypes.c_void_p()
        flags = (SERVER_BATCHED if batched else 0) | (0 if graph else SERVER_NO_GRAPH) | (SERVER_GROUP if group else 0)
        rc = lib().ddm_server_create2(device_budget, max_streams, flags, ctypes.byref(h))
        if rc != DDM_OK:
            raise DDMError(rc, lib().ddm_server_last_error(None).decode())
        self._h = h
        self._keep = {}
        self._shape = {}
        if hold_ms:
            self._check(lib().ddm_server_set_hold(self._h, hold_ms))

    def _check(self, rc):
        if rc < 0:
            raise DDMError(rc, lib().ddm_server_last_error(self._h).decode())
        return rc

    def open(self, width, height, scales, lags, lambdas, tolerance=1.2, chunk=30, bpp=1, pixfmt=PIX_U8,
             x_off=0, y_off=0, norm=NORM_FRAMES, lag_block=0, accumulation=ACC_COMPENSATED,
             spectrum=SPECTRUM_MASKED) -> int:
        """Returns the stream id; raises DDMError with code DDM_EBUSY when the
        stream does not fit yet."""
        d, keep = make_desc(width, height, scales, lags, lambdas, tolerance, chunk, bpp, pixfmt, x_off, y_off,
                            norm, lag_block, accumulation, spectrum)
        sid = ctypes.c_int()
        self._check(lib().ddm_server_open(self._h, ctypes.byref(d), ctypes.byref(sid)))
        self._keep[sid.value] = keep
        self._shape[sid.value] = (int(width), int(height), int(bpp))
        return sid.value

    def submit(self, sid: int, frames: np.ndarray):
        w, h, b = self._shape[sid]
        frames = np.ascontiguousarray(frames)
        n = frames.nbytes // (w * h * b)
        self._check(lib().ddm_server_submit(self._h, sid, frames.ctypes.data_as(ctypes.c_void_p), n))

    def engine_info(self, sid: int) -> Info:
        e = ctypes.c_void_p()
        self._check(lib().ddm_server_engine(self._h, sid, ctypes.byref(e)))
        i = Info()
        self._check(lib().ddm_get_info(e, ctypes.byref(i)))
        return i

    def results(self, sid: int) -> np.ndarray:
        out = np.zeros(self.engine_info(sid).result_count, dtype=np.float64)
        self._check(lib().ddm_server_results(self._h, sid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def drain(self):
        self._check(lib().ddm_server_drain(self._h))

    def close_stream(self, sid: int):
        self._check(lib().ddm_server_close(self._h, sid))
        self._keep.pop(sid, None)
        self._shape.pop(sid, None)

    def latencies(self, sid: int = -1, max_n: int = 1 << 22) -> np.ndarray:
        buf = np.zeros(max_n, dtype=np.float32)
        n = self._check(lib().ddm_server_latencies(self._h, sid, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                   max_n))
        return buf[:n].copy()

    def stats(self) -> ServerInfo:
        s = ServerInfo()
        self._check(lib().ddm_server_stats(self._h, ctypes.byref(s)))
        return s

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ddm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run_ddm(frames: np.ndarray, width: int, height: int, scales: Sequence[int], lags: Sequence[int],
            lambdas: Sequence[float], tolerance: float = 1.2, chunk: int = 30, bpp: int = 1,
            pixfmt: int = PIX_U8, x_off: int = 0, y_off: int = 0, norm: int = NORM_FRAMES,
            device_input: bool = False) -> Dict[int, np.ndarray]:
    """One whole analysis, as runDDM performs it: {scale: I[tile, q, lag]}."""
    with Engine(width, height, scales, lags, lambdas, tolerance, chunk, bpp, pixfmt, x_off, y_off, norm) as e:
        if device_input:
            import torch
            t = torch.from_numpy(np.ascontiguousarray(frames).view(np.uint8).reshape(-1)).cuda()
            torch.cuda.synchronize()  # the engine's stream does not wait for torch's
            e.submit_device(t.data_ptr(), t.numel() // (width * height * bpp))
            e.sync()
            del t
        else:
            e.submit_host(frames)
        return e.results_by_scale()


# ---------------------------------------------------------------------------
# Stage entry points on caller-owned device buffers (torch tensors).
# ---------------------------------------------------------------------------
def k_tile_fft(frames, n_frames, frame_bytes, width, height, bpp, pixfmt, x_off, y_off, largest_scale, scale,
               ring, ring_slots, elems_per_frame, elem_off, first_frame, stream=None):
    rc = lib().ddm_k_tile_fft(ctypes.c_void_p(frames.data_ptr()), n_frames, frame_bytes, width, height, bpp,
                              pixfmt, x_off, y_off, largest_scale, scale, ctypes.c_void_p(ring.data_ptr()),
                              ring_slots, elems_per_frame, elem_off, first_frame,
                              ctypes.c_void_p(stream) if stream else None)
    if rc != DDM_OK:
        raise DDMError(rc, "ddm_k_tile_fft")


def k_tile_fft_noalloc(frames, n_frames, frame_bytes, width, height, bpp, pixfmt, x_off, y_off, largest_scale,
                       scale, ring, ring_slots, elems_per_frame, elem_off, first_frame, stream=None):
    """k_tile_fft without device allocation (static pass buffer, shared per GPU)."""
    rc = lib().ddm_k_tile_fft_noalloc(ctypes.c_void_p(frames.data_ptr()), n_frames, frame_bytes, width, height, bpp,
                                      pixfmt, x_off, y_off, largest_scale, scale, ctypes.c_void_p(ring.data_ptr()),
                                      ring_slots, elems_per_frame, elem_off, first_frame,
                                      ctypes.c_void_p(stream) if stream else None)
    if rc != DDM_OK:
        raise DDMError(rc, lib().ddm_last_error(None).decode())


def k_lag_accumulate(ring, ring_slots, elems_per_frame, first_frame, n_new, lags, max_lag, acc, lag_block=32,
                     stream=None):
    rc = lib().ddm_k_lag_accumulate(ctypes.c_void_p(ring.data_ptr()), ring_slots, elems_per_frame, first_frame,
                                    n_new, ctypes.c_void_p(lags.data_ptr()), lags.numel(), max_lag,
                                    ctypes.c_void_p(acc.data_ptr()), lag_block,
                                    ctypes.c_void_p(stream) if stream else None)
    if rc != DDM_OK:
        raise DDMError(rc, "ddm_k_lag_accumulate")


def k_fit(curves, tau, out, iters=200, stream=None):
    """curves: device float64 [n, L]; tau: device float64 [L]; out: device float64 [n, 4]."""
    n, L = curves.shape
    rc = lib().ddm_k_fit(ctypes.c_void_p(curves.data_ptr()), ctypes.c_void_p(tau.data_ptr()), L, n, iters,
                         ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream) if stream else None)
    if rc != DDM_OK:
        raise DDMError(rc, "ddm_k_fit")


# ---------------------------------------------------------------------------
# Host utilities
# ---------------------------------------------------------------------------
def ring_mask(scale: int, lam: float, tol: float) -> np.ndarray:
    m = np.zeros((scale, scale // 2 + 1), dtype=np.uint8)
    rc = lib().ddm_ring_mask(scale, ctypes.c_float(lam), ctypes.c_float(tol), m.ctypes.data_as(ctypes.c_void_p))
    if rc < 0:
        raise DDMError(rc, "ddm_ring_mask")
    return m


def synth_uniform(n_frames, width, height, bits=8, seed=0, first_frame=0) -> np.ndarray:
    dt = np.uint8 if bits <= 8 else np.uint16
    out = np.zeros((n_frames, height, width), dtype=dt)
    rc = lib().ddm_synth_uniform(out.ctypes.data_as(ctypes.c_void_p), n_frames, width, height, bits, seed,
                                 first_frame)
    if rc:
        raise DDMError(rc, "ddm_synth_uniform")
    return out


def synth_brownian(n_frames, width, height, bits=8, n_spots=2000, sigma=2.0, amplitude=80.0, diffusion=0.5,
                   background=100.0, noise=5.0, seed=0, nthreads=0) -> np.ndarray:
    dt = np.uint8 if bits <= 8 else np.uint16
    out = np.zeros((n_frames, height, width), dtype=dt)
    rc = lib().ddm_synth_brownian(out.ctypes.data_as(ctypes.c_void_p), n_frames, width, height, bits, n_spots,
                                  sigma, amplitude, diffusion, background, noise, seed, nthreads)
    if rc:
        raise DDMError(rc, "ddm_synth_brownian")
    return out


def write_iqt_file(path: str, iqt_q_lag: np.ndarray, lambdas, lags, fps=1.0):
    iqt = np.ascontiguousarray(iqt_q_lag, dtype=np.float64)
    lm = _arr(ctypes.c_float, [float(x) for x in lambdas])
    lg = _arr(ctypes.c_int, [int(x) for x in lags])
    rc = lib().ddm_write_iqt_file(path.encode(), iqt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), lm,
                                  len(lambdas), lg, len(lags), fps)
    if rc:
        raise DDMError(rc, path)


def log_lags(tau_max: int, count: int):
    """Integer log-spaced lags: tau_k = max(round(tau_max^(k/(count-1))), tau_{k-1} + 1)."""
    out = []
    for k in range(count):
        v = int(round(tau_max ** (k / (count - 1)))) if count > 1 else tau_max
        if out and v <= out[-1]:
            v = out[-1] + 1
        out.append(v)
    return out


def log_lambdas(lo: float, hi: float, count: int):
    return [float(np.float32(lo * (hi / lo) ** (k / (count - 1)))) for k in range(count)]


def lin_lambdas(lo: float, hi: float, count: int):
    return [float(np.float32(lo + (hi - lo) * k / (count - 1))) for k in range(count)]
