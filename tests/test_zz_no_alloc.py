"""ddm_k_tile_fft_noalloc (the tile FFT entry point that allocates no device
memory: its pass buffer is a static device array) against ddm_k_tile_fft
through the C ABI, bit for bit, on two calls in a row, the second with a
smaller input; frames whose pass buffer exceeds the array are transformed in
several batches, and a frame too large for it is refused."""
import ctypes

import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from cuda_multiddm_amd import engine

pytestmark = pytest.mark.gpu

G = 64
SCRATCH_ELEMS = 1 << 23  # complex values in the static pass buffer


def _fn(name):
    f = getattr(engine.lib(), name)
    vp, ll = ctypes.c_void_p, ctypes.c_longlong
    f.argtypes = [vp, ctypes.c_int, ll, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ll, ll, ll, vp]
    f.restype = ctypes.c_int
    return f


def _case(M, s, n, bits=8, x_off=0, y_off=0, seed=0):
    import torch
    w, h = M + x_off + 3, M + y_off + 1
    frames = ddm.synth_uniform(n, w, h, bits, seed=seed)
    raw = np.ascontiguousarray(frames).view(np.uint8).reshape(n, -1)
    bpp = 1 if bits <= 8 else 2
    fmt = ddm.PIX_U8 if bits <= 8 else ddm.PIX_U16LE
    T = (M // s) ** 2
    Ep = (T * s * (s // 2 + 1) + G - 1) // G * G
    R, first = n + 2, n + 1  # the frames wrap around the ring's end
    d = torch.from_numpy(raw.copy()).cuda()
    rings = []
    for name in ("ddm_k_tile_fft", "ddm_k_tile_fft_noalloc"):
        ring = torch.zeros(R * Ep * 2, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        print(f"{name} M={M} s={s} n={n} bits={bits}", flush=True)
        rc = _fn(name)(d.data_ptr(), n, raw.shape[1], w, h, bpp, fmt, x_off, y_off, M, s, ring.data_ptr(), R, Ep, 0,
                       first, None)
        assert rc == 0, (name, rc)
        torch.cuda.synchronize()
        rings.append(ring.cpu().numpy())
    return rings


@pytest.mark.parametrize("first,second", [
    ((1024, 512, 20, 8, 3, 2), (256, 256, 3, 8, 0, 0)),      # 16 + 4 frames per pass buffer, then 1 batch
    ((2048, 1024, 6, 12, 1, 5), (512, 256, 5, 12, 2, 0)),    # 4 + 2 frames, 12-bit little endian
    ((1024, 256, 9, 8, 0, 0), (1024, 1024, 2, 8, 0, 0)),     # 256 tiles of a 1024 region, then one 1024 tile
    ((256, 64, 10, 8, 1, 1), (128, 32, 4, 8, 0, 0)),         # scales without a pass buffer
])
def test_noalloc_equals_allocating_entry_point(gpu, first, second):
    for k, (M, s, n, bits, xo, yo) in enumerate((first, second)):
        ref, got = _case(M, s, n, bits, xo, yo, seed=M + s + n + k)
        assert np.any(ref != 0)
        assert np.array_equal(ref, got), (M, s, n)


def test_noalloc_refuses_a_frame_larger_than_the_array(gpu):
    import torch
    M, s = 8192, 1024  # one frame's pass buffer: M^2 / 2 complex values > the array
    assert M * M // 2 > SCRATCH_ELEMS
    T = (M // s) ** 2
    Ep = (T * s * (s // 2 + 1) + G - 1) // G * G
    d = torch.zeros(M * M, dtype=torch.uint8, device="cuda")
    ring = torch.zeros(Ep * 2, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = _fn("ddm_k_tile_fft_noalloc")(d.data_ptr(), 1, M * M, M, M, 1, ddm.PIX_U8, 0, 0, M, s, ring.data_ptr(), 1,
                                       Ep, 0, 0, None)
    assert rc == -9  # DDM_ERANGE
    assert "scratch" in engine.lib().ddm_last_error(None).decode()
    torch.cuda.synchronize()
    assert not torch.any(ring != 0).item()
