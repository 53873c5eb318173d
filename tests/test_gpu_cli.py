"""./ddm end to end on the GPU: output files in the reference layout
(writeIqtToFile, src/azimuthal_average.cu:112-150) parsed the way the
reference's consumer parses them (plot_averaged.py:27-58), against the FP64
oracle on the same frames.  Values are printed with 6 significant digits,
so the comparison allows that rounding on top of the 1e-5 engine bound."""
import os
import subprocess

import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o
from tests.movie_util import write_movie
from tests.test_host import _params, read_file

pytestmark = pytest.mark.gpu

PRINT_TOL = 1e-5 + 6e-6  # engine bound + 6-significant-digit printing


def _run(*args):
    cmd = [ddm.CLI_PATH, *map(str, args)]
    print("cli:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _check_files(prefix, cfg, ref, fps=1.0):
    res = o.split_results(cfg, ref)
    lams = cfg.sorted_lambdas()
    for s, T in zip(cfg.sorted_scales(), cfg.tiles()):
        for t in range(T):
            lam, tau, isf = read_file(f"{prefix}{s}-{t}")
            assert np.allclose(lam, lams, rtol=6e-6)
            # labels float(tau) / fps, printed with 6 significant digits
            assert np.allclose(tau, np.float32(np.array(cfg.sorted_lags(), np.float32) / np.float32(fps)),
                               rtol=6e-6)
            want = res[s][t]
            nz = want != 0
            assert np.all(isf[~nz] == 0)
            assert np.max(np.abs(isf[nz] - want[nz]) / want[nz]) <= PRINT_TOL, (s, t)


def test_cli_benchmark_mode_uniform128(gpu, tmp_path):
    """-B: 64 frames of 128^2 uniform random bytes, tau 1..16, 20 lambdas 4..64, -t 20, -C 30."""
    lams = ddm.lin_lambdas(4, 64, 20)
    out = _run("-B", "-N", 64, "-C", 30, "-t", 20, "-o", tmp_path / "out", *_params(tmp_path, (128,), range(1, 17), lams))
    assert "[Time elapsed]" in out and out.strip().endswith("DDM End")
    frames = ddm.synth_uniform(64, 128, 128, 8, seed=0x5EED)
    cfg = o.Config(128, 128, 64, [128], list(range(1, 17)), lams, 1.2)
    _check_files(str(tmp_path / "out"), cfg, o.run(cfg, frames))


def test_cli_movie_roi_offset_fps(gpu, tmp_path):
    rng = np.random.default_rng(9)
    frames = rng.integers(0, 4096, (50, 140, 150)).astype(np.uint16)
    mv = tmp_path / "v.movie"
    write_movie(mv, frames, camera=1, pixelmode=3, fps=100.0)
    lags, lams = [1, 2, 4, 9, 20, 33], [2.0, 3.5, 6.0, 11.0, 20.0]
    _run("-M", "-f", mv, "-N", 45, "-s", 3, "-x", 7, "-y", 5, "-C", 16, "-F", 25, "-o", tmp_path / "m",
         *_params(tmp_path, (32, 128, 64), lags, lams))
    cfg = o.Config(150, 140, 45, [128, 64, 32], lags, lams, 1.2, bpp=2, pixfmt=o.PIX_U16LE, x_off=7, y_off=5)
    _check_files(str(tmp_path / "m"), cfg, o.run(cfg, frames[3:48].view(np.uint8)), fps=25.0)


def test_cli_rolling_windows_and_nonave(gpu, tmp_path):
    """-G 2 -C 10 over 50 frames: windows of 20 frames written as _t0_, _t1_,
    the last 10 frames as the final output; pairs reach back across a purge."""
    frames = ddm.synth_uniform(50, 32, 32, 8, seed=0x5EED)
    lags, lams = [1, 3, 12], [2.0, 3.0, 6.0]
    _run("-B", "-N", 50, "-C", 10, "-G", 2, "-K", "-o", tmp_path / "g", *_params(tmp_path, (32, 16), lags, lams))
    I = {n: o.run(o.Config(32, 32, n, [32, 16], lags, lams), frames[:n]) for n in (20, 40, 50)}
    windows = {"g_t0_": I[20], "g_t1_": (40 * I[40] - 20 * I[20]) / 20, "g": (50 * I[50] - 40 * I[40]) / 10}
    cfg = o.Config(32, 32, 50, [32, 16], lags, lams)
    for name, want in windows.items():
        res = o.split_results(cfg, want)
        for s, T in zip(cfg.sorted_scales(), cfg.tiles()):
            for t in range(T):
                _, _, isf = read_file(str(tmp_path / f"{name}{s}-{t}"))
                assert np.max(np.abs(isf - res[s][t]) / np.abs(res[s][t])) <= 1e-4, (name, s, t)
    # -K: lag sums of the final window, one row of s(s/2+1) values per lag
    _, acc50 = o.run(o.Config(32, 32, 50, [32, 16], lags, lams), frames, acc_scale=1)
    _, acc40 = o.run(o.Config(32, 32, 40, [32, 16], lags, lams), frames[:40], acc_scale=1)
    want = acc50 - acc40
    for t in range(4):
        rows = open(tmp_path / f"g16-{t}-nonAve").read().split("\n")
        vals = np.array([[float(x) for x in r.split()] for r in rows[2:2 + len(lags)]])
        assert vals.shape == (len(lags), 16 * 9)
        assert np.allclose(vals, want[:, t].reshape(len(lags), -1), rtol=2e-5, atol=1e-12)


def test_cli_fft_dump(gpu, tmp_path):
    frames = ddm.synth_uniform(12, 32, 32, 8, seed=0x5EED)
    _run("-B", "-N", 12, "-C", 6, "-A", "-o", tmp_path / "a", *_params(tmp_path, (32, 16), [1, 2], [3.0]))
    rows = open(tmp_path / "a_fft_chunk_1-16").read().strip().split("\n")
    assert len(rows) == 6 * 4  # (frame, tile) rows of the second chunk
    vals = np.array([[float(x) for x in r.split()] for r in rows])
    for f in range(6):
        for t in range(4):
            ty, tx = divmod(t, 2)
            F = np.fft.rfft2(frames[6 + f, ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16].astype(float)) / 256
            assert np.allclose(vals[f * 4 + t], (np.abs(F) ** 2).ravel(), rtol=2e-5, atol=1e-9)


def test_cli_reference_example1(gpu, tmp_path):
    """resources/example1/run_example: -Q/-T/-S from the reference's files
    (scales 256, 512, 1024 in ascending order), -C 60 -F 60 -v, on a
    synthetic 1024^2 .movie (120 frames instead of 1500)."""
    from tests.test_host import example_files, read_params
    q, t, s = example_files("example1")
    lams, lags, scales = read_params(q, t, s)
    frames = ddm.synth_brownian(120, 1024, 1024, 8, 8000, seed=41, nthreads=16)
    mv = tmp_path / "ex1.movie"
    write_movie(mv, frames, camera=1, pixelmode=1, fps=60.0)
    out = _run("-M", "-f", mv, "-o", tmp_path / "out", "-Q", q, "-T", t, "-S", s, "-N", 120, "-C", 60, "-v",
               "-F", 60)
    assert "[Time elapsed]" in out
    cfg = o.Config(1024, 1024, 120, scales, lags, lams, 1.2)
    _check_files(str(tmp_path / "out"), cfg, o.run(cfg, frames), fps=60.0)


def test_cli_reference_example2(gpu, tmp_path):
    """resources/example2/run_example (a web camera there; a 1280x720 .movie
    here): -Q/-T/-S from the reference's files (scales 256, 512 ascending,
    lambdas 1..16, lags 1..58), -C 75 -x 600 -y 100 -v, 150 frames."""
    from tests.test_host import example_files, read_params
    q, t, s = example_files("example2")
    lams, lags, scales = read_params(q, t, s)
    frames = ddm.synth_brownian(150, 1280, 720, 8, 6000, seed=42, nthreads=16)
    mv = tmp_path / "ex2.movie"
    write_movie(mv, frames, camera=1, pixelmode=1, fps=30.0)
    _run("-M", "-f", mv, "-o", tmp_path / "out", "-Q", q, "-T", t, "-S", s, "-N", 150, "-C", 75, "-v", "-x", 600,
         "-y", 100, "-I")
    cfg = o.Config(1280, 720, 150, scales, lags, lams, 1.2, x_off=600, y_off=100)
    _check_files(str(tmp_path / "out"), cfg, o.run(cfg, frames), fps=1.0)
