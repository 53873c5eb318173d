"""Host-side tests that need no GPU: the C ABI library loads and exports
every declared symbol, descriptor validation (before any device access),
the q-ring mask against the oracle, the synthetic generators, the output
writer's text format, and the ./ddm CLI's flag parsing, error messages and
exit codes (src/main.cpp:113-313, src/DDM.cu:319-350, 459-463)."""
import os
import re
import subprocess

import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from cuda_multiddm_amd import engine
from oracle import oracle as o
from tests.layout_util import kept_mask
from tests.movie_util import write_movie

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# --------------------------------------------------------------------------- C ABI
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "ddm_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(ddm_\w+)\s*\(", hdr))
    assert declared == set(engine.EXPORTED)
    lib = ddm.lib()
    for name in declared:
        assert hasattr(lib, name), name


def _v(**kw):
    base = dict(width=128, height=128, scales=[128], lags=list(range(1, 17)),
                lambdas=ddm.lin_lambdas(4, 64, 20), tolerance=1.2, chunk=30)
    base.update(kw)
    return ddm.validate(**base)


def _kept(s, lambdas, tol):
    """Elements per tile a masked engine keeps for scale s."""
    return int(kept_mask(s, lambdas, tol).sum())


def test_validate_accepts_uniform128_params_and_sorts():
    rc, msg, info = _v(spectrum=ddm.SPECTRUM_FULL)
    assert rc == 0 and msg == ""
    # ring: one chunk + one lag block (256 by default) + the largest lag
    assert info.largest_scale == 128 and info.lag_block == 256 and info.ring_slots == 30 + 256 + 16
    assert info.elems_per_frame == 128 * 65 and info.elems_per_frame_padded == 8320
    assert info.result_count == 20 * 16
    # masked (the default): the rows' prefixes holding the rings q = 128 /
    # lambda .. 1.2 q, lambda = 4 .. 64 (a disc of radius 38: 28 % of the half
    # spectrum)
    rc, msg, info = _v()
    n = _kept(128, [float(np.float32(x)) for x in ddm.lin_lambdas(4, 64, 20)], float(np.float32(1.2)))
    assert rc == 0 and info.elems_per_frame == n and n < 0.29 * 128 * 65
    assert info.elems_per_frame_padded == (n + 63) // 64 * 64 and info.result_count == 20 * 16
    rc, _, info = _v(width=1100, height=1030, scales=[256, 1024, 512], lags=[30, 1, 2],
                     lambdas=[16.0, 1.0, 3.0])
    assert rc == 0 and info.largest_scale == 1024
    assert info.result_count == (1 + 4 + 16) * 3 * 3
    assert info.ring_slots == 30 + 256 + 30
    rc, _, info = _v(lag_block=128, chunk=7)
    assert rc == 0 and info.lag_block == 128 and info.ring_slots == 7 + 128 + 16


@pytest.mark.parametrize("name", ["uniform128", "brownian512", "brownian2048_u16", "serving"])
def test_masked_spectrum_sizes_of_the_workloads(name):
    """Elements per frame a masked engine keeps, per workload, against the
    union of the oracle's q-ring masks; the full-spectrum size is unchanged."""
    from cuda_multiddm_amd.workloads import WORKLOADS, serving_streams
    if name == "serving":
        descs = [s.desc() for s in serving_streams(6)]
    else:
        w = WORKLOADS[name]
        descs = [dict(width=w.size, height=w.size, scales=w.scales, lags=w.lags, lambdas=w.lambdas,
                      tolerance=1.2, chunk=30)]
    for d in descs:
        rc, _, masked = ddm.validate(**d)
        rc2, _, full = ddm.validate(spectrum=ddm.SPECTRUM_FULL, **d)
        assert rc == 0 and rc2 == 0
        M = max(d["scales"])
        tol = float(np.float32(d.get("tolerance", 1.2)))
        lams = [float(np.float32(x)) for x in d["lambdas"]]
        assert full.elems_per_frame == sum((M // s) ** 2 * s * (s // 2 + 1) for s in d["scales"])
        assert masked.elems_per_frame == sum((M // s) ** 2 * _kept(s, lams, tol) for s in d["scales"])
        print(name, M, masked.elems_per_frame / full.elems_per_frame)


def test_masked_scale_without_ring_members_keeps_one_element():
    """8 x 8 tiles with rings q = 8 / lambda .. 1.05 q that hold no element:
    the scale keeps exactly one element per tile (its tables are never empty)."""
    d = dict(width=256, height=256, scales=[256, 64, 8], lags=[1, 4, 9], lambdas=[3.0, 11.0, 40.0],
             tolerance=1.05, chunk=20)
    for lam in d["lambdas"]:
        assert not o.mask(8, lam, float(np.float32(1.05))).any()
    rc, _, info = ddm.validate(**d)
    lams = [float(np.float32(x)) for x in d["lambdas"]]
    tol = float(np.float32(1.05))
    assert rc == 0 and _kept(8, lams, tol) == 1
    assert info.elems_per_frame == _kept(256, lams, tol) + 16 * _kept(64, lams, tol) + 1024


def test_validate_rejects_unknown_spectrum_mode():
    rc, msg, _ = _v(spectrum=2)
    assert rc == -1 and msg == "unknown spectrum mode"


@pytest.mark.parametrize("kw,code,text", [
    (dict(scales=[100]), -5, "scales must be powers of two (> 0)"),
    (dict(scales=[128, 128]), -5, "scales should be descending order"),
    (dict(lags=[3, -1]), -6, "Tau values should be positive"),
    (dict(lags=[1, 2, 2]), -6, "Tau vector should be ascending order"),
    (dict(lambdas=[4.0, -2.0]), -7, "q-vector values should be positive"),
    (dict(lambdas=[4.0, 4.0]), -7, "q-vector vector should be ascending order"),
    (dict(x_off=1), -4, "the specified out dimensions must be smaller than actual image size"),
    (dict(width=127), -4, "the specified out dimensions must be smaller than actual image size"),
    (dict(chunk=0), -1, "chunk frame count must be positive"),
    (dict(lag_block=16), -1, "lag block must be 0 (automatic), 32, 64, 128 or 256"),
    (dict(scales=[4096], width=4096, height=4096), -5, "scales above 2048 are not supported"),
])
def test_validate_rejects(kw, code, text):
    rc, msg, _ = _v(**kw)
    assert rc == code and msg == text


def test_validate_lifts_reference_hard_limits_with_warnings():
    # tau > chunk (src/DDM.cu:462) is allowed: the ring grows to chunk + max tau
    rc, msg, info = _v(lags=[1, 100])
    assert rc == 0 and info.ring_slots == 30 + 256 + 100
    # lambda > smallest scale (src/DDM.cu:349) is a warning: the ring is empty
    rc, msg, _ = _v(lambdas=[4.0, 200.0])
    assert rc == 0 and msg == "The largest q-vector should be smaller than the smallest scale."
    rc, msg, _ = _v(tolerance=12.0)
    assert rc == 0 and "mask_tolerance" in msg


def test_engine_create_reports_validation_errors_before_device():
    with pytest.raises(ddm.DDMError) as ex:
        ddm.Engine(64, 64, [48], [1], [4.0])
    assert ex.value.code == -5 and "powers of two" in str(ex.value)


# --------------------------------------------------------------------------- mask
@pytest.mark.parametrize("s", [32, 64, 128, 256, 512, 1024])
def test_ring_mask_bit_exact_with_oracle(s):
    lams = sorted(set(ddm.log_lambdas(4, 256, 30) + ddm.lin_lambdas(4, 64, 20) + [7, 8, 10, 12, 28]))
    for tol in (np.float32(1.2), np.float32(1.0 + 5 / 100.0)):
        for lam in lams:
            a = ddm.ring_mask(s, lam, tol)
            b = o.mask(s, lam, tol)
            assert np.array_equal(a, b), (s, lam, tol)


# --------------------------------------------------------------------------- synthetic video
def test_synth_uniform_deterministic_and_streamable():
    a = ddm.synth_uniform(6, 40, 30, 8, seed=3)
    b = ddm.synth_uniform(6, 40, 30, 8, seed=3)
    assert np.array_equal(a, b) and a.dtype == np.uint8
    c = ddm.synth_uniform(3, 40, 30, 8, seed=3, first_frame=3)  # chunked generation
    assert np.array_equal(a[3:], c)
    assert not np.array_equal(a, ddm.synth_uniform(6, 40, 30, 8, seed=4))
    big = ddm.synth_uniform(4, 256, 256, 8, seed=1)
    counts = np.bincount(big.ravel(), minlength=256)
    assert counts.min() > 0 and abs(big.mean() - 127.5) < 1.0
    w = ddm.synth_uniform(2, 64, 64, 12, seed=1)
    assert w.dtype == np.uint16 and w.max() <= 4095 and w.max() > 4000


def test_synth_brownian_statistics():
    f = ddm.synth_brownian(40, 128, 128, 8, n_spots=0, background=100.0, noise=5.0, seed=1)
    assert abs(f.mean() - 100) < 0.1 and abs(f.std() - 5) < 0.2
    g = ddm.synth_brownian(30, 128, 128, 8, n_spots=50, sigma=2.0, amplitude=80.0, diffusion=0.5,
                           background=100.0, noise=0.0, seed=2)
    # each spot adds amplitude * 2 pi sigma^2 of intensity
    excess = g.astype(float).sum(axis=(1, 2)) - 100.0 * 128 * 128
    assert np.allclose(excess, 50 * 80 * 2 * np.pi * 4, rtol=0.02)
    assert np.array_equal(g, ddm.synth_brownian(30, 128, 128, 8, n_spots=50, sigma=2.0, amplitude=80.0,
                                                diffusion=0.5, background=100.0, noise=0.0, seed=2))
    h = ddm.synth_brownian(3, 64, 64, 12, n_spots=20, amplitude=2000.0, background=1000.0, noise=30.0, seed=3)
    assert h.dtype == np.uint16 and h.max() <= 4095


def test_log_spaced_parameter_lists():
    assert ddm.log_lags(512, 32) == list(range(1, 15)) + [17, 20, 25, 31, 37, 46, 56, 68, 84, 102, 125, 153,
                                                          187, 229, 280, 342, 419, 512]
    assert ddm.log_lags(300, 24) == list(range(1, 11)) + [12, 15, 20, 25, 32, 41, 53, 68, 87, 111, 143, 183,
                                                          234, 300]
    l3 = ddm.log_lags(999, 64)
    assert l3[:33] == list(range(1, 34)) and l3[-5:] == [644, 719, 802, 895, 999] and len(l3) == 64


# --------------------------------------------------------------------------- writer
def read_file(path):
    """The reference consumer's parsing (resources/python_tools/plot_averaged.py:27-58)."""
    with open(path) as f:
        lam = [float(n) for n in f.readline().split()]
        tau = [float(n) for n in f.readline().split()]
        isf = [np.array([float(n) for n in f.readline().split()]) for _ in lam]
    return np.array(lam), np.array(tau), np.array(isf)


def test_writer_layout_and_formatting(tmp_path):
    lams = [4.0, 7.157895, 10.31579, 64.0]
    lags = [1, 2, 3, 10, 100]
    iqt = np.arange(20, dtype=np.float64).reshape(4, 5) * 1.234567891e-3 + 1e-9
    p = str(tmp_path / "out128-0")
    ddm.write_iqt_file(p, iqt, lams, lags, fps=60.0)
    text = open(p).read().split("\n")
    assert text[0] == "".join("%g " % float(np.float32(x)) for x in lams)
    assert text[1] == "".join("%g " % float(np.float32(np.float32(t) / np.float32(60.0))) for t in lags)
    for q in range(4):
        assert text[2 + q] == "".join("%g " % float(np.float32(v)) for v in iqt[q])
    lam, tau, isf = read_file(p)
    assert np.allclose(lam, lams, rtol=1e-6) and isf.shape == (4, 5)
    assert np.allclose(isf, iqt, rtol=1e-5)


# --------------------------------------------------------------------------- CLI
def _params(tmp_path, scales=(128,), lags=range(1, 17), lams=None):
    lams = lams if lams is not None else ddm.lin_lambdas(4, 64, 20)
    q, t, s = tmp_path / "lambda.txt", tmp_path / "tau.txt", tmp_path / "scale.txt"
    q.write_text("\n".join("%.9g" % x for x in lams) + "\n")
    t.write_text("\n".join(str(x) for x in lags) + "\n")
    s.write_text("\n".join(str(x) for x in scales) + "\n")
    return ["-Q", str(q), "-T", str(t), "-S", str(s)]


def _cli(*args):
    r = subprocess.run([ddm.CLI_PATH, *map(str, args)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def test_cli_help_and_unknown_flag():
    rc, out, err = _cli("-h")
    assert rc == 255 and "Usage ./ddm" in err and out.startswith("DDM Start")
    rc, _, err = _cli("-q")
    assert rc == 255 and "Usage ./ddm" in err


def test_cli_stray_argument_and_missing_input(tmp_path):
    rc, _, err = _cli("-o", tmp_path / "o", "-N", 10, *_params(tmp_path), "-B", "extra")
    assert rc == 1 and "[Error] An unexpected option was found." in err
    rc, _, err = _cli("-o", tmp_path / "o", "-N", 10, *_params(tmp_path))
    assert rc == 1 and err.strip().endswith("[Error] Must specify input.")


def test_cli_parameter_files(tmp_path):
    p = _params(tmp_path)
    rc, _, err = _cli("-B", "-N", 10, "-o", tmp_path / "o", "-Q", tmp_path / "nope", *p[2:])
    assert rc == 1 and "[Error] cannot open lambda-file." in err
    rc, _, err = _cli("-B", "-N", 10, "-o", tmp_path / "o", *p[:2], "-T", tmp_path / "nope", *p[4:])
    assert rc == 1 and "[Error] cannot open tau-file" in err
    rc, _, err = _cli("-B", "-N", 10, "-o", tmp_path / "o", *p[:4], "-S", tmp_path / "nope")
    assert rc == 1 and "[Error] cannot open scales-file." in err


@pytest.mark.parametrize("scales,lags,lams,extra,text", [
    ((96,), (1, 2), (4.0,), [], "[Error] scales must be powers of two (> 0)"),
    ((128,), (1, 2, 2), (4.0,), [], "[Error] Tau vector should be ascending order"),
    ((128,), (1, 2), (4.0, 4.0), [], "[Error] q-vector vector should be ascending order"),
    ((128,), (1, 2), (4.0,), ["-x", 1], "[Error] the specified out dimensions must be smaller than actual image size"),
    ((128,), (1, 2), (4.0,), ["-N", 0], "[Error] Number of frames to analyse (-N) must be positive."),
])
def test_cli_validation_before_device(tmp_path, scales, lags, lams, extra, text):
    args = ["-B", "-o", tmp_path / "o", *_params(tmp_path, scales, lags, lams)]
    if "-N" not in extra:
        args += ["-N", 10]
    rc, out, err = _cli(*args, *extra)
    assert rc == 1 and text in err, err


def test_cli_inputs_not_in_this_build(tmp_path):
    rc, _, err = _cli("-f", tmp_path / "video.mp4", "-N", 10, "-o", tmp_path / "o", *_params(tmp_path))
    assert rc == 1 and "error opening video file with openCV" in err
    rc, _, err = _cli("-W", "-N", 10, "-o", tmp_path / "o", *_params(tmp_path))
    assert rc == 1 and "web-camera" in err
    rc, _, err = _cli("-f", "a", "-W", "-N", 10, "-o", tmp_path / "o", *_params(tmp_path))
    assert rc == 1 and "Cannot use both in-filepath and web-cam option at same time." in err


def test_cli_warnings(tmp_path):
    # tolerance 1 + 950/100 = 10.5 is outside (1, 10): a warning, not an error
    rc, _, err = _cli("-B", "-N", 4, "-o", tmp_path / "o", *_params(tmp_path), "-t", 950)
    assert "[Warning] mask_tolerance is likely undesired value" in err


@pytest.mark.parametrize("camera,pixelmode,bpp", [(1, 1, 1), (2, 3, 2), (3, 2, 2)])
def test_cli_movie_header(tmp_path, camera, pixelmode, bpp):
    """.movie parsing happens before device use: size, bytes per pixel and
    frame rate (from the first two headers) are reported in verbose mode."""
    rng = np.random.default_rng(camera)
    dt = np.uint8 if pixelmode == 1 else np.uint16
    frames = rng.integers(0, 200, (5, 140, 150)).astype(dt)
    path = tmp_path / "v.movie"
    write_movie(path, frames, camera, pixelmode, fps=40.0, junk_prefix=b"\x01\x02\x03", pad=8)
    rc, out, err = _cli("-M", "-f", path, "-v", "-N", 4, "-s", 1, "-o", tmp_path / "o", *_params(tmp_path))
    assert "[Video info - (150 x 140), 4 Frames (offset 1), 40.0000 FPS, %d bytes per pixel]" % bpp in out, out


def test_cli_movie_errors(tmp_path):
    f = np.zeros((2, 130, 130), np.uint8)
    bad = tmp_path / "bad.movie"
    bad.write_bytes(b"\0" * 1000)
    rc, _, err = _cli("-M", "-f", bad, "-N", 1, "-o", tmp_path / "o", *_params(tmp_path))
    assert rc == 1 and "Wrong magic" in err
    from tests.movie_util import frame_header
    v2 = tmp_path / "v2.movie"
    v2.write_bytes(frame_header(1, 1, 130, 130, 130 * 130, 0, version=2) + b"\0" * (130 * 130))
    rc, _, err = _cli("-M", "-f", v2, "-N", 1, "-o", tmp_path / "o", *_params(tmp_path))
    assert rc == 1 and "Unsupported movie version 2" in err
    missing = tmp_path / "missing.movie"
    rc, _, err = _cli("-M", "-f", missing, "-N", 1, "-o", tmp_path / "o", *_params(tmp_path))
    assert rc == 1 and "couldn't open .movie file" in err
    ok = tmp_path / "ok.movie"
    write_movie(ok, f)
    rc, out, err = _cli("-M", "-f", ok, "-v", "-N", 1, "-o", tmp_path / "o", *_params(tmp_path, scales=(256,)))
    assert rc == 1 and "the specified out dimensions must be smaller than actual image size" in err


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXAMPLES = {  # the reference's example parameter files (resources/example1, example2), kept as data
    "example1": ("lambda_vector.txt", "tau_vector.txt", "scale_vector.txt"),
    "example2": ("lamda_vector.txt", "tau_vector.txt", "scale_vector.txt"),
}


def example_files(name):
    return [os.path.join(GOLDEN, name, f) for f in EXAMPLES[name]]


def read_params(lam_path, tau_path, scale_path):
    """Whitespace-separated float / int / int, as src/main.cpp:246-313 reads them."""
    lams = [float(np.float32(x)) for x in open(lam_path).read().split()]
    lags = [int(x) for x in open(tau_path).read().split()]
    scales = [int(x) for x in open(scale_path).read().split()]
    return lams, lags, scales


@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_reference_example_parameter_files(name):
    lams, lags, scales = read_params(*example_files(name))
    assert scales == sorted(scales)  # ascending in the files: the engine sorts them descending
    M = max(scales)
    rc, msg, info = ddm.validate(width=M + 700, height=M + 120, scales=scales, lags=lags, lambdas=lams,
                                 chunk=60, x_off=600, y_off=100)
    assert rc == 0, msg
    assert info.largest_scale == M and info.scale_count == len(scales)
    assert info.lag_count == len(lags) and info.lambda_count == len(lams)
    assert info.result_count == sum((M // s) ** 2 for s in scales) * len(lams) * len(lags)


def test_cli_example2_offsets_outside_benchmark_frame(tmp_path):
    """example2's run line (-x 600 -y 100) with its files on a benchmark frame
    (the largest scale, 512^2): refused before any device access."""
    q, t, s = example_files("example2")
    rc, _, err = _cli("-B", "-N", 10, "-C", 75, "-x", 600, "-y", 100, "-o", tmp_path / "o", "-Q", q, "-T", t, "-S", s)
    assert rc == 1 and "[Error] the specified out dimensions must be smaller than actual image size" in err


def test_accumulation_mode_is_validated():
    """ddm_desc.accumulation is checked before any device access, and the
    compensated mode's terms (FP32 compensation + FP64 accumulators) are part
    of the footprint a server admits by."""
    with pytest.raises(ddm.DDMError) as ei:
        ddm.Engine(32, 32, [32], [1], [4.0], accumulation=7)
    assert ei.value.code == -1
    rc, _, info_c = ddm.validate(width=64, height=64, scales=[64], lags=[1, 2], lambdas=[4.0])
    rc2, _, info_p = ddm.validate(width=64, height=64, scales=[64], lags=[1, 2], lambdas=[4.0],
                                  accumulation=ddm.ACC_PLAIN)
    assert rc == 0 and rc2 == 0
    assert info_c.device_bytes - info_p.device_bytes == 2 * info_c.elems_per_frame_padded * 12
