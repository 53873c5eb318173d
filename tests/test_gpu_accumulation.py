"""Accumulation of the lag sums over long streams (ddm_desc.accumulation):
compensated FP32 (each launch's partial sums added with their exact rounding
error kept) folded into FP64 accumulators every 10^4 frame pairs, against
plain FP32 accumulation and the FP64 oracle.  The reference adds every term
in FP32 (src/DDM_kernel.cuh:85-101) and merges in FP32
(src/DDM_kernel.cuh:127-139)."""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-5
FOLD = 10000  # DDM_ACC_FOLD_PAIRS


def _rel(got, ref):
    return o.max_rel_err(got, ref)


def test_compensated_and_plain_against_oracle(gpu):
    frames = ddm.synth_brownian(300, 96, 96, 8, 60, seed=3, nthreads=4)
    args = (96, 96, [64, 32], [1, 2, 5, 17, 60, 200], ddm.log_lambdas(2, 48, 10), 1.2, 30)
    out = {}
    for mode in (ddm.ACC_COMPENSATED, ddm.ACC_PLAIN):
        with ddm.Engine(*args, lag_block=32, accumulation=mode) as e:
            e.submit_host(frames)
            out[mode] = e.results()
    cfg = o.Config(96, 96, 300, [64, 32], args[3], args[4], np.float32(1.2).item())
    ref = o.run(cfg, frames)
    ec, ep = _rel(out[ddm.ACC_COMPENSATED], ref), _rel(out[ddm.ACC_PLAIN], ref)
    print(f"compensated {ec:.3e}, plain {ep:.3e}")
    assert ec <= TOL and ep <= TOL


def test_fold_into_fp64_over_a_long_stream(gpu):
    """12 200 frames of 64^2 in 30-frame chunks, lag block 32: the FP32 sums
    fold into FP64 after 10 016 frames.  Results before and after the fold,
    the accumulator dump after it, and the window after a reset all match the
    oracle; the compensated error does not grow with the stream."""
    N1, N2, N3 = 9000, 12000, 12200
    frames = ddm.synth_uniform(N3, 64, 64, 8, seed=17)
    scales, lags, lams = [64, 32], [1, 3, 10, 100, 1000], [2.0, 4.0, 8.0, 16.0]
    ref = {n: o.run(o.Config(64, 64, n, scales, lags, lams), frames[:n]) for n in (N1, N2, N3)}
    _, acc_ref = o.run(o.Config(64, 64, N2, scales, lags, lams), frames[:N2], acc_scale=0)
    res = {}
    for mode in (ddm.ACC_COMPENSATED, ddm.ACC_PLAIN):
        with ddm.Engine(64, 64, scales, lags, lams, 1.2, 30, lag_block=32, accumulation=mode,
                        spectrum=ddm.SPECTRUM_FULL) as e:
            print(f"mode {mode}: {N2} frames", flush=True)
            for i in range(0, N2, 30):
                e.submit_host(frames[i:min(i + 30, N2)])
                if i + 30 == N1:
                    res[mode, N1] = e.results()
            res[mode, N2] = e.results()
            if mode == ddm.ACC_COMPENSATED:
                acc = e.accumulators(0)
                err = np.max(np.abs(acc - acc_ref) / np.maximum(np.abs(acc_ref), 1e-30))
                print(f"accumulator dump after the fold: max rel err {err:.3e}")
                assert err <= 2e-6
            e.reset()
            e.submit_host(frames[N2:N3])
            res[mode, "window"] = e.results()
    window = (N3 * ref[N3] - N2 * ref[N2]) / (N3 - N2)
    for mode in (ddm.ACC_COMPENSATED, ddm.ACC_PLAIN):
        e1, e2 = _rel(res[mode, N1], ref[N1]), _rel(res[mode, N2], ref[N2])
        ew = o.max_rel_err(res[mode, "window"], window)
        print(f"mode {mode}: {N1} frames {e1:.3e}, {N2} frames {e2:.3e}, window after reset {ew:.3e}")
        assert e1 <= TOL and e2 <= TOL and ew <= 1e-4
    # compensated and folded: the error at 12 000 frames stays at the level of
    # one launch's partial sums
    assert _rel(res[ddm.ACC_COMPENSATED, N2], ref[N2]) <= 1e-6
