"""On-device ISF fitting (ddm_k_fit, ddm_fit_results) through the C ABI.

The fitter is checked three ways:
  * curves built from known (F, A, B): the parameters come back to 1e-6;
  * noisy curves: the same least-squares minimum as scipy's curve_fit
    (MINPACK Levenberg-Marquardt, the routine the reference's analysis script
    calls, resources/python_tools/plot_averaged.py:79), F to 1e-6 relative;
  * a Brownian video of known diffusion coefficient D0: every fitted curve
    matches curve_fit on the same I(q, tau), the curves match those fitted to
    the FP64 oracle's I(q, tau) (F to 1e-4), and D lands where the ring
    geometry puts it (q runs from 2 pi / lambda to 1.2 x that, so the
    estimate with q = 2 pi / lambda lies in [D0, 1.44 D0] up to noise).
Pair-mean normalisation is used for the video: with the reference's 1/N
normalisation I(q, tau) carries a (N - tau) / N factor that a
1 - exp(-F tau) model reads as a faster decay.
"""
import numpy as np
import pytest

import cuda_multiddm_amd as ddm
from oracle import oracle as o

pytestmark = pytest.mark.gpu

scipy_opt = pytest.importorskip("scipy.optimize")


def _model(t, F, A, B):
    return A * (1.0 - np.exp(-F * t)) + B


def _curve_fit(tau, y, p0):
    ref, _ = scipy_opt.curve_fit(_model, tau, y, p0=p0, maxfev=20000, xtol=1e-14, ftol=1e-14, gtol=1e-14)
    return ref


def _k_fit(curves, tau, iters=200):
    import torch
    c = torch.from_numpy(np.ascontiguousarray(curves, dtype=np.float64)).cuda()
    t = torch.from_numpy(np.ascontiguousarray(tau, dtype=np.float64)).cuda()
    out = torch.zeros((curves.shape[0], 4), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ddm.k_fit(c, t, out, iters)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _known_curves(n, seed):
    rng = np.random.default_rng(seed)
    tau = np.array(ddm.log_lags(512, 32), dtype=np.float64)
    F = 10.0 ** rng.uniform(-2.3, -0.3, n)
    A = rng.uniform(0.5, 5.0, n)
    B = rng.uniform(0.0, 1.0, n)
    curves = _model(tau[None, :], F[:, None], A[:, None], B[:, None])
    return tau, F, A, B, curves


def test_k_fit_recovers_known_parameters(gpu):
    tau, F, A, B, curves = _known_curves(1000, 1)
    print("k_fit exact curves", curves.shape, flush=True)
    p = _k_fit(curves, tau)
    assert np.max(np.abs(p[:, 0] - F) / F) <= 1e-6
    assert np.max(np.abs(p[:, 1] - A) / A) <= 1e-6
    assert np.max(np.abs(p[:, 2] - B)) <= 1e-6 * np.max(A)
    assert np.max(p[:, 3]) <= 1e-9 * np.max(A)


def test_k_fit_matches_curve_fit_on_noisy_curves(gpu):
    tau, F, A, B, curves = _known_curves(200, 2)
    rng = np.random.default_rng(3)
    noisy = curves * (1.0 + 0.02 * rng.standard_normal(curves.shape))
    print("k_fit noisy curves", noisy.shape, flush=True)
    p = _k_fit(noisy, tau)
    for i in range(noisy.shape[0]):
        ref = _curve_fit(tau, noisy[i], [F[i], A[i], B[i]])
        assert abs(p[i, 0] - ref[0]) <= 1e-6 * ref[0], (i, p[i], ref)
        assert abs(p[i, 1] - ref[1]) <= 1e-6 * abs(ref[1]), (i, p[i], ref)
        rms = np.sqrt(np.mean((noisy[i] - _model(tau, *ref)) ** 2))
        assert abs(p[i, 3] - rms) <= 1e-6 * rms


def test_k_fit_zero_curve_and_bad_arguments(gpu):
    import torch
    tau = np.arange(1, 9, dtype=np.float64)
    p = _k_fit(np.zeros((3, 8)), tau)
    assert np.all(p == 0.0)
    c = torch.zeros((2, 2), dtype=torch.float64, device="cuda")
    out = torch.zeros((2, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(ddm.DDMError):
        ddm.k_fit(c, c[0], out)  # L < 3
    c = torch.zeros((2, 8), dtype=torch.float64, device="cuda")
    t = torch.ones(8, dtype=torch.float64, device="cuda")
    with pytest.raises(ddm.DDMError):
        ddm.k_fit(c, t, out, iters=0)


def test_fit_results_brownian_video(gpu):
    N, S, D0 = 256, 128, 0.5
    frames = ddm.synth_brownian(N, S, S, 8, n_spots=200, sigma=2.0, amplitude=80.0, diffusion=D0,
                                background=100.0, noise=5.0, seed=3)
    lags, lams = ddm.log_lags(N // 2, 24), ddm.log_lambdas(4, 64, 12)
    print("fit brownian", N, S, flush=True)
    with ddm.Engine(S, S, [S, S // 2], lags, lams, 1.2, 32, norm=ddm.NORM_PAIRS) as e:
        e.submit_host(frames)
        res = e.results_by_scale()
        params, D = e.fit(fps=1.0)
        params10, _ = e.fit(fps=10.0)
        lam_sorted, lag_sorted = np.array(e.lambdas, dtype=np.float64), np.array(e.lags, dtype=np.float64)
    # fps only rescales tau: F scales with it, A and B do not
    assert np.allclose(params10[S][..., 0], 10.0 * params[S][..., 0], rtol=1e-6)
    assert np.allclose(params10[S][..., 1:3], params[S][..., 1:3], rtol=1e-6, atol=1e-12)
    # every well-posed curve (decay time inside the lag range) against
    # curve_fit on the same data
    checked = 0
    for s in (S, S // 2):
        for t in range(res[s].shape[0]):
            for q in range(len(lam_sorted)):
                y, got = res[s][t, q], params[s][t, q]
                if not (got[0] > 0 and got[1] > 0 and lag_sorted[0] <= 1.0 / got[0] <= lag_sorted[-1]):
                    continue
                ref = _curve_fit(lag_sorted, y, got[:3])
                assert abs(got[0] - ref[0]) <= 1e-5 * ref[0], (s, t, q, got, ref)
                checked += 1
    assert checked >= 40
    # the same fit of the FP64 oracle's I(q, tau)
    cfg = o.Config(S, S, N, [S, S // 2], lags, lams, np.float32(1.2).item(), norm_mode=o.NORM_PAIRS)
    ref_res = o.split_results(cfg, o.run(cfg, frames))
    for q in range(len(lam_sorted)):
        ref = _curve_fit(lag_sorted, ref_res[S][0, q], params[S][0, q, :3])
        assert abs(params[S][0, q, 0] - ref[0]) <= 1e-4 * ref[0], (q, params[S][0, q], ref)
    # D from the returned parameters by the documented rule, and its value
    qv = 2.0 * np.pi / lam_sorted
    for s in (S, S // 2):
        for t in range(len(D[s])):
            F, A = params[s][t, :, 0], params[s][t, :, 1]
            decay = 1.0 / np.where(F > 0, F, np.inf)
            ok = (F > 0) & (A > 0) & (decay >= lag_sorted[0]) & (decay <= lag_sorted[-1])
            want = np.exp(np.mean(np.log(F[ok]) - 2.0 * np.log(qv[ok]))) if ok.any() else 0.0
            assert abs(D[s][t] - want) <= 1e-12 * want
    print("D", D[S], "per-tile mean at", S // 2, D[S // 2].mean())
    assert 0.85 * D0 <= D[S][0] <= 1.5 * D0
    assert 0.7 * D0 <= np.median(D[S // 2]) <= 1.7 * D0


def test_fit_results_rejects_bad_fps(gpu):
    with ddm.Engine(64, 64, [64], [1, 2, 4], [8.0]) as e:
        with pytest.raises(ddm.DDMError):
            e.fit(fps=0.0)
    with ddm.Engine(64, 64, [64], [1, 2], [8.0]) as e:
        with pytest.raises(ddm.DDMError):
            e.fit(fps=1.0)
