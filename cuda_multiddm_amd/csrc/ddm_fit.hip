// On-device fit of the image structure function.  For every curve
// I(tau) of a (scale, tile, q) the Brownian model of the reference's
// analysis script, I = A (1 - exp(-F tau)) + B
// (resources/python_tools/plot_averaged.py:61-62), is fitted by
// Levenberg-Marquardt in FP64, one thread per curve.  The data are scaled to
// max |I| = 1 and tau to tau_max = 1 before the fit; F is carried as
// exp(phi) so it stays positive.
#include <hip/hip_runtime.h>

#include "ddm_kernels.h"

namespace ddm {

__device__ __forceinline__ void model(double phi, double A, double B, double t, double *f, double *dphi, double *dA) {
    const double F = exp(phi);
    const double e = exp(-F * t);
    *f = A * (1.0 - e) + B;
    *dA = 1.0 - e;
    *dphi = A * t * e * F;  // d f / d phi = d f / d F * F
}

// curves[c][L] results, tau[L] lags (any positive unit), out[c][4] = F, A, B, rms.
__global__ void fit_kernel(const double *__restrict__ curves, const double *__restrict__ tau, int L, long long n,
                           int iters, double *__restrict__ out) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *y = curves + c * L;
    double ymax = 0.0, tmax = 0.0;
    for (int i = 0; i < L; ++i) {
        ymax = fmax(ymax, fabs(y[i]));
        tmax = fmax(tmax, tau[i]);
    }
    double *o = out + c * 4;
    if (!(ymax > 0.0) || !(tmax > 0.0) || L < 3) {
        o[0] = o[1] = o[2] = o[3] = 0.0;
        return;
    }
    const double sy = 1.0 / ymax, st = 1.0 / tmax;
    // initial guess: B from the shortest lag, A from the plateau, F from the
    // lag where the curve crosses half-way
    double B = y[0] * sy, A = y[L - 1] * sy - B;
    double thalf = tau[L - 1] * st;
    for (int i = 0; i < L; ++i)
        if (y[i] * sy >= B + 0.5 * A) {
            thalf = tau[i] * st;
            break;
        }
    double phi = log(0.693147180559945309 / fmax(thalf, 1e-6));
    auto rss_of = [&](double ph, double a, double b) {
        double s = 0.0;
        for (int i = 0; i < L; ++i) {
            double f, d1, d2;
            model(ph, a, b, tau[i] * st, &f, &d1, &d2);
            const double r = y[i] * sy - f;
            s += r * r;
        }
        return s;
    };
    double rss = rss_of(phi, A, B), mu = 1e-3;
    for (int it = 0; it < iters; ++it) {
        double JTJ[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, JTr[3] = {0, 0, 0};
        for (int i = 0; i < L; ++i) {
            double f, dphi, dA;
            model(phi, A, B, tau[i] * st, &f, &dphi, &dA);
            const double J[3] = {dphi, dA, 1.0};
            const double r = y[i] * sy - f;
            for (int a = 0; a < 3; ++a) {
                JTr[a] += J[a] * r;
                for (int b = 0; b < 3; ++b) JTJ[a][b] += J[a] * J[b];
            }
        }
        bool improved = false;
        for (int tries = 0; tries < 8 && !improved; ++tries) {
            double M[3][4];
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) M[a][b] = JTJ[a][b] + (a == b ? mu * (JTJ[a][a] + 1e-12) : 0.0);
                M[a][3] = JTr[a];
            }
            // Gaussian elimination with partial pivoting on the 3x3 system
            bool ok = true;
            for (int k = 0; k < 3 && ok; ++k) {
                int p = k;
                for (int r = k + 1; r < 3; ++r)
                    if (fabs(M[r][k]) > fabs(M[p][k])) p = r;
                if (!(fabs(M[p][k]) > 1e-300)) {
                    ok = false;
                    break;
                }
                if (p != k)
                    for (int cc = 0; cc < 4; ++cc) {
                        const double t = M[k][cc];
                        M[k][cc] = M[p][cc];
                        M[p][cc] = t;
                    }
                for (int r = k + 1; r < 3; ++r) {
                    const double m = M[r][k] / M[k][k];
                    for (int cc = k; cc < 4; ++cc) M[r][cc] -= m * M[k][cc];
                }
            }
            if (!ok) {
                mu *= 10.0;
                continue;
            }
            double d[3];
            for (int k = 2; k >= 0; --k) {
                double s = M[k][3];
                for (int cc = k + 1; cc < 3; ++cc) s -= M[k][cc] * d[cc];
                d[k] = s / M[k][k];
            }
            const double dphi = fmax(fmin(d[0], 2.0), -2.0);  // at most a factor e^2 in F per step
            const double r2 = rss_of(phi + dphi, A + d[1], B + d[2]);
            if (r2 < rss) {
                phi += dphi;
                A += d[1];
                B += d[2];
                rss = r2;
                mu = fmax(mu * 0.3, 1e-12);
                improved = true;
            } else {
                mu *= 10.0;
            }
        }
        if (!improved) break;
    }
    o[0] = exp(phi) * st;  // back to the caller's tau unit
    o[1] = A * ymax;
    o[2] = B * ymax;
    o[3] = sqrt(rss / L) * ymax;
}

hipError_t launch_fit(const double *curves, const double *tau, int L, long long n, int iters, double *out,
                      hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int nt = 64;
    hipLaunchKernelGGL(fit_kernel, dim3((unsigned)((n + nt - 1) / nt)), dim3(nt), 0, st, curves, tau, L, n, iters,
                       out);
    return hipGetLastError();
}

}  // namespace ddm
