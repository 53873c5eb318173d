// Lag engine of the DDM pipeline (processFFT / analyseChunk,
// src/DDM_kernel.cuh:85-101, src/DDM.cu:222-262): for a block of n new
// frames t0 .. t0+n-1 held in the spectrum ring and every lag tau,
//   acc[l][e] += sum_{j < n, t0+j-tau >= 0} |F_{t0+j}[e] - F_{t0+j-tau}[e]|^2
// (no 1/s^2 factor: it is a power of two, applied exactly in the average).
//
// Mapping: one workgroup per ring group (RING_GROUP = 64 consecutive
// spectrum elements, one per lane); its 16 waves take the lags in turn
// (wave w: lags w, w+16, ...).  The block's new frames are staged once in
// LDS.  For a lag tau, new frame j pairs with frame j - tau: for j < tau that
// frame is in the history, and the slots of j = jh0 .. min(tau, n)-1 are
// consecutive in the ring (one wrap at most), so each lag reads its
// partners as at most two contiguous runs; for j >= tau both frames are in
// LDS.  Runs are summed eight pairs at a time with two partial sums, which
// keeps the scalar unit out of the inner loop.  The sum per lag is added to
// the accumulator once per block, so results do not depend on timing.
#include <hip/hip_runtime.h>

#include "ddm_internal.h"
#include "ddm_kernels.h"

namespace ddm {

constexpr int LAG_NT = 1024;                    // threads per workgroup: 16 waves
constexpr int LAG_WAVES = LAG_NT / RING_GROUP;  // waves sharing one ring group
constexpr int LAG_UNROLL = 8;                   // pairs per step of a run

// sum_{k < count} |a[k*G] - b[k*G]|^2, G = RING_GROUP (consecutive ring slots
// or consecutive LDS rows are RING_GROUP complex values apart)
__device__ __forceinline__ float run_sum(const float2 *a, const float2 *b, int count) {
    float s0 = 0.f, s1 = 0.f;
    int k = 0;
    for (; k + LAG_UNROLL <= count; k += LAG_UNROLL) {
        float2 va[LAG_UNROLL], vb[LAG_UNROLL];
#pragma unroll
        for (int u = 0; u < LAG_UNROLL; ++u) {
            va[u] = a[(k + u) * RING_GROUP];
            vb[u] = b[(k + u) * RING_GROUP];
        }
#pragma unroll
        for (int u = 0; u < LAG_UNROLL; ++u) {
            const float dx = va[u].x - vb[u].x, dy = va[u].y - vb[u].y;
            if (u & 1)
                s1 += dx * dx + dy * dy;
            else
                s0 += dx * dx + dy * dy;
        }
    }
    for (; k < count; ++k) {
        const float2 va = a[k * RING_GROUP], vb = b[k * RING_GROUP];
        const float dx = va.x - vb.x, dy = va.y - vb.y;
        s0 += dx * dx + dy * dy;
    }
    return s0 + s1;
}

// E: ring / accumulator row stride (padded), R: ring slots, slot0: slot of
// t0, dmax = min(max lag, t0): oldest history frame the block may pair with.
template <int J>
__global__ __launch_bounds__(LAG_NT) void lag_kernel(const float2 *__restrict__ ring, float *__restrict__ acc,
                                                    long long E, int R, int slot0, int n, int dmax,
                                                    const int *__restrict__ lags, int L) {
    __shared__ float2 nb[J * RING_GROUP];
    const int tid = threadIdx.x;
    const int lane = tid % RING_GROUP;
    const int w = __builtin_amdgcn_readfirstlane(tid / RING_GROUP);
    const float2 *g = ring + (long long)blockIdx.x * R * RING_GROUP + lane;  // slot s at g[s * RING_GROUP]
    {
        // stage the new frames: all loads in flight before the LDS stores
        constexpr int K = (J + LAG_WAVES - 1) / LAG_WAVES;
        float2 v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = w + k * LAG_WAVES;
            if (j < n) {
                int s = slot0 + j;
                s -= (s >= R) ? R : 0;
                v[k] = g[(long long)s * RING_GROUP];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = w + k * LAG_WAVES;
            if (j < n) nb[j * RING_GROUP + lane] = v[k];
        }
    }
    __syncthreads();
    const float2 *nl = nb + lane;
    float *accg = acc + (long long)blockIdx.x * RING_GROUP + lane;
    for (int l = w; l < L; l += LAG_WAVES) {
        const int tau = lags[l];
        float sum = 0.f;
        // history partners: j in [jh0, jh1), partner slot slot0 + j - tau (+R if negative)
        const int jh0 = (tau - dmax) > 0 ? (tau - dmax) : 0;
        const int jh1 = tau < n ? tau : n;
        if (jh0 < jh1) {
            int jw = tau - slot0;  // j < jw: the partner slot wrapped
            jw = jw < jh0 ? jh0 : (jw > jh1 ? jh1 : jw);
            if (jh0 < jw)
                sum += run_sum(g + (long long)(slot0 + jh0 - tau + R) * RING_GROUP, nl + jh0 * RING_GROUP, jw - jh0);
            if (jw < jh1)
                sum += run_sum(g + (long long)(slot0 + jw - tau) * RING_GROUP, nl + jw * RING_GROUP, jh1 - jw);
        }
        // partners inside the block: j in [tau, n) pairs with j - tau
        if (tau < n) sum += run_sum(nl, nl + tau * RING_GROUP, n - tau);
        accg[(long long)l * E] += sum;
    }
}

template <int J>
static void launch_lag_t(const LagArgs &a, hipStream_t st) {
    const unsigned grid = (unsigned)(a.elems_per_frame / RING_GROUP);
    hipLaunchKernelGGL(lag_kernel<J>, dim3(grid), dim3(LAG_NT), 0, st, (const float2 *)a.ring, a.acc,
                       a.elems_per_frame, a.ring_slots, a.slot0, a.n_new, a.dmax, a.lags, a.lag_count);
}

hipError_t launch_lag(const LagArgs &a, hipStream_t st) {
    if (a.elems_per_frame % RING_GROUP != 0 || a.n_new < 1 || a.n_new > a.lag_block || a.dmax < 0 ||
        a.slot0 < 0 || a.slot0 >= a.ring_slots)
        return hipErrorInvalidValue;
    switch (a.lag_block) {
    case 32: launch_lag_t<32>(a, st); break;
    case 64: launch_lag_t<64>(a, st); break;
    case 128: launch_lag_t<128>(a, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ddm
