// HIP kernels of the multi-scale DDM engine for gfx950 (CDNA4, wave64).
//
//   tile_fft_small<S>  S <= 128: tile cut + 2-D R2C FFT of whole tiles in LDS
//   tile_fft_rows<S>   S >= 256: tile cut + row R2C FFTs (first half of the 2-D FFT)
//   tile_fft_cols<S>   S >= 256: column FFTs, in place in the ring slot
//   azimuthal_kernel   masked q-ring average of the lag sums, FP64
// (the lag engine is in ddm_lag.hip)
//
// Reference counterparts: parseBufferScalePow2 (src/DDM_kernel.cuh:43-78),
// cufftExecR2C (src/DDM.cu:189-192), processFFT (src/DDM_kernel.cuh:85-101),
// kernelMaskReduce (src/azimuthal_average_kernel.cuh:68-135).
//
// FFT scheme (all sizes): the real s x s tile is read as s/2 complex rows
// z_r = x_{2r} + i x_{2r+1}; each gets one complex Stockham FFT of length s
// (radix-4 passes, one radix-2 pass for odd log2 s); the two real rows'
// spectra are separated as X_a = (Z_k + conj Z_{s-k}) / 2,
// X_b = -i (Z_k - conj Z_{s-k}) / 2.  Columns kx = 0 and kx = s/2 are real
// after the row pass, so they travel packed as one complex column
// (DC + i Nyquist); this leaves exactly s/2 complex columns of length s,
// and the packed column is separated the same way when the result is
// written.  Twiddles come from a table rounded once from double precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "ddm_device.h"
#include "ddm_internal.h"
#include "ddm_kernels.h"

namespace ddm {

constexpr int NT = 256;  // threads per workgroup (4 waves of 64)


// LDS index with one padding slot per 32 complex values: Stockham writes
// whose lanes step by 4 or 16 values then fall on distinct banks.
__device__ __forceinline__ int lp(int i) { return i + (i >> 5); }
constexpr int lds_len(int n) { return n + (n >> 5) + 1; }

template <int R>
struct Bfly;
template <>
struct Bfly<2> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <>
struct Bfly<4> {
    // X0 = (v0+v2)+(v1+v3), X1 = (v0-v2)-i(v1-v3), X2 = (v0+v2)-(v1+v3), X3 = (v0-v2)+i(v1-v3)
    __device__ __forceinline__ static void run(float2 *v) {
        float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = make_float2(d02.x + d13.y, d02.y - d13.x);
        v[3] = make_float2(d02.x - d13.y, d02.y + d13.x);
    }
};

// Sequence q (of NSEQ) starts at (q / SPT) * TSTR + (q % SPT) * SSTR; element
// i of a sequence is ESTR further.  SEQ_FAST: consecutive threads walk
// sequences first (column layouts), else elements first (row layouts).
template <int N, int NSEQ, int SPT, int SSTR, int TSTR, int ESTR, bool SEQ_FAST>
struct Layout {
    static constexpr int n = N, nseq = NSEQ, spt = SPT, sstr = SSTR, tstr = TSTR, estr = ESTR;
    static constexpr bool seq_fast = SEQ_FAST;
    __device__ __forceinline__ static int base(int q) { return (q / SPT) * TSTR + (q % SPT) * SSTR; }
};

// One in-place Stockham pass of radix R over all sequences of the layout.
// Every thread loads all of its butterflies' inputs before the barrier and
// writes after it, so the pass needs no second buffer.
template <class LY, int R, int NS, int NTH>
__device__ __forceinline__ void stockham_pass(float2 *buf, const float2 *tw) {
    constexpr int NT = NTH;
    constexpr int N = LY::n;
    constexpr int BPS = N / R;
    constexpr int TOTAL = BPS * LY::nseq;
    constexpr int KB = (TOTAL + NT - 1) / NT;
    constexpr int TWS = N / (NS * R);
    const int tid = threadIdx.x;
    float2 v[KB][R];
    int dst[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int b = tid + k * NT;
        dst[k] = -1;
        if ((TOTAL % NT == 0) || b < TOTAL) {
            const int q = LY::seq_fast ? (b % LY::nseq) : (b / BPS);
            const int j = LY::seq_fast ? (b / LY::nseq) : (b % BPS);
            const int kk = j % NS;
            const int base = LY::base(q);
#pragma unroll
            for (int r = 0; r < R; ++r) v[k][r] = buf[lp(base + (j + r * BPS) * LY::estr)];
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[k][r] = cmul(v[k][r], tw[kk * r * TWS]);
            }
            Bfly<R>::run(v[k]);
            dst[k] = base + ((j / NS) * NS * R + kk) * LY::estr;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        if ((TOTAL % NT == 0) || dst[k] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lp(dst[k] + r * NS * LY::estr)] = v[k][r];
        }
    }
    __syncthreads();
}

template <class LY, int NS, int NTH>
__device__ __forceinline__ void fft_passes(float2 *buf, const float2 *tw) {
    if constexpr (NS < LY::n) {
        constexpr int REM = LY::n / NS;
        constexpr int R = (REM % 4 == 0) ? 4 : 2;
        stockham_pass<LY, R, NS, NTH>(buf, tw);
        fft_passes<LY, NS * R, NTH>(buf, tw);
    }
}

// ---------------------------------------------------------------------------
// S <= 128: whole tiles in LDS.  TPW tiles per workgroup so that each
// workgroup holds at least 1024 complex values.
// ---------------------------------------------------------------------------
template <int S>
struct SmallCfg {
    static constexpr int H = S / 2;
    static constexpr int TPW = (2048 / (S * S)) > 0 ? (2048 / (S * S)) : 1;
    static constexpr int TILE = S * H;  // complex values per tile in LDS
    // 128x128 tiles: 1024 threads, two butterflies per thread and pass
    static constexpr int NTH = S >= 128 ? 1024 : 256;
};

template <int S>
__global__ __launch_bounds__(SmallCfg<S>::NTH) void tile_fft_small(PixSrc ps, RingDst rd, int tps, int tiles,
                                                                  int groups, const float2 *__restrict__ twg) {
    using C = SmallCfg<S>;
    constexpr int H = C::H, TPW = C::TPW, TILE = C::TILE, NT = C::NTH;
    __shared__ float2 buf[lds_len(TPW * TILE)];
    __shared__ float2 tw[S];
    const int tid = threadIdx.x;
    // frames vary fastest across the grid: concurrent workgroups write
    // neighbouring ring slots of the same element groups
    const int nf = gridDim.x / groups;
    const int f = blockIdx.x % nf;
    const int tile0 = (blockIdx.x / nf) * TPW;
    for (int i = tid; i < S; i += NT) tw[i] = twg[S + i];

    // load: z[t][r][x] = x_{2r}[x] + i x_{2r+1}[x]; every pixel load is in
    // flight before the first LDS store
    const uint8_t *fr = ps.frames + (long long)f * ps.frame_bytes;
    {
        constexpr int NL = TPW * TILE, KL = (NL + NT - 1) / NT;
        float2 v[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int i = tid + k * NT;
            const int t = i / TILE, rem = i % TILE;
            const int r = rem / S, x = rem % S;
            const int tile = tile0 + t;
            v[k] = make_float2(0.f, 0.f);
            if ((NL % NT == 0 || i < NL) && tile < tiles) {
                const int px = ps.x_off + (tile % tps) * S + x;
                const int py = ps.y_off + (tile / tps) * S + 2 * r;
                v[k] = make_float2(load_px(fr, ps, px, py), load_px(fr, ps, px, py + 1));
            }
        }
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int i = tid + k * NT;
            if (NL % NT == 0 || i < NL) buf[lp(i)] = v[k];
        }
    }
    __syncthreads();

    // row FFTs: TPW * H sequences of length S, contiguous
    using Rows = Layout<S, TPW * H, TPW * H, S, 0, 1, false>;
    fft_passes<Rows, 1, NT>(buf, tw);

    // separate the row pairs: Y[2r][k] at r*S + k, Y[2r+1][k] at r*S + H + k
    {
        constexpr int NU = TPW * H * H;
        constexpr int KU = (NU + NT - 1) / NT;
        float2 za[KU], zb[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int i = tid + u * NT;
            if ((NU % NT == 0) || i < NU) {
                const int k = i % H, rr = i / H;  // rr = t * H + r
                const int base = rr * S;
                za[u] = buf[lp(base + k)];
                zb[u] = buf[lp(base + (k == 0 ? H : S - k))];
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int i = tid + u * NT;
            if ((NU % NT == 0) || i < NU) {
                const int k = i % H, rr = i / H;
                const int base = rr * S;
                float2 ya, yb;
                if (k == 0) {
                    ya = make_float2(za[u].x, zb[u].x);  // (DC, Nyquist) of row 2r
                    yb = make_float2(za[u].y, zb[u].y);  // (DC, Nyquist) of row 2r+1
                } else {
                    ya = make_float2(0.5f * (za[u].x + zb[u].x), 0.5f * (za[u].y - zb[u].y));
                    yb = make_float2(0.5f * (za[u].y + zb[u].y), -0.5f * (za[u].x - zb[u].x));
                }
                buf[lp(base + k)] = ya;
                buf[lp(base + H + k)] = yb;
            }
        }
        __syncthreads();
    }

    // column FFTs: per tile H columns of length S, stride H
    using Cols = Layout<S, TPW * H, H, 1, TILE, H, true>;
    fft_passes<Cols, 1, NT>(buf, tw);

    // write [tile][ky][kx], separating the packed DC / Nyquist column
    constexpr int OUT = S * (H + 1);
    for (int i = tid; i < TPW * OUT; i += NT) {
        const int t = i / OUT, rem = i % OUT;
        const int tile = tile0 + t;
        if (tile >= tiles) continue;
        const int ky = rem / (H + 1), kx = rem % (H + 1);
        const int y0 = t * TILE;
        float2 val;
        if (kx == 0 || kx == H) {
            const float2 w = buf[lp(y0 + ky * H)];
            const float2 v = buf[lp(y0 + ((S - ky) & (S - 1)) * H)];
            val = (kx == 0) ? make_float2(0.5f * (w.x + v.x), 0.5f * (w.y - v.y))
                            : make_float2(0.5f * (w.y + v.y), -0.5f * (w.x - v.x));
        } else {
            val = buf[lp(y0 + ky * H + kx)];
        }
        tile_out(rd, f, tile, S).put(ky, kx, val);
    }
}

// ---------------------------------------------------------------------------
// S >= 256, pass 1: P row pairs per workgroup; row FFT, separation, and the
// half-spectrum rows written to the ring slot (columns 0..H-1, with the
// packed DC/Nyquist value in column 0).
// ---------------------------------------------------------------------------
template <int S>
struct LargeCfg {
    static constexpr int H = S / 2;
    static constexpr int P = (2048 / S) > 0 ? (2048 / S) : 1;  // row pairs per workgroup
    static constexpr int CB = 8192 / S;                        // columns per workgroup
    static constexpr int NTH_COLS = 1024;                      // two butterflies per thread and pass
};

template <int S>
__global__ __launch_bounds__(NT) void tile_fft_rows(PixSrc ps, RingDst rd, int tps, int tiles,
                                                   const float2 *__restrict__ twg) {
    using C = LargeCfg<S>;
    constexpr int H = C::H, P = C::P, GROUPS = H / P;
    __shared__ float2 buf[lds_len(P * S)];
    __shared__ float2 tw[S];
    const int tid = threadIdx.x;
    const int nf = gridDim.x / (GROUPS * tiles);  // frames vary fastest
    const int f = blockIdx.x % nf;
    const int g = (blockIdx.x / nf) % GROUPS;
    const int tile = blockIdx.x / (nf * GROUPS);
    for (int i = tid; i < S; i += NT) tw[i] = twg[S + i];
    const uint8_t *fr = ps.frames + (long long)f * ps.frame_bytes;
    const int x0 = ps.x_off + (tile % tps) * S;
    const int y0 = ps.y_off + (tile / tps) * S + 2 * g * P;
    {
        constexpr int KL = P * S / NT;  // P * S = 2048 = 8 * NT
        float2 v[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int i = tid + k * NT;
            const int r = i / S, x = i % S;
            v[k] = make_float2(load_px(fr, ps, x0 + x, y0 + 2 * r), load_px(fr, ps, x0 + x, y0 + 2 * r + 1));
        }
#pragma unroll
        for (int k = 0; k < KL; ++k) buf[lp(tid + k * NT)] = v[k];
    }
    __syncthreads();
    using Rows = Layout<S, P, P, S, 0, 1, false>;
    fft_passes<Rows, 1, NT>(buf, tw);
    const TileOut out = tile_out(rd, f, tile, S);  // the full row-order layout (the host checks)
    for (int i = tid; i < P * H; i += NT) {
        const int r = i / H, k = i % H;
        const float2 za = buf[lp(r * S + k)];
        const float2 zb = buf[lp(r * S + (k == 0 ? H : S - k))];
        float2 ya, yb;
        if (k == 0) {
            ya = make_float2(za.x, zb.x);
            yb = make_float2(za.y, zb.y);
        } else {
            ya = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
            yb = make_float2(0.5f * (za.y + zb.y), -0.5f * (za.x - zb.x));
        }
        const int ly = 2 * (g * P + r);
        out.full(ly, k) = ya;
        out.full(ly + 1, k) = yb;
    }
}

// S >= 256, pass 2: CB columns per workgroup, FFT along ky, in place.
template <int S>
__global__ __launch_bounds__(LargeCfg<S>::NTH_COLS) void tile_fft_cols(RingDst rd, int tiles,
                                                                      const float2 *__restrict__ twg) {
    using C = LargeCfg<S>;
    constexpr int H = C::H, CB = C::CB, NB = H / CB, NT = C::NTH_COLS;
    __shared__ float2 buf[lds_len(S * CB)];
    __shared__ float2 tw[S];
    const int tid = threadIdx.x;
    const int nf = gridDim.x / (NB * tiles);  // frames vary fastest
    const int f = blockIdx.x % nf;
    const int cb = (blockIdx.x / nf) % NB;
    const int tile = blockIdx.x / (nf * NB);
    for (int i = tid; i < S; i += NT) tw[i] = twg[S + i];
    const TileOut t = tile_out(rd, f, tile, S);  // the full row-order layout (the host checks)
    const int c0 = cb * CB;
    {
        constexpr int KL = S * CB / NT;  // S * CB = 8192 = 8 * NT
        float2 v[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int i = tid + k * NT;
            const int ly = i / CB, c = i % CB;
            v[k] = t.full(ly, c0 + c);
        }
#pragma unroll
        for (int k = 0; k < KL; ++k) buf[lp(tid + k * NT)] = v[k];
    }
    __syncthreads();
    using Cols = Layout<S, CB, CB, 1, 0, CB, true>;
    fft_passes<Cols, 1, NT>(buf, tw);
    for (int i = tid; i < S * CB; i += NT) {
        const int ky = i / CB, c = i % CB;
        if (c0 + c == 0) {
            const float2 w = buf[lp(ky * CB)];
            const float2 v = buf[lp(((S - ky) & (S - 1)) * CB)];
            t.full(ky, 0) = make_float2(0.5f * (w.x + v.x), 0.5f * (w.y - v.y));
            t.full(ky, H) = make_float2(0.5f * (w.y + v.y), -0.5f * (w.x - v.x));
        } else {
            t.full(ky, c0 + c) = buf[lp(i)];
        }
    }
}

// ---------------------------------------------------------------------------
// Masked azimuthal average of results[scale][tile][q][lag]: the q ring's
// elements summed in FP64 in a fixed order and normalised,
// I = 2 * sum * s^-4 / count * norm[lag].  WPR waves per value: four (the
// whole workgroup) for scales with large rings, one for the rest, so that a
// workgroup then takes AZ_WAVES consecutive values sharing the ring's
// element list.
// ---------------------------------------------------------------------------
constexpr int AZ_WAVES = 4;

// The lag sum of one accumulator entry in FP64: the FP32 accumulator, plus
// (SRC >= 1) its compensation term, plus (SRC >= 2) the FP64 part folded
// out of them earlier.
template <int SRC>
struct AccSrc {
    const float *a;
    const float *c;
    const double *d;
    __device__ __forceinline__ double operator[](int i) const {
        double v = (double)a[i];
        if constexpr (SRC >= 1) v += (double)c[i];
        if constexpr (SRC >= 2) v += d[i];
        return v;
    }
    __device__ __forceinline__ AccSrc shift(long long k) const {
        return AccSrc{a + k, SRC >= 1 ? c + k : c, SRC >= 2 ? d + k : d};
    }
};

// sum_{i = first, first + step, ... < cnt} a[ix[i]] in FP64: four independent
// partial sums, so four element loads per lane are in flight at once; the
// order is fixed, so the result does not depend on timing.
template <class A>
__device__ __forceinline__ double ring_sum(const A a, const int *__restrict__ ix, int first, int step, int cnt) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = first;
    for (; i + 3 * step < cnt; i += 4 * step) {
        const int i0 = ix[i], i1 = ix[i + step], i2 = ix[i + 2 * step], i3 = ix[i + 3 * step];
        s0 += a[i0];
        s1 += a[i1];
        s2 += a[i2];
        s3 += a[i3];
    }
    for (; i < cnt; i += step) s0 += a[ix[i]];
    return (s0 + s1) + (s2 + s3);
}

template <int WPR, int SRC>
__global__ __launch_bounds__(AZ_WAVES * 64) void azimuthal_kernel(AccSrc<SRC> acc, long long E,
                                                                  const AzScale *__restrict__ scales, int S,
                                                                  const int *__restrict__ q_off,
                                                                  const int *__restrict__ q_cnt,
                                                                  const int *__restrict__ idx,
                                                                  const double *__restrict__ norm, int Q, int L,
                                                                  double *__restrict__ out, long long first_item,
                                                                  long long count) {
    constexpr int NTR = WPR * 64;  // threads per value
    __shared__ double part[AZ_WAVES];
    const int tid = threadIdx.x % NTR;
    const long long item = first_item + (long long)blockIdx.x * (AZ_WAVES / WPR) + threadIdx.x / NTR;
    if (item >= first_item + count) return;  // whole values only (the WPR = 4 grid is exact)
    int si = 0;
    while (si + 1 < S && item >= scales[si + 1].out_base) ++si;
    const AzScale sc = scales[si];
    const long long r = item - sc.out_base;
    const int l = (int)(r % L);
    const int q = (int)((r / L) % Q);
    const long long tile = r / ((long long)L * Q);
    const int cnt = q_cnt[si * Q + q];
    if (cnt == 0) {
        if (tid == 0) out[item] = 0.0;
        return;
    }
    const int *ix = idx + q_off[si * Q + q];
    const AccSrc<SRC> a = acc.shift((long long)l * E + sc.elem_off + tile * sc.tile_elems);
    double sum = ring_sum(a, ix, tid, NTR, cnt);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, 64);
    if constexpr (WPR > 1) {
        if ((tid & 63) == 0) part[tid >> 6] = sum;
        __syncthreads();
        if (tid != 0) return;
        sum = 0.0;
#pragma unroll
        for (int w = 0; w < WPR; ++w) sum += part[w];
    } else if (tid != 0) {
        return;
    }
    out[item] = 2.0 * sum * sc.inv_s4 / (double)cnt * norm[l];
}

// Narrow-ring scales: a wave per (tile, q, group of lags).  A ring of cnt
// elements gets P = az_lanes(cnt) lanes per lag and the wave 64 / P lags;
// lane (g, p) sums elements p, p + P, ... of lag 64 / P * group + g in FP64,
// the P lanes of each lag are added by an xor butterfly, and the lane with
// p = 0 writes the value.  Small rings thus fill the wave instead of one
// lag leaving most lanes idle.
template <int SRC>
__global__ __launch_bounds__(AZ_WAVES * 64) void azimuthal_narrow_kernel(
    AccSrc<SRC> acc, long long E, const AzScale *__restrict__ scales, int S, int s0,
    const int *__restrict__ q_off, const int *__restrict__ q_cnt, const int *__restrict__ idx,
    const int *__restrict__ wl, const double *__restrict__ norm, int Q, int L, double *__restrict__ out,
    long long nwaves) {
    const long long gw = (long long)blockIdx.x * AZ_WAVES + threadIdx.x / 64;
    if (gw >= nwaves) return;
    const int lane = threadIdx.x % 64;
    int si = s0;
    while (si + 1 < S && gw >= scales[si + 1].wave_base) ++si;
    const AzScale sc = scales[si];
    const long long local = gw - sc.wave_base;
    const long long tile = local / sc.wl_n;
    const int ent = wl[sc.wl_off + (int)(local - tile * sc.wl_n)];
    const int q = ent & 0xffff, grp = ent >> 16;
    const int cnt = q_cnt[si * Q + q];
    const int P = az_lanes(cnt), G = 64 / P;
    const int g = lane / P, p = lane % P;
    const int l = grp * G + g;
    double sum = 0.0;
    if (l < L) {
        const int *ix = idx + q_off[si * Q + q];
        const AccSrc<SRC> a = acc.shift((long long)l * E + sc.elem_off + tile * sc.tile_elems);
        sum = ring_sum(a, ix, p, P, cnt);
    }
    for (int o = P / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (p == 0 && l < L)
        out[sc.out_base + (tile * Q + q) * L + l] = cnt ? 2.0 * sum * sc.inv_s4 / (double)cnt * norm[l] : 0.0;
}

// acc64 += acc + comp, then acc = comp = 0: the FP32 pair's value moves to
// the FP64 accumulators (the two FP32 terms add exactly in FP64).
__global__ __launch_bounds__(256) void acc_fold_kernel(float4 *__restrict__ acc, float4 *__restrict__ comp,
                                                       double2 *__restrict__ acc64, long long n4) {
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) {
        const float4 a = acc[i], c = comp[i];
        double2 d0 = acc64[2 * i], d1 = acc64[2 * i + 1];
        d0.x += (double)a.x + (double)c.x;
        d0.y += (double)a.y + (double)c.y;
        d1.x += (double)a.z + (double)c.z;
        d1.y += (double)a.w + (double)c.w;
        acc64[2 * i] = d0;
        acc64[2 * i + 1] = d1;
        acc[i] = z;
        comp[i] = z;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// Twiddle table (fill_twiddles) in static device memory, filled once per
// device on first use: no launch path allocates device memory for it.
static __device__ __align__(256) float2 g_twiddles[2 * kMaxScale];
static std::mutex g_tw_mu;
static float2 *g_tw[64] = {nullptr};

hipError_t twiddle_table(const float2 **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (!g_tw[dev]) {
        std::vector<float> h;
        fill_twiddles(&h);
        float2 *d = nullptr;
        // the array's device address on this device (its name in host code
        // is only a stand-in)
        e = hipGetSymbolAddress((void **)&d, HIP_SYMBOL(g_twiddles));
        if (e != hipSuccess) return e;
        e = hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        g_tw[dev] = d;
    }
    *out = g_tw[dev];
    return hipSuccess;
}

template <int S>
static void launch_small(const PixSrc &ps, const RingDst &rd, int tps, int tiles, int n_frames,
                         const float2 *tw, hipStream_t st) {
    constexpr int TPW = SmallCfg<S>::TPW;
    const int groups = (tiles + TPW - 1) / TPW;
    hipLaunchKernelGGL(tile_fft_small<S>, dim3((unsigned)(groups * n_frames)), dim3(SmallCfg<S>::NTH), 0, st, ps, rd,
                       tps,
                       tiles, groups, tw);
}

template <int S>
static void launch_large(const PixSrc &ps, const RingDst &rd, int tps, int tiles, int n_frames,
                         const float2 *tw, hipStream_t st) {
    using C = LargeCfg<S>;
    const unsigned g1 = (unsigned)((C::H / C::P) * tiles * n_frames);
    hipLaunchKernelGGL(tile_fft_rows<S>, dim3(g1), dim3(NT), 0, st, ps, rd, tps, tiles, tw);
    const unsigned g2 = (unsigned)((C::H / C::CB) * tiles * n_frames);
    hipLaunchKernelGGL(tile_fft_cols<S>, dim3(g2), dim3(C::NTH_COLS), 0, st, rd, tiles, tw);
}

hipError_t launch_tile_fft(const TileFftArgs &a, hipStream_t st) {
    const float2 *tw = nullptr;
    hipError_t e = twiddle_table(&tw);
    if (e != hipSuccess) return e;
    PixSrc ps{(const uint8_t *)a.frames, a.frame_bytes, (long long)a.width * a.bpp, a.bpp, a.fmt, a.x_off,
              a.y_off};
    const TileLayout lay = a.lay.elems ? a.lay : TileLayout::full(a.scale);
    RingDst rd{(float2 *)a.ring, a.elem_off, a.ring_slots, a.first_slot, lay};
    const int tps = a.largest_scale / a.scale, tiles = tps * tps;
    if (a.precision != DDM_PREC_FP32 && (a.scale == 64 || a.scale == 128))
        return launch_tile_dft_mfma(ps, rd, a.scale, tps, tiles, a.n_frames, a.precision, st);
    const bool lds_family = a.family == DDM_FFT_LDS && a.scale <= 128;
    if (!lds_family && launch_tile_fft_reg(ps, rd, a.scale, tps, tiles, a.n_frames, tw, (float2 *)a.xt, st))
        return hipGetLastError();
    // the 256 .. 2048 transforms below go through the ring slot in place:
    // full row-order layout only
    if (lay.kind != LAY_FULL && (a.scale >= 256 || a.scale < 8)) return hipErrorInvalidValue;
    switch (a.scale) {
    case 2: launch_small<2>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 4: launch_small<4>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 8: launch_small<8>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 16: launch_small<16>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 32: launch_small<32>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 64: launch_small<64>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 128: launch_small<128>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 256: launch_large<256>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 512: launch_large<512>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 1024: launch_large<1024>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    case 2048: launch_large<2048>(ps, rd, tps, tiles, a.n_frames, tw, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static bool tiler_block_arg(const TileFftArgs &a) {
    const bool mfma = a.precision != DDM_PREC_FP32 && (a.scale == 64 || a.scale == 128);
    return !mfma && a.family != DDM_FFT_LDS && tiler_block_scale(a.scale, a.largest_scale);
}

hipError_t launch_tile_fft_scales(const TileFftArgs *a, int n, bool fused, hipStream_t st, int *launches) {
    int nl = 0;
    std::vector<int> blk, band;
    if (fused) {
        for (int i = 0; i < n; ++i) {
            if (tiler_block_arg(a[i])) blk.push_back(i);
            else if (a[i].xt && tiler_band_scale(a[i].scale, a[i].largest_scale)) band.push_back(i);
        }
        if (blk.size() < 2) blk.clear();
        if (band.size() < 2) band.clear();
    }
    for (int i = 0; i < n; ++i) {
        if (std::find(blk.begin(), blk.end(), i) != blk.end() || std::find(band.begin(), band.end(), i) != band.end())
            continue;
        const hipError_t e = launch_tile_fft(a[i], st);
        if (e != hipSuccess) return e;
        nl += a[i].scale >= 256 ? 2 : 1;
    }
    if (blk.empty() && band.empty()) {
        if (launches) *launches += nl;
        return hipSuccess;
    }
    const float2 *tw = nullptr;
    hipError_t e = twiddle_table(&tw);
    if (e != hipSuccess) return e;
    const TileFftArgs &a0 = a[blk.empty() ? band[0] : blk[0]];
    const PixSrc ps{(const uint8_t *)a0.frames, a0.frame_bytes, (long long)a0.width * a0.bpp, a0.bpp, a0.fmt,
                    a0.x_off, a0.y_off};
    const int M = a0.largest_scale;
    auto scale_of = [&](int i, float2 *xt) {
        const TileLayout lay = a[i].lay.elems ? a[i].lay : TileLayout::full(a[i].scale);
        return TilerScale{RingDst{(float2 *)a[i].ring, a[i].elem_off, a[i].ring_slots, a[i].first_slot, lay},
                          a[i].scale, xt};
    };
    if (!blk.empty()) {
        TilerScale sc[5];
        for (size_t k = 0; k < blk.size(); ++k) sc[k] = scale_of(blk[k], nullptr);
        e = launch_tiler_blocks(ps, sc, (int)blk.size(), M, a0.n_frames, tw, st);
        if (e != hipSuccess) return e;
        nl += 1;
    }
    if (!band.empty()) {
        // one pass-buffer region per scale
        const long long region = big_fft_scratch_elems(M, a[band[0]].scale, a0.n_frames);
        TilerScale sc[3];
        for (size_t k = 0; k < band.size(); ++k) sc[k] = scale_of(band[k], (float2 *)a[band[k]].xt + k * region);
        e = launch_tiler_bands(ps, sc, (int)band.size(), M, a0.n_frames, tw, st);
        if (e != hipSuccess) return e;
        nl += 1 + (int)band.size();
    }
    if (launches) *launches += nl;
    return hipSuccess;
}

template <int WPR, int SRC>
static void launch_az_range(const AzArgs &a, long long first, long long end, hipStream_t st) {
    constexpr long long per_block = AZ_WAVES / WPR;
    const AccSrc<SRC> acc{a.acc, a.comp, a.acc64};
    // grid.x is limited to 2^31 - 1 workgroups; split larger result sets
    const long long maxn = (1LL << 30) * per_block;
    for (; first < end; first += maxn) {
        const long long n = (end - first) < maxn ? (end - first) : maxn;
        const unsigned blocks = (unsigned)((n + per_block - 1) / per_block);
        hipLaunchKernelGGL((azimuthal_kernel<WPR, SRC>), dim3(blocks), dim3(AZ_WAVES * 64), 0, st, acc,
                           a.elems_per_frame, a.scales, a.scale_count, a.q_off, a.q_cnt, a.idx, a.norm,
                           a.lambda_count, a.lag_count, a.out, first, n);
    }
}

template <int SRC>
static hipError_t launch_azimuthal_t(const AzArgs &a, hipStream_t st) {
    const long long split = a.wide_items < a.result_count ? a.wide_items : a.result_count;
    if (split > 0) launch_az_range<AZ_WAVES, SRC>(a, 0, split, st);
    if (split < a.result_count) {
        if (a.wave_list && a.narrow_waves > 0) {
            const long long blocks = (a.narrow_waves + AZ_WAVES - 1) / AZ_WAVES;
            if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
            const AccSrc<SRC> acc{a.acc, a.comp, a.acc64};
            hipLaunchKernelGGL((azimuthal_narrow_kernel<SRC>), dim3((unsigned)blocks), dim3(AZ_WAVES * 64), 0, st,
                               acc, a.elems_per_frame, a.scales, a.scale_count, a.narrow_first_scale, a.q_off,
                               a.q_cnt, a.idx, a.wave_list, a.norm, a.lambda_count, a.lag_count, a.out,
                               a.narrow_waves);
        } else {
            launch_az_range<1, SRC>(a, split, a.result_count, st);
        }
    }
    return hipGetLastError();
}

hipError_t launch_azimuthal(const AzArgs &a, hipStream_t st) {
    if (a.acc64 && !a.comp) return hipErrorInvalidValue;
    if (a.acc64) return launch_azimuthal_t<2>(a, st);
    if (a.comp) return launch_azimuthal_t<1>(a, st);
    return launch_azimuthal_t<0>(a, st);
}

hipError_t launch_acc_fold(float *acc, float *comp, double *acc64, long long n, hipStream_t st) {
    if (n % 4 != 0) return hipErrorInvalidValue;
    const long long n4 = n / 4;
    const long long want = (n4 + 255) / 256;
    const unsigned blocks = (unsigned)(want < 4096 ? (want > 0 ? want : 1) : 4096);
    hipLaunchKernelGGL(acc_fold_kernel, dim3(blocks), dim3(256), 0, st, (float4 *)acc, (float4 *)comp,
                       (double2 *)acc64, n4);
    return hipGetLastError();
}

}  // namespace ddm
