// Mixed-precision tile transform for 64 x 64 and 128 x 128 tiles: the 2-D
// R2C DFT as dense matrix products on the matrix cores (gfx950's
// v_mfma_f32_16x16x32 bf16 / f16: eight 16-bit values of A and of B per lane,
// A[i][8 (lane / 16) ..] with i = lane % 16, B[8 (lane / 16) ..][lane % 16],
// D[4 (lane / 16) + r][lane % 16] out), every operand split into a high and a low 16-bit part
// (a = a_h + a_l) and each product formed as a_h b_h + a_h b_l + a_l b_h with
// FP32 accumulation.  It replaces the FFT of those scales when an engine is
// switched to DDM_PREC_BF16X2 or DDM_PREC_FP16X2; the output layout is the
// same as the FP32 FFT's (the reference's cufftExecR2C, src/DDM.cu:188-192).
//
// Stage 1 (rows):    Y = X W1,  W1[x][c] = cos(2 pi x c / S)        c <  H
//                                        = (-1)^x                    c == H
//                                        = -sin(2 pi x (c - H) / S)  c >  H
//   so Y[:, :H] + i Y[:, H:] is the row half spectrum with the Nyquist value
//   parked in the (zero) imaginary part of column 0.
// Stage 2 (columns): Z = C2 Y + S2 Y',  C2 = cos(2 pi ky y / S), S2 = sin(..),
//   Y'[:, c] = Y[:, c + H] (c < H), -Y[:, c - H] (c >= H): Z[:, :H] + i Z[:, H:]
//   is the 2-D spectrum of the packed columns; column 0 (DC + i Nyquist) is
//   separated when written.
// Pixels are scaled by 2^-8 and Y by 2^-8 again before splitting (exact
// powers of two), which keeps every FP16 part in range; Z is scaled back by
// 2^16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ddm_device.h"
#include "ddm_kernels.h"

namespace ddm {

namespace {

typedef __bf16 b16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MF_NTH = 256;  // 4 waves

// 16-bit formats: BF16 (round to nearest even) or IEEE half
template <bool BF>
struct F16 {
    __device__ __host__ static uint16_t from(float v) {
        if constexpr (BF) {
            uint32_t u;
            memcpy(&u, &v, 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            return (uint16_t)(u >> 16);
        } else {
            _Float16 h = (_Float16)v;
            uint16_t r;
            memcpy(&r, &h, 2);
            return r;
        }
    }
    __device__ __host__ static float to(uint16_t b) {
        if constexpr (BF) {
            const uint32_t u = (uint32_t)b << 16;
            float v;
            memcpy(&v, &u, 4);
            return v;
        } else {
            _Float16 h;
            memcpy(&h, &b, 2);
            return (float)h;
        }
    }
};

// v = hi + lo with hi, lo 16-bit values (round to nearest even; gfx950
// converts two values per instruction), two values per 32-bit word.
template <bool BF>
__device__ __forceinline__ void split2(float a, float b, uint32_t *hi, uint32_t *lo) {
    if constexpr (BF) {
        const __bf16 ha = (__bf16)a, hb = (__bf16)b;
        const __bf16 la = (__bf16)(a - (float)ha), lb = (__bf16)(b - (float)hb);
        uint16_t u[4];
        memcpy(&u[0], &ha, 2);
        memcpy(&u[1], &hb, 2);
        memcpy(&u[2], &la, 2);
        memcpy(&u[3], &lb, 2);
        *hi = u[0] | ((uint32_t)u[1] << 16);
        *lo = u[2] | ((uint32_t)u[3] << 16);
    } else {
        const _Float16 ha = (_Float16)a, hb = (_Float16)b;
        const _Float16 la = (_Float16)(a - (float)ha), lb = (_Float16)(b - (float)hb);
        uint16_t u[4];
        memcpy(&u[0], &ha, 2);
        memcpy(&u[1], &hb, 2);
        memcpy(&u[2], &la, 2);
        memcpy(&u[3], &lb, 2);
        *hi = u[0] | ((uint32_t)u[1] << 16);
        *lo = u[2] | ((uint32_t)u[3] << 16);
    }
}

template <bool BF>
__device__ __forceinline__ f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
    if constexpr (BF) {
        b16x8 av, bv;
        memcpy(&av, &a, 16);
        memcpy(&bv, &b, 16);
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
    } else {
        h16x8 av, bv;
        memcpy(&av, &a, 16);
        memcpy(&bv, &b, 16);
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, c, 0, 0, 0);
    }
}

// a b with both split: a_h b_h + a_h b_l + a_l b_h
template <bool BF>
__device__ __forceinline__ f32x4 mfma3(uint4 ah, uint4 al, uint4 bh, uint4 bl, f32x4 c) {
    c = mfma<BF>(ah, bh, c);
    c = mfma<BF>(ah, bl, c);
    return mfma<BF>(al, bh, c);
}

__device__ __forceinline__ uint4 neg8(uint4 v) {
    return make_uint4(v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u, v.w ^ 0x80008000u);
}

template <int S>
struct MfCfg {
    static constexpr int H = S / 2;
    static constexpr int TPW = S == 128 ? 1 : 4;  // tiles per workgroup
    static constexpr int RT = S == 128 ? 2 : 4;   // 16-row tiles per wave
    static constexpr int CT = S / 16;             // 16-column tiles per wave
    static constexpr int KS = S / 32;             // k steps
    static constexpr int TILE = S * S;            // 16-bit values per matrix per tile
};

// LDS rows of X and Y^T: the 16-byte chunk c of row y is stored at chunk
// c ^ key(y), so that the 16 rows one operand read of a wave touches (lanes
// i = 0 .. 15, each reading eight values of its row) land in distinct banks
// (128 tiles: every row starts at bank 0; 64 tiles: every other row does).
template <int S>
__device__ __forceinline__ int chunk_key(int y) {
    return S == 128 ? (y & 15) : ((y >> 1) & 7);
}
template <int S>
__device__ __forceinline__ int swz(int y, int x) {  // x: a multiple of 4 below S
    return ((((x >> 3) ^ chunk_key<S>(y)) << 3) | (x & 7));
}

// Pixel loads of the transform: 8 consecutive pixels from 8- (u8) or
// 16-byte (u16) aligned rows, or one by one.
enum { MPIX_ANY = 0, MPIX_U8 = 1, MPIX_U16 = 2 };

}  // namespace

// mats: six S x S matrices of 16-bit values, each row-major:
// W1T_h, W1T_l ([c][x]), C2_h, C2_l, S2_h, S2_l ([ky][y]).
template <int S, bool BF, int PIX>
__global__ __launch_bounds__(MF_NTH) void tile_dft_mfma(PixSrc ps, RingDst rd, int tps, int tiles, int groups,
                                                       const uint16_t *__restrict__ mats) {
    using C = MfCfg<S>;
    constexpr int H = C::H, TPW = C::TPW, RT = C::RT, CT = C::CT, KS = C::KS, TILE = C::TILE;
    __shared__ __align__(16) uint16_t xs[2][TPW * TILE];  // hi / lo: X[y][x], later Y^T[c][y]
    __shared__ float2 zc[TPW][S];           // packed column 0 of the result
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nf = gridDim.x / groups;  // frames vary fastest
    const int f = blockIdx.x % nf;
    const int t = S == 128 ? 0 : w;  // this wave's tile within the workgroup
    const int tile = (blockIdx.x / nf) * TPW + t;
    const int r0 = S == 128 ? 32 * w : 0;  // first output row of this wave
    // operand k of this lane (8 values), output rows (4 values)
    const int li = lane & 15, lk = 8 * (lane >> 4), lo = 4 * (lane >> 4);
    const uint16_t *W1h = mats, *W1l = mats + TILE, *C2h = mats + 2 * TILE, *C2l = mats + 3 * TILE,
                   *S2h = mats + 4 * TILE, *S2l = mats + 5 * TILE;
    uint16_t *xh = xs[0], *xl = xs[1];

    // ---- pixels -> X (scaled by 2^-8, split), 8 pixels per item ----
    {
        const uint8_t *fr = ps.frames + (long long)f * ps.frame_bytes;
        for (int i = tid; i < TPW * TILE / 8; i += MF_NTH) {
            const int tt = i / (TILE / 8), y = (i / (S / 8)) % S, x = 8 * (i % (S / 8));
            const int tl = (blockIdx.x / nf) * TPW + tt;
            float v[8];
            if (tl < tiles) {
                const int gx = ps.x_off + (tl % tps) * S + x, gy = ps.y_off + (tl / tps) * S + y;
                const uint8_t *p = fr + (long long)gy * ps.row_bytes + (long long)gx * ps.bpp;
                if constexpr (PIX == MPIX_U8) {
                    const uint2 u = *reinterpret_cast<const uint2 *>(p);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        v[k] = (float)((u.x >> (8 * k)) & 0xffu);
                        v[4 + k] = (float)((u.y >> (8 * k)) & 0xffu);
                    }
                } else if constexpr (PIX == MPIX_U16) {
                    const uint4 u = *reinterpret_cast<const uint4 *>(p);
                    const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        v[2 * k] = (float)(wv[k] & 0xffffu);
                        v[2 * k + 1] = (float)(wv[k] >> 16);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 8; ++k) v[k] = load_px(fr, ps, gx + k, gy);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = 0.f;
            }
            uint32_t hv[4], lv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                split2<BF>(v[2 * k] * (1.0f / 256.0f), v[2 * k + 1] * (1.0f / 256.0f), &hv[k], &lv[k]);
            const int o = tt * TILE + y * S + swz<S>(y, x);
            *reinterpret_cast<uint4 *>(xh + o) = make_uint4(hv[0], hv[1], hv[2], hv[3]);
            *reinterpret_cast<uint4 *>(xl + o) = make_uint4(lv[0], lv[1], lv[2], lv[3]);
        }
    }
    __syncthreads();

    f32x4 acc[RT][CT];
    // ---- stage 1: Y = X W1 ----
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    {
        const uint16_t *xt_h = xh + t * TILE, *xt_l = xl + t * TILE;
        for (int ks = 0; ks < KS; ++ks) {
            const int k0 = 32 * ks + lk;
            uint4 ah[RT], al[RT];
#pragma unroll
            for (int a = 0; a < RT; ++a) {
                const int row = r0 + 16 * a + li;
                ah[a] = *reinterpret_cast<const uint4 *>(xt_h + row * S + swz<S>(row, k0));
                al[a] = *reinterpret_cast<const uint4 *>(xt_l + row * S + swz<S>(row, k0));
            }
#pragma unroll
            for (int b = 0; b < CT; ++b) {
                const int col = 16 * b + li;
                const uint4 bh = *reinterpret_cast<const uint4 *>(W1h + col * S + k0);
                const uint4 bl = *reinterpret_cast<const uint4 *>(W1l + col * S + k0);
#pragma unroll
                for (int a = 0; a < RT; ++a) acc[a][b] = mfma3<BF>(ah[a], al[a], bh, bl, acc[a][b]);
            }
        }
    }
    __syncthreads();  // X no longer needed
    // ---- Y (scaled by 2^-8, split) -> Y^T[c][y] ----
    {
        uint16_t *yt_h = xh + t * TILE, *yt_l = xl + t * TILE;
#pragma unroll
        for (int a = 0; a < RT; ++a)
#pragma unroll
            for (int b = 0; b < CT; ++b) {
                // lane holds Y[row .. row + 3][col]
                const int row = r0 + 16 * a + lo, col = 16 * b + li;
                uint2 ph, pl;
                split2<BF>(acc[a][b][0] * (1.0f / 256.0f), acc[a][b][1] * (1.0f / 256.0f), &ph.x, &pl.x);
                split2<BF>(acc[a][b][2] * (1.0f / 256.0f), acc[a][b][3] * (1.0f / 256.0f), &ph.y, &pl.y);
                *reinterpret_cast<uint2 *>(yt_h + col * S + swz<S>(col, row)) = ph;
                *reinterpret_cast<uint2 *>(yt_l + col * S + swz<S>(col, row)) = pl;
            }
    }
    __syncthreads();
    // ---- stage 2: Z = C2 Y + S2 Y' ----
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    {
        const uint16_t *yt_h = xh + t * TILE, *yt_l = xl + t * TILE;
        for (int ks = 0; ks < KS; ++ks) {
            const int k0 = 32 * ks + lk;
            uint4 ch[RT], cl[RT], sh[RT], sl[RT];
#pragma unroll
            for (int a = 0; a < RT; ++a) {
                const int row = r0 + 16 * a + li;
                ch[a] = *reinterpret_cast<const uint4 *>(C2h + row * S + k0);
                cl[a] = *reinterpret_cast<const uint4 *>(C2l + row * S + k0);
                sh[a] = *reinterpret_cast<const uint4 *>(S2h + row * S + k0);
                sl[a] = *reinterpret_cast<const uint4 *>(S2l + row * S + k0);
            }
#pragma unroll
            for (int b = 0; b < CT; ++b) {
                const int col = 16 * b + li;
                const uint4 yh = *reinterpret_cast<const uint4 *>(yt_h + col * S + swz<S>(col, k0));
                const uint4 yl = *reinterpret_cast<const uint4 *>(yt_l + col * S + swz<S>(col, k0));
                // Y'[:, col]: column col + H, or minus column col - H
                const int pc = col < H ? col + H : col - H;
                uint4 qh = *reinterpret_cast<const uint4 *>(yt_h + pc * S + swz<S>(pc, k0));
                uint4 ql = *reinterpret_cast<const uint4 *>(yt_l + pc * S + swz<S>(pc, k0));
                if (col >= H) {
                    qh = neg8(qh);
                    ql = neg8(ql);
                }
#pragma unroll
                for (int a = 0; a < RT; ++a) {
                    acc[a][b] = mfma3<BF>(ch[a], cl[a], yh, yl, acc[a][b]);
                    acc[a][b] = mfma3<BF>(sh[a], sl[a], qh, ql, acc[a][b]);
                }
            }
        }
    }
    // ---- write: Z[ky][k] = (acc[..][k], acc[..][k + H]) * 2^16 ----
    const bool valid = tile < tiles;
    constexpr float BACK = 65536.0f;
    if (li == 0) {
        // column 0 (packed) of rows ky: b = 0 (real part) and b = CT/2 (imaginary)
#pragma unroll
        for (int a = 0; a < RT; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                zc[t][r0 + 16 * a + lo + r] = make_float2(acc[a][0][r] * BACK, acc[a][CT / 2][r] * BACK);
    }
    __syncthreads();
    if (!valid) return;
    const TileOut out = tile_out(rd, f, tile, S);
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < CT / 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ky = r0 + 16 * a + lo + r, k = 16 * b + li;
                if (k != 0) out.put(ky, k, make_float2(acc[a][b][r] * BACK, acc[a][b + CT / 2][r] * BACK));
            }
    // separate the packed column: DC (kx = 0) and Nyquist (kx = H)
    for (int ky = lane; ky < S; ky += 64) {
        if (S == 128 && w != 0) break;  // one wave per tile writes it
        const float2 wv = zc[t][ky], pv = zc[t][(S - ky) & (S - 1)];
        out.put(ky, 0, make_float2(0.5f * (wv.x + pv.x), 0.5f * (wv.y - pv.y)));
        out.put(ky, H, make_float2(0.5f * (wv.y + pv.y), -0.5f * (wv.x - pv.x)));
    }
}

// ---------------------------------------------------------------------------
// Matrices per (scale, format), built once per device from double precision.
// ---------------------------------------------------------------------------
static std::mutex g_mat_mu;
static uint16_t *g_mats[64][2][2] = {};  // [device][scale 64/128][bf16/fp16]

template <bool BF>
static void build_mats(int S, std::vector<uint16_t> *m) {
    const int H = S / 2, T = S * S;
    m->assign(6 * (size_t)T, 0);
    auto put = [&](int which, int i, double v) {
        const float fv = (float)v;
        const uint16_t h = F16<BF>::from(fv);
        (*m)[(size_t)(2 * which) * T + i] = h;
        (*m)[(size_t)(2 * which + 1) * T + i] = F16<BF>::from(fv - F16<BF>::to(h));
    };
    const double w = 2.0 * M_PI / S;
    for (int c = 0; c < S; ++c)
        for (int x = 0; x < S; ++x) {
            double v;
            if (c < H)
                v = std::cos(w * (double)((long long)x * c % S));
            else if (c == H)
                v = (x & 1) ? -1.0 : 1.0;
            else
                v = -std::sin(w * (double)((long long)x * (c - H) % S));
            put(0, c * S + x, v);
        }
    for (int ky = 0; ky < S; ++ky)
        for (int y = 0; y < S; ++y) {
            const double a = w * (double)((long long)ky * y % S);
            put(1, ky * S + y, std::cos(a));
            put(2, ky * S + y, std::sin(a));
        }
}

static hipError_t mats_for(int S, bool bf, const uint16_t **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    const int si = S == 128 ? 1 : 0, fi = bf ? 0 : 1;
    std::lock_guard<std::mutex> lk(g_mat_mu);
    if (!g_mats[dev][si][fi]) {
        std::vector<uint16_t> h;
        if (bf)
            build_mats<true>(S, &h);
        else
            build_mats<false>(S, &h);
        uint16_t *d = nullptr;
        e = hipMalloc(&d, h.size() * sizeof(uint16_t));
        if (e != hipSuccess) return e;
        e = hipMemcpy(d, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return e;
        }
        g_mats[dev][si][fi] = d;
    }
    *out = g_mats[dev][si][fi];
    return hipSuccess;
}

template <int S, bool BF>
static hipError_t launch_mf(const PixSrc &ps, const RingDst &rd, int tps, int tiles, int n_frames,
                            hipStream_t st) {
    const uint16_t *mats = nullptr;
    hipError_t e = mats_for(S, BF, &mats);
    if (e != hipSuccess) return e;
    constexpr int TPW = MfCfg<S>::TPW;
    const int groups = (tiles + TPW - 1) / TPW;
    const dim3 grid((unsigned)(groups * n_frames)), block(MF_NTH);
    // 8-pixel loads need every row piece aligned to 8 pixels' bytes
    const long long bits = ps.row_bytes | ps.frame_bytes | (long long)(uintptr_t)ps.frames | (long long)ps.x_off * ps.bpp;
    if (ps.fmt == DDM_PIX_U8 && ps.bpp == 1 && (bits & 7) == 0)
        hipLaunchKernelGGL((tile_dft_mfma<S, BF, MPIX_U8>), grid, block, 0, st, ps, rd, tps, tiles, groups, mats);
    else if (ps.fmt == DDM_PIX_U16LE && ps.bpp == 2 && (bits & 15) == 0)
        hipLaunchKernelGGL((tile_dft_mfma<S, BF, MPIX_U16>), grid, block, 0, st, ps, rd, tps, tiles, groups, mats);
    else
        hipLaunchKernelGGL((tile_dft_mfma<S, BF, MPIX_ANY>), grid, block, 0, st, ps, rd, tps, tiles, groups, mats);
    return hipGetLastError();
}

hipError_t launch_tile_dft_mfma(const PixSrc &ps, const RingDst &rd, int scale, int tps, int tiles, int n_frames,
                                int precision, hipStream_t st) {
    const bool bf = precision == DDM_PREC_BF16X2;
    if (scale == 128) return bf ? launch_mf<128, true>(ps, rd, tps, tiles, n_frames, st)
                                : launch_mf<128, false>(ps, rd, tps, tiles, n_frames, st);
    if (scale == 64) return bf ? launch_mf<64, true>(ps, rd, tps, tiles, n_frames, st)
                               : launch_mf<64, false>(ps, rd, tps, tiles, n_frames, st);
    return hipErrorInvalidValue;
}

}  // namespace ddm
