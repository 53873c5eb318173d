// Register-resident 2-D R2C FFT of 64 x 64 and 128 x 128 tiles (tile cut +
// FFT, the parseBufferScalePow2 + cufftExecR2C step of the reference,
// src/DDM_kernel.cuh:43-78 and src/DDM.cu:188-192).
//
// Same packing as the LDS kernels of ddm_kernels.hip: the real S x S tile is
// read as H = S/2 complex rows z_r = x_{2r} + i x_{2r+1}; after the row FFTs
// the two real rows are separated, X_a = (Z_k + conj Z_{S-k}) / 2 and
// X_b = -i (Z_k - conj Z_{S-k}) / 2, and columns kx = 0 and kx = H travel
// packed as one complex column (DC + i Nyquist), leaving H complex columns of
// length S that are separated again when the result is written.
//
// Every 1-D FFT of length S = 16 L runs on L lanes in two register phases:
//   phase 1: lane q holds x[16 elements at stride L, offset q], a radix-16
//            DFT in registers, then the twiddles W_S^(q k1);
//   exchange: the 16 x L values of the sequence go through LDS (a region
//            private to the wave, so no workgroup barrier);
//   phase 2: each lane takes the 16 / L values of k1 closed under
//            k1 -> 16 - k1 and runs length-L DFTs, so the spectrum values at
//            k and S - k, which the real-pair separation combines, are in the
//            same lane.
// The row pass writes the separated half spectra into the tile buffer in
// LDS (one transpose), the column pass reads them back and writes the
// spectrum straight to the ring.  LDS addresses are swizzled so that every
// group of 16 lanes touches 16 distinct 8-byte bank slots in each access.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ddm_batch.h"
#include "ddm_device.h"
#include "ddm_kernels.h"

namespace ddm {

// Stores of spectra and pass-buffer columns that no workgroup of the same
// launch reads again: nontemporal (streamed through the caches), measured
// 2 % faster for the 256 .. 1024 transforms than ordinary stores.
typedef float nt_f32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) nt_f32x2 g_nt_f32x2;
__device__ __forceinline__ void store_nt(float2 *p, float2 v) {
    __builtin_nontemporal_store(nt_f32x2{v.x, v.y}, (g_nt_f32x2 *)p);
}

namespace {

constexpr int REG_NTH = 512;  // 8 waves
constexpr int REG_WAVE_SLOTS = 1024;  // complex values of LDS per wave

// log2 of a power of two
__host__ __device__ constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

constexpr float C16 = 0.92387953251128674f;  // cos(pi/8)
constexpr float S16 = 0.38268343236508977f;  // sin(pi/8)
constexpr float R2 = 0.70710678118654752f;   // sqrt(2)/2

// v * (c + i s), fused multiply-adds written out as in cmul
__device__ __forceinline__ float2 mulw(float2 v, float c, float s) {
    return make_float2(__builtin_fmaf(v.x, c, -(v.y * s)), __builtin_fmaf(v.x, s, v.y * c));
}
__device__ __forceinline__ float2 mul_mi(float2 v) { return make_float2(v.y, -v.x); }  // v * (-i)

// In-place DFT of four values, natural order.
__device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d) {
    const float2 s02 = cadd(a, c), d02 = csub(a, c), s13 = cadd(b, d), d13 = csub(b, d);
    a = cadd(s02, s13);
    c = csub(s02, s13);
    b = make_float2(d02.x + d13.y, d02.y - d13.x);  // d02 - i d13
    d = make_float2(d02.x - d13.y, d02.y + d13.x);  // d02 + i d13
}

// X[k] = sum_n v[n] exp(-2 pi i n k / 16), natural order in and out.
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // v[n2 + 4 k1] *= W16^(n2 k1)
    v[5] = mulw(v[5], C16, -S16);
    v[9] = mulw(v[9], R2, -R2);
    v[13] = mulw(v[13], S16, -C16);
    v[6] = mulw(v[6], R2, -R2);
    v[10] = mul_mi(v[10]);
    v[14] = mulw(v[14], -R2, -R2);
    v[7] = mulw(v[7], S16, -C16);
    v[11] = mulw(v[11], -R2, -R2);
    v[15] = mulw(v[15], -C16, S16);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    // v[4 k1 + k2] = X[k1 + 4 k2]
    float2 t[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) t[k] = v[4 * (k % 4) + k / 4];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = t[k];
}

__device__ __forceinline__ void dft8(float2 (&u)[8]) {
    float2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
    float2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = mulw(o1, R2, -R2);
    o2 = mul_mi(o2);
    o3 = mulw(o3, -R2, -R2);
    u[0] = cadd(e0, o0);
    u[4] = csub(e0, o0);
    u[1] = cadd(e1, o1);
    u[5] = csub(e1, o1);
    u[2] = cadd(e2, o2);
    u[6] = csub(e2, o2);
    u[3] = cadd(e3, o3);
    u[7] = csub(e3, o3);
}

// cos and sin of 2 pi k / 32, k = 0..7
constexpr float W32C[8] = {1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
                                    0.70710678118654757f, 0.55557023301960229f, 0.38268343236508984f,
                                    0.19509032201612833f};

// X[k] = sum_n v[n] exp(-2 pi i n k / 32): radix-2 split into two length-16
// DFTs, natural order in and out.
__device__ __forceinline__ void dft32(float2 (&v)[32]) {
    float2 e[16], o[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft16(e);
    dft16(o);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        float2 t;
        if (k == 0) {
            t = o[0];
        } else if (k == 8) {
            t = mul_mi(o[8]);
        } else if (k < 8) {
            t = mulw(o[k], W32C[k], -W32C[8 - k]);  // W32^k = cos - i sin, sin(2 pi k/32) = cos(2 pi (8-k)/32)
        } else {
            t = mulw(o[k], -W32C[16 - k], -W32C[k - 8]);  // k = 8 + m: W32^k = -i W32^m
        }
        v[k] = cadd(e[k], t);
        v[k + 16] = csub(e[k], t);
    }
}

template <int L>
__device__ __forceinline__ void dftL(float2 (&u)[L]);
template <>
__device__ __forceinline__ void dftL<8>(float2 (&u)[8]) {
    dft8(u);
}
template <>
__device__ __forceinline__ void dftL<4>(float2 (&u)[4]) {
    dft4(u[0], u[1], u[2], u[3]);
}

template <int S>
struct RegCfg {
    static constexpr int H = S / 2;
    static constexpr int L = S / 16;          // lanes per sequence
    static constexpr int P = 16 / L;          // k1 values per lane in phase 2
    static constexpr int Q = 64 / L;          // sequences per wave
    static constexpr int WPT = H / Q;         // waves per tile
    static constexpr int TPW = (REG_NTH / 64) / WPT;  // tiles per workgroup
    static constexpr int TILE = S * H;        // complex values per tile in LDS
    static_assert(WPT * REG_WAVE_SLOTS == TILE, "a wave's rows are its exchange region");
};

// The k1 values lane j of a sequence takes in phase 2 (p = 0 .. P-1); each
// set is closed under k1 -> (16 - k1) mod 16.
template <int L>
__device__ __forceinline__ int kappa(int j, int p) {
    if (L == 8) return p == 0 ? j : (j == 0 ? 8 : 16 - j);
    // L == 4: {j, 16 - j, 8 - j, 8 + j}, lane 0: {0, 8, 4, 12}
    switch (p) {
    case 0: return j;
    case 1: return j == 0 ? 8 : 16 - j;
    case 2: return 8 - j - (j == 0 ? 4 : 0);
    default: return 8 + j + (j == 0 ? 4 : 0);
    }
}

// Exchange slot of value (sequence sq, k1, n2) inside a wave's region.
template <int L>
__device__ __forceinline__ int xslot(int sq, int k1, int n2) {
    constexpr int P = 16 / L;
    return sq * 16 * L + (k1 ^ (sq & (P - 1))) * L + ((n2 + k1) & (L - 1));
}

// Column swizzle of tile row y.
__device__ __forceinline__ int tswz(int y) { return ((y & 1) * 6) ^ (((y >> 1) & 1) * 8) ^ (((y >> 2) & 1) * 4); }

// Register of the partner value Z[S - k] for the value at (p, k2): returns
// (p', k2') for lanes j != 0 in (*pa, *ka) and for lane 0 in (*pb, *kb).
template <int L>
__device__ __forceinline__ void partner(int p, int k2, int *pa, int *ka, int *pb, int *kb) {
    *ka = L - 1 - k2;
    *pa = (L == 8) ? (1 - p) : (p ^ 1);
    if (p == 0) {
        *pb = 0;
        *kb = (L - k2) & (L - 1);
    } else if (L == 8 || p == 1) {
        *pb = p;
        *kb = L - 1 - k2;
    } else {
        *pb = p ^ 1;  // L == 4: 4 <-> 12
        *kb = L - 1 - k2;
    }
}

// zp[p][k2] = the value of z at S - k for k = kappa(q, p) + 16 k2.  Lane 0
// of a sequence pairs its registers differently from the others, so the two
// cases are separate branches (a select between array elements would send
// z to scratch memory).
template <int L, int P>
__device__ __forceinline__ void partners(const float2 (&z)[P][L], bool lane0, float2 (&zp)[P][L]) {
    if (lane0) {
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int k2 = 0; k2 < L; ++k2) {
                int pa, ka, pb, kb;
                partner<L>(p, k2, &pa, &ka, &pb, &kb);
                zp[p][k2] = z[pb][kb];
            }
    } else {
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int k2 = 0; k2 < L; ++k2) {
                int pa, ka, pb, kb;
                partner<L>(p, k2, &pa, &ka, &pb, &kb);
                zp[p][k2] = z[pa][ka];
            }
    }
}

// Pixel access modes of the register FFT kernels (see pixel_mode).
enum { PIX_ANY = 0, PIX_U8_1 = 1, PIX_U16LE_2 = 2 };

// The u16 mode needs every pixel 2-byte aligned.
__host__ __device__ inline int pixel_mode(const PixSrc &ps) {
    if (ps.fmt == DDM_PIX_U8 && ps.bpp == 1) return PIX_U8_1;
    if (ps.fmt == DDM_PIX_U16LE && ps.bpp == 2 && ((ps.row_bytes | ps.frame_bytes | (long long)(uintptr_t)ps.frames) & 1) == 0)
        return PIX_U16LE_2;
    return PIX_ANY;
}

// Pixels (x0 + n * STEP, y0) and (x0 + n * STEP, y0 + 1), n < N, as complex
// values x_{2r} + i x_{2r+1}.
template <int PIX, int N, int STEP>
__device__ __forceinline__ void load_row_pair(const uint8_t *fr, const PixSrc &ps, int x0, int y0, float2 (&v)[N]) {
    if constexpr (PIX == PIX_U8_1) {
        const uint8_t *r0 = fr + (long long)y0 * ps.row_bytes + x0;
        const uint8_t *r1 = r0 + ps.row_bytes;
#pragma unroll
        for (int n = 0; n < N; ++n) v[n] = make_float2((float)r0[n * STEP], (float)r1[n * STEP]);
    } else if constexpr (PIX == PIX_U16LE_2) {
        const uint16_t *r0 = reinterpret_cast<const uint16_t *>(fr + (long long)y0 * ps.row_bytes) + x0;
        const uint16_t *r1 = reinterpret_cast<const uint16_t *>(reinterpret_cast<const uint8_t *>(r0) + ps.row_bytes);
#pragma unroll
        for (int n = 0; n < N; ++n) v[n] = make_float2((float)r0[n * STEP], (float)r1[n * STEP]);
    } else {
#pragma unroll
        for (int n = 0; n < N; ++n)
            v[n] = make_float2(load_px(fr, ps, x0 + n * STEP, y0), load_px(fr, ps, x0 + n * STEP, y0 + 1));
    }
}

// S consecutive pixels of row y from x0: with A4 every row start is 4-byte
// aligned, so 1- and 2-byte pixels come in 4-byte words.
template <int PIX, int S, bool A4>
__device__ __forceinline__ void load_row(const uint8_t *fr, const PixSrc &ps, int x0, int y, float (&out)[S]) {
    const uint8_t *row = fr + (long long)y * ps.row_bytes;
    if constexpr (PIX == PIX_U8_1 && A4 && S % 4 == 0) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(row + x0);
#pragma unroll
        for (int i = 0; i < S / 4; ++i) {
            const uint32_t u = w[i];
            out[4 * i] = (float)(u & 0xff);
            out[4 * i + 1] = (float)((u >> 8) & 0xff);
            out[4 * i + 2] = (float)((u >> 16) & 0xff);
            out[4 * i + 3] = (float)(u >> 24);
        }
    } else if constexpr (PIX == PIX_U16LE_2 && A4 && S % 2 == 0) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(row + 2 * x0);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) {
            const uint32_t u = w[i];
            out[2 * i] = (float)(u & 0xffff);
            out[2 * i + 1] = (float)(u >> 16);
        }
    } else if constexpr (PIX == PIX_U8_1) {
#pragma unroll
        for (int i = 0; i < S; ++i) out[i] = (float)row[x0 + i];
    } else if constexpr (PIX == PIX_U16LE_2) {
        const uint16_t *h = reinterpret_cast<const uint16_t *>(row) + x0;
#pragma unroll
        for (int i = 0; i < S; ++i) out[i] = (float)h[i];
    } else {
#pragma unroll
        for (int i = 0; i < S; ++i) out[i] = load_px(fr, ps, x0 + i, y);
    }
}

// Pixel sources of the transform bodies; (x, y) are coordinates in the
// region of interest.
// A frame in device memory (PIX as in load_row_pair / load_row).
template <int PIX>
struct FramePix {
    const uint8_t *fr;
    PixSrc ps;
    // pixels (x + n STEP, y) and (x + n STEP, y + 1) as x_{2r} + i x_{2r+1}
    template <int N, int STEP>
    __device__ __forceinline__ void row_pair(int x, int y, float2 (&v)[N]) const {
        load_row_pair<PIX, N, STEP>(fr, ps, ps.x_off + x, ps.y_off + y, v);
    }
    // S consecutive pixels of row y from x
    template <int S, bool A4>
    __device__ __forceinline__ void row(int x, int y, float (&out)[S]) const {
        load_row<PIX, S, A4>(fr, ps, ps.x_off + x, ps.y_off + y, out);
    }
};

// A 128 x 128 block of 16-bit pixel values staged in LDS (multi-scale
// tiler): row pitch 132 values (66 dwords), so that the rows 2r that the
// lanes of one LDS access read fall in distinct banks for every body's
// access pattern; (ox, oy) is the block's origin.
constexpr int BLK = 128;
constexpr int BLK_PITCH = 132;
struct BlockPix {
    const uint16_t *blk;
    int ox, oy;
    template <int N, int STEP>
    __device__ __forceinline__ void row_pair(int x, int y, float2 (&v)[N]) const {
        const uint16_t *r0 = blk + (y - oy) * BLK_PITCH + (x - ox);
#pragma unroll
        for (int n = 0; n < N; ++n) v[n] = make_float2((float)r0[n * STEP], (float)r0[BLK_PITCH + n * STEP]);
    }
    // x - ox is a multiple of S >= 8, so every row piece is 8-byte aligned
    template <int S, bool A4>
    __device__ __forceinline__ void row(int x, int y, float (&out)[S]) const {
        const uint2 *r = reinterpret_cast<const uint2 *>(blk + (y - oy) * BLK_PITCH + (x - ox));
#pragma unroll
        for (int i = 0; i < S / 4; ++i) {
            const uint2 u = r[i];
            out[4 * i] = (float)(u.x & 0xffff);
            out[4 * i + 1] = (float)(u.x >> 16);
            out[4 * i + 2] = (float)(u.y & 0xffff);
            out[4 * i + 3] = (float)(u.y >> 16);
        }
    }
};

}  // namespace

// 64 x 64 and 128 x 128 tiles: the TPW tiles of ts on REG_NTH threads (tid),
// pixels from src, for batch frame f.  lds holds TPW * S * (S/2 + 1)
// complex values (row/column transpose and exchange regions; at the end
// the tiles' spectra in ring element order), rtab S + 1 ints.
template <int S, class Src>
__device__ __forceinline__ void reg_tiles(const Src &src, const RingDst &rd, int tps, const TileSet &ts,
                                          const float2 *__restrict__ twg, float2 *lds, int *rtab, int tid, int f) {
    using C = RegCfg<S>;
    constexpr int H = C::H, L = C::L, P = C::P, Q = C::Q, WPT = C::WPT, TILE = C::TILE;
    constexpr int TE = S * (H + 1);  // spectrum elements per tile
    const int lane = tid & 63;
    const int w = tid >> 6;
    load_row_starts<S>(rd, rtab, tid, REG_NTH);
    const int t = w / WPT, wit = w % WPT;
    const int tile = ts.tile(t, tps);
    const bool valid = t < ts.ntile;
    const int q = lane % L, sq = lane / L;  // lane within the sequence, sequence within the wave
    float2 *reg = lds + w * REG_WAVE_SLOTS;  // this wave's region
    float2 *tb = lds + t * TILE;             // this wave's tile

    // twiddles W_S^(q k1), k1 = 1..15 (the same in both passes)
    float2 tw[16];
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) tw[k1] = twg[S + __mul24(q, k1)];

    float2 v[16];
    // ---------------- rows: sequence r = row pair ----------------
    {
        const int r = wit * Q + sq;
        if (valid) {
            const int x0 = (tile % tps) * S + q;
            const int y0 = (tile / tps) * S + 2 * r;
            src.template row_pair<16, L>(x0, y0, v);
        } else {
#pragma unroll
            for (int n1 = 0; n1 < 16; ++n1) v[n1] = make_float2(0.f, 0.f);
        }
        dft16(v);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul(v[k1], tw[k1]);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) reg[xslot<L>(sq, k1, q)] = v[k1];
        // phase 2: lane j = q takes k1 in kappa(j, .)
        float2 z[P][L];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k1 = kappa<L>(q, p);
#pragma unroll
            for (int n2 = 0; n2 < L; ++n2) z[p][n2] = reg[xslot<L>(sq, k1, n2)];
            dftL<L>(z[p]);  // z[p][k2] = Z[k1 + 16 k2]
        }
        // separate the real row pair; write rows 2r, 2r+1, columns k < H
        const int ya = 2 * r, yb = 2 * r + 1;
        const int sa = tswz(ya), sb = tswz(yb);
        float2 zq[P][L];
        partners<L, P>(z, q == 0, zq);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k1 = kappa<L>(q, p);
#pragma unroll
            for (int k2 = 0; k2 < H / 16; ++k2) {
                const float2 zk = z[p][k2];
                const float2 zp = zq[p][k2];
                float2 xa = make_float2(0.5f * (zk.x + zp.x), 0.5f * (zk.y - zp.y));
                float2 xb = make_float2(0.5f * (zk.y + zp.y), -0.5f * (zk.x - zp.x));
                if (p == 0 && k2 == 0 && q == 0) {
                    const float2 zh = z[0][H / 16];  // Z[H]: DC and Nyquist of both rows are real
                    xa = make_float2(zk.x, zh.x);
                    xb = make_float2(zk.y, zh.y);
                }
                const int c = k1 + 16 * k2;
                tb[ya * H + (c ^ sa)] = xa;
                tb[yb * H + (c ^ sb)] = xb;
            }
        }
    }
    __syncthreads();
    // ---------------- columns: sequence c = packed column ----------------
    const int c = wit * Q + sq;
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
        const int y = n1 * L + q;
        v[n1] = tb[y * H + (c ^ tswz(y))];
    }
    dft16(v);
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul(v[k1], tw[k1]);
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) reg[xslot<L>(sq, k1, q)] = v[k1];
    float2 z[P][L];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k1 = kappa<L>(q, p);
#pragma unroll
        for (int n2 = 0; n2 < L; ++n2) z[p][n2] = reg[xslot<L>(sq, k1, n2)];
        dftL<L>(z[p]);  // z[p][k2] = V[k1 + 16 k2], ky = k1 + 16 k2
    }
    // the tile's spectrum [ky][kx] into LDS, then out to the ring in element
    // order: whole 512-byte element groups per wave store
    __syncthreads();
    float2 *ob = lds + t * TE;
    if (c != 0) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k1 = kappa<L>(q, p);
#pragma unroll
            for (int k2 = 0; k2 < L; ++k2) ob[(k1 + 16 * k2) * (H + 1) + c] = z[p][k2];
        }
    } else {
        // packed column: DC (kx = 0) and Nyquist (kx = H) of every row ky
        float2 zq[P][L];
        partners<L, P>(z, q == 0, zq);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k1 = kappa<L>(q, p);
#pragma unroll
            for (int k2 = 0; k2 < L; ++k2) {
                const int row = (k1 + 16 * k2) * (H + 1);
                const float2 wv = z[p][k2], pv = zq[p][k2];  // V[ky], V[S - ky]
                ob[row] = make_float2(0.5f * (wv.x + pv.x), 0.5f * (wv.y - pv.y));
                ob[row + H] = make_float2(0.5f * (wv.y + pv.y), -0.5f * (wv.x - pv.x));
            }
        }
    }
    __syncthreads();
    store_tiles<S>(rd, lds, rtab, f, ts, tps, tid, REG_NTH);
}

// A per-scale launch: workgroup bid of nblk takes TPW consecutive tiles of
// one frame.  PIX: PIX_U8_1 (bpp 1, DDM_PIX_U8) and PIX_U16LE_2 (bpp 2,
// DDM_PIX_U16LE) load pixels with constant offsets from two row pointers;
// PIX_ANY sends every pixel through load_px.
template <int S, int PIX>
__device__ __forceinline__ void tile_fft_reg_body(const PixSrc &ps, const RingDst &rd, int tps, int tiles,
                                                  int groups, const float2 *__restrict__ twg, int bid, int nblk) {
    constexpr int TPW = RegCfg<S>::TPW;
    __shared__ float2 lds[TPW * S * (S / 2 + 1)];
    __shared__ int rtab[S + 1];  // row starts of a LAY_ROWS layout
    // frames vary fastest across the grid: concurrent workgroups write
    // neighbouring ring slots of the same element groups
    const int nf = nblk / groups;
    const int f = bid % nf;
    const int first = (bid / nf) * TPW;  // the workgroup's first tile
    const FramePix<PIX> src{ps.frames + (long long)f * ps.frame_bytes, ps};
    reg_tiles<S>(src, rd, tps, TileSet{first, ilog2(TPW), min(TPW, tiles - first)}, twg, lds,
                 rtab, threadIdx.x, f);
}

template <int S, int PIX>
__global__ __launch_bounds__(REG_NTH) void tile_fft_reg(PixSrc ps, RingDst rd, int tps, int tiles, int groups,
                                                       const float2 *__restrict__ twg) {
    tile_fft_reg_body<S, PIX>(ps, rd, tps, tiles, groups, twg, blockIdx.x, gridDim.x);
}

template <int S>
static void launch_reg(const PixSrc &ps, const RingDst &rd, int tps, int tiles, int n_frames, const float2 *tw,
                       hipStream_t st) {
    constexpr int TPW = RegCfg<S>::TPW;
    const int groups = (tiles + TPW - 1) / TPW;
    const dim3 grid((unsigned)(groups * n_frames)), block(REG_NTH);
    switch (pixel_mode(ps)) {
    case PIX_U8_1:
        hipLaunchKernelGGL((tile_fft_reg<S, PIX_U8_1>), grid, block, 0, st, ps, rd, tps, tiles, groups, tw);
        break;
    case PIX_U16LE_2:
        hipLaunchKernelGGL((tile_fft_reg<S, PIX_U16LE_2>), grid, block, 0, st, ps, rd, tps, tiles, groups, tw);
        break;
    default:
        hipLaunchKernelGGL((tile_fft_reg<S, PIX_ANY>), grid, block, 0, st, ps, rd, tps, tiles, groups, tw);
    }
}

// ---------------------------------------------------------------------------
// 8, 16 and 32 tiles: one lane per row pair, then one lane per packed column.
// A tile takes R = S/2 lanes, a wave 64/R tiles, a workgroup SMALL_NTH/R.
// Row phase: lane r holds z[n] = x_{2r}[n] + i x_{2r+1}[n] and does the
// whole length-S DFT in registers, separates the two real rows' spectra and
// writes columns 0 .. R-1 of rows 2r, 2r+1 (column 0 = DC + i Nyquist) to the
// tile's LDS block.  Column phase: lane c reads packed column c, transforms
// it in registers and puts it back (lane 0 separates the packed column into
// the DC and Nyquist columns); the workgroup's spectra then go to the ring
// in element order.
// ---------------------------------------------------------------------------
constexpr int SMALL_NTH = 256;

template <int S>
__device__ __forceinline__ void dftS(float2 (&v)[S]);
template <>
__device__ __forceinline__ void dftS<8>(float2 (&v)[8]) {
    dft8(v);
}
template <>
__device__ __forceinline__ void dftS<16>(float2 (&v)[16]) {
    dft16(v);
}
template <>
__device__ __forceinline__ void dftS<32>(float2 (&v)[32]) {
    dft32(v);
}

// 8, 16 and 32 tiles: the tiles of ts (at most NTH / (S/2)) on threads
// tid < NTH of a workgroup of WG threads (WG >= NTH: the others only meet
// the barriers), pixels from src, for batch frame f.  lds holds
// NTH / (S/2) * S * (S/2 + 1) complex values, rtab S + 1 ints.
template <int S, bool A4, int NTH, int WG, class Src>
__device__ __forceinline__ void lane_tiles(const Src &src, const RingDst &rd, int tps, const TileSet &ts,
                                           float2 *lds, int *rtab, int tid, int f) {
    constexpr int R = S / 2;   // lanes per tile
    constexpr int LS = R + 1;  // LDS row stride (complex values)
    const int lt = tid / R, r = tid % R;
    const bool act = WG == NTH || tid < NTH;
    load_row_starts<S>(rd, rtab, tid, WG);
    const int tile = ts.tile(lt, tps);
    const bool live = act && lt < ts.ntile;
    float2 *blk = lds + lt * S * LS;
    if (live) {
        const int x0 = (tile % tps) * S;
        const int y0 = (tile / tps) * S + 2 * r;
        float2 z[S];
        {
            float a[S], b[S];
            src.template row<S, A4>(x0, y0, a);
            src.template row<S, A4>(x0, y0 + 1, b);
#pragma unroll
            for (int n = 0; n < S; ++n) z[n] = make_float2(a[n], b[n]);
        }
        dftS<S>(z);
        // rows 2r (x_a) and 2r+1 (x_b): X_a = (Z_k + conj Z_{S-k}) / 2,
        // X_b = -i (Z_k - conj Z_{S-k}) / 2; column 0 packs DC + i Nyquist
        blk[(2 * r) * LS] = make_float2(z[0].x, z[R].x);
        blk[(2 * r + 1) * LS] = make_float2(z[0].y, z[R].y);
#pragma unroll
        for (int k = 1; k < R; ++k) {
            const float2 zk = z[k], zp = z[S - k];
            blk[(2 * r) * LS + k] = make_float2(0.5f * (zk.x + zp.x), 0.5f * (zk.y - zp.y));
            blk[(2 * r + 1) * LS + k] = make_float2(0.5f * (zk.y + zp.y), -0.5f * (zk.x - zp.x));
        }
    }
    __syncthreads();
    // column c = r of the tile
    float2 v[S];
    if (act) {
#pragma unroll
        for (int ky = 0; ky < S; ++ky) v[ky] = blk[ky * LS + r];
        dftS<S>(v);
    }
    __syncthreads();
    // the spectrum [ky][kx <= R] back into the tile's block (LS = R + 1 is
    // its row length), then out to the ring in element order
    if (act) {
        if (r != 0) {
#pragma unroll
            for (int ky = 0; ky < S; ++ky) blk[ky * LS + r] = v[ky];
        } else {
            // packed column: DC (kx = 0) and Nyquist (kx = R) columns
#pragma unroll
            for (int ky = 0; ky < S; ++ky) {
                const float2 w = v[ky], pv = v[(S - ky) & (S - 1)];
                blk[ky * LS] = make_float2(0.5f * (w.x + pv.x), 0.5f * (w.y - pv.y));
                blk[ky * LS + R] = make_float2(0.5f * (w.y + pv.y), -0.5f * (w.x - pv.x));
            }
        }
    }
    __syncthreads();
    if (act) store_tiles<S>(rd, lds, rtab, f, ts, tps, tid, NTH);
}

// A per-scale launch: workgroup bid of nblk takes TPG consecutive tiles of
// one frame.
template <int S, int PIX, bool A4>
__device__ __forceinline__ void tile_fft_lane_body(const PixSrc &ps, const RingDst &rd, int tps, int tiles,
                                                   int groups, int bid, int nblk) {
    constexpr int R = S / 2, TPG = SMALL_NTH / R;  // lanes per tile, tiles per workgroup
    __shared__ float2 lds[TPG * S * (R + 1)];
    __shared__ int rtab[S + 1];  // row starts of a LAY_ROWS layout
    const int nf = nblk / groups;  // frames vary fastest
    const int f = bid % nf;
    const int first = (bid / nf) * TPG;  // the workgroup's tiles are consecutive
    const FramePix<PIX> src{ps.frames + (long long)f * ps.frame_bytes, ps};
    lane_tiles<S, A4, SMALL_NTH, SMALL_NTH>(src, rd, tps, TileSet{first, ilog2(TPG), min(TPG, tiles - first)}, lds,
                                            rtab, threadIdx.x, f);
}

template <int S, int PIX, bool A4>
__global__ __launch_bounds__(SMALL_NTH) void tile_fft_lane(PixSrc ps, RingDst rd, int tps, int tiles, int groups) {
    tile_fft_lane_body<S, PIX, A4>(ps, rd, tps, tiles, groups, blockIdx.x, gridDim.x);
}

template <int S>
static void launch_lane(const PixSrc &ps, const RingDst &rd, int tps, int tiles, int n_frames, hipStream_t st) {
    constexpr int TPG = SMALL_NTH / (S / 2);
    const int groups = (tiles + TPG - 1) / TPG;
    const dim3 grid((unsigned)(groups * n_frames)), block(SMALL_NTH);
    const int mode = pixel_mode(ps);
    // every row start 4-byte aligned: frames, frame stride, row stride and the ROI's first byte
    const bool a4 = ((ps.row_bytes | ps.frame_bytes | (long long)(uintptr_t)ps.frames |
                      (long long)ps.x_off * ps.bpp) & 3) == 0;
    if (mode == PIX_U8_1 && a4)
        hipLaunchKernelGGL((tile_fft_lane<S, PIX_U8_1, true>), grid, block, 0, st, ps, rd, tps, tiles, groups);
    else if (mode == PIX_U8_1)
        hipLaunchKernelGGL((tile_fft_lane<S, PIX_U8_1, false>), grid, block, 0, st, ps, rd, tps, tiles, groups);
    else if (mode == PIX_U16LE_2 && a4)
        hipLaunchKernelGGL((tile_fft_lane<S, PIX_U16LE_2, true>), grid, block, 0, st, ps, rd, tps, tiles, groups);
    else if (mode == PIX_U16LE_2)
        hipLaunchKernelGGL((tile_fft_lane<S, PIX_U16LE_2, false>), grid, block, 0, st, ps, rd, tps, tiles, groups);
    else
        hipLaunchKernelGGL((tile_fft_lane<S, PIX_ANY, false>), grid, block, 0, st, ps, rd, tps, tiles, groups);
}

// ---------------------------------------------------------------------------
// 256, 512 and 1024 tiles: rows and columns in two kernels, the row half
// spectra parked in the batch's pass buffer in between.  Sequences of length
// S = E1 x L on L lanes (L = 16, or 32 for S = 1024): phase 1 a radix-E1
// DFT of each lane's E1 values at stride L (radix-16 or radix-32 in
// registers) and the twiddles W_S^(q k1); phase 2 after an exchange through
// LDS private to the sequence: E1/L length-L DFTs per lane.  The real-pair
// separation and the packed DC/Nyquist column go through LDS.
// ---------------------------------------------------------------------------
template <int S>
struct BigCfg {
    static constexpr int H = S / 2;
    static constexpr int L = S >= 1024 ? 32 : 16;  // lanes per sequence
    static constexpr int E1 = S / L;               // phase-1 values per lane
    static constexpr int P = E1 / L;               // length-L DFTs per lane in phase 2
    static constexpr int NTH = 256;                // 4 waves
    static constexpr int SEQ = NTH / L;            // sequences (rows pairs / columns) per workgroup
    // column pass: 16 columns per workgroup, so that its xt reads and ring
    // writes move 128-byte row segments; 8 for S >= 512 (66 KB of LDS for
    // S = 1024: two workgroups per CU, whose phases overlap, instead of one;
    // four for S = 512), two workgroups reading the halves of each 16-column
    // block of the pass buffer
    static constexpr int SEQ_C = S >= 512 ? 8 : 16;
    static constexpr int NTH_C = SEQ_C * L;
    static_assert(E1 == 16 || E1 == 32, "phase 1 is radix 16 or 32");
};

template <int E>
__device__ __forceinline__ void dftE(float2 (&v)[E]);
template <>
__device__ __forceinline__ void dftE<16>(float2 (&v)[16]) {
    dft16(v);
}
template <>
__device__ __forceinline__ void dftE<32>(float2 (&v)[32]) {
    dft32(v);
}

// One length-S FFT per group of L lanes (q = lane index in the group): v
// holds x[n1 * L + q] on entry; on return the natural-order spectrum of the
// sequence is at seqbuf[0 .. S) (LDS, private to the sequence).  x is
// exchanged through seqbuf as [k1][n2] with a rotation of n2 by k1 so that
// the lanes of each access hit distinct 8-byte bank slots.
template <int S>
__device__ __forceinline__ void big_fft(float2 (&v)[BigCfg<S>::E1], float2 *seqbuf, int q,
                                        const float2 *__restrict__ twg) {
    constexpr int L = BigCfg<S>::L, E1 = BigCfg<S>::E1, P = BigCfg<S>::P;
    dftE<E1>(v);
#pragma unroll
    for (int k1 = 1; k1 < E1; ++k1) v[k1] = cmul(v[k1], twg[S + __mul24(q, k1)]);
#pragma unroll
    for (int k1 = 0; k1 < E1; ++k1) seqbuf[k1 * L + ((q + k1) & (L - 1))] = v[k1];
    float2 u[P][L];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k1 = q + L * p;
#pragma unroll
        for (int n2 = 0; n2 < L; ++n2) u[p][n2] = seqbuf[k1 * L + ((n2 + k1) & (L - 1))];
        dftE<L>(u[p]);  // u[p][k2] = X[k1 + E1 k2]
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k2 = 0; k2 < L; ++k2) seqbuf[q + L * p + E1 * k2] = u[p][k2];
}

// Rows: SEQ row pairs per workgroup; columns 0 .. H-1 of rows 2r, 2r+1 of the
// tile's half spectrum (column 0 = DC + i Nyquist) go to the batch's
// column-block buffer xt: frame f, tile t, column block b = kx / 16, row ky,
// column kx % 16 at xt[(((f * tiles + t) * H / 16 + b) * S + ky) * 16 + kx % 16],
// so every 16 lanes of a row piece write 128 aligned bytes and the column
// pass reads each block as one contiguous run.  Only columns kx < xcols (a
// multiple of 16: the blocks the column pass reads) are written.
// The row pair r of tile `tile` (frame f) transformed by big_fft into
// seqbuf, separated and written to the pass buffer by the sequence's lane q.
template <int S>
__device__ __forceinline__ void rows_to_xt(const float2 *seqbuf, float2 *__restrict__ xt, int f, int tiles, int tile,
                                           int r, int q, int xcols) {
    using C = BigCfg<S>;
    constexpr int H = C::H, L = C::L;
    // lane q writes columns k = q + L m, block k / 16 = (L / 16) m + q / 16,
    // column q % 16 of it
    float2 *xa = xt + ((long long)f * tiles + tile) * ((long long)S * H) + (2 * r) * 16 + (q & 15) +
                 (long long)(q >> 4) * S * 16;
#pragma unroll
    for (int m = 0; m < H / L; ++m) {
        const int k = q + L * m;
        const float2 zk = seqbuf[k];
        const float2 zp = seqbuf[(S - k) & (S - 1)];
        float2 a = make_float2(0.5f * (zk.x + zp.x), 0.5f * (zk.y - zp.y));
        float2 b = make_float2(0.5f * (zk.y + zp.y), -0.5f * (zk.x - zp.x));
        if (k == 0) {
            const float2 zh = seqbuf[H];  // DC and Nyquist of both rows are real
            a = make_float2(zk.x, zh.x);
            b = make_float2(zk.y, zh.y);
        }
        if (k < xcols) {
            store_nt(&xa[(long long)m * (L / 16) * S * 16], a);
            store_nt(&xa[(long long)m * (L / 16) * S * 16 + 16], b);
        }
    }
}

template <int S, int PIX>
__device__ __forceinline__ void tile_fft_rows_reg_body(const PixSrc &ps, int tps, int tiles, float2 *__restrict__ xt,
                                                       const float2 *__restrict__ twg, int xcols, int bid,
                                                       int nblk) {
    using C = BigCfg<S>;
    constexpr int L = C::L, E1 = C::E1, SEQ = C::SEQ, GROUPS = C::H / SEQ;
    __shared__ float2 lds[SEQ * S];
    const int tid = threadIdx.x, q = tid & (L - 1), sq = tid / L;
    const int nf = nblk / (GROUPS * tiles);  // frames vary fastest
    const int f = bid % nf;
    const int g = (bid / nf) % GROUPS;
    const int tile = bid / (nf * GROUPS);
    const int r = g * SEQ + sq;  // row pair within the tile
    const uint8_t *fr = ps.frames + (long long)f * ps.frame_bytes;
    const int x0 = ps.x_off + (tile % tps) * S + q;
    const int y0 = ps.y_off + (tile / tps) * S + 2 * r;
    float2 v[E1];
    load_row_pair<PIX, E1, L>(fr, ps, x0, y0, v);
    float2 *seqbuf = lds + sq * S;
    big_fft<S>(v, seqbuf, q, twg);
    rows_to_xt<S>(seqbuf, xt, f, tiles, tile, r, q, xcols);
}

template <int S, int PIX>
__global__ __launch_bounds__(BigCfg<S>::NTH) void tile_fft_rows_reg(PixSrc ps, int tps, int tiles,
                                                                   float2 *__restrict__ xt,
                                                                   const float2 *__restrict__ twg, int xcols) {
    tile_fft_rows_reg_body<S, PIX>(ps, tps, tiles, xt, twg, xcols, blockIdx.x, gridDim.x);
}

// Columns: SEQ consecutive packed columns per workgroup, read from the pass
// buffer's column block and written to the ring slot; the packed column is
// separated into the DC (kx = 0) and Nyquist (kx = H) columns on the way out.
// In the column-order layout (LAY_COLS) each column's kept elements are one
// contiguous run of the tile's ring elements, and the workgroup's columns
// are adjacent; blocks with no kept element end at once.
template <int S>
__device__ __forceinline__ void tile_fft_cols_reg_body(const RingDst &rd, int tiles, const float2 *__restrict__ xt,
                                                       const float2 *__restrict__ twg, int bid, int nblk) {
    using C = BigCfg<S>;
    constexpr int H = C::H, L = C::L, E1 = C::E1, SEQ = C::SEQ_C, NB = H / SEQ, NTH = C::NTH_C;
    constexpr int RP = NTH / SEQ;  // rows per pass of the staging loops (= L)
    static_assert(RP == L && RP * E1 == S && H % SEQ == 0, "staging covers the tile block");
    // tile block [y][c ^ (y % SEQ)], SEQ columns per row; later the columns'
    // spectra, S + 1 apart so that reading one row across them is conflict free
    __shared__ float2 lds[SEQ * (S + 1)];
    // LAY_COLS: starts of columns c0 .. c0 + SEQ, bands of c0 .. c0 + SEQ - 1,
    // then start, end and band of the Nyquist column
    __shared__ int ctab[2 * SEQ + 4];
    const int tid = threadIdx.x;
    int f, cb, tile;
    if constexpr (SEQ < 16) {
        // columns c0 and c0 + 8 share a 16-column pass-buffer block: they go
        // to workgroups b and b + 8, which land on the same XCD (workgroups
        // are dealt to the 8 XCDs in turn) and so share its L2 (nblk is a
        // multiple of 16)
        static_assert(2 * SEQ == 16, "two workgroups per 16-column block");
        const int pair = (bid >> 4) * 8 + (bid & 7), sib = (bid >> 3) & 1;
        const int nf = nblk / (NB * tiles);  // frames vary fastest
        f = pair % nf;
        cb = 2 * ((pair / nf) % (NB / 2)) + sib;
        tile = pair / (nf * (NB / 2));
    } else {
        const int nf = nblk / (NB * tiles);  // frames vary fastest
        f = bid % nf;
        cb = (bid / nf) % NB;
        tile = bid / (nf * NB);
    }
    const int c0 = cb * SEQ;
    // the same for the whole workgroup (block 0 also holds the Nyquist column)
    if (c0 >= rd.lay.pass_cols) return;
    const TileOut t = tile_out(rd, f, tile, S);
    if (rd.lay.kind == LAY_COLS && tid < 2 * SEQ + 4) {
        const int k = tid;
        ctab[k] = k <= SEQ ? rd.lay.start[c0 + k]
                           : (k < 2 * SEQ + 1 ? rd.lay.band[c0 + k - SEQ - 1]
                                              : (k == 2 * SEQ + 3 ? rd.lay.band[H] : rd.lay.start[H + k - 2 * SEQ - 1]));
    }
    const int col = tid % SEQ, row = tid / SEQ;  // staging: SEQ consecutive columns of RP rows per pass
    {
        float2 v[E1];
        const float2 *xb = xt + (((long long)f * tiles + tile) * (H / 16) + c0 / 16) * ((long long)S * 16) +
                           row * 16 + (c0 & 15) + col;
#pragma unroll
        for (int k = 0; k < E1; ++k) v[k] = xb[k * RP * 16];
#pragma unroll
        for (int k = 0; k < E1; ++k) {
            const int y = row + RP * k;
            lds[y * SEQ + (col ^ (y & (SEQ - 1)))] = v[k];
        }
    }
    __syncthreads();
    // column sq of the block, lane q: rows n1 * L + q
    const int q = tid & (L - 1), sq = tid / L;
    float2 v[E1];
#pragma unroll
    for (int n1 = 0; n1 < E1; ++n1) {
        const int y = n1 * L + q;
        v[n1] = lds[y * SEQ + (sq ^ (y & (SEQ - 1)))];
    }
    __syncthreads();
    float2 *seqbuf = lds + sq * (S + 1);
    big_fft<S>(v, seqbuf, q, twg);
    __syncthreads();
    // column c's spectrum is at lds[c * (S + 1) + ky]; the packed column's
    // DC and Nyquist of row ky come from V[ky] and V[S - ky]
    if (t.lay.kind == LAY_COLS) {
        // column by column: the kept rows of column kx are one contiguous
        // run from start[kx], rows ky < band[kx] first, then the bottom ones
        const int a0 = ctab[0];
        if (ctab[SEQ] - a0 == SEQ * S) {
            // every column of the block keeps every row
            for (int j = tid; j < SEQ * S; j += NTH) {
                const int c = j / S, ky = j & (S - 1);
                float2 val;
                if (c0 + c == 0) {
                    const float2 w = lds[ky], pv = lds[(S - ky) & (S - 1)];
                    val = make_float2(0.5f * (w.x + pv.x), 0.5f * (w.y - pv.y));
                } else {
                    val = lds[c * (S + 1) + ky];
                }
                store_nt(&t.ring[ring_index(t.base + a0 + j, t.slot, t.R)], val);
            }
        } else {
        for (int i = a0 + tid; i < ctab[SEQ]; i += NTH) {
            const int c = find_start(ctab, SEQ, i), a = ctab[c], n = ctab[c + 1] - a, r = i - a;
            const int lo = ctab[SEQ + 1 + c];
            const int ky = r < lo ? r : r + (S - n);
            float2 val;
            if (c0 + c == 0) {
                const float2 w = lds[ky], pv = lds[(S - ky) & (S - 1)];
                val = make_float2(0.5f * (w.x + pv.x), 0.5f * (w.y - pv.y));
            } else {
                val = lds[c * (S + 1) + ky];
            }
            store_nt(&t.ring[ring_index(t.base + i, t.slot, t.R)], val);
        }
        }
        if (c0 == 0) {
            // the Nyquist column kx = H
            const int a = ctab[2 * SEQ + 1], n = ctab[2 * SEQ + 2] - a, lo = ctab[2 * SEQ + 3];
            for (int r = tid; r < n; r += NTH) {
                const int ky = r < lo ? r : r + (S - n);
                const float2 w = lds[ky], pv = lds[(S - ky) & (S - 1)];
                store_nt(&t.ring[ring_index(t.base + a + r, t.slot, t.R)],
                         make_float2(0.5f * (w.y + pv.y), -0.5f * (w.x - pv.x)));
            }
        }
        return;
    }
    // row-order layout: rows of SEQ columns
    if (c0 == 0) {
        for (int ky = tid; ky < S; ky += NTH) {
            const float2 w = lds[ky], pv = lds[(S - ky) & (S - 1)];
            t.put(ky, 0, make_float2(0.5f * (w.x + pv.x), 0.5f * (w.y - pv.y)));
            t.put(ky, H, make_float2(0.5f * (w.y + pv.y), -0.5f * (w.x - pv.x)));
        }
    }
#pragma unroll
    for (int k = 0; k < E1; ++k) {
        const int ky = row + RP * k;
        if (c0 + col != 0) t.put(ky, c0 + col, lds[col * (S + 1) + ky]);
    }
}

template <int S>
__global__ __launch_bounds__(BigCfg<S>::NTH_C) void tile_fft_cols_reg(RingDst rd, int tiles,
                                                                   const float2 *__restrict__ xt,
                                                                   const float2 *__restrict__ twg) {
    tile_fft_cols_reg_body<S>(rd, tiles, xt, twg, blockIdx.x, gridDim.x);
}

template <int S>
static void launch_cols(const RingDst &rd, int tiles, int n_frames, const float2 *tw, const float2 *xt,
                        hipStream_t st) {
    using C = BigCfg<S>;
    const dim3 g2((unsigned)((C::H / C::SEQ_C) * tiles * n_frames)), block2(C::NTH_C);
    hipLaunchKernelGGL(tile_fft_cols_reg<S>, g2, block2, 0, st, rd, tiles, xt, tw);
}

template <int S>
static void launch_big(const PixSrc &ps, const RingDst &rd, int tps, int tiles, int n_frames, const float2 *tw,
                       float2 *xt, hipStream_t st) {
    using C = BigCfg<S>;
    const dim3 g1((unsigned)((C::H / C::SEQ) * tiles * n_frames)), block(C::NTH);
    const int xcols = rd.lay.pass_cols;
    switch (pixel_mode(ps)) {
    case PIX_U8_1:
        hipLaunchKernelGGL((tile_fft_rows_reg<S, PIX_U8_1>), g1, block, 0, st, ps, tps, tiles, xt, tw, xcols);
        break;
    case PIX_U16LE_2:
        hipLaunchKernelGGL((tile_fft_rows_reg<S, PIX_U16LE_2>), g1, block, 0, st, ps, tps, tiles, xt, tw, xcols);
        break;
    default: hipLaunchKernelGGL((tile_fft_rows_reg<S, PIX_ANY>), g1, block, 0, st, ps, tps, tiles, xt, tw, xcols);
    }
    launch_cols<S>(rd, tiles, n_frames, tw, xt, st);
}

// ---------------------------------------------------------------------------
// Batched forms: one launch per kernel kind over a job table of many
// streams' chunks, a workgroup per queue item (the batch's blocks are the
// concatenation of the jobs' own grids).  Each block runs the body of its
// job's own launch with the job's block index, so a chunk's spectra are the
// same bit for bit whether it is transformed alone or in a batch.  (A
// persistent loop over the items was measured out: the loop kept the
// bodies' loop-invariant values live across it, 15-90 % more registers.)
// ---------------------------------------------------------------------------
template <int S, int PIX>
__global__ __launch_bounds__(REG_NTH) void tile_fft_reg_batch(const BatchHdr *__restrict__ hdr,
                                                             const FftJob *__restrict__ jobs,
                                                             const float2 *__restrict__ twg) {
    const long long b = blockIdx.x;
    const FftJob &j = jobs[batch_job(jobs, hdr->njobs, b)];
    tile_fft_reg_body<S, PIX>(j.ps, j.rd, j.tps, j.tiles, j.groups, twg, (int)(b - j.first), j.nblk);
}

template <int S, int PIX>
__global__ __launch_bounds__(SMALL_NTH) void tile_fft_lane_batch(const BatchHdr *__restrict__ hdr,
                                                                const FftJob *__restrict__ jobs) {
    const long long b = blockIdx.x;
    const FftJob &j = jobs[batch_job(jobs, hdr->njobs, b)];
    tile_fft_lane_body<S, PIX, true>(j.ps, j.rd, j.tps, j.tiles, j.groups, (int)(b - j.first), j.nblk);
}

template <int S, int PIX>
__global__ __launch_bounds__(BigCfg<S>::NTH) void tile_fft_rows_reg_batch(const BatchHdr *__restrict__ hdr,
                                                                         const FftJob *__restrict__ jobs,
                                                                         const float2 *__restrict__ twg) {
    const long long b = blockIdx.x;
    const FftJob &j = jobs[batch_job(jobs, hdr->njobs, b)];
    tile_fft_rows_reg_body<S, PIX>(j.ps, j.tps, j.tiles, j.xt, twg, j.rd.lay.pass_cols, (int)(b - j.first),
                                   j.nblk);
}

template <int S>
__global__ __launch_bounds__(BigCfg<S>::NTH_C) void tile_fft_cols_reg_batch(const BatchHdr *__restrict__ hdr,
                                                                           const FftJob *__restrict__ jobs,
                                                                           const float2 *__restrict__ twg) {
    const long long b = blockIdx.x;
    const FftJob &j = jobs[batch_job(jobs, hdr->njobs, b)];
    tile_fft_cols_reg_body<S>(j.rd, j.tiles, j.xt, twg, (int)(b - j.first), j.nblk);
}

bool fft_batch_jobs(const PixSrc &ps, const RingDst &rd, int scale, int tps, int tiles, int n_frames, float2 *xt,
                    std::vector<std::pair<FftKey, FftJob>> *out) {
    const int pix = pixel_mode(ps);
    if (pix == PIX_ANY) return false;
    FftJob j{};
    j.ps = ps;
    j.rd = rd;
    j.xt = xt;
    j.tps = tps;
    j.tiles = tiles;
    switch (scale) {
    case 8:
    case 16:
    case 32: {
        // the batched lane kernel reads 4-byte aligned rows
        const bool a4 = ((ps.row_bytes | ps.frame_bytes | (long long)(uintptr_t)ps.frames |
                          (long long)ps.x_off * ps.bpp) & 3) == 0;
        if (!a4) return false;
        const int tpg = SMALL_NTH / (scale / 2);
        j.groups = (tiles + tpg - 1) / tpg;
        j.nblk = j.groups * n_frames;
        out->push_back({FftKey{FFT_LANE, scale, pix}, j});
        return true;
    }
    case 64:
    case 128: {
        const int tpw = scale == 64 ? RegCfg<64>::TPW : RegCfg<128>::TPW;
        j.groups = (tiles + tpw - 1) / tpw;
        j.nblk = j.groups * n_frames;
        out->push_back({FftKey{FFT_REG, scale, pix}, j});
        return true;
    }
    case 256:
    case 512:
    case 1024: {
        if (!xt) return false;
        const int h = scale / 2, seq = scale >= 1024 ? BigCfg<1024>::SEQ : BigCfg<256>::SEQ;
        j.groups = 0;
        j.nblk = (h / seq) * tiles * n_frames;
        out->push_back({FftKey{FFT_ROWS, scale, pix}, j});
        const int seqc = scale >= 512 ? BigCfg<512>::SEQ_C : BigCfg<256>::SEQ_C;
        j.nblk = (h / seqc) * tiles * n_frames;
        out->push_back({FftKey{FFT_COLS, scale, 0}, j});
        return true;
    }
    default: return false;
    }
}

template <int S>
static void launch_fft_batch_s(const FftKey &k, const BatchHdr *hdr, const FftJob *jobs, long long blocks,
                               const float2 *tw, hipStream_t st) {
    const dim3 grid((unsigned)blocks);
    if (k.kind == FFT_COLS) {
        if constexpr (S >= 256)
            hipLaunchKernelGGL(tile_fft_cols_reg_batch<S>, grid, dim3(BigCfg<S>::NTH_C), 0, st, hdr, jobs, tw);
        return;
    }
    if constexpr (S <= 32) {
        if (k.pix == PIX_U8_1)
            hipLaunchKernelGGL((tile_fft_lane_batch<S, PIX_U8_1>), grid, dim3(SMALL_NTH), 0, st, hdr, jobs);
        else
            hipLaunchKernelGGL((tile_fft_lane_batch<S, PIX_U16LE_2>), grid, dim3(SMALL_NTH), 0, st, hdr, jobs);
    } else if constexpr (S <= 128) {
        if (k.pix == PIX_U8_1)
            hipLaunchKernelGGL((tile_fft_reg_batch<S, PIX_U8_1>), grid, dim3(REG_NTH), 0, st, hdr, jobs, tw);
        else
            hipLaunchKernelGGL((tile_fft_reg_batch<S, PIX_U16LE_2>), grid, dim3(REG_NTH), 0, st, hdr, jobs, tw);
    } else {
        if (k.pix == PIX_U8_1)
            hipLaunchKernelGGL((tile_fft_rows_reg_batch<S, PIX_U8_1>), grid, dim3(BigCfg<S>::NTH), 0, st, hdr, jobs,
                               tw);
        else
            hipLaunchKernelGGL((tile_fft_rows_reg_batch<S, PIX_U16LE_2>), grid, dim3(BigCfg<S>::NTH), 0, st, hdr,
                               jobs, tw);
    }
}

hipError_t launch_fft_batch(const FftKey &k, const BatchHdr *hdr, const FftJob *jobs, long long blocks,
                            const float2 *tw, hipStream_t st) {
    if (blocks < 1 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (k.scale) {
    case 8: launch_fft_batch_s<8>(k, hdr, jobs, blocks, tw, st); break;
    case 16: launch_fft_batch_s<16>(k, hdr, jobs, blocks, tw, st); break;
    case 32: launch_fft_batch_s<32>(k, hdr, jobs, blocks, tw, st); break;
    case 64: launch_fft_batch_s<64>(k, hdr, jobs, blocks, tw, st); break;
    case 128: launch_fft_batch_s<128>(k, hdr, jobs, blocks, tw, st); break;
    case 256: launch_fft_batch_s<256>(k, hdr, jobs, blocks, tw, st); break;
    case 512: launch_fft_batch_s<512>(k, hdr, jobs, blocks, tw, st); break;
    case 1024: launch_fft_batch_s<1024>(k, hdr, jobs, blocks, tw, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool launch_tile_fft_reg(const PixSrc &ps, const RingDst &rd, int scale, int tps, int tiles, int n_frames,
                         const float2 *tw, float2 *xt, hipStream_t st) {
    switch (scale) {
    case 8: launch_lane<8>(ps, rd, tps, tiles, n_frames, st); return true;
    case 16: launch_lane<16>(ps, rd, tps, tiles, n_frames, st); return true;
    case 32: launch_lane<32>(ps, rd, tps, tiles, n_frames, st); return true;
    case 64: launch_reg<64>(ps, rd, tps, tiles, n_frames, tw, st); return true;
    case 128: launch_reg<128>(ps, rd, tps, tiles, n_frames, tw, st); return true;
    case 256:
        if (!xt) return false;
        launch_big<256>(ps, rd, tps, tiles, n_frames, tw, xt, st);
        return true;
    case 512:
        if (!xt) return false;
        launch_big<512>(ps, rd, tps, tiles, n_frames, tw, xt, st);
        return true;
    case 1024:
        if (!xt) return false;
        launch_big<1024>(ps, rd, tps, tiles, n_frames, tw, xt, st);
        return true;
    default: return false;
    }
}

// ---------------------------------------------------------------------------
// Multi-scale tiler: each frame's pixels are read once for all 8 .. 128 tile
// scales and once for the row passes of all 256 .. 1024 scales (the
// reference cuts and transforms every scale from the frame separately,
// src/DDM.cu:171-192 with parseBufferScalePow2, src/DDM_kernel.cuh:43-78).
//
// tile_fft_blocks: a workgroup per (frame, 128 x 128 block of the region):
// the block's pixels are staged once in LDS as 16-bit values, then the
// register transforms of every scale run on it in turn (the 128 tile, the
// four 64 tiles, the sixteen 32 tiles, ...) and store to the ring exactly as
// their per-scale launches do, so the spectra are the same bit for bit.
//
// tile_fft_bands: a workgroup per (frame, band of 16384 / M rows across the
// whole M-pixel-wide region): each thread loads 64 pixels once and keeps them
// in registers; for every 256 .. 1024 scale (and every pass of 256 threads)
// the band is restaged in LDS with that scale's row padding, each sequence
// reads its row pair, and the row FFT and pass-buffer writes are those of
// tile_fft_rows_reg.  The column passes then run per scale.
// ---------------------------------------------------------------------------
constexpr int BLK_NTH = 512;
// LDS of the bodies run on a block (complex values): the largest is the
// 16-tile lane body (64 tiles of 16 x 9)
constexpr int BLK_SLOTS = 64 * 16 * 9;
static_assert(BLK_SLOTS >= 128 * 65 && BLK_SLOTS >= 4 * 64 * 33 && BLK_SLOTS >= 16 * 32 * 17 &&
                  BLK_SLOTS >= 128 * 8 * 5,
              "every body's LDS fits");

struct BlockScale {
    RingDst rd;
    int s, tps;
};
constexpr int MAX_BLOCK_SCALES = 5;  // 128, 64, 32, 16, 8
struct BlockArgs {
    PixSrc ps;
    int bpr;   // blocks per region row: M / 128
    int nblk;  // blocks per frame
    BlockScale sc[MAX_BLOCK_SCALES];  // slot k: tile size 128 >> k, tps = 0 when absent
};

// Global loads of the tiler: 16-byte and 4-byte words (every row piece
// aligned to that size), or pixel by pixel through load_px.
enum { LD_PX = 0, LD_4 = 1, LD_16 = 2 };

// 16 8-bit pixels (four little-endian words) as 16-bit values in eight words
__device__ __forceinline__ uint2 widen_u8(uint32_t w) {
    return make_uint2((w & 0xffu) | ((w & 0xff00u) << 8), ((w >> 16) & 0xffu) | ((w >> 8) & 0xff0000u));
}

// The block at region coordinates (X0, Y0) of frame fr into blk (BLK rows of
// BLK_PITCH 16-bit values), by BLK_NTH threads.
template <int PIX, int LD>
__device__ __forceinline__ void stage_block(const uint8_t *fr, const PixSrc &ps, int X0, int Y0, uint16_t *blk,
                                            int tid) {
    const uint8_t *base = fr + (long long)(ps.y_off + Y0) * ps.row_bytes + (long long)(ps.x_off + X0) * ps.bpp;
    if constexpr (PIX == PIX_U16LE_2 && LD == LD_16) {
        // 16 pieces of 16 bytes per row, 4 per thread
        uint4 u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 4, c = i & 15;
            u[k] = *reinterpret_cast<const uint4 *>(base + (long long)y * ps.row_bytes + c * 16);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 4, c = i & 15;
            uint2 *d = reinterpret_cast<uint2 *>(blk + y * BLK_PITCH + c * 8);
            d[0] = make_uint2(u[k].x, u[k].y);
            d[1] = make_uint2(u[k].z, u[k].w);
        }
    } else if constexpr (PIX == PIX_U8_1 && LD == LD_16) {
        // 8 pieces of 16 bytes per row, 2 per thread
        uint4 u[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 3, c = i & 7;
            u[k] = *reinterpret_cast<const uint4 *>(base + (long long)y * ps.row_bytes + c * 16);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 3, c = i & 7;
            uint2 *d = reinterpret_cast<uint2 *>(blk + y * BLK_PITCH + c * 16);
            d[0] = widen_u8(u[k].x);
            d[1] = widen_u8(u[k].y);
            d[2] = widen_u8(u[k].z);
            d[3] = widen_u8(u[k].w);
        }
    } else if constexpr (PIX == PIX_U16LE_2 && LD == LD_4) {
        // 64 words per row, 16 per thread
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 6, c = i & 63;
            const uint32_t w = *reinterpret_cast<const uint32_t *>(base + (long long)y * ps.row_bytes + c * 4);
            *reinterpret_cast<uint32_t *>(blk + y * BLK_PITCH + c * 2) = w;
        }
    } else if constexpr (PIX == PIX_U8_1 && LD == LD_4) {
        // 32 words per row, 8 per thread
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 5, c = i & 31;
            const uint32_t w = *reinterpret_cast<const uint32_t *>(base + (long long)y * ps.row_bytes + c * 4);
            *reinterpret_cast<uint2 *>(blk + y * BLK_PITCH + c * 4) = widen_u8(w);
        }
    } else {
        // 128 pixels per row, 32 per thread, consecutive threads on consecutive pixels
#pragma unroll 8
        for (int k = 0; k < 32; ++k) {
            const int i = tid + BLK_NTH * k, y = i >> 7, x = i & 127;
            blk[y * BLK_PITCH + x] = (uint16_t)load_px(fr, ps, ps.x_off + X0 + x, ps.y_off + Y0 + y);
        }
    }
}

// Every scale's tiles of block (bx, by) of batch frame f in turn, largest
// first, pixels from src; slot k of a.sc holds tile size 128 >> k (tps = 0:
// scale absent).
template <bool A4, class Src>
__device__ __forceinline__ void block_scales(const Src &src, const BlockArgs &a, const float2 *__restrict__ twg,
                                             float2 *lds, int *rtab, int tid, int f, int bx, int by) {
    auto tiles_of = [&](int k) { return TileSet{by * (1 << k) * a.sc[k].tps + bx * (1 << k), k, 1 << (2 * k)}; };
    if (a.sc[0].tps) {
        reg_tiles<128>(src, a.sc[0].rd, a.sc[0].tps, tiles_of(0), twg, lds, rtab, tid, f);
        __syncthreads();  // the next scale reuses lds and rtab
    }
    if (a.sc[1].tps) {
        reg_tiles<64>(src, a.sc[1].rd, a.sc[1].tps, tiles_of(1), twg, lds, rtab, tid, f);
        __syncthreads();
    }
    if (a.sc[2].tps) {
        // sixteen 32 tiles take 256 of the 512 threads
        lane_tiles<32, A4, 256, BLK_NTH>(src, a.sc[2].rd, a.sc[2].tps, tiles_of(2), lds, rtab, tid, f);
        __syncthreads();
    }
    if (a.sc[3].tps) {
        lane_tiles<16, A4, BLK_NTH, BLK_NTH>(src, a.sc[3].rd, a.sc[3].tps, tiles_of(3), lds, rtab, tid, f);
        __syncthreads();
    }
    if (a.sc[4].tps) {
        // 256 tiles: two rounds of 128 (8 rows of 16 tiles each)
        const TileSet ts = tiles_of(4);
        lane_tiles<8, A4, BLK_NTH, BLK_NTH>(src, a.sc[4].rd, a.sc[4].tps, TileSet{ts.first, 4, 128}, lds, rtab, tid,
                                            f);
        __syncthreads();
        lane_tiles<8, A4, BLK_NTH, BLK_NTH>(src, a.sc[4].rd, a.sc[4].tps, TileSet{ts.first + 8 * a.sc[4].tps, 4, 128},
                                            lds, rtab, tid, f);
    }
}

// The block's pixels staged once in LDS, every scale cut from there.
template <int PIX, int LD>
__global__ __launch_bounds__(BLK_NTH) void tile_fft_blocks(BlockArgs a, const float2 *__restrict__ twg) {
    __shared__ __align__(16) uint16_t blk[BLK * BLK_PITCH];
    __shared__ float2 lds[BLK_SLOTS];
    __shared__ int rtab[BLK + 1];
    const int tid = threadIdx.x;
    // frames vary fastest across the grid, as in the per-scale launches
    const int nf = gridDim.x / a.nblk;
    const int f = blockIdx.x % nf, b = blockIdx.x / nf;
    const int bx = b % a.bpr, by = b / a.bpr;
    const int X0 = bx * BLK, Y0 = by * BLK;
    stage_block<PIX, LD>(a.ps.frames + (long long)f * a.ps.frame_bytes, a.ps, X0, Y0, blk, tid);
    __syncthreads();
    block_scales<true>(BlockPix{blk, X0, Y0}, a, twg, lds, rtab, tid, f, bx, by);
}

constexpr int BAND_NTH = 256;
constexpr int BAND_PX = 16384;  // pixels per band: M columns x 16384 / M rows
struct BandScale {
    float2 *xt;  // this scale's pass buffer region
    int s, tps, xcols;
};
constexpr int MAX_BAND_SCALES = 3;  // 1024, 512, 256
struct BandArgs {
    PixSrc ps;
    int M;
    int nband;  // bands per frame
    BandScale sc[MAX_BAND_SCALES];  // slot k: tile size 1024 >> k, tps = 0 when absent
};

// This thread's 64 pixels (row y, columns x .. x + 63 of the region) as
// 16-bit values, two per word.
template <int PIX, int LD>
__device__ __forceinline__ void load_band(const uint8_t *fr, const PixSrc &ps, int x, int y, uint32_t (&raw)[32]) {
    const uint8_t *p = fr + (long long)(ps.y_off + y) * ps.row_bytes + (long long)(ps.x_off + x) * ps.bpp;
    if constexpr (PIX == PIX_U16LE_2 && LD == LD_16) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint4 u = reinterpret_cast<const uint4 *>(p)[k];
            raw[4 * k] = u.x;
            raw[4 * k + 1] = u.y;
            raw[4 * k + 2] = u.z;
            raw[4 * k + 3] = u.w;
        }
    } else if constexpr (PIX == PIX_U8_1 && LD == LD_16) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 u = reinterpret_cast<const uint4 *>(p)[k];
            const uint2 a = widen_u8(u.x), b = widen_u8(u.y), c = widen_u8(u.z), d = widen_u8(u.w);
            raw[8 * k] = a.x;
            raw[8 * k + 1] = a.y;
            raw[8 * k + 2] = b.x;
            raw[8 * k + 3] = b.y;
            raw[8 * k + 4] = c.x;
            raw[8 * k + 5] = c.y;
            raw[8 * k + 6] = d.x;
            raw[8 * k + 7] = d.y;
        }
    } else if constexpr (PIX == PIX_U16LE_2 && LD == LD_4) {
#pragma unroll
        for (int k = 0; k < 32; ++k) raw[k] = reinterpret_cast<const uint32_t *>(p)[k];
    } else if constexpr (PIX == PIX_U8_1 && LD == LD_4) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint2 a = widen_u8(reinterpret_cast<const uint32_t *>(p)[k]);
            raw[2 * k] = a.x;
            raw[2 * k + 1] = a.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const uint32_t lo = (uint32_t)load_px(fr, ps, ps.x_off + x + 2 * k, ps.y_off + y);
            const uint32_t hi = (uint32_t)load_px(fr, ps, ps.x_off + x + 2 * k + 1, ps.y_off + y);
            raw[k] = lo | (hi << 16);
        }
    }
}

// The row passes of scale S for band rows Y0 .. Y0 + 16384 / M of frame f.
// The band goes to LDS with rows of M + PAD 16-bit values, PAD chosen so
// that the sequences of one wave (consecutive row pairs) read distinct
// banks: 8 dwords apart for 16-lane sequences, 16 for 32-lane ones.
template <int S>
__device__ __forceinline__ void band_rows(const uint32_t (&raw)[32], int M, const BandScale &sc, float2 *lds,
                                          const float2 *__restrict__ twg, int tid, int f, int Y0, int ry, int cx) {
    using C = BigCfg<S>;
    constexpr int L = C::L, E1 = C::E1;
    constexpr int PAD = L == 32 ? 16 : 8;
    constexpr int SEQ = BAND_NTH / L;  // sequences per pass
    // sequences of the band: NP row pairs x M / S tiles = 8192 / S, in
    // passes of SEQ
    constexpr int PASSES = 8192 / S / SEQ;
    static_assert(PASSES >= 1 && PASSES * SEQ * S == 8192, "whole passes");
    const int PW = M + PAD;
    const int NP = BAND_PX / M / 2;  // row pairs per band
    uint16_t *band = reinterpret_cast<uint16_t *>(lds);
    const int q = tid & (L - 1);
#pragma unroll
    for (int p = 0; p < PASSES; ++p) {
        __syncthreads();  // the previous pass is done with lds
        {
            uint4 *d = reinterpret_cast<uint4 *>(band + ry * PW + cx);
#pragma unroll
            for (int k = 0; k < 8; ++k) d[k] = make_uint4(raw[4 * k], raw[4 * k + 1], raw[4 * k + 2], raw[4 * k + 3]);
        }
        __syncthreads();
        const int sq = p * SEQ + tid / L;
        const int r = sq % NP, tx = sq / NP;  // row pair within the band, tile column
        float2 v[E1];
        {
            const uint16_t *r0 = band + (2 * r) * PW + tx * S + q;
#pragma unroll
            for (int n1 = 0; n1 < E1; ++n1) v[n1] = make_float2((float)r0[n1 * L], (float)r0[PW + n1 * L]);
        }
        __syncthreads();  // every sequence has its pixels before the FFTs use lds
        float2 *seqbuf = lds + (tid / L) * S;
        big_fft<S>(v, seqbuf, q, twg);
        const int tile = (Y0 / S) * sc.tps + tx;
        rows_to_xt<S>(seqbuf, sc.xt, f, sc.tps * sc.tps, tile, (Y0 % S) / 2 + r, q, sc.xcols);
    }
}

template <int PIX, int LD>
__global__ __launch_bounds__(BAND_NTH) void tile_fft_bands(BandArgs a, const float2 *__restrict__ twg) {
    __shared__ __align__(16) float2 lds[8192];  // the band, then the sequences (64 KB)
    const int tid = threadIdx.x;
    const int M = a.M;
    const int nf = gridDim.x / a.nband;
    const int f = blockIdx.x % nf, band = blockIdx.x / nf;
    const int Y0 = band * (BAND_PX / M);
    const int tpr = M / 64;  // threads per band row
    const int ry = tid / tpr, cx = (tid % tpr) * 64;
    uint32_t raw[32];
    load_band<PIX, LD>(a.ps.frames + (long long)f * a.ps.frame_bytes, a.ps, cx, Y0 + ry, raw);
    // slot k of a.sc holds tile size 1024 >> k (tps = 0: scale absent)
    if (a.sc[0].tps) band_rows<1024>(raw, M, a.sc[0], lds, twg, tid, f, Y0, ry, cx);
    if (a.sc[1].tps) band_rows<512>(raw, M, a.sc[1], lds, twg, tid, f, Y0, ry, cx);
    if (a.sc[2].tps) band_rows<256>(raw, M, a.sc[2], lds, twg, tid, f, Y0, ry, cx);
}

static int tiler_load_mode(const PixSrc &ps, int pix) {
    if (pix == PIX_ANY) return LD_PX;
    const long long bits = ps.row_bytes | ps.frame_bytes | (long long)(uintptr_t)ps.frames | (long long)ps.x_off * ps.bpp;
    return (bits & 15) == 0 ? LD_16 : ((bits & 3) == 0 ? LD_4 : LD_PX);
}

template <class K>
static void tiler_dispatch(const PixSrc &ps, K launch) {
    const int pix = pixel_mode(ps), ld = tiler_load_mode(ps, pix);
    if (pix == PIX_U8_1 && ld == LD_16) launch(std::integral_constant<int, PIX_U8_1>{}, std::integral_constant<int, LD_16>{});
    else if (pix == PIX_U8_1 && ld == LD_4) launch(std::integral_constant<int, PIX_U8_1>{}, std::integral_constant<int, LD_4>{});
    else if (pix == PIX_U16LE_2 && ld == LD_16) launch(std::integral_constant<int, PIX_U16LE_2>{}, std::integral_constant<int, LD_16>{});
    else if (pix == PIX_U16LE_2 && ld == LD_4) launch(std::integral_constant<int, PIX_U16LE_2>{}, std::integral_constant<int, LD_4>{});
    else launch(std::integral_constant<int, PIX_ANY>{}, std::integral_constant<int, LD_PX>{});
}

bool tiler_block_scale(int scale, int M) { return scale >= 8 && scale <= BLK && M >= BLK; }
bool tiler_band_scale(int scale, int M) {
    return (scale == 256 || scale == 512 || scale == 1024) && M <= BAND_PX / 2;
}

hipError_t launch_tiler_blocks(const PixSrc &ps, const TilerScale *sc, int n, int M, int n_frames, const float2 *tw,
                               hipStream_t st) {
    if (n < 1 || n > MAX_BLOCK_SCALES || M < BLK || M % BLK) return hipErrorInvalidValue;
    BlockArgs a{};
    a.ps = ps;
    a.bpr = M / BLK;
    a.nblk = a.bpr * a.bpr;
    for (int i = 0; i < n; ++i) {
        if (!tiler_block_scale(sc[i].s, M) || (i && sc[i].s >= sc[i - 1].s)) return hipErrorInvalidValue;
        a.sc[ilog2(BLK / sc[i].s)] = BlockScale{sc[i].rd, sc[i].s, M / sc[i].s};
    }
    const long long blocks = (long long)a.nblk * n_frames;
    if (blocks < 1 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    tiler_dispatch(ps, [&](auto pix, auto ld) {
        hipLaunchKernelGGL((tile_fft_blocks<decltype(pix)::value, decltype(ld)::value>), dim3((unsigned)blocks),
                           dim3(BLK_NTH), 0, st, a, tw);
    });
    return hipGetLastError();
}

hipError_t launch_tiler_bands(const PixSrc &ps, const TilerScale *sc, int n, int M, int n_frames, const float2 *tw,
                              hipStream_t st) {
    if (n < 1 || n > MAX_BAND_SCALES || M < 256 || M > BAND_PX / 2 || (M & (M - 1))) return hipErrorInvalidValue;
    BandArgs a{};
    a.ps = ps;
    a.M = M;
    a.nband = M / (BAND_PX / M);
    for (int i = 0; i < n; ++i) {
        if (!tiler_band_scale(sc[i].s, M) || sc[i].s > M || !sc[i].xt || (i && sc[i].s >= sc[i - 1].s))
            return hipErrorInvalidValue;
        a.sc[ilog2(1024 / sc[i].s)] = BandScale{sc[i].xt, sc[i].s, M / sc[i].s, sc[i].rd.lay.pass_cols};
    }
    const long long blocks = (long long)a.nband * n_frames;
    if (blocks < 1 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    tiler_dispatch(ps, [&](auto pix, auto ld) {
        hipLaunchKernelGGL((tile_fft_bands<decltype(pix)::value, decltype(ld)::value>), dim3((unsigned)blocks),
                           dim3(BAND_NTH), 0, st, a, tw);
    });
    // the column passes, per scale
    for (int i = 0; i < n; ++i) {
        const int tps = M / sc[i].s, tiles = tps * tps;
        switch (sc[i].s) {
        case 256: launch_cols<256>(sc[i].rd, tiles, n_frames, tw, sc[i].xt, st); break;
        case 512: launch_cols<512>(sc[i].rd, tiles, n_frames, tw, sc[i].xt, st); break;
        case 1024: launch_cols<1024>(sc[i].rd, tiles, n_frames, tw, sc[i].xt, st); break;
        default: break;
        }
    }
    return hipGetLastError();
}

}  // namespace ddm
