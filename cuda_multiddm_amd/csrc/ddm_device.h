// Device helpers shared by the tile / FFT kernels: complex arithmetic, pixel
// addressing, and the blocked spectrum-ring layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ddm_amd.h"
#include "ddm_internal.h"

namespace ddm {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

struct PixSrc {
    const uint8_t *frames;
    long long frame_bytes;
    long long row_bytes;
    int bpp, fmt, x_off, y_off;
};

__device__ __forceinline__ float load_px(const uint8_t *fr, const PixSrc &ps, int x, int y) {
    const uint8_t *p = fr + (long long)y * ps.row_bytes + (long long)x * ps.bpp;
    if (ps.fmt == DDM_PIX_U8) return (float)p[0];
    const unsigned b0 = p[0], b1 = p[1];
    return (float)(ps.fmt == DDM_PIX_U16LE ? (b0 | (b1 << 8)) : ((b0 << 8) | b1));
}

// Ring layout: the spectrum elements of a frame are cut into groups of
// RING_GROUP consecutive elements, and each group keeps all ring slots
// together: element e of slot s lives at ((e / RING_GROUP) * R + s) *
// RING_GROUP + e % RING_GROUP.  A lag-engine workgroup (one group) then
// walks the frame history as one contiguous stream.
__device__ __forceinline__ long long ring_index(long long e, int slot, int R) {
    return ((e >> RING_GROUP_LOG2) * R + slot) * RING_GROUP + (e & (RING_GROUP - 1));
}

struct RingDst {
    float2 *ring;
    long long elem_off;  // of this scale
    int ring_slots;
    int first_slot;      // slot of frame 0 of this launch
    TileLayout lay;      // the part of each tile's half spectrum the ring holds
};

// floor(n / d) for 0 <= n < 2^22 and d >= 1 with rcp = 1.0f / d: the float
// product is within one of the quotient, and one step corrects it.
__device__ __forceinline__ int div_small(int n, int d, float rcp) {
    int q = (int)((float)n * rcp);
    const int r = n - q * d;
    q += (r >= d) ? 1 : 0;
    q -= (r < 0) ? 1 : 0;
    return q;
}

// Writes the elements of tile `tile` of batch frame f into the ring,
// addressed by (ky, kx) of the tile's half spectrum; elements the layout
// leaves out are dropped.
struct TileOut {
    float2 *ring;
    long long base;  // element index of the tile's first element
    int slot, R, S;
    TileLayout lay;
    __device__ __forceinline__ void put(int ky, int kx, float2 v) const {
        const long long at = lay.pos(S, ky, kx);
        if (at >= 0) ring[ring_index(base + at, slot, R)] = v;
    }
    // element (ky, kx) of a tile in the full row-order layout
    __device__ __forceinline__ float2 &full(int ky, int kx) const {
        return ring[ring_index(base + (long long)ky * (S / 2 + 1) + kx, slot, R)];
    }
};

__device__ __forceinline__ int ring_slot(const RingDst &rd, int f) {
    const int slot = rd.first_slot + f;
    return slot - ((slot >= rd.ring_slots) ? rd.ring_slots : 0);  // f < ring_slots
}

__device__ __forceinline__ TileOut tile_out(const RingDst &rd, int f, int tile, int S) {
    return TileOut{rd.ring, rd.elem_off + (long long)tile * rd.lay.elems(), ring_slot(rd, f), rd.ring_slots, S,
                   rd.lay};
}

// The spectra of ntile consecutive tiles (first, first + 1, ...) of batch
// frame f, staged in LDS at lds[t * S * (S/2 + 1) + ky * (S/2 + 1) + kx],
// stored to the ring in the row-order layout rd.lay (full or a band) by nth
// threads: consecutive threads write consecutive ring elements.
template <int S>
__device__ __forceinline__ void store_tiles(const RingDst &rd, const float2 *lds, int f, int first, int ntile,
                                            int tid, int nth) {
    constexpr int H = S / 2, TE = S * (H + 1);
    const int slot = ring_slot(rd, f);
    const long long gstride = (long long)rd.ring_slots * RING_GROUP;
    if (rd.lay.is_full(S)) {
        for (int i = tid; i < ntile * TE; i += nth) {
            const long long e = rd.elem_off + (long long)first * TE + i;
            rd.ring[(e >> RING_GROUP_LOG2) * gstride + (long long)slot * RING_GROUP + (e & (RING_GROUP - 1))] = lds[i];
        }
        return;
    }
    const int cols = rd.lay.cols, nr = rd.lay.rows(), CE = cols * nr;
    const float rce = 1.0f / (float)CE, rcols = 1.0f / (float)cols;
    for (int i = tid; i < ntile * CE; i += nth) {
        const int t = div_small(i, CE, rce), ci = i - t * CE;
        const int r = div_small(ci, cols, rcols), kx = ci - r * cols;
        const int ky = r < rd.lay.row_lo ? r : r + (S - nr);
        const long long e = rd.elem_off + (long long)first * CE + i;
        rd.ring[(e >> RING_GROUP_LOG2) * gstride + (long long)slot * RING_GROUP + (e & (RING_GROUP - 1))] =
            lds[t * TE + ky * (H + 1) + kx];
    }
}

// Register-resident tile FFT (ddm_fft.hip) for the scales it covers;
// returns false for other scales.  xt (optional) holds the row spectra of
// the 256 and 512 tiles between their two passes: n_frames * (M/s)^2 * s^2/2
// complex values (see big_fft_scratch_elems).
bool launch_tile_fft_reg(const PixSrc &ps, const RingDst &rd, int scale, int tps, int tiles, int n_frames,
                         const float2 *tw, float2 *xt, hipStream_t st);

// 64 x 64 / 128 x 128 tile DFT on the matrix cores with split 16-bit
// operands (ddm_dft_mfma.hip); precision is DDM_PREC_BF16X2 or _FP16X2.
hipError_t launch_tile_dft_mfma(const PixSrc &ps, const RingDst &rd, int scale, int tps, int tiles, int n_frames,
                                int precision, hipStream_t st);

}  // namespace ddm
