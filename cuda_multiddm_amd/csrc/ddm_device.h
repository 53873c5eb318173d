// Device helpers shared by the tile / FFT kernels: complex arithmetic, pixel
// addressing, and the blocked spectrum-ring layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ddm_amd.h"
#include "ddm_internal.h"

namespace ddm {

// Complex product with its fused multiply-adds written out: the rounding is
// then the same in every kernel the transforms are inlined into (left to the
// compiler, a.x * b.x - a.y * b.y may fuse either product), so a tile's
// spectrum does not depend on which launch computed it.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
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
// addressed by (ky, kx) of the tile's half spectrum, for the row-order
// layouts (LAY_FULL, LAY_ROWS); elements the layout leaves out are dropped.
struct TileOut {
    float2 *ring;
    long long base;  // element index of the tile's first element
    int slot, R, S;
    TileLayout lay;
    __device__ __forceinline__ void put(int ky, int kx, float2 v) const {
        if (lay.kind == LAY_FULL) {
            ring[ring_index(base + ky * (S / 2 + 1) + kx, slot, R)] = v;
            return;
        }
        const int a = lay.start[ky];
        if (kx < lay.start[ky + 1] - a) ring[ring_index(base + a + kx, slot, R)] = v;
    }
    // element (ky, kx) of a tile in the LAY_FULL layout
    __device__ __forceinline__ float2 &full(int ky, int kx) const {
        return ring[ring_index(base + (long long)ky * (S / 2 + 1) + kx, slot, R)];
    }
};

__device__ __forceinline__ int ring_slot(const RingDst &rd, int f) {
    const int slot = rd.first_slot + f;
    return slot - ((slot >= rd.ring_slots) ? rd.ring_slots : 0);  // f < ring_slots
}

__device__ __forceinline__ TileOut tile_out(const RingDst &rd, int f, int tile, int S) {
    return TileOut{rd.ring, rd.elem_off + (long long)tile * rd.lay.elems, ring_slot(rd, f), rd.ring_slots, S,
                   rd.lay};
}

// The largest k < n with tab[k] <= x, for tab ascending in LDS and
// tab[0] <= x: the row (column) of kept element x when tab holds the row
// (column) starts; empty rows share their start with the next row and are
// skipped.
__device__ __forceinline__ int find_start(const int *tab, int n, int x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Copies the row starts of a LAY_ROWS layout (S + 1 entries) into LDS at
// the start of a tile kernel, so that its store loop (after a barrier) does
// not wait for them.
template <int S>
__device__ __forceinline__ void load_row_starts(const RingDst &rd, int *rtab, int tid, int nth) {
    if (rd.lay.kind == LAY_ROWS)
        for (int i = tid; i <= S; i += nth) rtab[i] = rd.lay.start[i];
}

// The tiles one workgroup transforms: local tile t < ntile is tile
// first + t % 2^wsh + (t >> wsh) * tps of its scale (tps tiles per side), so
// a per-scale launch takes runs of consecutive tiles (2^wsh >= ntile) and
// the multi-scale tiler the (2^wsh)^2 tiles of one block.
struct TileSet {
    int first, wsh, ntile;
    __device__ __forceinline__ int tile(int t, int tps) const {
        return first + (t & ((1 << wsh) - 1)) + (t >> wsh) * tps;
    }
};

// The spectra of the tiles of ts for batch frame f, staged in LDS at
// lds[t * S * (S/2 + 1) + ky * (S/2 + 1) + kx], stored to the ring in the
// row-order layout rd.lay (LAY_FULL or LAY_ROWS, whose row starts rtab holds
// in LDS) by nth threads (a multiple of 64): consecutive lanes write
// consecutive ring elements.
template <int S>
__device__ __forceinline__ void store_tiles(const RingDst &rd, const float2 *lds, const int *rtab, int f,
                                            const TileSet &ts, int tps, int tid, int nth) {
    constexpr int TE = S * (S / 2 + 1);
    const int slot = ring_slot(rd, f);
    const long long gstride = (long long)rd.ring_slots * RING_GROUP;
    const int CE = rd.lay.elems;
    const int ntile = ts.ntile;
    if (rd.lay.kind == LAY_FULL) {
        if (ntile <= (1 << ts.wsh)) {
            // consecutive tiles: one run of ring elements
            const long long e0 = rd.elem_off + (long long)ts.first * CE;
            for (int i = tid; i < ntile * TE; i += nth) {
                const long long e = e0 + i;
                rd.ring[(e >> RING_GROUP_LOG2) * gstride + (long long)slot * RING_GROUP + (e & (RING_GROUP - 1))] =
                    lds[i];
            }
            return;
        }
        for (int t = 0; t < ntile; ++t) {
            const long long e0 = rd.elem_off + (long long)ts.tile(t, tps) * CE;
            for (int i = tid; i < TE; i += nth) {
                const long long e = e0 + i;
                rd.ring[(e >> RING_GROUP_LOG2) * gstride + (long long)slot * RING_GROUP + (e & (RING_GROUP - 1))] =
                    lds[t * TE + i];
            }
        }
        return;
    }
    // LAY_ROWS: row_lanes lanes per row (64 / row_lanes rows per wave), each
    // row's kept columns a contiguous run
    const int lpr = rd.lay.row_lanes, rpw = 64 / lpr;
    const int lane = tid & 63, sub = lane / lpr, kl = lane - sub * lpr;
    const int step = (nth >> 6) * rpw;
    for (int j = (tid >> 6) * rpw + sub; j < ntile * S; j += step) {
        const int t = j / S, ky = j % S;
        const int a = rtab[ky], n = rtab[ky + 1] - a;
        const long long eb = rd.elem_off + (long long)ts.tile(t, tps) * CE + a;
        const float2 *src = lds + t * TE + ky * (S / 2 + 1);
        for (int kx = kl; kx < n; kx += lpr) {
            const long long e = eb + kx;
            rd.ring[(e >> RING_GROUP_LOG2) * gstride + (long long)slot * RING_GROUP + (e & (RING_GROUP - 1))] =
                src[kx];
        }
    }
}

// Multi-scale tiler (ddm_fft.hip).  One scale of a batch: its ring
// destination, tile size and (256 .. 1024) its own pass-buffer region.
struct TilerScale {
    RingDst rd;
    int s;
    float2 *xt;
};
// Scales the tiler's launches cover in an M x M region: 8 .. 128 tiles
// (register transforms) from 128 x 128 blocks; 256 .. 1024 row passes from
// row bands.
bool tiler_block_scale(int scale, int M);
bool tiler_band_scale(int scale, int M);
// The 8 .. 128 scales sc[0 .. n) (descending) of n_frames frames in one
// launch, every 128 x 128 block's pixels read once.
hipError_t launch_tiler_blocks(const PixSrc &ps, const TilerScale *sc, int n, int M, int n_frames, const float2 *tw,
                               hipStream_t st);
// The row passes of the 256 .. 1024 scales sc[0 .. n) (descending) in one
// launch, every row band's pixels read once, then each scale's column pass.
hipError_t launch_tiler_bands(const PixSrc &ps, const TilerScale *sc, int n, int M, int n_frames, const float2 *tw,
                              hipStream_t st);

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
