// Internal declarations shared by the host-side translation units of the
// DDM engine (validation, q-ring tables, synthetic video, writers) and the
// HIP engine/kernels.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "ddm_amd.h"

namespace ddm {

// Largest tile size the FFT kernels are instantiated for.
constexpr int kMaxScale = 2048;

// Spectrum ring layout: elements in groups of RING_GROUP, each group holding
// all slots contiguously (see ring_index in ddm_kernels.hip).  Frames are
// padded to a whole number of groups.
constexpr int RING_GROUP_LOG2 = 6;
constexpr int RING_GROUP = 1 << RING_GROUP_LOG2;
inline long long pad_elems(long long e) { return (e + RING_GROUP - 1) / RING_GROUP * RING_GROUP; }

// Device buffers host frames are uploaded into, used in turn.
constexpr int UPLOAD_STAGES = 3;
// At most this many (a batched server that groups chunks gives its streams more).
constexpr int MAX_UPLOAD_STAGES = 8;

// Complex values of the tile FFT's between-pass buffer (TileFftArgs::xt) a
// batch of n frames needs for tiles of size `scale` cut from an M x M region
// (0 for scales that do not use it).
inline long long big_fft_scratch_elems(int M, int scale, int n) {
    return (scale == 256 || scale == 512 || scale == 1024) ? (long long)n * M * M / 2 : 0;
}

// The whole pass buffer of a batch of n frames: one scale's region when the
// 256 .. 1024 scales take turns (per-scale launches), one region per scale
// when the multi-scale tiler's row bands fill them all before the column
// passes (two or more such scales and fused = true).
inline long long fft_pass_elems(int M, const std::vector<int> &scales, int n, bool fused) {
    int big = 0;
    for (int s : scales) big += big_fft_scratch_elems(M, s, 1) > 0;
    const long long one = (long long)n * M * M / 2;
    return big == 0 ? 0 : (fused && big >= 2 ? big * one : one);
}

// The part of an s x s tile's half spectrum [ky][kx <= s/2] a spectrum ring
// holds, and its order.  Only elements some q ring contains reach the
// averages, so an engine may keep just a superset of those:
//  * LAY_FULL: every element, row after row (index ky * (s/2 + 1) + kx);
//  * LAY_ROWS: row after row, row ky keeping columns kx < start[ky + 1] -
//    start[ky], the first of them at start[ky] (s + 1 entries);
//  * LAY_COLS: column after column, column kx keeping rows ky < band[kx]
//    and the last start[kx + 1] - start[kx] - band[kx] rows (the low |ky|,
//    both signs), the first of them at start[kx] (s/2 + 2 entries; band has
//    s/2 + 1).  The 256 .. 1024 transforms end in a column pass, which then
//    writes one contiguous run per column.
// A launch's layout holds device pointers, LayoutTables host vectors.
enum LayoutKind { LAY_FULL = 0, LAY_ROWS = 1, LAY_COLS = 2 };
struct TileLayout {
    int kind = LAY_FULL;
    int elems = 0;      // kept elements per tile
    int pass_cols = 0;  // row-pass columns the column pass reads (multiple of 16, <= s/2)
    int row_lanes = 64; // LAY_ROWS: lanes storing one row (a power of two >= the widest row, at most 64)
    const int *start = nullptr;
    const int *band = nullptr;
    static TileLayout full(int s) {
        TileLayout l;
        l.elems = s * (s / 2 + 1);
        l.pass_cols = s / 2;
        return l;
    }
};

// Host tables of one scale's layout (shared by every engine with the same
// scale, lambdas, tolerance and spectrum mode).  pos[full index] is the
// kept index or -1, for every kind.
struct LayoutTables {
    int kind = LAY_FULL;
    int elems = 0;
    int pass_cols = 0;
    int row_lanes = 64;
    std::vector<int> start, band, pos;
};

// Frames of device-resident input (or of one large host call) transformed
// per FFT batch by default: whole chunks, up to about 128 frames.  A 30-frame
// chunk of one scale is too little work to fill the device; ~128 frames is,
// while the 512-tile row spectra of a batch still fit in the last-level
// cache for the column pass.
inline int default_fft_batch(int chunk) { return chunk * (128 / chunk > 1 ? 128 / chunk : 1); }

// Ring slots are addressed per ring group with 32-bit byte offsets:
// slots * RING_GROUP * 8 bytes must stay below 2^31.
constexpr long long kMaxRingSlots = (1LL << 31) / (RING_GROUP * 8);

// Largest region (M) whose engine uses the multi-scale tiler by default.
constexpr int kTilerFusedMaxRegion = 256;

// Frames per lag-engine pass when the descriptor leaves it to the engine.
constexpr int kDefaultLagBlock = 256;
// Lag values above this are rejected (ring memory would be absurd).
constexpr int kMaxLag = 1 << 20;

struct ScaleGeom {
    int s = 0;            // tile size
    int tps = 0;          // tiles per side, M / s
    int tiles = 0;        // tps * tps
    long long tile_elems = 0;  // lay->elems: s * (s/2 + 1) for the full spectrum
    long long elem_off = 0;    // offset of this scale inside one frame's spectrum
    std::shared_ptr<const LayoutTables> lay;  // the part of each tile's spectrum the ring holds
};

// Descriptor after validation and sorting: scales descending, lags and
// lambdas ascending (src/DDM.cu:319-321).
struct Params {
    int width = 0, height = 0, bpp = 1, fmt = DDM_PIX_U8;
    int x_off = 0, y_off = 0;
    std::vector<int> scales;
    std::vector<int> lags;
    std::vector<float> lambdas;
    float tol = 1.2f;
    int chunk = 30;
    int norm_mode = DDM_NORM_FRAMES;
    int lag_block = 0;
    int acc_mode = DDM_ACC_COMPENSATED;
    int spectrum = DDM_SPECTRUM_MASKED;
    // derived
    int M = 0;
    int max_lag = 0;
    int ring_slots = 0;
    long long elems_per_frame = 0;
    long long elems_padded = 0;  // ring / accumulator row stride
    long long result_count = 0;
    std::vector<ScaleGeom> geom;
    std::vector<std::string> warnings;
};

// Validates a descriptor without touching any device.  Returns DDM_OK or a
// negative code with the reference's message text where one exists.
int check_desc(const ddm_desc *d, Params *out, std::string *msg);

// Device memory an engine for these parameters allocates (an upper bound:
// the q-ring tables are counted at their largest possible size).
size_t device_footprint(const Params &p);

// Element lists of every q ring of every scale: for scale i and q j the
// member elements (ky * (s/2+1) + kx, ascending), as places in the scale's
// stored tile layout (ScaleGeom::lay), are
// idx[off[i*Q + j] .. off[i*Q + j] + cnt[i*Q + j]).
struct RingTables {
    std::vector<int> idx;
    std::vector<int> off;
    std::vector<int> cnt;
};
int build_ring_tables(const Params &p, RingTables *out);

// Pixel format helpers.
inline int pixel_bytes_needed(int fmt) { return fmt == DDM_PIX_U8 ? 1 : 2; }

// Twiddle table for every power-of-two length n <= kMaxScale: entries
// [n, 2n) hold exp(-2 pi i m / n), m < n, as (re, im) floats rounded from
// double precision.
void fill_twiddles(std::vector<float> *tw);

}  // namespace ddm
