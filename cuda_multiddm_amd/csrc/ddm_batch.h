// Batched launches: the tile transforms and lag-engine passes of many
// streams' chunks run as one launch per kernel kind over a device-resident
// job table (a work queue of (stream, scale, tile block, frame) and (stream,
// ring group) items), a workgroup per item.  The table pointers are fixed, so
// a tick's sequence of launches is captured once as a HIP graph and replayed
// with new table contents and grid sizes.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "ddm_device.h"

namespace ddm {

// Header of a job table: njobs jobs covering `total` blocks of work.
struct BatchHdr {
    int njobs;
    int pad;
    long long total;
};

// Block b of a batch belongs to the last job whose first block is <= b
// (jobs are sorted by `first`; b is uniform across the workgroup).
template <class Job>
__device__ __forceinline__ int batch_job(Job *__restrict__ jobs, int nj, long long b) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Tile transforms
// ---------------------------------------------------------------------------
enum FftKind { FFT_LANE = 0, FFT_REG = 1, FFT_ROWS = 2, FFT_COLS = 3 };

// One chunk of one scale of one stream, for one kernel kind: the arguments
// its own launch would take, and its place in the batch.
struct FftJob {
    PixSrc ps;        // the chunk's frames
    RingDst rd;       // the stream's ring, the scale's element offset, slot of frame 0
    float2 *xt;       // pass buffer region of 256 .. 1024 tiles
    int tps, tiles, groups;
    int nblk;         // workgroups of the job's own launch
    long long first;  // first batch block of the job
};

// Launch order of a tick: the one-kernel kinds first, then for every
// 256 .. 1024 tile size its row pass directly followed by its column pass
// (the scales of one chunk share one pass-buffer region, as the scales of a
// per-stream submission do).
struct FftKey {
    int kind, scale, pix;
    bool two_pass() const { return kind == FFT_ROWS || kind == FFT_COLS; }
    bool operator<(const FftKey &o) const {
        if (two_pass() != o.two_pass()) return !two_pass();
        if (scale != o.scale) return two_pass() ? scale > o.scale : scale < o.scale;
        return kind != o.kind ? kind < o.kind : pix < o.pix;
    }
    bool operator==(const FftKey &o) const { return kind == o.kind && scale == o.scale && pix == o.pix; }
};

// Appends the batched launches the transform of n_frames frames of one scale
// takes (one kind, or rows then columns for 256 .. 1024 tiles, which need
// xt: n_frames * M^2 / 2 complex values) as (key, job) pairs in launch
// order, `first` left 0; false when the scale or the frames' pixel layout
// has no batched form (then the caller launches it on its own).
bool fft_batch_jobs(const PixSrc &ps, const RingDst &rd, int scale, int tps, int tiles, int n_frames, float2 *xt,
                    std::vector<std::pair<FftKey, FftJob>> *out);

// One launch of `key` over a job table whose jobs hold `blocks` blocks.
hipError_t launch_fft_batch(const FftKey &key, const BatchHdr *hdr, const FftJob *jobs, long long blocks,
                            const float2 *tw, hipStream_t st);

// ---------------------------------------------------------------------------
// Lag engine
// ---------------------------------------------------------------------------
constexpr int LAG_NT = 1024;                    // threads per workgroup: 16 waves
constexpr int LAG_WAVES = LAG_NT / RING_GROUP;  // waves sharing one ring group

// Work split of one lag launch, the same for every workgroup of a stream's
// pass, computed on the host from the lag list, the block size n and the
// history depth dmax (make_lag_plan).
constexpr int PLAN_MAX_BLOCKS = 252;
struct LagPlan {
    int dense_lags;  // leading consecutive lags on the register window
    int wd;          // window waves: chunk w / sd, segment w % sd
    int sd;          // segments per window chunk
    int ss;          // units per other lag; 0: one wave per lag, direct accumulation
    int hist;        // the other lags go history-major
    int pmin, pmax;  // history-major frame range
    int listed;      // blk_start / blk_list hold each history wave's blocks
    int all_hist;    // the lists cover all LAG_WAVES waves (window waves take blocks after their window)
    int pre;         // frames before the block the register window reads (staged in LDS up to the block size's limit)
    unsigned char blk_start[LAG_WAVES + 1];   // wave k: blk_list[blk_start[k] .. blk_start[k + 1])
    unsigned char blk_list[PLAN_MAX_BLOCKS];  // block b: frames pmin + b * HIST_BLOCK ...
};

// The plan of one pass of n new frames (lag block J) against dmax history
// frames for ascending host lags[L].
void make_lag_plan(int J, const int *lags, int L, int n, int dmax, LagPlan *pl);

// One stream's pass: every ring group of its spectrum.
struct LagJob {
    const float2 *ring;
    float *acc;
    float *comp;      // compensation terms, or nullptr (the batch's kernel decides which it uses)
    const int *lags;  // device, ascending
    long long E;      // padded elements per frame
    int R, slot0, n, dmax, L;
    int plan;         // index into the batch's plan table
    long long first;  // first batch block (ring group) of the job
};

// One launch over a job table of passes with lag block J = 32, 64 or 128
// (all compensated or all plain) holding `groups` ring groups in all.
hipError_t launch_lag_batch(int J, bool comp, const BatchHdr *hdr, const LagJob *jobs, const LagPlan *plans,
                            long long groups, hipStream_t st);

}  // namespace ddm
