// Launch interface of the DDM kernels (ddm_kernels.hip), used by the engine
// and by the stage entry points of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "ddm_amd.h"
#include "ddm_internal.h"

namespace ddm {

struct TileFftArgs {
    const void *frames;        // device pointer, frame f at frames + f * frame_bytes
    int n_frames;
    long long frame_bytes;
    int width, bpp, fmt;       // pixel addressing: row stride = width * bpp
    int x_off, y_off;          // ROI origin inside the frame
    int largest_scale, scale;
    float *ring;               // [slot][elems_per_frame] complex
    int ring_slots;
    int first_slot;            // slot of frame 0 of this batch
    long long elems_per_frame;
    long long elem_off;        // offset of this scale within a frame's spectrum
    int precision = DDM_PREC_FP32;  // enum ddm_precision (64 / 128 tiles only)
    // optional buffer for the row spectra of 256 / 512 tiles between their
    // row and column passes (big_fft_scratch_elems values); without it the
    // row spectra go through the ring slot
    float *xt = nullptr;
    // the part of each tile's spectrum the ring holds (device tables);
    // elems == 0: the full spectrum in row order
    TileLayout lay{};
    int family = DDM_FFT_REGISTER;  // enum ddm_fft_family (8 .. 128 tiles)
};


struct LagArgs {
    const float *ring;
    float *acc;                // [lag][elems_per_frame]
    long long elems_per_frame; // padded to a multiple of RING_GROUP
    int ring_slots;
    int slot0;                 // slot of the block's first frame t0
    int n_new;                 // frames in the block, <= lag_block
    int dmax;                  // history frames paired with: min(max lag, t0)
    const int *lags;           // device, ascending
    const int *lags_host;      // the same lags in host memory (the launch plan is made on the host)
    int lag_count;
    int lag_block;             // 32, 64, 128 or 256
    // compensated accumulation: the rounding error of every addition into
    // acc goes to comp ([lag][elems_per_frame], same layout); nullptr: plain
    float *comp = nullptr;
};

struct AzScale {
    long long out_base;        // first result index of this scale
    long long elem_off;
    long long tile_elems;
    double inv_s4;             // 1 / s^4: the reference's (1/s^2)^2 per-term factor
    // scales with small rings: waves [wave_base, wave_base + tiles * wl_n)
    // of the narrow launch, wave k of a tile taking entry wl[wl_off + k]
    long long wave_base = 0;
    int wl_off = 0, wl_n = 0;
};

// Narrow-ring averages: lanes per lag for a ring of cnt elements (the
// smallest power of two >= cnt, 8 .. 64); a wave takes 64 / lanes lags.
__host__ __device__ inline int az_lanes(int cnt) { return cnt <= 8 ? 8 : (cnt <= 16 ? 16 : (cnt <= 32 ? 32 : 64)); }

struct AzArgs {
    const float *acc;
    const float *comp = nullptr;    // compensation terms of acc (same layout), or none
    const double *acc64 = nullptr;  // FP64 accumulators folded out of acc + comp, or none (needs comp)
    long long elems_per_frame;
    const AzScale *scales;     // device
    int scale_count;
    const int *q_off, *q_cnt, *idx;  // device ring tables
    const double *norm;        // device [lag]
    int lambda_count, lag_count;
    double *out;               // device [result_count]
    long long result_count;
    // values [0, wide_items) (the scales with large rings, which come first)
    // take a workgroup each; the others go to narrow_waves waves, each taking
    // one (tile, q) and a group of lags (entries of wave_list: q | group << 16)
    long long wide_items = 0;
    int narrow_first_scale = 0;
    long long narrow_waves = 0;
    const int *wave_list = nullptr;  // device
};

// Ring size from which a scale's averages take a whole workgroup per value.
constexpr int AZ_WIDE_RING = 2048;

hipError_t twiddle_table(const float2 **out);

// The HIP stream an engine launches its kernels on (ddm_engine.hip).
hipStream_t engine_stream(ddm_engine *e);

struct Params;
// What the batched server (ddm_server.hip) uses of an engine: its buffers,
// geometry and the index of its next frame.
struct BatchStream {
    const Params *p;
    float *ring;
    float *acc;
    float *comp;       // nullptr: plain accumulation
    const int *lags;   // device
    int lag_block;
    long long frames;  // index of the next frame
    int precision;
    const TileLayout *lay;  // per scale, device tables
    const int *family;      // per scale, enum ddm_fft_family
};
void engine_batch_stream(ddm_engine *e, BatchStream *bs);
// Uploads n host frames (n <= chunk, <= lag block) into the engine's next
// staging buffer on its copy stream; *b is the buffer, *ev_up its upload
// event, *stage its device address.  Pinned frames are waited for.
int engine_batch_upload(ddm_engine *e, const void *frames, int n, int *b, hipEvent_t *ev_up, const uint8_t **stage);
// Before the first frame: k staging buffers instead of UPLOAD_STAGES (at
// most MAX_UPLOAD_STAGES), so that more chunks can wait for a batched tick.
int engine_batch_stages(ddm_engine *e, int k);
// The n frames of staging buffer b were transformed into the ring and
// through the lag engine by work queued on st: the buffer is free once st
// gets past this point, counts advance, and the accumulators are folded into
// FP64 (on st) when due.
int engine_batch_commit(ddm_engine *e, int n, int b, hipStream_t st);
hipError_t launch_tile_fft(const TileFftArgs &a, hipStream_t st);
// The transforms of every scale of one batch (a[0 .. n), descending, the
// same frames; a[i].xt the start of a pass buffer of fft_pass_elems values).
// With fused, the multi-scale tiler takes the scales it covers when two or
// more of a kind are present (8 .. 128 register transforms; 256 .. 1024),
// the others are launched per scale.  *launches (optional) counts kernel
// launches.  The spectra are the same bit for bit either way.
hipError_t launch_tile_fft_scales(const TileFftArgs *a, int n, bool fused, hipStream_t st, int *launches);
hipError_t launch_lag(const LagArgs &a, hipStream_t st);
hipError_t launch_azimuthal(const AzArgs &a, hipStream_t st);
// acc64[i] += acc[i] + comp[i]; acc[i] = comp[i] = 0 for i < n (n a multiple of 4).
hipError_t launch_acc_fold(float *acc, float *comp, double *acc64, long long n, hipStream_t st);
// Fits A (1 - exp(-F tau)) + B to n curves of L points: out[c][4] = F, A, B, rms.
hipError_t launch_fit(const double *curves, const double *tau, int L, long long n, int iters, double *out,
                      hipStream_t st);

}  // namespace ddm
