/*
 * ddm_amd.h -- C ABI of the MI355X (gfx950) multi-scale DDM engine.
 *
 * The engine replaces the analysis path of zhangyuyi99/cuda_multiDDM:
 * runDDM(...) (declared src/main.cpp:40-64, defined src/DDM.cu:268-292),
 * i.e. tile cut -> 2-D R2C FFT -> lag difference -> masked azimuthal average,
 * behind the reference's ./ddm CLI (src/main.cpp:104-342).  runDDM owns every
 * buffer and runs a whole video in one call; this ABI splits the same work
 * into create / submit frames / read results so that the CLI, tests and a
 * serving loop can drive it.  Plain C types only.
 *
 * Conventions
 *   - Every function returns DDM_OK (0) or a negative DDM_E* code.  The
 *     message of the last failure is available from ddm_last_error().  No
 *     function terminates the process (the reference calls exit(), see
 *     src/debug.cpp:10-15, 33-45).
 *   - The engine owns all device memory and its pinned staging buffers; the
 *     caller owns descriptors, parameter arrays and output arrays.
 *     Parameter arrays are copied and sorted inside ddm_create; the caller's
 *     arrays are never modified (the reference sorts them in place,
 *     src/DDM.cu:319-321).
 *   - One engine is used from one host thread at a time.  Several engines
 *     may live in one process and share the GPU.
 *   - Result layout: results[scale][tile][q][lag] with scales descending,
 *     tiles in the reference order (tile = ty * (M/s) + tx, M = largest
 *     scale), q in ascending-lambda order, lags ascending.
 */
#ifndef DDM_AMD_H
#define DDM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDM_ABI_VERSION 3

/* status codes */
#define DDM_OK 0
#define DDM_EINVAL (-1)   /* invalid argument */
#define DDM_ENOMEM (-2)   /* host or device allocation failed */
#define DDM_EDEVICE (-3)  /* HIP runtime error / no device */
#define DDM_EROI (-4)     /* ROI (largest scale at x_off, y_off) outside the frame */
#define DDM_ESCALE (-5)   /* scale not a power of two, or repeated */
#define DDM_ELAG (-6)     /* lag negative or repeated */
#define DDM_ELAMBDA (-7)  /* lambda negative or repeated */
#define DDM_ESTATE (-8)   /* call not valid in the engine's current state */
#define DDM_ERANGE (-9)   /* frame or index outside what the engine holds */
#define DDM_EBUSY (-10)   /* server: no stream slot or not enough device budget left */

/* Pixel formats.  A pixel is bytes_per_pixel bytes; only channel 0 is used.
 * DDM_PIX_U8 reads byte 0 of each pixel (parseBufferScalePow2,
 * src/DDM_kernel.cuh:74-75).  The 16-bit formats read bytes 0-1 as one
 * value (the reference truncates 16-bit input to its low byte). */
enum ddm_pixel_format { DDM_PIX_U8 = 0, DDM_PIX_U16LE = 1, DDM_PIX_U16BE = 2 };

/* Normalisation of I(q, tau).  DDM_NORM_FRAMES multiplies by 1/N with N the
 * frames submitted since the last reset, as the reference does
 * (src/DDM.cu:92, 875).  DDM_NORM_PAIRS divides by the number of pairs
 * (t, t + tau) summed for each lag instead. */
enum ddm_norm_mode { DDM_NORM_FRAMES = 0, DDM_NORM_PAIRS = 1 };

/* Arithmetic of the 64 x 64 and 128 x 128 tile transforms (ddm_set_precision).
 * DDM_PREC_FP32: the FP32 FFT (default; every other scale always uses it).
 * DDM_PREC_BF16X2 / DDM_PREC_FP16X2: the 2-D DFT as matrix products on the
 * matrix cores, each operand split into a high and a low BF16 / FP16 part,
 * FP32 accumulation. */
enum ddm_precision { DDM_PREC_FP32 = 0, DDM_PREC_BF16X2 = 1, DDM_PREC_FP16X2 = 2 };

/* How the lag sums accumulate over a stream (ddm_desc.accumulation).
 * DDM_ACC_COMPENSATED (default): every lag-engine launch adds its FP32
 * partial sums to FP32 accumulators with the exact rounding error of each
 * addition kept in FP32 compensation terms, and once the accumulators hold
 * 10^4 frame pairs they are folded into FP64 accumulators on the device, so
 * the error does not grow with the stream's length.  DDM_ACC_PLAIN: FP32
 * accumulators only, as the reference adds each term in FP32
 * (src/DDM_kernel.cuh:85-101, merge src/DDM_kernel.cuh:127-139). */
enum ddm_accumulation { DDM_ACC_COMPENSATED = 0, DDM_ACC_PLAIN = 1 };

/* Pairs of frames the FP32 accumulators take before they are folded into
 * FP64 (DDM_ACC_COMPENSATED). */
#define DDM_ACC_FOLD_PAIRS 10000

/* Which spectrum elements an engine keeps (ddm_desc.spectrum).
 * DDM_SPECTRUM_MASKED (default): only the I(q, tau) averages read the lag
 * sums, and they read only elements inside some q ring, so for each scale
 * the ring and accumulators hold only the part of a tile's half spectrum
 * that reaches the rings (for tiles up to 128, each row's columns up to its
 * last ring member; for 256 .. 1024, each column's low |ky| rows down to
 * its last ring member), when that saves at least a tenth of the elements
 * (scales 8 .. 1024).
 * Results are bit-identical to DDM_SPECTRUM_FULL; per-element reads
 * (ddm_get_accumulators, ddm_get_spectra) return NaN for elements left
 * out.  DDM_SPECTRUM_FULL: every element, as the reference computes them
 * (needed for its non-averaged -K output and the -A spectrum dump). */
enum ddm_spectrum { DDM_SPECTRUM_MASKED = 0, DDM_SPECTRUM_FULL = 1 };

/* Tile transform family of a scale (ddm_set_fft_family, ddm_autotune).
 * DDM_FFT_REGISTER (default): register FFTs, radix 16 / 32 with a lane per
 * row pair for 8 .. 32 tiles and 16 / 32 lanes per sequence for 64 .. 1024.
 * DDM_FFT_LDS: Stockham radix 4 / 2 passes in LDS, several tiles per
 * workgroup (available for 8 .. 128 tiles; 2, 4 and 2048 tiles always use
 * it).  Both compute the same transform; the autotuner keeps the faster. */
enum ddm_fft_family { DDM_FFT_REGISTER = 0, DDM_FFT_LDS = 1 };

/* Multi-scale tiler (ddm_set_tiler).  DDM_TILER_FUSED: each frame's pixels
 * are read once for all 8 .. 128 tile scales (128 x 128 blocks staged in
 * LDS, every scale's tiles cut from there) and once for the row passes of
 * all 256 .. 1024 scales (row bands), instead of once per scale; the FFT
 * pass buffer then holds one region per 256 .. 1024 scale.
 * DDM_TILER_PER_SCALE: one transform launch per scale, each reading the
 * frame, as the reference does (src/DDM.cu:171-192).  The spectra are the
 * same bit for bit.  Default: fused for regions up to 256 x 256 (short,
 * launch-bound transforms), per scale above (bandwidth-bound); ddm_autotune
 * times both. */
enum ddm_tiler { DDM_TILER_FUSED = 0, DDM_TILER_PER_SCALE = 1 };

typedef struct ddm_desc {
    int width;            /* frame width in pixels */
    int height;           /* frame height in pixels */
    int bytes_per_pixel;  /* pixel stride in bytes (info.bpp) */
    int pixel_format;     /* enum ddm_pixel_format */
    int x_off, y_off;     /* ROI origin (-x / -y) */
    const int *scales;    /* tile sizes, powers of two (-S file) */
    int scale_count;
    const int *lags;      /* tau values in frames (-T file); may exceed chunk_frames */
    int lag_count;
    const float *lambdas; /* wavelengths in pixels (-Q file); q = scale / lambda */
    int lambda_count;
    float mask_tolerance; /* ring [q, tol * q] (-t: 1 + percent/100) */
    int chunk_frames;     /* frames per upload / FFT batch (-C) */
    int norm_mode;        /* enum ddm_norm_mode */
    int lag_block;        /* frames per lag-engine pass: 0 = automatic (256), else 32, 64, 128 or 256 */
    int accumulation;     /* enum ddm_accumulation (0 = compensated with FP64 folding) */
    int spectrum;         /* enum ddm_spectrum (0 = keep the q-ring band only) */
} ddm_desc;

/* Sizes the engine derived from a descriptor (after sorting). */
typedef struct ddm_info {
    int largest_scale;         /* M */
    int scale_count, lag_count, lambda_count;
    int ring_slots;            /* spectra kept resident: chunk_frames + lag_block + max lag */
    int lag_block;             /* frames per lag-engine pass (after ddm_autotune: its choice) */
    long long elems_per_frame; /* complex spectrum elements per frame, all scales */
    long long elems_per_frame_padded; /* rounded up to a multiple of 64: ring and accumulator row stride */
    long long result_count;    /* sum over scales of tiles * lambda_count * lag_count */
    long long frames_submitted;   /* since creation */
    long long frames_since_reset; /* since creation or the last ddm_reset */
    size_t device_bytes;       /* device memory held by the engine */
    size_t pinned_bytes;       /* pinned host memory held by the engine */
} ddm_info;

/* Per-stage device time, collected with HIP events on the engine's stream
 * when timing is on (ddm_set_timing). */
typedef struct ddm_stage_stats {
    double fft_ms;        /* tiler + FFT kernels */
    double lag_ms;        /* lag-engine kernels */
    double azimuthal_ms;  /* masked azimuthal average */
    long long fft_launches, lag_launches, azimuthal_launches;
    long long chunks;     /* chunk latencies recorded */
} ddm_stage_stats;

typedef struct ddm_engine ddm_engine;

/* Validates the descriptor (before touching the device), then allocates the
 * spectrum ring, accumulators and q-ring tables on the current device.
 * Replaces the setup half of runDDM (src/DDM.cu:294-762). */
int ddm_create(const ddm_desc *desc, ddm_engine **out);

/* Validation only, no device access: returns the code ddm_create would
 * return for this descriptor and fills info (may be NULL). */
int ddm_validate(const ddm_desc *desc, ddm_info *info, char *msg, size_t msg_len);

void ddm_destroy(ddm_engine *eng);

/* Appends n_frames frames (n_frames * width * height * bytes_per_pixel bytes,
 * row-major) from host memory.  Copies the ROI rows into the engine's pinned
 * double buffers and uploads them asynchronously.  Replaces the chunk loop
 * of runDDM (src/DDM.cu:765-843): H2D, parseChunk, analyseChunk. */
int ddm_submit_host(ddm_engine *eng, const void *frames, int n_frames);

/* Same as ddm_submit_host for frames already in device memory.  The buffer
 * must stay valid until ddm_sync returns. */
int ddm_submit_device(ddm_engine *eng, const void *frames, int n_frames);

/* Waits for all submitted work. */
int ddm_sync(ddm_engine *eng);

/* Runs the lag engine on frames still waiting for a full lag block
 * (asynchronous).  ddm_results, ddm_reset and ddm_get_accumulators flush by
 * themselves; latency-sensitive callers flush after every chunk. */
int ddm_flush(ddm_engine *eng);

/* Masked azimuthal average of everything accumulated since the last reset:
 * results[info.result_count] doubles, layout [scale][tile][q][lag].
 * Replaces analyse_accums / analyseFFTDevice (src/DDM.cu:72-146,
 * src/azimuthal_average.cu:292-346). */
int ddm_results(ddm_engine *eng, double *results);

/* Fits the Brownian model I(q, tau) = A (1 - exp(-F tau)) + B to every
 * curve of ddm_results on the device (FP64 Levenberg-Marquardt, one thread
 * per curve), tau = lag / fps in seconds.  fit[curve][4] = F, A, B, rms,
 * curve = (scale, tile, q) in result order; D[scale][tile] =
 * exp(mean(ln F - 2 ln q)) with q = 2 pi / lambda, over the curves with
 * F > 0, A > 0 and 1/F inside the lag range (0 if none).  Either output
 * may be NULL.  Replaces fit_ISF and plot_diff_coeff of
 * resources/python_tools/plot_averaged.py:61-136. */
int ddm_fit_results(ddm_engine *eng, float fps, double *fit, double *D);

/* Zeroes the lag accumulators and the frame count used for normalisation;
 * the spectrum ring is kept, so pairs reaching back across the reset are
 * still formed.  The -G purge of src/DDM.cu:800-826. */
int ddm_reset(ddm_engine *eng);

int ddm_get_info(const ddm_engine *eng, ddm_info *info);

/* Sorted parameters as the engine uses them. */
int ddm_get_params(const ddm_engine *eng, int *scales, int *lags, float *lambdas);

/* Lag sums of one scale, [lag][tile][ky][kx] (kx <= s/2), multiplied by
 * 1/s^4 like the reference accumulator (src/DDM_kernel.cuh:95-98).
 * Backs the -K dump. */
int ddm_get_accumulators(ddm_engine *eng, int scale_idx, float *out);

/* Spectrum of one frame still held in the ring, one scale,
 * [tile][ky][kx][re, im] unnormalised.  Backs the -A dump and FFT tests. */
int ddm_get_spectra(ddm_engine *eng, long long frame, int scale_idx, float *out);

int ddm_set_timing(ddm_engine *eng, int on);

/* Selects the arithmetic of the 64 x 64 and 128 x 128 tile transforms for the
 * frames submitted from now on (enum ddm_precision).  No counterpart in the
 * reference, which always uses cuFFT in FP32. */
int ddm_set_precision(ddm_engine *eng, int mode);

/* Autotuner: before the first frame, times the lag-engine block (64, 128 or
 * 256 frames, at most the descriptor's lag_block), the FFT batch of
 * device-resident input (1, 2, 4 or 8 chunks), for every 8 .. 128 tile
 * scale the tile transform family (enum ddm_fft_family: radix 16 / 32 in
 * registers or radix 4 / 2 in LDS), and the tiler mode (enum ddm_tiler) on
 * this engine's buffers with synthetic frames, and keeps the fastest per
 * frame (a candidate must beat the default by 2 %).  cache_path (may be NULL): text file of earlier choices, keyed by
 * device and geometry; a hit skips the timing, a miss appends the new
 * choice.  Writes the choice to *lag_block / *fft_batch_frames (either may
 * be NULL; ddm_get_fft_families reads the families).  Results do not depend
 * on the choice beyond FP32 rounding of the transforms.  No counterpart in
 * the reference, whose launch shapes are compile-time constants
 * (src/constants.hpp:11-13). */
int ddm_autotune(ddm_engine *eng, const char *cache_path, int *lag_block, int *fft_batch_frames);

/* The tile transform family of scale scale_idx (sorted descending), before
 * the first frame: DDM_EINVAL for a family the scale has no kernel for.  No
 * counterpart in the reference (cuFFT picks its own kernels). */
int ddm_set_fft_family(ddm_engine *eng, int scale_idx, int family);

/* Each scale's transform family (scale_count entries, scales descending). */
int ddm_get_fft_families(const ddm_engine *eng, int *families);

/* The multi-scale tiler mode (enum ddm_tiler), before the first frame.
 * Replaces the per-scale cut of parseBufferScalePow2 in the chunk loop
 * (src/DDM.cu:171-192, src/DDM_kernel.cuh:43-78).  The device footprint
 * (ddm_validate, server admission) counts the default mode's pass buffer;
 * fusing two or more 256 .. 1024 scales allocates one region per scale. */
int ddm_set_tiler(ddm_engine *eng, int mode);

/* The engine's tiler mode (enum ddm_tiler), or DDM_EINVAL. */
int ddm_get_tiler(const ddm_engine *eng);

/* Allocates now the transient device buffers that a submission of up to
 * max_frames frames per call would otherwise allocate on first use (the FFT
 * pass buffer of 256 .. 1024 tiles for one FFT batch, and batch-sized upload
 * staging when max_frames exceeds a chunk), so that such a submission starts
 * its kernels at once.  Results do not change.  No counterpart in the
 * reference, which allocates everything up front for its fixed chunk
 * (src/DDM.cu:493-634). */
int ddm_reserve(ddm_engine *eng, int max_frames);
int ddm_get_stage_stats(ddm_engine *eng, ddm_stage_stats *stats);
/* Device time of each submitted chunk (tiler+FFT+lag), in submission order;
 * returns the number written (<= max). */
int ddm_get_chunk_times(ddm_engine *eng, float *ms, int max);

/* Message of the last failure of eng; with eng == NULL, of the last
 * ddm_create (or refused ddm_k_tile_fft_noalloc) on the calling thread. */
const char *ddm_last_error(const ddm_engine *eng);
const char *ddm_status_string(int code);

/* -------------------------------------------------------------------------
 * Serving: many streams on one device.  No counterpart in the reference
 * (it analyses one video per process, src/DDM.cu:268-896).  A stream is
 * admitted when its device footprint (ddm_info.device_bytes) fits in the
 * budget left; each stream runs on its own HIP stream, so chunks of
 * different streams overlap on the device.  Every submitted chunk is
 * flushed through the lag engine and its chunk-to-result latency (host
 * submission -> chunk accumulated on the device) is recorded.  Results of a
 * stream depend only on its own frames.  The server functions may be called
 * from several host threads at once (chunks of one stream keep their order
 * when that stream is fed from one thread); ddm_server_close waits for calls
 * still using the stream, and ddm_server_destroy must not overlap any call.
 * ------------------------------------------------------------------------- */
typedef struct ddm_server ddm_server;

typedef struct ddm_server_info {
    size_t budget_bytes, used_bytes;
    int streams, max_streams;
    long long frames, chunks, chunks_in_flight;
    long long ticks;        /* batched: ticks run */
    long long graph_ticks;  /* batched: ticks run as a HIP graph */
    long long tick_items;   /* batched: queue items (workgroups) of all ticks */
} ddm_server_info;

/* Flags of ddm_server_create2.
 * DDM_SERVER_BATCHED: submissions only upload and queue their chunk; a
 *   scheduler thread runs the waiting chunks of all streams (the oldest of
 *   each) as one tick: every tile-transform kind (per tile size) and every
 *   lag-block size is one launch over the tick's work queue of (stream,
 *   scale, tile block, frame) and (stream, ring group) items, the tick's
 *   launches captured as a HIP graph.  Chunks arriving while a tick runs go
 *   into the next (at most two ticks in flight); each stream keeps at most
 *   three chunks waiting (further submissions block).  A stream's chunks are
 *   submitted in pieces of at most min(chunk, lag block) frames.  Results
 *   are bit-identical to the per-stream mode.
 * DDM_SERVER_NO_GRAPH: batched ticks are launched directly, not as graphs.
 * DDM_SERVER_GROUP: a batched tick takes all of a stream's waiting chunks up
 *   to its lag block and runs them as one lag pass (a chunk that arrives
 *   alone still runs alone): under load the stream's history is read once
 *   for several chunks.  Results then depend on how chunks happened to be
 *   grouped, within the FP32 path's error bound, instead of being
 *   bit-identical to the per-stream mode. */
#define DDM_SERVER_BATCHED 1
#define DDM_SERVER_NO_GRAPH 2
#define DDM_SERVER_GROUP 4

/* With DDM_SERVER_GROUP: a stream's waiting chunks may wait up to hold_ms
 * (from the arrival of the oldest) for more of the stream's chunks, so that a
 * lag pass takes several; they go at once when a pass is full, when the
 * stream's staging buffers are all in use, or when results, close or drain
 * wait for the stream.  0 (the default) never holds a chunk back. */
int ddm_server_set_hold(ddm_server *srv, double hold_ms);

/* device_budget 0 = 90 % of the device memory free at creation (a batched
 * server takes 1 GiB of it for the tick pass buffer). */
int ddm_server_create(size_t device_budget, int max_streams, ddm_server **out);
int ddm_server_create2(size_t device_budget, int max_streams, int flags, ddm_server **out);
void ddm_server_destroy(ddm_server *srv);
/* DDM_EBUSY when no slot is free or the footprint does not fit yet. */
int ddm_server_open(ddm_server *srv, const ddm_desc *desc, int *stream_id);
/* Host frames of one chunk (any count); returns once the chunk is queued. */
int ddm_server_submit(ddm_server *srv, int stream_id, const void *frames, int n_frames);
int ddm_server_results(ddm_server *srv, int stream_id, double *results);
/* The stream's engine, for ddm_get_info / ddm_get_params / ddm_reset. */
int ddm_server_engine(ddm_server *srv, int stream_id, ddm_engine **eng);
/* Waits until every submitted chunk of every stream is done. */
int ddm_server_drain(ddm_server *srv);
int ddm_server_close(ddm_server *srv, int stream_id);
/* Chunk-to-result latencies (ms) of finished chunks of one stream, or of
 * all streams with stream_id < 0; returns the number written. */
int ddm_server_latencies(ddm_server *srv, int stream_id, float *ms, int max);
int ddm_server_stats(ddm_server *srv, ddm_server_info *info);
/* Message of the last failing server call (ddm_server_create included) on
 * the calling thread; srv is not used. */
const char *ddm_server_last_error(const ddm_server *srv);

/* -------------------------------------------------------------------------
 * Stage entry points on caller-owned device buffers (the kernels the
 * engine launches).  `stream` is a hipStream_t (NULL = default stream).
 * Ring layout: complex floats (interleaved re, im).  elem runs over all
 * scales' tiles of a frame, elems_per_frame is a multiple of 64 (padded),
 * and slot = frame index mod ring_slots.  Elements are grouped by 64 and a
 * group keeps all its slots together:
 *   ring[((elem / 64) * ring_slots + slot) * 64 + elem % 64].
 * Accumulators: acc[lag][elem], row stride elems_per_frame.
 * ------------------------------------------------------------------------- */

/* Tile cut + unnormalised 2-D R2C FFT of n_frames frames for one scale s
 * (parseBufferScalePow2 + cufftExecR2C, src/DDM.cu:188-192).  Frame f is at
 * frames + f * frame_bytes; its spectra go to ring slot (first_frame + f) mod
 * ring_slots at element offset elem_off, layout [tile][ky][kx]. */
int ddm_k_tile_fft(const void *frames, int n_frames, long long frame_bytes,
                   int width, int height, int bytes_per_pixel, int pixel_format,
                   int x_off, int y_off, int largest_scale, int scale,
                   float *ring, int ring_slots, long long elems_per_frame,
                   long long elem_off, long long first_frame, void *stream);

/* ddm_k_tile_fft without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops or harnesses that forbid hipMalloc).  The
 * pass buffer of 256 .. 1024 tiles is a static 64 MiB device array per GPU
 * (held until the process ends) shared by every call on that GPU, so calls
 * that overlap on one GPU race on it; frames whose pass buffers exceed it are
 * transformed in several batches in stream order, and a call whose single
 * frame needs more (largest_scale^2 / 2 complex values > 2^23) is refused
 * with DDM_ERANGE (message in ddm_last_error(NULL)).  Arguments and results
 * as ddm_k_tile_fft. */
int ddm_k_tile_fft_noalloc(const void *frames, int n_frames, long long frame_bytes,
                           int width, int height, int bytes_per_pixel, int pixel_format,
                           int x_off, int y_off, int largest_scale, int scale,
                           float *ring, int ring_slots, long long elems_per_frame,
                           long long elem_off, long long first_frame, void *stream);

/* Lag engine (processFFT / analyseChunk, src/DDM_kernel.cuh:85-101,
 * src/DDM.cu:222-262): for the n_new frames first_frame .. first_frame+n_new-1
 * held in the ring and every lag l, acc[l][e] += sum |F_t[e] - F_{t-lag}[e]|^2
 * over those t with t - lag >= 0.  No 1/s^2 factor is applied.  lags is a
 * device array of lag_count strictly ascending ints in [0, max_lag] (read
 * back to the host for the launch plan, so the call waits for the stream);
 * the ring must hold frames
 * first_frame - max_lag .. first_frame + n_new - 1 (max_lag + n_new <=
 * ring_slots).  n_new <= lag_block (32, 64, 128 or 256). */
int ddm_k_lag_accumulate(const float *ring, int ring_slots, long long elems_per_frame,
                         long long first_frame, int n_new, const int *lags, int lag_count,
                         int max_lag, float *acc, int lag_block, void *stream);

/* ISF fit (the fitting step of ddm_fit_results) on device arrays:
 * curves[n][L] doubles, tau[L] > 0 ascending, L >= 3, out[n][4] =
 * F, A, B, rms of A (1 - exp(-F tau)) + B.  iters bounds the
 * Levenberg-Marquardt iterations (1 .. 10000).  A curve that is all zero
 * gives all-zero parameters.  fit_func / fit_ISF of
 * resources/python_tools/plot_averaged.py:61-89. */
int ddm_k_fit(const double *curves, const double *tau, int L, long long n, int iters, double *out,
              void *stream);

/* -------------------------------------------------------------------------
 * Host utilities (no device access).
 * ------------------------------------------------------------------------- */

/* q-ring membership of scale s for one lambda, computed in FP32 as
 * buildAzimuthMask does (src/azimuthal_average.cu:229-283) with
 * q = float(s) / lambda (src/DDM.cu:102).  mask[ky][kx], kx <= s/2.
 * Returns the member count (>= 0) or a negative code. */
int ddm_ring_mask(int scale, float lambda, float tolerance, uint8_t *mask);

/* Deterministic synthetic video.  Frames are written as n_frames * h * w
 * pixels of 1 byte (bits <= 8) or 2 bytes little-endian (bits > 8).
 * Uniform: i.i.d. values in [0, 2^bits - 1] from a counter-based hash of
 * (seed, frame, pixel). */
int ddm_synth_uniform(void *out, int n_frames, int width, int height, int bits,
                      uint64_t seed, long long first_frame);

/* Brownian spots: n_spots Gaussian spots (sigma_px, amplitude) on
 * `background`, per-axis step N(0, 2*D) per frame, periodic wrap, plus
 * N(0, noise_sigma) pixel noise, rounded and clipped to [0, 2^bits - 1]. */
int ddm_synth_brownian(void *out, int n_frames, int width, int height, int bits,
                       int n_spots, float sigma_px, float amplitude, float diffusion,
                       float background, float noise_sigma, uint64_t seed, int nthreads);

/* Writes one I(lambda, tau) file in the reference layout (writeIqtToFile,
 * src/azimuthal_average.cu:112-150): lambdas, then tau/fps, then one line
 * of lag values per lambda; values space-separated with a trailing space. */
int ddm_write_iqt_file(const char *path, const double *iqt_q_lag,
                       const float *lambdas, int lambda_count,
                       const int *lags, int lag_count, float fps);

#ifdef __cplusplus
}
#endif

#endif /* DDM_AMD_H */
