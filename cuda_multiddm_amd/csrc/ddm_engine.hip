// DDM engine: owns the spectrum ring, lag accumulators and q-ring tables of
// one analysis on one device, and implements the C ABI of ddm_amd.h.
//
// Per chunk of <= chunk_frames frames (runDDM's chunk loop,
// src/DDM.cu:765-843):
//   1. (host input) ROI rows -> pinned staging buffer -> async H2D on the
//      copy stream; the compute stream waits on its event;
//   2. tile cut + FFT of every scale into ring slots frame % ring_slots;
//   3. lag engine over the chunk in blocks of lag_block frames.
// The ring holds chunk_frames + max_lag spectra, so every lag of every
// frame finds its partner (the reference's 3-chunk ring limits
// lag <= chunk, src/DDM.cu:462).  Results (analyse_accums,
// src/DDM.cu:72-146) run one azimuthal launch for every
// (scale, tile, q, lag).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ddm_internal.h"
#include "ddm_kernels.h"

using ddm::Params;

namespace {

enum MarkKind { MK_FFT = 0, MK_LAG = 1, MK_AZ = 2, MK_CHUNK = 3 };

struct Mark {
    hipEvent_t a, b;
    int kind;
};

using ddm::default_fft_batch;  // ddm_autotune may choose another multiple of the chunk
constexpr int NSTAGE = ddm::MAX_UPLOAD_STAGES;  // capacity; an engine uses nstage of them

}  // namespace

struct ddm_engine {
    Params p;
    ddm::RingTables rt;
    int device = 0;
    int lag_block = 32;
    int fft_batch = 30;  // frames of device-resident input transformed per batch
    float *d_ring = nullptr;
    float *d_acc = nullptr;
    float *d_comp = nullptr;     // compensation terms of d_acc (DDM_ACC_COMPENSATED)
    double *d_acc64 = nullptr;   // FP64 accumulators d_acc + d_comp are folded into (allocated on first fold)
    bool acc64_live = false;     // d_acc64 holds sums of the current window
    long long frames_since_fold = 0;  // frames (pairs per lag and element, at most) in d_acc since the last fold
    int *d_lags = nullptr;
    ddm::AzScale *d_az = nullptr;
    int *d_wl = nullptr;            // narrow-ring average waves: q | lag group << 16
    long long az_wide_items = 0;    // averages of the large-ring scales (a workgroup each)
    int az_s0 = 0;                  // first narrow-ring scale
    long long az_narrow_waves = 0;
    int *d_qoff = nullptr, *d_qcnt = nullptr, *d_idx = nullptr;
    int *d_lay = nullptr;              // tile layout tables of every scale
    std::vector<ddm::TileLayout> lay;  // per scale, pointing into d_lay
    std::vector<int> family;           // per scale, enum ddm_fft_family
    int tiler = DDM_TILER_PER_SCALE;   // enum ddm_tiler (default set by region size at creation)
    double *d_norm = nullptr, *d_out = nullptr;
    double *d_tau = nullptr, *d_fit = nullptr;  // allocated by the first ddm_fit_results
    float *d_xt = nullptr;    // row spectra of 256 / 512 tiles between FFT passes
    long long xt_elems = 0;   // complex values d_xt holds (grown on demand)
    // upload staging for host frames: NSTAGE device buffers used in turn (an
    // upload need not wait for the frames two uploads back to be consumed
    // while a lag launch holds the stream), stage_frames frames each
    uint8_t *d_stage[NSTAGE] = {};
    uint8_t *h_stage[NSTAGE] = {};
    bool stage_used[NSTAGE] = {};
    int next_stage = 0;
    int nstage = ddm::UPLOAD_STAGES;  // staging buffers in use
    int stage_frames = 0;
    hipStream_t st = nullptr, st_copy = nullptr;
    hipEvent_t ev_up[NSTAGE] = {}, ev_free[NSTAGE] = {};
    long long frames = 0, frames_since_reset = 0;
    long long frames_lagged = 0;  // frames the lag engine has processed (<= frames)
    std::vector<long long> pairs;
    std::vector<double> h_norm;  // per-lag normalisation, kept alive until its upload is done
    size_t dev_bytes = 0, pinned_bytes = 0;
    std::string err;
    bool timing = false;
    int precision = DDM_PREC_FP32;
    std::vector<Mark> pending;
    std::vector<hipEvent_t> pool;
    ddm_stage_stats stats{};
    std::vector<float> chunk_ms;
};

#define HIP_TRY(eng, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (eng)->err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return (e_ == hipErrorOutOfMemory) ? DDM_ENOMEM : DDM_EDEVICE;                  \
        }                                                                                   \
    } while (0)

// Message of the last ddm_create failure on this thread (no engine exists
// to hold it); returned by ddm_last_error(NULL).
static thread_local std::string g_create_err;

static int fail(ddm_engine *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}

static hipEvent_t take_event(ddm_engine *e) {
    if (!e->pool.empty()) {
        hipEvent_t ev = e->pool.back();
        e->pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

static void mark_begin(ddm_engine *e, Mark *m, int kind) {
    m->a = m->b = nullptr;
    m->kind = kind;
    if (!e->timing) return;
    m->a = take_event(e);
    m->b = take_event(e);
    if (m->a) (void)hipEventRecord(m->a, e->st);
}

static void mark_end(ddm_engine *e, Mark *m) {
    if (!e->timing || !m->a || !m->b) return;
    (void)hipEventRecord(m->b, e->st);
    e->pending.push_back(*m);
}

// Called after the compute stream is idle.
static void harvest_marks(ddm_engine *e) {
    for (const Mark &m : e->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
            switch (m.kind) {
            case MK_FFT: e->stats.fft_ms += ms; break;
            case MK_LAG: e->stats.lag_ms += ms; break;
            case MK_AZ: e->stats.azimuthal_ms += ms; break;
            case MK_CHUNK:
                e->chunk_ms.push_back(ms);
                e->stats.chunks += 1;
                break;
            }
        }
        e->pool.push_back(m.a);
        e->pool.push_back(m.b);
    }
    e->pending.clear();
}

static void release(ddm_engine *e) {
    if (!e) return;
    if (e->st) (void)hipStreamSynchronize(e->st);
    if (e->st_copy) (void)hipStreamSynchronize(e->st_copy);
    for (const Mark &m : e->pending) {
        e->pool.push_back(m.a);
        e->pool.push_back(m.b);
    }
    e->pending.clear();
    for (hipEvent_t ev : e->pool)
        if (ev) (void)hipEventDestroy(ev);
    e->pool.clear();
    // device buffers come from the stream-ordered pool: freeing them does not
    // synchronise the device (other engines keep running)
    void *dev[] = {e->d_ring, e->d_acc, e->d_comp, e->d_acc64, e->d_norm, e->d_lags, e->d_az,  e->d_qoff,
                   e->d_qcnt, e->d_idx, e->d_out,  e->d_tau,   e->d_fit,  e->d_xt,   e->d_wl, e->d_lay};
    for (void *ptr : dev)
        if (ptr && e->st) (void)hipFreeAsync(ptr, e->st);
    for (uint8_t *ptr : e->d_stage)
        if (ptr && e->st) (void)hipFreeAsync(ptr, e->st);
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (int i = 0; i < NSTAGE; ++i) {
        if (e->h_stage[i]) (void)hipHostFree(e->h_stage[i]);
        if (e->ev_up[i]) (void)hipEventDestroy(e->ev_up[i]);
        if (e->ev_free[i]) (void)hipEventDestroy(e->ev_free[i]);
    }
    if (e->st) (void)hipStreamDestroy(e->st);
    if (e->st_copy) (void)hipStreamDestroy(e->st_copy);
}

template <class T>
static int dev_alloc(ddm_engine *e, T **ptr, size_t count, const char *what, hipStream_t st = nullptr) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t rc = hipMallocAsync((void **)ptr, bytes, st ? st : e->st);
    if (rc != hipSuccess) {
        *ptr = nullptr;
        e->err = std::string("device allocation of ") + what + " (" + std::to_string(bytes) +
                 " bytes) failed: " + hipGetErrorString(rc);
        return rc == hipErrorOutOfMemory ? DDM_ENOMEM : DDM_EDEVICE;
    }
    e->dev_bytes += bytes;
    return DDM_OK;
}

extern "C" int ddm_create(const ddm_desc *desc, ddm_engine **out) {
    if (!out) return DDM_EINVAL;
    *out = nullptr;
    ddm_engine *e = new (std::nothrow) ddm_engine();
    if (!e) return DDM_ENOMEM;
    std::string msg;
    int rc = ddm::check_desc(desc, &e->p, &msg);  // validate before touching the device
    if (rc != DDM_OK) {
        g_create_err = msg;
        delete e;
        return rc;
    }
    const Params &p = e->p;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_err = "no HIP device available";
        delete e;
        return DDM_EDEVICE;
    }
    auto bail = [&](int code) {
        g_create_err = e->err;
        release(e);
        delete e;
        return code;
    };
#define TRY(x)                 \
    do {                       \
        int r_ = (x);          \
        if (r_) return bail(r_); \
    } while (0)
#define HTRY(call)                                                              \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            e->err = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return bail(e_ == hipErrorOutOfMemory ? DDM_ENOMEM : DDM_EDEVICE);  \
        }                                                                       \
    } while (0)
    HTRY(hipGetDevice(&e->device));
    e->lag_block = p.lag_block;
    e->fft_batch = default_fft_batch(p.chunk);
    TRY(ddm::build_ring_tables(p, &e->rt));
    const long long E = p.elems_padded;
    const size_t L = p.lags.size();
    HTRY(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
    HTRY(hipStreamCreateWithFlags(&e->st_copy, hipStreamNonBlocking));
    for (int i = 0; i < NSTAGE; ++i) {
        HTRY(hipEventCreateWithFlags(&e->ev_up[i], hipEventDisableTiming));
        HTRY(hipEventCreateWithFlags(&e->ev_free[i], hipEventDisableTiming));
    }
    TRY(dev_alloc(e, &e->d_ring, (size_t)p.ring_slots * (size_t)E * 2, "spectrum ring"));
    TRY(dev_alloc(e, &e->d_acc, L * (size_t)E, "lag accumulators"));
    if (p.acc_mode == DDM_ACC_COMPENSATED) TRY(dev_alloc(e, &e->d_comp, L * (size_t)E, "compensation terms"));
    TRY(dev_alloc(e, &e->d_lags, L, "lags"));
    TRY(dev_alloc(e, &e->d_az, p.geom.size(), "scale table"));
    TRY(dev_alloc(e, &e->d_qoff, e->rt.off.size(), "ring offsets"));
    TRY(dev_alloc(e, &e->d_qcnt, e->rt.cnt.size(), "ring counts"));
    TRY(dev_alloc(e, &e->d_idx, e->rt.idx.size(), "ring elements"));
    TRY(dev_alloc(e, &e->d_norm, L, "normalisation"));
    TRY(dev_alloc(e, &e->d_out, (size_t)p.result_count, "results"));
    // tile layout tables: the kept elements of each scale's tiles
    size_t lay_ints = 0;
    for (const ddm::ScaleGeom &g : p.geom)
        lay_ints += g.lay->start.size() + g.lay->band.size();
    if (lay_ints) TRY(dev_alloc(e, &e->d_lay, lay_ints, "tile layout tables"));
    {
        int *at = e->d_lay;
        hipError_t up = hipSuccess;
        // the host tables are shared and outlive the upload
        auto put = [&](const std::vector<int> &v) -> const int * {
            if (v.empty()) return nullptr;
            const int *d = at;
            const hipError_t r = hipMemcpyAsync(at, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, e->st);
            if (up == hipSuccess) up = r;
            at += v.size();
            return d;
        };
        for (const ddm::ScaleGeom &g : p.geom) {
            const ddm::LayoutTables &t = *g.lay;
            ddm::TileLayout l;
            l.kind = t.kind;
            l.elems = t.elems;
            l.pass_cols = t.pass_cols;
            l.row_lanes = t.row_lanes;
            l.start = put(t.start);
            l.band = put(t.band);
            e->lay.push_back(l);
        }
        HTRY(up);
    }
    e->family.assign(p.geom.size(), DDM_FFT_REGISTER);
    // small regions: a chunk's transforms are short launches, and fusing them
    // pays (256^2 serving streams +7 %); larger regions are bandwidth-bound
    // and keep one launch per scale (512^2 -5 to -13 %, 1024^2 -4 to -11 %)
    e->tiler = p.M <= ddm::kTilerFusedMaxRegion ? DDM_TILER_FUSED : DDM_TILER_PER_SCALE;
    const size_t roi_frame = (size_t)p.M * p.M * p.bpp;
    // pinned host staging is allocated on the first host submission that needs it
    for (int i = 0; i < e->nstage; ++i) TRY(dev_alloc(e, &e->d_stage[i], roi_frame * p.chunk, "upload staging"));
    e->stage_frames = p.chunk;
    std::vector<ddm::AzScale> az;
    long long out_base = 0;
    for (const ddm::ScaleGeom &g : p.geom) {
        ddm::AzScale a;
        a.out_base = out_base;
        a.elem_off = g.elem_off;
        a.tile_elems = g.tile_elems;
        const double s2 = (double)g.s * (double)g.s;
        a.inv_s4 = 1.0 / (s2 * s2);
        az.push_back(a);
        out_base += (long long)g.tiles * (long long)p.lambdas.size() * (long long)L;
    }
    // Averages: the scales whose largest ring has AZ_WIDE_RING elements or
    // more (the largest scales, first) take a workgroup per value; the others
    // go to the narrow kernel, a wave per (tile, q, group of lags).
    std::vector<int> wl;
    {
        const size_t Q = p.lambdas.size();
        size_t s0 = 0;
        for (; s0 < p.geom.size(); ++s0) {
            int big = 0;
            for (size_t q = 0; q < Q; ++q) big = std::max(big, e->rt.cnt[s0 * Q + q]);
            if (big < ddm::AZ_WIDE_RING) break;
            e->az_wide_items += (long long)p.geom[s0].tiles * (long long)Q * (long long)L;
        }
        e->az_s0 = (int)s0;
        if (Q < 0x10000) {
            for (size_t si = s0; si < p.geom.size(); ++si) {
                az[si].wave_base = e->az_narrow_waves;
                az[si].wl_off = (int)wl.size();
                for (size_t q = 0; q < Q; ++q) {
                    const int G = 64 / ddm::az_lanes(e->rt.cnt[si * Q + q]);
                    for (int grp = 0; grp * G < (int)L; ++grp) wl.push_back((int)q | (grp << 16));
                }
                az[si].wl_n = (int)wl.size() - az[si].wl_off;
                e->az_narrow_waves += (long long)p.geom[si].tiles * az[si].wl_n;
            }
            if (!wl.empty()) TRY(dev_alloc(e, &e->d_wl, wl.size(), "average wave list"));
        }
    }
    // stream-ordered after the allocations; the host arrays live until the
    // synchronisation below
    HTRY(hipMemcpyAsync(e->d_lags, p.lags.data(), L * sizeof(int), hipMemcpyHostToDevice, e->st));
    HTRY(hipMemcpyAsync(e->d_az, az.data(), az.size() * sizeof(ddm::AzScale), hipMemcpyHostToDevice, e->st));
    if (e->d_wl) HTRY(hipMemcpyAsync(e->d_wl, wl.data(), wl.size() * sizeof(int), hipMemcpyHostToDevice, e->st));
    HTRY(hipMemcpyAsync(e->d_qoff, e->rt.off.data(), e->rt.off.size() * sizeof(int), hipMemcpyHostToDevice, e->st));
    HTRY(hipMemcpyAsync(e->d_qcnt, e->rt.cnt.data(), e->rt.cnt.size() * sizeof(int), hipMemcpyHostToDevice, e->st));
    if (!e->rt.idx.empty())
        HTRY(hipMemcpyAsync(e->d_idx, e->rt.idx.data(), e->rt.idx.size() * sizeof(int), hipMemcpyHostToDevice,
                            e->st));
    // The ring needs no clearing: the lag engine only pairs frames that were
    // written, and the padding elements of the last ring group feed only
    // accumulator entries that no result reads.
    HTRY(hipMemsetAsync(e->d_acc, 0, L * (size_t)E * sizeof(float), e->st));
    if (e->d_comp) HTRY(hipMemsetAsync(e->d_comp, 0, L * (size_t)E * sizeof(float), e->st));
    const float2 *tw = nullptr;
    HTRY(ddm::twiddle_table(&tw));
    HTRY(hipStreamSynchronize(e->st));
    e->pairs.assign(L, 0);
#undef TRY
#undef HTRY
    *out = e;
    return DDM_OK;
}

extern "C" void ddm_destroy(ddm_engine *e) {
    if (!e) return;
    release(e);
    delete e;
}

// Moves the compensated FP32 sums into the FP64 accumulators (allocated and
// zeroed on first use) and clears them.
static int fold_acc(ddm_engine *e, hipStream_t st) {
    const size_t n = e->p.lags.size() * (size_t)e->p.elems_padded;
    if (!e->d_acc64) {
        int rc = dev_alloc(e, &e->d_acc64, n, "FP64 accumulators", st);
        if (rc) return rc;
        HIP_TRY(e, hipMemsetAsync(e->d_acc64, 0, n * sizeof(double), st));
    }
    HIP_TRY(e, ddm::launch_acc_fold(e->d_acc, e->d_comp, e->d_acc64, (long long)n, st));
    e->acc64_live = true;
    e->frames_since_fold = 0;
    return DDM_OK;
}

// Lag engine over the next n frames waiting in the ring.
static int lag_block(ddm_engine *e, int n) {
    const Params &p = e->p;
    const long long t0 = e->frames_lagged;
    ddm::LagArgs a;
    a.ring = e->d_ring;
    a.acc = e->d_acc;
    a.elems_per_frame = p.elems_padded;
    a.ring_slots = p.ring_slots;
    a.slot0 = (int)(t0 % p.ring_slots);
    a.n_new = n;
    a.dmax = (int)std::min<long long>(p.max_lag, t0);
    a.lags = e->d_lags;
    a.lags_host = p.lags.data();
    a.lag_count = (int)p.lags.size();
    a.lag_block = e->lag_block;
    a.comp = e->d_comp;
    Mark lag;
    mark_begin(e, &lag, MK_LAG);
    HIP_TRY(e, ddm::launch_lag(a, e->st));
    mark_end(e, &lag);
    e->stats.lag_launches += 1;
    e->frames_lagged += n;
    if (e->d_comp) {
        e->frames_since_fold += n;
        if (e->frames_since_fold >= DDM_ACC_FOLD_PAIRS) return fold_acc(e, e->st);
    }
    return DDM_OK;
}

// Runs the lag engine on frames still waiting for a full block.
static int flush_lags(ddm_engine *e) {
    const long long pending = e->frames - e->frames_lagged;
    return pending > 0 ? lag_block(e, (int)pending) : DDM_OK;
}

// The between-pass buffer of the 256 / 512 tile FFTs, large enough for a
// batch of n frames.  Grown on the engine's stream, so work already queued
// on it keeps the old buffer until it is done.
static int ensure_xt(ddm_engine *e, int n, bool fused) {
    const long long need = ddm::fft_pass_elems(e->p.M, e->p.scales, n, fused);
    if (need <= e->xt_elems) return DDM_OK;
    if (e->d_xt) {
        HIP_TRY(e, hipFreeAsync(e->d_xt, e->st));
        e->dev_bytes -= (size_t)e->xt_elems * 2 * sizeof(float);
        e->d_xt = nullptr;
        e->xt_elems = 0;
    }
    int rc = dev_alloc(e, &e->d_xt, (size_t)need * 2, "FFT pass buffer");
    if (rc) return rc;
    e->xt_elems = need;
    return DDM_OK;
}

// Tile + FFT for n frames at device address src (n <= the ring room left,
// see batch_room), then the lag engine over every full block of lag_block
// frames now waiting.
static int process_chunk(ddm_engine *e, const uint8_t *src, int n, long long frame_bytes, int width,
                         int x_off, int y_off, hipEvent_t consumed) {
    const Params &p = e->p;
    const long long t0 = e->frames;
    const int R = p.ring_slots;
    const int first_slot = (int)(t0 % R);
    int rc = ensure_xt(e, n, e->tiler == DDM_TILER_FUSED);
    if (rc) return rc;
    Mark chunk, fft;
    mark_begin(e, &chunk, MK_CHUNK);
    mark_begin(e, &fft, MK_FFT);
    std::vector<ddm::TileFftArgs> args(p.geom.size());
    for (size_t i = 0; i < p.geom.size(); ++i) {
        const ddm::ScaleGeom &g = p.geom[i];
        ddm::TileFftArgs &a = args[i];
        a.frames = src;
        a.n_frames = n;
        a.frame_bytes = frame_bytes;
        a.width = width;
        a.bpp = p.bpp;
        a.fmt = p.fmt;
        a.x_off = x_off;
        a.y_off = y_off;
        a.largest_scale = p.M;
        a.scale = g.s;
        a.ring = e->d_ring;
        a.ring_slots = R;
        a.first_slot = first_slot;
        a.elems_per_frame = p.elems_padded;
        a.elem_off = g.elem_off;
        a.lay = e->lay[i];
        a.family = e->family[i];
        a.precision = e->precision;
        a.xt = e->d_xt;
    }
    int launches = 0;
    HIP_TRY(e, ddm::launch_tile_fft_scales(args.data(), (int)args.size(), e->tiler == DDM_TILER_FUSED, e->st,
                                           &launches));
    e->stats.fft_launches += launches;
    mark_end(e, &fft);
    if (consumed) HIP_TRY(e, hipEventRecord(consumed, e->st));
    // pairs (t - tau, t) formed for t in [t0, t0 + n)
    for (size_t l = 0; l < p.lags.size(); ++l) {
        const long long lo = std::max<long long>(t0, p.lags[l]);
        const long long hi = t0 + n;
        if (hi > lo) e->pairs[l] += hi - lo;
    }
    e->frames += n;
    e->frames_since_reset += n;
    while (e->frames - e->frames_lagged >= e->lag_block) {
        int rc = lag_block(e, e->lag_block);
        if (rc) return rc;
    }
    mark_end(e, &chunk);
    return DDM_OK;
}


// Frames the next batch may hold without overwriting a ring slot still
// needed: frame t goes to slot t % R, and the slots of frames not yet lag
// processed, with their max-lag history, must survive.  Since the ring holds
// chunk + lag_block + max lag slots and fewer than lag_block frames wait for
// the lag engine between batches, this is at least chunk frames.
static int batch_room(const ddm_engine *e) {
    const Params &p = e->p;
    const long long waiting = e->frames - e->frames_lagged;
    return (int)std::max<long long>(1, (long long)p.ring_slots - p.max_lag - waiting);
}

// Largest batch of device-resident frames: an FFT batch, or up to a lag
// block when that is at most twice the FFT batch (see device_batch).
static int device_batch_max(const ddm_engine *e) {
    return std::max(e->fft_batch, std::min(e->lag_block, 2 * e->fft_batch));
}

// Frames of the next batch of device-resident input: up to the next lag-block
// boundary when that is between half and twice an FFT batch (each batch then
// completes one lag pass, and no short batches are left between passes),
// otherwise an FFT batch; never more than the ring room or the frames left.
// Lag blocks fall on the same frames either way.
static int device_batch(const ddm_engine *e, int left) {
    const long long waiting = e->frames - e->frames_lagged;
    const int to_block = e->lag_block - (int)(waiting % e->lag_block);
    const int room = batch_room(e);
    int n = e->fft_batch;
    if (2 * to_block >= e->fft_batch && to_block <= 2 * e->fft_batch && to_block <= room) n = to_block;
    return std::min({n, left, room});
}

extern "C" int ddm_submit_device(ddm_engine *e, const void *frames, int n_frames) {
    if (!e) return DDM_EINVAL;
    if (n_frames < 0 || (n_frames > 0 && !frames)) return fail(e, DDM_EINVAL, "bad frame buffer");
    const Params &p = e->p;
    const long long fb = (long long)p.width * p.height * p.bpp;
    const uint8_t *src = (const uint8_t *)frames;
    for (int i = 0; i < n_frames;) {
        const int n = device_batch(e, n_frames - i);
        int rc = process_chunk(e, src + (long long)i * fb, n, fb, p.width, p.x_off, p.y_off, nullptr);
        if (rc) return rc;
        i += n;
    }
    return DDM_OK;
}

// Staging buffers of `frames` frames each (a call that hands over many frames
// uploads and transforms them a whole FFT batch at a time).  The old buffers
// are released once the work that reads them is done.
static int grow_stage(ddm_engine *e, int frames) {
    const size_t roi_frame = (size_t)e->p.M * e->p.M * e->p.bpp;
    for (int b = 0; b < e->nstage; ++b) {
        if (e->stage_used[b]) HIP_TRY(e, hipEventSynchronize(e->ev_free[b]));
        if (e->d_stage[b]) {
            HIP_TRY(e, hipFreeAsync(e->d_stage[b], e->st));
            e->dev_bytes -= roi_frame * e->stage_frames;
            e->d_stage[b] = nullptr;
        }
        if (e->h_stage[b]) {
            HIP_TRY(e, hipHostFree(e->h_stage[b]));
            e->pinned_bytes -= roi_frame * e->stage_frames;
            e->h_stage[b] = nullptr;
        }
    }
    for (int b = 0; b < e->nstage; ++b) {
        int rc = dev_alloc(e, &e->d_stage[b], roi_frame * frames, "upload staging");
        if (rc) return rc;
    }
    e->stage_frames = frames;
    return DDM_OK;
}

extern "C" int ddm_reserve(ddm_engine *e, int max_frames) {
    if (!e) return DDM_EINVAL;
    if (max_frames < 0) return fail(e, DDM_EINVAL, "bad frame count");
    const int chunk = e->p.chunk;
    const int batch = std::min(device_batch_max(e), std::max(max_frames, 1));
    int rc = ensure_xt(e, batch, e->tiler == DDM_TILER_FUSED);
    if (rc) return rc;
    if (max_frames > e->stage_frames && e->stage_frames < e->fft_batch) {
        rc = grow_stage(e, std::min(e->fft_batch, (max_frames + chunk - 1) / chunk * chunk));
        if (rc) return rc;
    }
    return DDM_OK;
}

// Whether host frames sit in pinned memory (then the copy engine gathers the
// ROI rows straight from them).
static bool pinned_host(const void *frames) {
    hipPointerAttribute_t attr;
    const bool pinned = hipPointerGetAttributes(&attr, frames) == hipSuccess && attr.type == hipMemoryTypeHost;
    (void)hipGetLastError();  // pageable memory reports an error here; do not leave it behind
    return pinned;
}

// Uploads the ROI rows of n host frames (frame i at src + i * frame bytes)
// into the next staging buffer, on the copy stream, and records its upload
// event; waits first until the work that read that buffer last is done.
// Returns the buffer index in *b_out.
static int stage_upload(ddm_engine *e, const uint8_t *src, int n, bool direct, int *b_out) {
    const Params &p = e->p;
    const size_t row = (size_t)p.M * p.bpp;
    const size_t roi_frame = row * p.M;
    const size_t fb = (size_t)p.width * p.height * p.bpp;
    const bool whole = p.x_off == 0 && p.y_off == 0 && p.width == p.M && p.height == p.M;
    const bool rows_stack = p.y_off == 0 && p.height == p.M;  // ROI rows of consecutive frames are evenly spaced
    const int b = e->next_stage;
    if (e->stage_used[b]) HIP_TRY(e, hipEventSynchronize(e->ev_free[b]));
    if (direct && whole) {
        HIP_TRY(e, hipMemcpyAsync(e->d_stage[b], src, roi_frame * n, hipMemcpyHostToDevice, e->st_copy));
    } else if (direct) {
        const size_t pitch = (size_t)p.width * p.bpp;
        const uint8_t *roi0 = src + (size_t)p.x_off * p.bpp;
        if (rows_stack) {
            HIP_TRY(e, hipMemcpy2DAsync(e->d_stage[b], row, roi0, pitch, row, (size_t)p.M * n, hipMemcpyHostToDevice,
                                        e->st_copy));
        } else {
            for (int f = 0; f < n; ++f)
                HIP_TRY(e, hipMemcpy2DAsync(e->d_stage[b] + (size_t)f * roi_frame, row,
                                            roi0 + (size_t)f * fb + (size_t)p.y_off * pitch, pitch, row, p.M,
                                            hipMemcpyHostToDevice, e->st_copy));
        }
    } else {
        if (!e->h_stage[b]) {
            HIP_TRY(e, hipHostMalloc((void **)&e->h_stage[b], roi_frame * e->stage_frames, hipHostMallocDefault));
            e->pinned_bytes += roi_frame * e->stage_frames;
        }
        // ROI-only copy into pinned memory (the reference uploads whole frames, src/DDM.cu:506)
        uint8_t *dst = e->h_stage[b];
        for (int f = 0; f < n; ++f) {
            const uint8_t *fr = src + (size_t)f * fb;
            for (int y = 0; y < p.M; ++y)
                std::memcpy(dst + (size_t)f * roi_frame + (size_t)y * row,
                            fr + ((size_t)(p.y_off + y) * p.width + p.x_off) * p.bpp, row);
        }
        HIP_TRY(e, hipMemcpyAsync(e->d_stage[b], dst, roi_frame * n, hipMemcpyHostToDevice, e->st_copy));
    }
    HIP_TRY(e, hipEventRecord(e->ev_up[b], e->st_copy));
    e->stage_used[b] = true;
    e->next_stage = (b + 1) % e->nstage;
    *b_out = b;
    return DDM_OK;
}

extern "C" int ddm_submit_host(ddm_engine *e, const void *frames, int n_frames) {
    if (!e) return DDM_EINVAL;
    if (n_frames < 0 || (n_frames > 0 && !frames)) return fail(e, DDM_EINVAL, "bad frame buffer");
    const Params &p = e->p;
    const size_t roi_frame = (size_t)p.M * p.bpp * p.M;
    const uint8_t *src = (const uint8_t *)frames;
    const size_t fb = (size_t)p.width * p.height * p.bpp;
    // Frames that already sit in pinned memory go to the device without the
    // staging copy: the ROI rows are gathered by the copy engine (2-D copy).
    const bool direct = n_frames > 0 && pinned_host(frames);
    if (n_frames > e->stage_frames && e->stage_frames < e->fft_batch) {
        const int want = std::min(e->fft_batch, (n_frames + p.chunk - 1) / p.chunk * p.chunk);
        int rc = grow_stage(e, want);
        if (rc) return rc;
    }
    int last = -1;
    for (int i = 0; i < n_frames;) {
        const int n = std::min({e->stage_frames, n_frames - i, batch_room(e)});
        int b = 0;
        int rc = stage_upload(e, src + (size_t)i * fb, n, direct, &b);
        if (rc) return rc;
        HIP_TRY(e, hipStreamWaitEvent(e->st, e->ev_up[b], 0));
        rc = process_chunk(e, e->d_stage[b], n, (long long)roi_frame, p.M, 0, 0, e->ev_free[b]);
        if (rc) return rc;
        last = b;
        i += n;
    }
    // the caller may reuse its buffer as soon as this call returns (uploads
    // run in order on the copy stream: the last one done means all are)
    if (direct && last >= 0) HIP_TRY(e, hipEventSynchronize(e->ev_up[last]));
    return DDM_OK;
}

extern "C" int ddm_flush(ddm_engine *e) {
    if (!e) return DDM_EINVAL;
    return flush_lags(e);
}

namespace ddm {
hipStream_t engine_stream(ddm_engine *e) { return e ? e->st : nullptr; }

void engine_batch_stream(ddm_engine *e, BatchStream *bs) {
    bs->p = &e->p;
    bs->ring = e->d_ring;
    bs->acc = e->d_acc;
    bs->comp = e->d_comp;
    bs->lags = e->d_lags;
    bs->lag_block = e->lag_block;
    bs->frames = e->frames;
    bs->precision = e->precision;
    bs->lay = e->lay.data();
    bs->family = e->family.data();
}

int engine_batch_upload(ddm_engine *e, const void *frames, int n, int *b, hipEvent_t *ev_up, const uint8_t **stage) {
    if (n < 1 || n > e->stage_frames || n > e->lag_block || n > batch_room(e))
        return fail(e, DDM_EINVAL, "a batched chunk must fit a staging buffer, a lag block and the ring room");
    const bool direct = pinned_host(frames);
    int rc = stage_upload(e, (const uint8_t *)frames, n, direct, b);
    if (rc) return rc;
    // the caller may reuse its buffer as soon as this returns
    if (direct) HIP_TRY(e, hipEventSynchronize(e->ev_up[*b]));
    *ev_up = e->ev_up[*b];
    *stage = e->d_stage[*b];
    return DDM_OK;
}

int engine_batch_stages(ddm_engine *e, int k) {
    if (e->frames != 0 || k < e->nstage || k > NSTAGE) return fail(e, DDM_ESTATE, "staging buffers are fixed once frames arrived");
    const size_t bytes = (size_t)e->p.M * e->p.M * e->p.bpp * e->stage_frames;
    for (int b = e->nstage; b < k; ++b) {
        int rc = dev_alloc(e, &e->d_stage[b], bytes, "upload staging");
        if (rc) return rc;
    }
    e->nstage = k;
    HIP_TRY(e, hipStreamSynchronize(e->st));
    return DDM_OK;
}

int engine_batch_commit(ddm_engine *e, int n, int b, hipStream_t st) {
    const long long t0 = e->frames;
    HIP_TRY(e, hipEventRecord(e->ev_free[b], st));
    for (size_t l = 0; l < e->p.lags.size(); ++l) {
        const long long lo = std::max<long long>(t0, e->p.lags[l]);
        if (t0 + n > lo) e->pairs[l] += t0 + n - lo;
    }
    e->frames += n;
    e->frames_since_reset += n;
    e->frames_lagged += n;
    e->stats.lag_launches += 1;
    if (e->d_comp) {
        e->frames_since_fold += n;
        if (e->frames_since_fold >= DDM_ACC_FOLD_PAIRS) return fold_acc(e, st);
    }
    return DDM_OK;
}
}  // namespace ddm

extern "C" int ddm_sync(ddm_engine *e) {
    if (!e) return DDM_EINVAL;
    HIP_TRY(e, hipStreamSynchronize(e->st));
    HIP_TRY(e, hipStreamSynchronize(e->st_copy));
    harvest_marks(e);
    return DDM_OK;
}

// Masked averages of the current window into d_out (asynchronous).
static int compute_results(ddm_engine *e) {
    const Params &p = e->p;
    const size_t L = p.lags.size();
    int rc = flush_lags(e);
    if (rc) return rc;
    std::vector<double> &norm = e->h_norm;
    // the previous upload from this vector has finished: results are read
    // after a stream synchronisation
    norm.assign(L, 0.0);
    for (size_t l = 0; l < L; ++l) {
        if (p.norm_mode == DDM_NORM_PAIRS)
            norm[l] = e->pairs[l] > 0 ? 1.0 / (double)e->pairs[l] : 0.0;
        else
            norm[l] = e->frames_since_reset > 0 ? 1.0 / (double)e->frames_since_reset : 0.0;
    }
    HIP_TRY(e, hipMemcpyAsync(e->d_norm, norm.data(), L * sizeof(double), hipMemcpyHostToDevice, e->st));
    ddm::AzArgs a;
    a.acc = e->d_acc;
    a.comp = e->d_comp;
    a.acc64 = e->acc64_live ? e->d_acc64 : nullptr;
    a.elems_per_frame = p.elems_padded;
    a.scales = e->d_az;
    a.scale_count = (int)p.geom.size();
    a.q_off = e->d_qoff;
    a.q_cnt = e->d_qcnt;
    a.idx = e->d_idx;
    a.norm = e->d_norm;
    a.lambda_count = (int)p.lambdas.size();
    a.lag_count = (int)L;
    a.out = e->d_out;
    a.result_count = p.result_count;
    a.wide_items = e->az_wide_items;
    a.narrow_first_scale = e->az_s0;
    a.narrow_waves = e->az_narrow_waves;
    a.wave_list = e->d_wl;
    Mark az;
    mark_begin(e, &az, MK_AZ);
    HIP_TRY(e, ddm::launch_azimuthal(a, e->st));
    mark_end(e, &az);
    e->stats.azimuthal_launches += 1;
    return DDM_OK;
}

extern "C" int ddm_results(ddm_engine *e, double *results) {
    if (!e || !results) return DDM_EINVAL;
    int rc = compute_results(e);
    if (rc) return rc;
    HIP_TRY(e, hipMemcpyAsync(results, e->d_out, (size_t)e->p.result_count * sizeof(double), hipMemcpyDeviceToHost,
                              e->st));
    return ddm_sync(e);
}

extern "C" int ddm_fit_results(ddm_engine *e, float fps, double *fit, double *D) {
    if (!e || !(fps > 0.0f)) return DDM_EINVAL;
    const Params &p = e->p;
    const int L = (int)p.lags.size(), Q = (int)p.lambdas.size();
    const long long curves = p.result_count / L;
    if (L < 3) return fail(e, DDM_EINVAL, "fitting needs at least 3 tau values");
    int rc = compute_results(e);
    if (rc) return rc;
    if (!e->d_fit) {
        rc = dev_alloc(e, &e->d_tau, (size_t)L, "fit lags");
        if (rc) return rc;
        rc = dev_alloc(e, &e->d_fit, (size_t)curves * 4, "fit parameters");
        if (rc) return rc;
    }
    std::vector<double> tau(L);
    for (int l = 0; l < L; ++l) tau[l] = (double)p.lags[l] / (double)fps;
    HIP_TRY(e, hipMemcpyAsync(e->d_tau, tau.data(), L * sizeof(double), hipMemcpyHostToDevice, e->st));
    HIP_TRY(e, ddm::launch_fit(e->d_out, e->d_tau, L, curves, 200, e->d_fit, e->st));
    std::vector<double> prm((size_t)curves * 4);
    HIP_TRY(e, hipMemcpyAsync(prm.data(), e->d_fit, prm.size() * sizeof(double), hipMemcpyDeviceToHost, e->st));
    rc = ddm_sync(e);
    if (rc) return rc;
    if (fit) std::copy(prm.begin(), prm.end(), fit);
    if (D) {
        // ln F = ln D + 2 ln q with q = 2 pi / lambda, over the curves whose
        // decay time 1/F lies inside the lag range
        const double tmin = tau.front(), tmax = tau.back();
        long long c = 0, t_out = 0;
        for (const ddm::ScaleGeom &g : p.geom) {
            for (int t = 0; t < g.tiles; ++t, ++t_out) {
                double sum = 0.0;
                int n = 0;
                for (int q = 0; q < Q; ++q, ++c) {
                    const double F = prm[(size_t)c * 4], A = prm[(size_t)c * 4 + 1];
                    if (!(F > 0.0) || !(A > 0.0) || 1.0 / F < tmin || 1.0 / F > tmax || !(p.lambdas[q] > 0.0f))
                        continue;
                    const double qv = 2.0 * M_PI / (double)p.lambdas[q];
                    sum += std::log(F) - 2.0 * std::log(qv);
                    ++n;
                }
                D[t_out] = n ? std::exp(sum / n) : 0.0;
            }
        }
    }
    return DDM_OK;
}

extern "C" int ddm_reset(ddm_engine *e) {
    if (!e) return DDM_EINVAL;
    const Params &p = e->p;
    int rc = flush_lags(e);  // waiting frames belong to the window being closed
    if (rc) return rc;
    const size_t n = p.lags.size() * (size_t)p.elems_padded;
    HIP_TRY(e, hipMemsetAsync(e->d_acc, 0, n * sizeof(float), e->st));
    if (e->d_comp) HIP_TRY(e, hipMemsetAsync(e->d_comp, 0, n * sizeof(float), e->st));
    if (e->acc64_live) HIP_TRY(e, hipMemsetAsync(e->d_acc64, 0, n * sizeof(double), e->st));
    e->acc64_live = false;
    e->frames_since_fold = 0;
    e->frames_since_reset = 0;
    std::fill(e->pairs.begin(), e->pairs.end(), 0);
    return DDM_OK;
}

extern "C" int ddm_get_info(const ddm_engine *e, ddm_info *info) {
    if (!e || !info) return DDM_EINVAL;
    const Params &p = e->p;
    std::memset(info, 0, sizeof(*info));
    info->largest_scale = p.M;
    info->scale_count = (int)p.scales.size();
    info->lag_count = (int)p.lags.size();
    info->lambda_count = (int)p.lambdas.size();
    info->ring_slots = p.ring_slots;
    info->lag_block = e->lag_block;  // the descriptor's, or the autotuner's choice
    info->elems_per_frame = p.elems_per_frame;
    info->elems_per_frame_padded = p.elems_padded;
    info->result_count = p.result_count;
    info->frames_submitted = e->frames;
    info->frames_since_reset = e->frames_since_reset;
    info->device_bytes = e->dev_bytes;
    info->pinned_bytes = e->pinned_bytes;
    return DDM_OK;
}

extern "C" int ddm_get_params(const ddm_engine *e, int *scales, int *lags, float *lambdas) {
    if (!e) return DDM_EINVAL;
    const Params &p = e->p;
    if (scales) std::copy(p.scales.begin(), p.scales.end(), scales);
    if (lags) std::copy(p.lags.begin(), p.lags.end(), lags);
    if (lambdas) std::copy(p.lambdas.begin(), p.lambdas.end(), lambdas);
    return DDM_OK;
}

// Element values of every tile of a scale, `comps` floats each, stored in
// the scale's tile layout (src), into out in the full row order
// [tile][ky][kx <= s/2]; elements the layout leaves out read NaN.
static void expand_tiles(const ddm::ScaleGeom &g, const float *src, int comps, float *out) {
    const int s = g.s, H = s / 2;
    const size_t full = (size_t)s * (H + 1);
    for (int t = 0; t < g.tiles; ++t) {
        const float *ts = src + (size_t)t * g.tile_elems * comps;
        float *to = out + (size_t)t * full * comps;
        for (int ky = 0; ky < s; ++ky)
            for (int kx = 0; kx <= H; ++kx) {
                const int at = g.lay->pos[(size_t)ky * (H + 1) + kx];
                float *o = to + ((size_t)ky * (H + 1) + kx) * comps;
                for (int c = 0; c < comps; ++c) o[c] = at < 0 ? NAN : ts[(size_t)at * comps + c];
            }
    }
}

extern "C" int ddm_get_accumulators(ddm_engine *e, int scale_idx, float *out) {
    if (!e || !out) return DDM_EINVAL;
    const Params &p = e->p;
    if (scale_idx < 0 || scale_idx >= (int)p.geom.size()) return fail(e, DDM_ERANGE, "scale index out of range");
    int rc = flush_lags(e);
    if (rc) return rc;
    rc = ddm_sync(e);
    if (rc) return rc;
    const ddm::ScaleGeom &g = p.geom[scale_idx];
    const size_t n = (size_t)g.tiles * g.tile_elems;             // stored
    const size_t nfull = (size_t)g.tiles * g.s * (g.s / 2 + 1);  // returned
    // the reference accumulates (|dF| / s^2)^2; 1/s^4 is a power of two, so the
    // scaling is exact
    const double inv = 1.0 / ((double)g.s * g.s * (double)g.s * g.s);
    std::vector<float> a(n), c(e->d_comp ? n : 0);
    std::vector<double> d(e->acc64_live ? n : 0);
    for (size_t l = 0; l < p.lags.size(); ++l) {
        const size_t off = l * (size_t)p.elems_padded + g.elem_off;
        HIP_TRY(e, hipMemcpy(a.data(), e->d_acc + off, n * sizeof(float), hipMemcpyDeviceToHost));
        if (e->d_comp) {
            // compensated: acc + comp (+ the FP64 part), rounded once to FP32
            HIP_TRY(e, hipMemcpy(c.data(), e->d_comp + off, n * sizeof(float), hipMemcpyDeviceToHost));
            if (e->acc64_live)
                HIP_TRY(e, hipMemcpy(d.data(), e->d_acc64 + off, n * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                a[i] = (float)((((double)a[i] + (double)c[i]) + (e->acc64_live ? d[i] : 0.0)) * inv);
        } else {
            for (size_t i = 0; i < n; ++i) a[i] *= (float)inv;
        }
        expand_tiles(g, a.data(), 1, out + l * nfull);
    }
    return DDM_OK;
}

extern "C" int ddm_get_spectra(ddm_engine *e, long long frame, int scale_idx, float *out) {
    if (!e || !out) return DDM_EINVAL;
    const Params &p = e->p;
    if (scale_idx < 0 || scale_idx >= (int)p.geom.size()) return fail(e, DDM_ERANGE, "scale index out of range");
    if (frame < 0 || frame >= e->frames || frame < e->frames - p.ring_slots)
        return fail(e, DDM_ERANGE, "frame is not held in the spectrum ring");
    int rc = ddm_sync(e);
    if (rc) return rc;
    const ddm::ScaleGeom &g = p.geom[scale_idx];
    const size_t slot = (size_t)(frame % p.ring_slots);
    const size_t n = (size_t)g.tiles * g.tile_elems;
    // gather the slot's rows of every ring group the scale touches
    const size_t G = ddm::RING_GROUP, g0 = (size_t)g.elem_off / G, g1 = ((size_t)g.elem_off + n + G - 1) / G;
    std::vector<float> rows((g1 - g0) * G * 2);
    HIP_TRY(e, hipMemcpy2D(rows.data(), G * 2 * sizeof(float),
                           e->d_ring + ((g0 * p.ring_slots + slot) * G) * 2, (size_t)p.ring_slots * G * 2 * sizeof(float),
                           G * 2 * sizeof(float), g1 - g0, hipMemcpyDeviceToHost));
    expand_tiles(g, rows.data() + ((size_t)g.elem_off - g0 * G) * 2, 2, out);
    return DDM_OK;
}

extern "C" int ddm_set_precision(ddm_engine *e, int mode) {
    if (!e) return DDM_EINVAL;
    if (mode != DDM_PREC_FP32 && mode != DDM_PREC_BF16X2 && mode != DDM_PREC_FP16X2)
        return fail(e, DDM_EINVAL, "unknown precision mode");
    e->precision = mode;
    return DDM_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: the lag block and the FFT batch of device-resident input are
// timed on this engine's own buffers (synthetic frames, the real geometry)
// and the fastest per frame is kept.  Results are cached per device and
// geometry in a text file, one "key<TAB>lag_block fft_batch tiler families" line each.
// ---------------------------------------------------------------------------
static std::string tune_key(const ddm_engine *e) {
    const Params &p = e->p;
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, e->device);
    uint64_t h = 1469598103934665603ull;  // FNV-1a of the lag list
    for (int l : p.lags) {
        h ^= (uint64_t)(uint32_t)l;
        h *= 1099511628211ull;
    }
    std::string k = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) +
                    " M" + std::to_string(p.M) + " bpp" + std::to_string(p.bpp) + " fmt" + std::to_string(p.fmt) +
                    " C" + std::to_string(p.chunk) + " J" + std::to_string(p.lag_block) + " S";
    for (const ddm::ScaleGeom &g : p.geom) k += std::to_string(g.s) + ",";
    char hx[64];
    std::snprintf(hx, sizeof hx, " L%zu:%016llx E%lld", p.lags.size(), (unsigned long long)h, p.elems_per_frame);
    return k + hx;
}

// A cache line is `key<TAB>J B T f0 f1 ...`: the lag block, the FFT batch,
// the tiler mode and the transform family of every scale (descending).
static bool tune_lookup(const char *path, const std::string &key, int *J, int *B, int *T, std::vector<int> *fam) {
    if (!path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[4096];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        const char *tab = std::strchr(line, '\t');
        if (!tab || std::string(line, tab - line) != key) continue;
        int j = 0, b = 0, t = 0, used = 0;
        if (std::sscanf(tab + 1, "%d %d %d%n", &j, &b, &t, &used) != 3) continue;
        std::vector<int> fl;
        const char *at = tab + 1 + used;
        int v = 0, n = 0;
        while (std::sscanf(at, "%d%n", &v, &n) == 1) {
            fl.push_back(v);
            at += n;
        }
        if (fl.size() != fam->size()) continue;
        *J = j;  // the last matching line wins
        *B = b;
        *T = t;
        *fam = fl;
        found = true;
    }
    std::fclose(f);
    return found;
}

// Scales whose transform family may be chosen: 8 .. 128 tiles.
static bool family_choice(int s) { return s >= 8 && s <= 128; }

// Milliseconds of `launch` on the engine's stream: best of three after one
// untimed run.
template <class F>
static int time_ms(ddm_engine *e, F launch, float *best) {
    hipEvent_t a = take_event(e), b = take_event(e);
    if (!a || !b) return fail(e, DDM_EDEVICE, "event creation failed");
    *best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
        HIP_TRY(e, hipEventRecord(a, e->st));
        HIP_TRY(e, launch());
        HIP_TRY(e, hipEventRecord(b, e->st));
        HIP_TRY(e, hipEventSynchronize(b));
        float ms = 0.f;
        HIP_TRY(e, hipEventElapsedTime(&ms, a, b));
        if (rep) *best = std::min(*best, ms);
    }
    e->pool.push_back(a);
    e->pool.push_back(b);
    return DDM_OK;
}

extern "C" int ddm_autotune(ddm_engine *e, const char *cache_path, int *lag_block_out, int *fft_batch_out) {
    if (!e) return DDM_EINVAL;
    if (e->frames != 0) return fail(e, DDM_ESTATE, "ddm_autotune must run before the first frame");
    int rc = ddm_sync(e);
    if (rc) return rc;
    const Params &p = e->p;
    const std::string key = tune_key(e);
    int J = 0, B = 0, T = DDM_TILER_PER_SCALE;
    std::vector<int> fam(p.geom.size(), DDM_FFT_REGISTER);
    bool cached = tune_lookup(cache_path, key, &J, &B, &T, &fam) && J >= 32 && J <= p.lag_block &&
                  (J & (J - 1)) == 0 && B >= p.chunk && B % p.chunk == 0 && B <= p.chunk + p.lag_block &&
                  (T == DDM_TILER_FUSED || T == DDM_TILER_PER_SCALE);
    for (size_t si = 0; cached && si < fam.size(); ++si)
        cached = fam[si] == DDM_FFT_REGISTER || (fam[si] == DDM_FFT_LDS && family_choice(p.geom[si].s));
    if (!cached) {
        std::fill(fam.begin(), fam.end(), DDM_FFT_REGISTER);
        // lag block: J frames against the full max-lag history, per frame; a
        // candidate must beat the descriptor's block by 2 % to replace it
        float best = 1e30f, def_lag = 1e30f;
        for (int cand : {64, 128, 256}) {
            if (cand > p.lag_block) continue;
            ddm::LagArgs a;
            a.ring = e->d_ring;
            a.acc = e->d_acc;
            a.elems_per_frame = p.elems_padded;
            a.ring_slots = p.ring_slots;
            // the block starts max_lag slots into the ring, so its history is
            // one run of slots, as for most blocks of a stream (a block at slot
            // 0 reads every partner across the ring's wrap)
            a.slot0 = p.max_lag;
            a.n_new = cand;
            a.dmax = p.max_lag;
            a.lags = e->d_lags;
            a.lags_host = p.lags.data();
            a.lag_count = (int)p.lags.size();
            a.lag_block = cand;
            a.comp = e->d_comp;
            float ms = 0.f;
            rc = time_ms(e, [&] { return ddm::launch_lag(a, e->st); }, &ms);
            if (rc) return rc;
            if (cand == p.lag_block) def_lag = ms / cand;
            if (ms / cand < best) {
                best = ms / cand;
                J = cand;
            }
        }
        if (!J || (def_lag < 1e30f && best > 0.98f * def_lag)) J = p.lag_block;
        // FFT batch: the same bmax synthetic frames through every scale in
        // batches of 1, 2, 4 or 8 chunks (per-launch costs included); a
        // candidate must beat the default by 2 % to replace it
        int bmax = p.chunk;
        for (int k : {2, 4, 8})
            if (p.chunk * k <= p.chunk + p.lag_block) bmax = p.chunk * k;
        const long long fb = (long long)p.width * p.height * p.bpp;
        uint8_t *frames = nullptr;
        rc = dev_alloc(e, &frames, (size_t)(fb * bmax), "autotune frames");
        if (rc) return rc;
        HIP_TRY(e, hipMemsetAsync(frames, 0x5a, (size_t)(fb * bmax), e->st));
        rc = ensure_xt(e, bmax, true);  // room for the tiler's regions too
        if (rc) return rc;
        best = 1e30f;
        float def_ms = 1e30f;
        const int def_b = std::min(default_fft_batch(p.chunk), bmax);
        for (int k : {1, 2, 4, 8}) {
            const int n = p.chunk * k;
            if (n > bmax) break;
            auto launch = [&]() -> hipError_t {
                for (int done = 0; done < bmax; done += n)
                for (const ddm::ScaleGeom &g : p.geom) {
                    ddm::TileFftArgs a;
                    a.frames = frames + (long long)done * fb;
                    a.n_frames = std::min(n, bmax - done);
                    a.frame_bytes = fb;
                    a.width = p.width;
                    a.bpp = p.bpp;
                    a.fmt = p.fmt;
                    a.x_off = p.x_off;
                    a.y_off = p.y_off;
                    a.largest_scale = p.M;
                    a.scale = g.s;
                    a.ring = e->d_ring;
                    a.ring_slots = p.ring_slots;
                    a.first_slot = 0;
                    a.elems_per_frame = p.elems_padded;
                    a.elem_off = g.elem_off;
                    a.lay = e->lay[&g - p.geom.data()];
                    a.precision = e->precision;
                    a.xt = e->d_xt;
                    const hipError_t err = ddm::launch_tile_fft(a, e->st);
                    if (err != hipSuccess) return err;
                }
                return hipSuccess;
            };
            float ms = 0.f;
            rc = time_ms(e, launch, &ms);
            if (rc) return rc;
            if (n == def_b) def_ms = ms;
            if (ms < best) {
                best = ms;
                B = n;
            }
        }
        if (def_ms < 1e30f && best > 0.98f * def_ms) B = def_b;
        // transform family per 8 .. 128 tile scale: that scale alone over
        // the chosen batches, registers (default) against LDS
        for (size_t si = 0; si < p.geom.size(); ++si) {
            const ddm::ScaleGeom &g = p.geom[si];
            if (!family_choice(g.s) || e->precision != DDM_PREC_FP32) continue;
            float t[2] = {1e30f, 1e30f};
            for (int fm : {DDM_FFT_REGISTER, DDM_FFT_LDS}) {
                auto launch = [&]() -> hipError_t {
                    for (int done = 0; done < bmax; done += B) {
                        ddm::TileFftArgs a;
                        a.frames = frames + (long long)done * fb;
                        a.n_frames = std::min(B, bmax - done);
                        a.frame_bytes = fb;
                        a.width = p.width;
                        a.bpp = p.bpp;
                        a.fmt = p.fmt;
                        a.x_off = p.x_off;
                        a.y_off = p.y_off;
                        a.largest_scale = p.M;
                        a.scale = g.s;
                        a.ring = e->d_ring;
                        a.ring_slots = p.ring_slots;
                        a.first_slot = 0;
                        a.elems_per_frame = p.elems_padded;
                        a.elem_off = g.elem_off;
                        a.lay = e->lay[si];
                        a.family = fm;
                        const hipError_t err = ddm::launch_tile_fft(a, e->st);
                        if (err != hipSuccess) return err;
                    }
                    return hipSuccess;
                };
                rc = time_ms(e, launch, &t[fm]);
                if (rc) return rc;
            }
            if (t[DDM_FFT_LDS] < 0.98f * t[DDM_FFT_REGISTER]) fam[si] = DDM_FFT_LDS;
        }
        // tiler: every scale (chosen families) over the chosen batches, one
        // launch per scale against the multi-scale tiler
        float tt[2] = {1e30f, 1e30f};
        for (int tm : {DDM_TILER_PER_SCALE, DDM_TILER_FUSED}) {
            auto launch = [&]() -> hipError_t {
                std::vector<ddm::TileFftArgs> args(p.geom.size());
                for (int done = 0; done < bmax; done += B) {
                    for (size_t si = 0; si < p.geom.size(); ++si) {
                        ddm::TileFftArgs &a = args[si];
                        a.frames = frames + (long long)done * fb;
                        a.n_frames = std::min(B, bmax - done);
                        a.frame_bytes = fb;
                        a.width = p.width;
                        a.bpp = p.bpp;
                        a.fmt = p.fmt;
                        a.x_off = p.x_off;
                        a.y_off = p.y_off;
                        a.largest_scale = p.M;
                        a.scale = p.geom[si].s;
                        a.ring = e->d_ring;
                        a.ring_slots = p.ring_slots;
                        a.first_slot = 0;
                        a.elems_per_frame = p.elems_padded;
                        a.elem_off = p.geom[si].elem_off;
                        a.lay = e->lay[si];
                        a.family = fam[si];
                        a.precision = e->precision;
                        a.xt = e->d_xt;
                    }
                    const hipError_t err = ddm::launch_tile_fft_scales(args.data(), (int)args.size(),
                                                                       tm == DDM_TILER_FUSED, e->st, nullptr);
                    if (err != hipSuccess) return err;
                }
                return hipSuccess;
            };
            rc = time_ms(e, launch, &tt[tm]);
            if (rc) return rc;
        }
        T = tt[DDM_TILER_FUSED] < 0.98f * tt[DDM_TILER_PER_SCALE] ? DDM_TILER_FUSED : DDM_TILER_PER_SCALE;
        HIP_TRY(e, hipFreeAsync(frames, e->st));
        e->dev_bytes -= (size_t)(fb * bmax);
        // the timing runs accumulated into the lag sums
        HIP_TRY(e, hipMemsetAsync(e->d_acc, 0, p.lags.size() * (size_t)p.elems_padded * sizeof(float), e->st));
        if (e->d_comp)
            HIP_TRY(e, hipMemsetAsync(e->d_comp, 0, p.lags.size() * (size_t)p.elems_padded * sizeof(float), e->st));
        HIP_TRY(e, hipStreamSynchronize(e->st));
        if (cache_path) {
            if (FILE *f = std::fopen(cache_path, "a")) {
                std::fprintf(f, "%s\t%d %d %d", key.c_str(), J, B, T);
                for (int v : fam) std::fprintf(f, " %d", v);
                std::fprintf(f, "\n");
                std::fclose(f);
            }
        }
    }
    e->lag_block = J;
    e->fft_batch = B;
    e->tiler = T;
    e->family = fam;
    if (lag_block_out) *lag_block_out = J;
    if (fft_batch_out) *fft_batch_out = B;
    return DDM_OK;
}

extern "C" int ddm_set_tiler(ddm_engine *e, int mode) {
    if (!e) return DDM_EINVAL;
    if (mode != DDM_TILER_FUSED && mode != DDM_TILER_PER_SCALE) return fail(e, DDM_EINVAL, "unknown tiler mode");
    if (e->frames != 0) return fail(e, DDM_ESTATE, "the tiler is fixed once frames arrived");
    e->tiler = mode;
    return DDM_OK;
}

extern "C" int ddm_get_tiler(const ddm_engine *e) { return e ? e->tiler : DDM_EINVAL; }

extern "C" int ddm_set_fft_family(ddm_engine *e, int scale_idx, int family) {
    if (!e) return DDM_EINVAL;
    if (scale_idx < 0 || scale_idx >= (int)e->p.geom.size()) return fail(e, DDM_ERANGE, "scale index out of range");
    if (e->frames != 0) return fail(e, DDM_ESTATE, "the transform family is fixed once frames arrived");
    if (family != DDM_FFT_REGISTER && !(family == DDM_FFT_LDS && family_choice(e->p.geom[scale_idx].s)))
        return fail(e, DDM_EINVAL, "no kernel of that family for this scale");
    e->family[scale_idx] = family;
    return DDM_OK;
}

extern "C" int ddm_get_fft_families(const ddm_engine *e, int *families) {
    if (!e || !families) return DDM_EINVAL;
    for (size_t i = 0; i < e->family.size(); ++i) families[i] = e->family[i];
    return DDM_OK;
}

extern "C" int ddm_set_timing(ddm_engine *e, int on) {
    if (!e) return DDM_EINVAL;
    int rc = ddm_sync(e);
    if (rc) return rc;
    e->timing = on != 0;
    return DDM_OK;
}

extern "C" int ddm_get_stage_stats(ddm_engine *e, ddm_stage_stats *s) {
    if (!e || !s) return DDM_EINVAL;
    int rc = ddm_sync(e);
    if (rc) return rc;
    *s = e->stats;
    return DDM_OK;
}

extern "C" int ddm_get_chunk_times(ddm_engine *e, float *ms, int max) {
    if (!e || (!ms && max > 0)) return DDM_EINVAL;
    int rc = ddm_sync(e);
    if (rc) return rc;
    const int n = std::min<int>(max, (int)e->chunk_ms.size());
    std::copy(e->chunk_ms.begin(), e->chunk_ms.begin() + n, ms);
    return n;
}

extern "C" const char *ddm_last_error(const ddm_engine *e) { return e ? e->err.c_str() : g_create_err.c_str(); }

// ---------------------------------------------------------------------------
// Stage entry points on caller-owned device buffers.
// ---------------------------------------------------------------------------
static bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Validates the arguments of the tile FFT entry points and fills `a`
// (without a pass buffer).
static int tile_fft_args(const void *frames, int n_frames, long long frame_bytes, int width, int height,
                         int bytes_per_pixel, int pixel_format, int x_off, int y_off, int largest_scale, int scale,
                         float *ring, int ring_slots, long long elems_per_frame, long long elem_off,
                         long long first_frame, ddm::TileFftArgs *a) {
    if (!frames || !ring || n_frames <= 0 || ring_slots <= 0 || n_frames > ring_slots || first_frame < 0)
        return DDM_EINVAL;
    if (!pow2(scale) || !pow2(largest_scale) || scale > largest_scale || scale > ddm::kMaxScale) return DDM_ESCALE;
    if (pixel_format < DDM_PIX_U8 || pixel_format > DDM_PIX_U16BE) return DDM_EINVAL;
    if (bytes_per_pixel < ddm::pixel_bytes_needed(pixel_format)) return DDM_EINVAL;
    if (x_off < 0 || y_off < 0 || (long long)x_off + largest_scale > width || (long long)y_off + largest_scale > height)
        return DDM_EROI;
    if (frame_bytes < (long long)width * height * bytes_per_pixel) return DDM_EINVAL;
    if (elems_per_frame % ddm::RING_GROUP != 0) return DDM_EINVAL;
    const long long tiles = (long long)(largest_scale / scale) * (largest_scale / scale);
    if (elem_off < 0 || elem_off + tiles * scale * (scale / 2 + 1) > elems_per_frame) return DDM_ERANGE;
    a->frames = frames;
    a->n_frames = n_frames;
    a->frame_bytes = frame_bytes;
    a->width = width;
    a->bpp = bytes_per_pixel;
    a->fmt = pixel_format;
    a->x_off = x_off;
    a->y_off = y_off;
    a->largest_scale = largest_scale;
    a->scale = scale;
    a->ring = ring;
    a->ring_slots = ring_slots;
    a->first_slot = (int)(first_frame % ring_slots);
    a->elems_per_frame = elems_per_frame;
    a->elem_off = elem_off;
    a->xt = nullptr;
    return DDM_OK;
}

extern "C" int ddm_k_tile_fft(const void *frames, int n_frames, long long frame_bytes, int width, int height,
                              int bytes_per_pixel, int pixel_format, int x_off, int y_off, int largest_scale,
                              int scale, float *ring, int ring_slots, long long elems_per_frame,
                              long long elem_off, long long first_frame, void *stream) {
    ddm::TileFftArgs a;
    int rc = tile_fft_args(frames, n_frames, frame_bytes, width, height, bytes_per_pixel, pixel_format, x_off, y_off,
                           largest_scale, scale, ring, ring_slots, elems_per_frame, elem_off, first_frame, &a);
    if (rc) return rc;
    // the between-pass buffer of 256 .. 1024 tiles lives for this call only
    // (stream-ordered allocation and release)
    const hipStream_t st = (hipStream_t)stream;
    const long long xt = ddm::big_fft_scratch_elems(largest_scale, scale, n_frames);
    if (xt && hipMallocAsync((void **)&a.xt, (size_t)xt * 2 * sizeof(float), st) != hipSuccess) return DDM_ENOMEM;
    const hipError_t err = ddm::launch_tile_fft(a, st);
    if (a.xt) (void)hipFreeAsync(a.xt, st);
    return err == hipSuccess ? DDM_OK : DDM_EDEVICE;
}

// ---------------------------------------------------------------------------
// ddm_k_tile_fft without device allocation: the pass buffer of the 256 ..
// 1024 tile transforms is this file-scope array, one copy per device, which
// every call on that device shares (calls that overlap on one device race on
// it).  Its device address is looked up once per device; in host code the
// array's name is only a stand-in and is never passed to a kernel.
// ---------------------------------------------------------------------------
constexpr long long kScratchElems = 1LL << 23;  // complex values: 64 MiB
static __device__ __align__(256) float2 scratch[kScratchElems];
static std::mutex g_scratch_mu;
static float2 *g_scratch_ptr[64] = {nullptr};

static hipError_t scratch_ptr(float2 **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        float2 *p = nullptr;
        e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

extern "C" int ddm_k_tile_fft_noalloc(const void *frames, int n_frames, long long frame_bytes, int width,
                                      int height, int bytes_per_pixel, int pixel_format, int x_off, int y_off,
                                      int largest_scale, int scale, float *ring, int ring_slots,
                                      long long elems_per_frame, long long elem_off, long long first_frame,
                                      void *stream) {
    ddm::TileFftArgs a;
    int rc = tile_fft_args(frames, n_frames, frame_bytes, width, height, bytes_per_pixel, pixel_format, x_off, y_off,
                           largest_scale, scale, ring, ring_slots, elems_per_frame, elem_off, first_frame, &a);
    if (rc) return rc;
    const long long per_frame = ddm::big_fft_scratch_elems(largest_scale, scale, 1);
    if (per_frame > kScratchElems) {
        g_create_err = "ddm_k_tile_fft_noalloc: the pass buffer of one frame (" + std::to_string(per_frame * 8) +
                       " bytes) exceeds the " + std::to_string(kScratchElems * 8) + "-byte scratch array";
        return DDM_ERANGE;
    }
    const hipStream_t st = (hipStream_t)stream;
    // the twiddle table lives in static device memory as well
    const float2 *tw = nullptr;
    if (ddm::twiddle_table(&tw) != hipSuccess) return DDM_EDEVICE;
    float2 *xt = nullptr;
    if (per_frame && scratch_ptr(&xt) != hipSuccess) return DDM_EDEVICE;
    // frames in batches whose pass buffer fits the array; the batches run in
    // stream order, so each reuses it after the previous one is done
    const int batch = per_frame ? (int)std::min<long long>(n_frames, kScratchElems / per_frame) : n_frames;
    const uint8_t *src = (const uint8_t *)frames;
    for (int i = 0; i < n_frames; i += batch) {
        ddm::TileFftArgs b = a;
        b.frames = src + (long long)i * frame_bytes;
        b.n_frames = std::min(batch, n_frames - i);
        b.first_slot = (int)((first_frame + i) % ring_slots);
        b.xt = (float *)xt;
        if (ddm::launch_tile_fft(b, st) != hipSuccess) return DDM_EDEVICE;
    }
    return DDM_OK;
}

extern "C" int ddm_k_lag_accumulate(const float *ring, int ring_slots, long long elems_per_frame,
                                    long long first_frame, int n_new, const int *lags, int lag_count, int max_lag,
                                    float *acc, int lag_block, void *stream) {
    if (!ring || !acc || !lags || ring_slots <= 0 || elems_per_frame <= 0 || first_frame < 0 || lag_count <= 0)
        return DDM_EINVAL;
    if (elems_per_frame % ddm::RING_GROUP != 0) return DDM_EINVAL;
    if (lag_block != 32 && lag_block != 64 && lag_block != 128 && lag_block != 256) return DDM_EINVAL;
    if (n_new <= 0 || n_new > lag_block) return DDM_EINVAL;
    if (max_lag < 0 || (long long)max_lag + n_new > ring_slots) return DDM_ERANGE;
    if (ring_slots >= ddm::kMaxRingSlots) return DDM_ERANGE;  // 32-bit offsets inside a ring group
    // the launch plan is made on the host: bring the lag list over (this
    // waits for the stream)
    std::vector<int> host_lags(lag_count);
    if (hipMemcpyAsync(host_lags.data(), lags, sizeof(int) * lag_count, hipMemcpyDeviceToHost,
                       (hipStream_t)stream) != hipSuccess ||
        hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        return DDM_EDEVICE;
    for (int l = 0; l < lag_count; ++l)
        if (host_lags[l] < 0 || host_lags[l] > max_lag || (l && host_lags[l] <= host_lags[l - 1])) return DDM_EINVAL;
    ddm::LagArgs a;
    a.ring = ring;
    a.acc = acc;
    a.elems_per_frame = elems_per_frame;
    a.ring_slots = ring_slots;
    a.slot0 = (int)(first_frame % ring_slots);
    a.n_new = n_new;
    a.dmax = (int)std::min<long long>(max_lag, first_frame);
    a.lags = lags;
    a.lags_host = host_lags.data();
    a.lag_count = lag_count;
    a.lag_block = lag_block;
    return ddm::launch_lag(a, (hipStream_t)stream) == hipSuccess ? DDM_OK : DDM_EDEVICE;
}

extern "C" int ddm_k_fit(const double *curves, const double *tau, int L, long long n, int iters, double *out,
                         void *stream) {
    if (!curves || !tau || !out || L < 3 || n < 0 || iters < 1 || iters > 10000) return DDM_EINVAL;
    return ddm::launch_fit(curves, tau, L, n, iters, out, (hipStream_t)stream) == hipSuccess ? DDM_OK
                                                                                           : DDM_EDEVICE;
}
