// Serving loop: many DDM streams on one device.
//
// A stream is an engine (its own spectrum ring, accumulators and staging)
// admitted when its device footprint fits in the server's budget.  The
// reference has no counterpart (it analyses one video per process,
// src/DDM.cu:268-896); its memory report (src/DDM.cu:493-634) is the seed of
// the footprint accounting.  Two ways to run the chunks:
//
// * per stream: ddm_server_submit uploads a chunk and queues its tile
//   transforms and lag pass on the stream's own HIP stream, so chunks of
//   different streams overlap on the device;
// * batched (DDM_SERVER_BATCHED): ddm_server_submit only uploads the chunk
//   and queues it; a scheduler thread gathers the waiting chunks of all
//   streams (the oldest of each, or with DDM_SERVER_GROUP all of a stream's
//   waiting chunks up to a lag block, which then share one lag pass) into a
//   tick, writes the tick's work
//   queue -- one job per (stream, scale, kernel kind) for the tile
//   transforms and one per stream for the lag engine, each job covering its
//   tile blocks and frames or its ring groups -- and runs every kernel kind
//   as one launch over that queue, a workgroup per item.  The launch
//   sequence of a tick is captured as a HIP graph (on a stream used only
//   for capture) and launched on the tick stream; ticks with the same
//   sequence of kinds replay one executable graph, updated to the tick's
//   grid sizes.  Requests that arrive while a tick runs go into the next one
//   (continuous batching); at most two ticks are in flight.
//
// Either way every chunk is flushed through the lag engine, its
// chunk-to-result latency (host submission -> chunk accumulated on the
// device) is recorded by a poller thread from an event recorded after it,
// and a stream's results depend only on its own frames: a chunk's kernels
// run the same code on the same data alone or in a batch, so results are
// bit-identical alone, in a mix and between the two modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ddm_batch.h"
#include "ddm_internal.h"
#include "ddm_kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
    hipEvent_t ev;
    Clock::time_point t_submit;
};

// A chunk uploaded to one of the stream's staging buffers, waiting for a tick.
struct Chunk {
    int b;                 // staging buffer
    int n;                 // frames
    hipEvent_t ev_up;      // upload done
    const uint8_t *stage;  // device address of the frames (ROI rows, M x M each)
    Clock::time_point t_submit;
};

struct Stream {
    ddm_engine *eng = nullptr;
    size_t footprint = 0;
    std::deque<Pending> pending;
    std::vector<float> latency_ms;
    long long frames = 0;
    int busy = 0;          // calls (and ticks) using the engine outside the lock
    bool closing = false;  // ddm_server_close waits for busy == 0
    std::deque<Chunk> queue;  // batched: uploaded chunks waiting for a tick
    int outstanding = 0;      // batched: chunks uploaded and not yet committed by a tick
    int lag_block = 0;        // batched: frames one lag pass of this stream may take
    int piece = 0;            // batched: frames of one queued chunk, at most
    int nstage = 0;           // batched: staging buffers (chunks uploaded and not yet committed, at most)
    int urgent = 0;           // batched: calls waiting for this stream's queue to drain (no holding)
    long long xt_need = 0;    // batched: pass-buffer values one chunk of this stream takes
    std::shared_ptr<std::mutex> submit_mu = std::make_shared<std::mutex>();  // one submission at a time
};

// Work queue of one tick in device memory (and its pinned host image): for
// every kernel kind a BatchHdr and its jobs, then the lag plans.  Two are
// used in turn, so a tick's tables are written while the previous tick runs.
struct TickBuf {
    uint8_t *h = nullptr, *d = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;  // the last tick that used this buffer has finished
    bool used = false;
};

struct FftLaunch {
    ddm::FftKey key;
    size_t hdr_off, jobs_off;
    long long blocks;
};
struct LagLaunch {
    int J;
    bool comp;
    size_t hdr_off, jobs_off;
    long long groups;
};

constexpr size_t TABLE_ALIGN = 256;
inline size_t align_up(size_t v) { return (v + TABLE_ALIGN - 1) / TABLE_ALIGN * TABLE_ALIGN; }

}  // namespace

struct ddm_server {
    size_t budget = 0, used = 0;
    int max_streams = 0;
    std::mutex mu;
    std::condition_variable idle;  // a stream's busy count dropped / a queue drained
    std::condition_variable work;  // batched: chunks queued or stop
    std::map<int, Stream> streams;
    int next_id = 0;
    std::vector<hipEvent_t> pool;
    std::thread poller;
    std::atomic<bool> stop{false};
    long long frames_total = 0, chunks_total = 0;
    // batched mode
    bool batched = false;
    bool use_graph = true;
    bool group = false;  // DDM_SERVER_GROUP
    long long hold_us = 0;  // grouping: how long a stream's chunk may wait for more of its chunks
    int draining = 0;       // ddm_server_drain calls waiting (no holding)
    std::thread sched;
    bool sched_stop = false;
    hipStream_t tick_st = nullptr;
    // ticks are captured on a stream of their own (never waited on or used
    // for events) and launched on tick_st: HIP refuses to wait for an event
    // recorded on a stream while that stream is capturing
    hipStream_t cap_st = nullptr;
    std::mutex tick_mu;  // the scheduler's tick vs results / close on the same engines
    TickBuf tb[2];
    long long ticks = 0, graph_ticks = 0, chunks_batched = 0, items = 0;
    float2 *d_xt = nullptr;
    long long xt_cap = 0;  // complex values
    std::map<std::string, hipGraphExec_t> graphs;
    const float2 *tw = nullptr;
    std::string tick_err;
};

// Message of the last failing server call on this thread (the server may be
// used from several threads, so it keeps no shared message).
static thread_local std::string g_server_err;

static int serr(int code, const std::string &msg) {
    g_server_err = msg;
    return code;
}

// Marks a stream in use so that ddm_server_close waits for the caller to
// finish with its engine.  Caller holds mu.
static int acquire(ddm_server *s, int stream_id, ddm_engine **eng) {
    auto it = s->streams.find(stream_id);
    if (it == s->streams.end() || it->second.closing) return serr(DDM_ERANGE, "unknown stream");
    it->second.busy += 1;
    *eng = it->second.eng;
    return DDM_OK;
}

static void release_stream(ddm_server *s, int stream_id) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->streams.find(stream_id);
    if (it != s->streams.end()) it->second.busy -= 1;
    s->idle.notify_all();
}

// Moves finished chunks from pending to the latency record.  Caller holds mu.
static void harvest(ddm_server *s, Stream &st) {
    while (!st.pending.empty()) {
        Pending &p = st.pending.front();
        if (hipEventQuery(p.ev) != hipSuccess) break;
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - p.t_submit).count();
        st.latency_ms.push_back((float)ms);
        s->pool.push_back(p.ev);
        st.pending.pop_front();
    }
}

static void poll_loop(ddm_server *s) {
    while (!s->stop.load()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &kv : s->streams) harvest(s, kv.second);
        }
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

// An event from the pool (caller holds mu).
static hipEvent_t take_event(ddm_server *s) {
    if (!s->pool.empty()) {
        hipEvent_t ev = s->pool.back();
        s->pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return nullptr;
    return ev;
}

// ---------------------------------------------------------------------------
// Batched ticks
// ---------------------------------------------------------------------------
// The chunks of one stream a tick takes: the oldest waiting one, or (with
// DDM_SERVER_GROUP) all its waiting chunks up to a lag block of frames,
// which then share one lag pass.
struct Taken {
    int id;
    ddm_engine *eng;
    std::vector<Chunk> cs;
};

static bool ensure_table(ddm_server *s, TickBuf &tb, size_t bytes) {
    if (bytes <= tb.cap) return true;
    // the previous user of this buffer has finished (checked by the caller)
    if (tb.h) (void)hipHostFree(tb.h);
    if (tb.d) (void)hipFree(tb.d);
    tb.h = nullptr;
    tb.d = nullptr;
    tb.cap = 0;
    const size_t cap = std::max(bytes * 2, (size_t)1 << 20);
    if (hipHostMalloc((void **)&tb.h, cap, hipHostMallocDefault) != hipSuccess) return false;
    if (hipMalloc((void **)&tb.d, cap) != hipSuccess) return false;
    tb.cap = cap;
    (void)s;
    return true;
}

static bool ensure_xt(ddm_server *s, long long need) {
    if (need <= s->xt_cap) return true;
    // earlier ticks may still read the old buffer
    if (hipStreamSynchronize(s->tick_st) != hipSuccess) return false;
    if (s->d_xt) (void)hipFree(s->d_xt);
    s->d_xt = nullptr;
    s->xt_cap = 0;
    if (hipMalloc((void **)&s->d_xt, (size_t)need * sizeof(float2)) != hipSuccess) return false;
    s->xt_cap = need;
    return true;
}

// Launches the tick's batched tile transforms on stream st.
static hipError_t launch_tick_fft(ddm_server *s, const TickBuf &tb, const std::vector<FftLaunch> &fl,
                                  hipStream_t st) {
    for (const FftLaunch &f : fl) {
        hipError_t e = ddm::launch_fft_batch(f.key, (const ddm::BatchHdr *)(tb.d + f.hdr_off),
                                             (const ddm::FftJob *)(tb.d + f.jobs_off), f.blocks, s->tw, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Launches the tick's batched lag passes on stream st (after every
// transform of the tick).
static hipError_t launch_tick_lag(ddm_server *s, const TickBuf &tb, const std::vector<LagLaunch> &ll,
                                  size_t plans_off, hipStream_t st) {
    for (const LagLaunch &l : ll) {
        hipError_t e = ddm::launch_lag_batch(l.J, l.comp, (const ddm::BatchHdr *)(tb.d + l.hdr_off),
                                             (const ddm::LagJob *)(tb.d + l.jobs_off),
                                             (const ddm::LagPlan *)(tb.d + plans_off), l.groups, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Builds and runs one tick over the taken chunks (one per stream).  Returns
// DDM_OK or an error code (s->tick_err).
static int run_tick(ddm_server *s, const std::vector<Taken> &taken) {
    std::lock_guard<std::mutex> g(s->tick_mu);
    const int k = (int)(s->ticks & 1);
    TickBuf &tb = s->tb[k];
    auto fail = [&](const std::string &m) {
        s->tick_err = m;
        return DDM_EDEVICE;
    };
    if (tb.used && hipEventSynchronize(tb.done) != hipSuccess) return fail("waiting for a tick");
    // ---- the work queue ----
    std::map<ddm::FftKey, std::vector<ddm::FftJob>> fft;  // in launch order (see FftKey)
    std::map<std::pair<int, int>, std::vector<ddm::LagJob>> lag;
    std::vector<ddm::LagPlan> plans;
    std::vector<ddm::TileFftArgs> direct_fft;
    std::vector<ddm::LagArgs> direct_lag;
    std::vector<std::vector<int>> direct_lag_lists;
    long long xt_off = 0;
    std::vector<std::pair<ddm::FftKey, ddm::FftJob>> tmp;
    for (const Taken &t : taken) {
        ddm::BatchStream bs;
        ddm::engine_batch_stream(t.eng, &bs);
        const ddm::Params &p = *bs.p;
        const int R = p.ring_slots;
        const int slot = (int)(bs.frames % R);  // of the tick's first frame of this stream
        const size_t roi_frame = (size_t)p.M * p.M * p.bpp;
        int n = 0;  // frames of the stream in this tick
        for (const Chunk &c : t.cs) {
        const int cslot = (int)((bs.frames + n) % R);
        n += c.n;
        ddm::PixSrc ps{c.stage, (long long)roi_frame, (long long)p.M * p.bpp, p.bpp, p.fmt, 0, 0};
        float2 *xt = nullptr;
        long long xt_need = 0;
        for (const ddm::ScaleGeom &gm : p.geom)
            xt_need = std::max(xt_need, ddm::big_fft_scratch_elems(p.M, gm.s, c.n));
        if (xt_need) {
            xt = s->d_xt + xt_off;
            xt_off += xt_need;
        }
        for (const ddm::ScaleGeom &gm : p.geom) {
            const ddm::TileLayout &lay = bs.lay[&gm - p.geom.data()];
            ddm::RingDst rd{(float2 *)bs.ring, gm.elem_off, R, cslot, lay};
            tmp.clear();
            const int fam = bs.family[&gm - p.geom.data()];
            if (bs.precision == DDM_PREC_FP32 && fam == DDM_FFT_REGISTER &&
                ddm::fft_batch_jobs(ps, rd, gm.s, gm.tps, gm.tiles, c.n, xt, &tmp)) {
                for (auto &kj : tmp) fft[kj.first].push_back(kj.second);
            } else {
                ddm::TileFftArgs a;
                a.frames = c.stage;
                a.n_frames = c.n;
                a.frame_bytes = (long long)roi_frame;
                a.width = p.M;
                a.bpp = p.bpp;
                a.fmt = p.fmt;
                a.x_off = a.y_off = 0;
                a.largest_scale = p.M;
                a.scale = gm.s;
                a.ring = bs.ring;
                a.ring_slots = R;
                a.first_slot = cslot;
                a.elems_per_frame = p.elems_padded;
                a.elem_off = gm.elem_off;
                a.lay = lay;
                a.family = fam;
                a.precision = bs.precision;
                a.xt = (float *)xt;
                direct_fft.push_back(a);
            }
        }
        }
        const int dmax = (int)std::min<long long>(p.max_lag, bs.frames);
        if (bs.lag_block <= 128) {
            ddm::LagJob j{};
            j.ring = (const float2 *)bs.ring;
            j.acc = bs.acc;
            j.comp = bs.comp;
            j.lags = bs.lags;
            j.E = p.elems_padded;
            j.R = R;
            j.slot0 = slot;
            j.n = n;
            j.dmax = dmax;
            j.L = (int)p.lags.size();
            j.plan = (int)plans.size();
            plans.emplace_back();
            ddm::make_lag_plan(bs.lag_block, p.lags.data(), j.L, n, dmax, &plans.back());
            lag[{bs.lag_block, bs.comp ? 1 : 0}].push_back(j);
        } else {
            ddm::LagArgs a;
            a.ring = bs.ring;
            a.acc = bs.acc;
            a.comp = bs.comp;
            a.elems_per_frame = p.elems_padded;
            a.ring_slots = R;
            a.slot0 = slot;
            a.n_new = n;
            a.dmax = dmax;
            a.lags = bs.lags;
            a.lags_host = p.lags.data();
            a.lag_count = (int)p.lags.size();
            a.lag_block = bs.lag_block;
            direct_lag.push_back(a);
        }
    }
    // ---- table layout ----
    std::vector<FftLaunch> fl;
    std::vector<LagLaunch> ll;
    size_t off = 0;
    std::string sig;
    for (auto &kv : fft) {
        FftLaunch f;
        f.key = kv.first;
        f.hdr_off = off;
        off = align_up(off + sizeof(ddm::BatchHdr));
        f.jobs_off = off;
        off = align_up(off + kv.second.size() * sizeof(ddm::FftJob));
        long long first = 0;
        for (ddm::FftJob &j : kv.second) {
            j.first = first;
            first += j.nblk;
        }
        f.blocks = first;
        fl.push_back(f);
        sig += "f" + std::to_string(f.key.kind) + "." + std::to_string(f.key.scale) + "." + std::to_string(f.key.pix);
    }
    for (auto &kv : lag) {
        LagLaunch l;
        l.J = kv.first.first;
        l.comp = kv.first.second != 0;
        l.hdr_off = off;
        off = align_up(off + sizeof(ddm::BatchHdr));
        l.jobs_off = off;
        off = align_up(off + kv.second.size() * sizeof(ddm::LagJob));
        long long first = 0;
        for (ddm::LagJob &j : kv.second) {
            j.first = first;
            first += j.E / ddm::RING_GROUP;
        }
        l.groups = first;
        ll.push_back(l);
        sig += "l" + std::to_string(l.J) + "." + std::to_string((int)l.comp);
    }
    const size_t plans_off = off;
    off = align_up(off + std::max<size_t>(plans.size(), 1) * sizeof(ddm::LagPlan));
    if (!ensure_table(s, tb, off)) return fail("tick table allocation failed");
    for (size_t i = 0; i < fl.size(); ++i) {
        const auto &jobs = fft[fl[i].key];
        ddm::BatchHdr h{(int)jobs.size(), 0, fl[i].blocks};
        std::memcpy(tb.h + fl[i].hdr_off, &h, sizeof h);
        std::memcpy(tb.h + fl[i].jobs_off, jobs.data(), jobs.size() * sizeof(ddm::FftJob));
    }
    {
        size_t i = 0;
        for (auto &kv : lag) {
            ddm::BatchHdr h{(int)kv.second.size(), 0, ll[i].groups};
            std::memcpy(tb.h + ll[i].hdr_off, &h, sizeof h);
            std::memcpy(tb.h + ll[i].jobs_off, kv.second.data(), kv.second.size() * sizeof(ddm::LagJob));
            ++i;
        }
    }
    if (!plans.empty()) std::memcpy(tb.h + plans_off, plans.data(), plans.size() * sizeof(ddm::LagPlan));
    // ---- run: uploads, table, kernels ----
    for (const Taken &t : taken)
        for (const Chunk &c : t.cs)
            if (hipStreamWaitEvent(s->tick_st, c.ev_up, 0) != hipSuccess) return fail("waiting for an upload");
    if (hipMemcpyAsync(tb.d, tb.h, off, hipMemcpyHostToDevice, s->tick_st) != hipSuccess)
        return fail("tick table upload failed");
    const bool graph = s->use_graph && direct_fft.empty() && direct_lag.empty();
    if (graph) {
        // the kernels of the tick as a graph; an executable graph of the same
        // launch sequence (and table buffer) is updated to this tick's grids
        const std::string key = sig + "#" + std::to_string(k);
        if (hipStreamBeginCapture(s->cap_st, hipStreamCaptureModeRelaxed) != hipSuccess)
            return fail("graph capture failed to start");
        hipError_t le = launch_tick_fft(s, tb, fl, s->cap_st);
        if (le == hipSuccess) le = launch_tick_lag(s, tb, ll, plans_off, s->cap_st);
        hipGraph_t gr = nullptr;
        const hipError_t ce = hipStreamEndCapture(s->cap_st, &gr);
        if (le != hipSuccess || ce != hipSuccess || !gr) {
            if (gr) (void)hipGraphDestroy(gr);
            return fail(std::string("tick capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
        }
        hipGraphExec_t ex = nullptr;
        auto it = s->graphs.find(key);
        if (it != s->graphs.end()) {
            hipGraphNode_t en = nullptr;
            hipGraphExecUpdateResult ur;
            if (hipGraphExecUpdate(it->second, gr, &en, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess) {
                ex = it->second;
            } else {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(it->second);
                s->graphs.erase(it);
            }
        }
        if (!ex) {
            if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(gr);
                return fail("graph instantiation failed");
            }
            s->graphs[key] = ex;
        }
        const hipError_t ge = hipGraphLaunch(ex, s->tick_st);
        (void)hipGraphDestroy(gr);
        if (ge != hipSuccess) return fail(std::string("graph launch failed: ") + hipGetErrorString(ge));
        s->graph_ticks += 1;
    } else {
        // every transform of the tick, then every lag pass
        hipError_t e = launch_tick_fft(s, tb, fl, s->tick_st);
        for (const ddm::TileFftArgs &a : direct_fft)
            if (e == hipSuccess) e = ddm::launch_tile_fft(a, s->tick_st);
        if (e == hipSuccess) e = launch_tick_lag(s, tb, ll, plans_off, s->tick_st);
        for (const ddm::LagArgs &a : direct_lag)
            if (e == hipSuccess) e = ddm::launch_lag(a, s->tick_st);
        if (e != hipSuccess) return fail(std::string("tick launch failed: ") + hipGetErrorString(e));
    }
    if (hipEventRecord(tb.done, s->tick_st) != hipSuccess) return fail("tick event");
    tb.used = true;
    for (const Taken &t : taken)
        for (const Chunk &c : t.cs) {
            const int rc = ddm::engine_batch_commit(t.eng, c.n, c.b, s->tick_st);
            if (rc) return fail(ddm_last_error(t.eng));
            s->chunks_batched += 1;
        }
    s->ticks += 1;
    for (const FftLaunch &f : fl) s->items += f.blocks;
    for (const LagLaunch &l : ll) s->items += l.groups;
    return DDM_OK;
}

static bool any_queued(const ddm_server *s) {
    for (const auto &kv : s->streams)
        if (!kv.second.queue.empty()) return true;
    return false;
}

// Whether a stream's waiting chunks go into the next tick: always, unless
// grouping holds them (for at most hold_us after the oldest arrived) while a
// fuller lag pass can still form and nobody waits for the stream.
static bool stream_ready(const ddm_server *s, const Stream &st, Clock::time_point now) {
    if (st.queue.empty()) return false;
    if (!s->group || s->hold_us <= 0 || s->sched_stop || s->draining || st.urgent) return true;
    int frames = 0;
    for (const Chunk &c : st.queue) frames += c.n;
    if (frames + st.piece > st.lag_block) return true;        // the pass is full
    if ((int)st.queue.size() + 1 >= st.nstage) return true;   // the next submission would block
    return now - st.queue.front().t_submit >= std::chrono::microseconds(s->hold_us);
}

static void sched_loop(ddm_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    int start = 0;  // streams take turns at being first when the pass buffer is short
    for (;;) {
        // wait for a stream whose chunks are due
        bool ready = false;
        for (;;) {
            const Clock::time_point now = Clock::now();
            Clock::time_point due = Clock::time_point::max();
            for (auto &kv : s->streams) {
                if (stream_ready(s, kv.second, now)) {
                    ready = true;
                    break;
                }
                if (!kv.second.queue.empty())
                    due = std::min(due, kv.second.queue.front().t_submit + std::chrono::microseconds(s->hold_us));
            }
            if (ready) break;
            if (s->sched_stop && !any_queued(s)) return;
            if (due == Clock::time_point::max())
                s->work.wait(lk);
            else
                s->work.wait_until(lk, due);
        }
        // the chunks of every due stream (the oldest, or with grouping all
        // of them up to a lag block), within the tick's pass buffer
        std::vector<Taken> taken;
        std::vector<std::pair<int, Stream *>> order;
        const Clock::time_point now = Clock::now();
        for (auto &kv : s->streams)
            if (stream_ready(s, kv.second, now)) order.emplace_back(kv.first, &kv.second);
        std::rotate(order.begin(), order.begin() + (start % order.size()), order.end());
        ++start;
        long long xt = 0;
        for (auto &o : order) {
            Stream &st = *o.second;
            Taken t{o.first, st.eng, {}};
            int frames = 0;
            while (!st.queue.empty()) {
                const Chunk &c = st.queue.front();
                if (!t.cs.empty() && (!s->group || frames + c.n > st.lag_block)) break;
                if (!(taken.empty() && t.cs.empty()) && xt + st.xt_need > s->xt_cap) break;
                xt += st.xt_need;
                frames += c.n;
                t.cs.push_back(c);
                st.queue.pop_front();
            }
            if (t.cs.empty()) continue;
            st.busy += 1;
            taken.push_back(std::move(t));
        }
        lk.unlock();
        int rc = DDM_OK;
        if (!ensure_xt(s, xt)) {
            s->tick_err = "pass buffer allocation failed";
            rc = DDM_ENOMEM;
        }
        if (rc == DDM_OK) rc = run_tick(s, taken);
        lk.lock();
        for (const Taken &t : taken) {
            auto it = s->streams.find(t.id);
            if (it == s->streams.end()) continue;
            Stream &st = it->second;
            for (const Chunk &c : t.cs) {
                if (rc == DDM_OK) {
                    hipEvent_t ev = take_event(s);
                    if (ev && hipEventRecord(ev, s->tick_st) == hipSuccess)
                        st.pending.push_back(Pending{ev, c.t_submit});
                }
                st.outstanding -= 1;
            }
            st.busy -= 1;
        }
        s->idle.notify_all();
    }
}

// Waits (holding lk on mu) until every chunk of the stream has been
// launched by a tick on the tick stream (taken from the queue is not enough:
// the scheduler takes chunks before it launches their tick), and returns the
// stream or nullptr when it is gone.
static Stream *wait_chunks_launched(ddm_server *s, std::unique_lock<std::mutex> &lk, int stream_id) {
    for (;;) {
        auto it = s->streams.find(stream_id);
        if (it == s->streams.end()) return nullptr;
        if (it->second.queue.empty() && it->second.outstanding == 0) return &it->second;
        s->idle.wait(lk);
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int ddm_server_create2(size_t device_budget, int max_streams, int flags, ddm_server **out) {
    if (!out || max_streams <= 0 || (flags & ~(DDM_SERVER_BATCHED | DDM_SERVER_NO_GRAPH | DDM_SERVER_GROUP)))
        return serr(DDM_EINVAL, "bad server arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return serr(DDM_EDEVICE, "no HIP device available");
    if (device_budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return serr(DDM_EDEVICE, "hipMemGetInfo failed");
        device_budget = free_b / 10 * 9;
    }
    // Engines allocate from the device's stream-ordered pool; keeping freed
    // memory in the pool makes admitting a stream after another left cheap.
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    ddm_server *s = new ddm_server();
    s->budget = device_budget;
    s->max_streams = max_streams;
    if (flags & DDM_SERVER_BATCHED) {
        s->batched = true;
        s->use_graph = !(flags & DDM_SERVER_NO_GRAPH);
        s->group = (flags & DDM_SERVER_GROUP) != 0;
        bool ok = hipStreamCreateWithFlags(&s->tick_st, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&s->cap_st, hipStreamNonBlocking) == hipSuccess;
        for (TickBuf &tb : s->tb) ok = ok && hipEventCreateWithFlags(&tb.done, hipEventDisableTiming) == hipSuccess;
        ok = ok && ddm::twiddle_table(&s->tw) == hipSuccess;
        // the tick pass buffer (256 .. 1024 tile row spectra of the chunks in
        // one tick) comes out of the budget
        const long long xt = (long long)1 << 27;  // 1 GiB of complex values
        ok = ok && s->budget > (size_t)xt * sizeof(float2) * 2 && ensure_xt(s, xt);
        if (!ok) {
            for (TickBuf &tb : s->tb)
                if (tb.done) (void)hipEventDestroy(tb.done);
            if (s->tick_st) (void)hipStreamDestroy(s->tick_st);
            if (s->cap_st) (void)hipStreamDestroy(s->cap_st);
            if (s->d_xt) (void)hipFree(s->d_xt);
            delete s;
            return serr(DDM_EDEVICE, "batched server set-up failed");
        }
        s->budget -= (size_t)xt * sizeof(float2);
        s->sched = std::thread(sched_loop, s);
    }
    s->poller = std::thread(poll_loop, s);
    *out = s;
    return DDM_OK;
}

extern "C" int ddm_server_set_hold(ddm_server *s, double hold_ms) {
    if (!s || !(hold_ms >= 0.0) || hold_ms > 10000.0) return serr(DDM_EINVAL, "bad hold time");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->hold_us = (long long)(hold_ms * 1000.0);
    }
    s->work.notify_all();
    return DDM_OK;
}

extern "C" int ddm_server_create(size_t device_budget, int max_streams, ddm_server **out) {
    return ddm_server_create2(device_budget, max_streams, 0, out);
}

extern "C" void ddm_server_destroy(ddm_server *s) {
    if (!s) return;
    if (s->batched) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->sched_stop = true;
        }
        s->work.notify_all();
        if (s->sched.joinable()) s->sched.join();  // runs the chunks still queued first
        (void)hipStreamSynchronize(s->tick_st);
    }
    s->stop.store(true);
    if (s->poller.joinable()) s->poller.join();
    for (auto &kv : s->streams) {
        ddm_sync(kv.second.eng);
        for (Pending &p : kv.second.pending) s->pool.push_back(p.ev);
        ddm_destroy(kv.second.eng);
    }
    for (hipEvent_t ev : s->pool) (void)hipEventDestroy(ev);
    if (s->batched) {
        for (auto &kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
        for (TickBuf &tb : s->tb) {
            if (tb.h) (void)hipHostFree(tb.h);
            if (tb.d) (void)hipFree(tb.d);
            if (tb.done) (void)hipEventDestroy(tb.done);
        }
        if (s->d_xt) (void)hipFree(s->d_xt);
        (void)hipStreamDestroy(s->tick_st);
        (void)hipStreamDestroy(s->cap_st);
    }
    delete s;
}

extern "C" int ddm_server_open(ddm_server *s, const ddm_desc *desc, int *stream_id) {
    if (!s || !desc || !stream_id) return DDM_EINVAL;
    ddm::Params p;
    std::string msg;
    int rc = ddm::check_desc(desc, &p, &msg);
    if (rc != DDM_OK) return serr(rc, msg);
    // grouping: enough staging buffers for a full lag pass of chunks to wait
    // while the next one uploads
    const int piece = std::min(p.chunk, p.lag_block);
    const int stages = (s->batched && s->group)
                           ? std::min(ddm::MAX_UPLOAD_STAGES, std::max(ddm::UPLOAD_STAGES, p.lag_block / piece + 2))
                           : ddm::UPLOAD_STAGES;
    const size_t need = ddm::device_footprint(p) +
                        (size_t)(stages - ddm::UPLOAD_STAGES) * p.chunk * p.M * p.M * p.bpp;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if ((int)s->streams.size() >= s->max_streams) return serr(DDM_EBUSY, "stream slots are all in use");
        if (s->used + need > s->budget)
            return serr(DDM_EBUSY, "stream footprint (" + std::to_string(need) +
                                       " bytes) exceeds the device budget left (" +
                                       std::to_string(s->budget - s->used) + " bytes)");
        s->used += need;  // reserved while the engine is created
    }
    ddm_engine *e = nullptr;
    rc = ddm_create(desc, &e);
    if (rc == DDM_ENOMEM) {
        // the device had less room than the budget assumed; the pool may be
        // holding freed blocks of other sizes: return them and try once more
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
            (void)hipMemPoolTrimTo(pool, 0);
        rc = ddm_create(desc, &e);
    }
    std::string why;
    if (rc == DDM_OK && stages > ddm::UPLOAD_STAGES) {
        rc = ddm::engine_batch_stages(e, stages);
        if (rc != DDM_OK) {
            why = ddm_last_error(e);
            ddm_destroy(e);
        }
    } else if (rc != DDM_OK) {
        why = ddm_last_error(nullptr);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != DDM_OK) {
        s->used -= need;
        // a refusal like any other over-budget request: the caller may try
        // again when streams have left
        if (rc == DDM_ENOMEM) return serr(DDM_EBUSY, "device memory exhausted below the budget (" + why + ")");
        return serr(rc, why);
    }
    const int id = s->next_id++;
    Stream &st = s->streams[id];
    st.eng = e;
    st.footprint = need;
    st.lag_block = p.lag_block;
    st.piece = piece;
    st.nstage = stages;
    for (const ddm::ScaleGeom &g : p.geom)
        st.xt_need = std::max(st.xt_need, ddm::big_fft_scratch_elems(p.M, g.s, piece));
    *stream_id = id;
    return DDM_OK;
}

// Batched submission: the frames go to the stream's staging buffers in
// pieces of at most a chunk and a lag block, each queued for a tick.
static int submit_batched(ddm_server *s, int stream_id, ddm_engine *e, const void *frames, int n_frames) {
    std::shared_ptr<std::mutex> smu;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        smu = s->streams.find(stream_id)->second.submit_mu;
    }
    std::lock_guard<std::mutex> sub(*smu);
    ddm::BatchStream bs;
    ddm::engine_batch_stream(e, &bs);
    const int piece = std::min(bs.p->chunk, bs.lag_block);
    const size_t fb = (size_t)bs.p->width * bs.p->height * bs.p->bpp;
    const Clock::time_point t0 = Clock::now();
    for (int i = 0; i < n_frames; i += piece) {
        Chunk c;
        c.n = std::min(piece, n_frames - i);
        c.t_submit = t0;
        // back-pressure: the staging buffer this chunk takes held the
        // stream's chunk nstage back, which must have been committed
        // by its tick (the upload then waits for that tick to finish)
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->idle.wait(lk, [&] {
                const Stream &st = s->streams.find(stream_id)->second;
                return st.outstanding < st.nstage;
            });
        }
        const int rc = ddm::engine_batch_upload(e, (const uint8_t *)frames + (size_t)i * fb, c.n, &c.b, &c.ev_up,
                                                &c.stage);
        if (rc) return serr(rc, ddm_last_error(e));
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->streams.find(stream_id);
        it->second.outstanding += 1;
        it->second.queue.push_back(c);
        it->second.frames += c.n;
        s->frames_total += c.n;
        s->chunks_total += 1;
        s->work.notify_one();
    }
    return DDM_OK;
}

extern "C" int ddm_server_submit(ddm_server *s, int stream_id, const void *frames, int n_frames) {
    if (!s) return DDM_EINVAL;
    if (n_frames < 0 || (n_frames > 0 && !frames)) return serr(DDM_EINVAL, "bad frame buffer");
    ddm_engine *e = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        int rc = acquire(s, stream_id, &e);
        if (rc) return rc;
    }
    if (s->batched) {
        const int rc = submit_batched(s, stream_id, e, frames, n_frames);
        release_stream(s, stream_id);
        return rc;
    }
    const Clock::time_point t0 = Clock::now();
    int rc = ddm_submit_host(e, frames, n_frames);
    if (rc == DDM_OK) rc = ddm_flush(e);
    hipEvent_t ev = nullptr;
    if (rc == DDM_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        ev = take_event(s);
    } else {
        g_server_err = ddm_last_error(e);
    }
    if (rc == DDM_OK && !ev) rc = serr(DDM_EDEVICE, "event creation failed");
    if (rc == DDM_OK && hipEventRecord(ev, ddm::engine_stream(e)) != hipSuccess)
        rc = serr(DDM_EDEVICE, "event record failed");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->streams.find(stream_id);  // held by `busy`: close waits for it
    if (rc == DDM_OK) {
        it->second.pending.push_back(Pending{ev, t0});
        it->second.frames += n_frames;
        s->frames_total += n_frames;
        s->chunks_total += 1;
    } else if (ev) {
        s->pool.push_back(ev);
    }
    it->second.busy -= 1;
    s->idle.notify_all();
    return rc;
}

extern "C" int ddm_server_results(ddm_server *s, int stream_id, double *results) {
    if (!s || !results) return DDM_EINVAL;
    ddm_engine *e = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        int rc = acquire(s, stream_id, &e);
        if (rc) return rc;
        if (s->batched) {
            auto it = s->streams.find(stream_id);
            it->second.urgent += 1;  // no holding its chunks back now
            s->work.notify_all();
            Stream *st = wait_chunks_launched(s, lk, stream_id);
            if (st) st->urgent -= 1;
        }
    }
    int rc;
    if (s->batched) {
        // the stream's ticks run on the tick stream: the engine's own stream
        // (which computes the averages) waits for everything queued there
        std::lock_guard<std::mutex> g(s->tick_mu);
        hipEvent_t ev = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            ev = take_event(s);
        }
        rc = (ev && hipEventRecord(ev, s->tick_st) == hipSuccess &&
              hipStreamWaitEvent(ddm::engine_stream(e), ev, 0) == hipSuccess)
                 ? ddm_results(e, results)
                 : serr(DDM_EDEVICE, "tick event failed");
        if (ev) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->pool.push_back(ev);
        }
    } else {
        rc = ddm_results(e, results);
    }
    if (rc) g_server_err = ddm_last_error(e);
    release_stream(s, stream_id);
    return rc;
}

extern "C" int ddm_server_engine(ddm_server *s, int stream_id, ddm_engine **eng) {
    if (!s || !eng) return DDM_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->streams.find(stream_id);
    if (it == s->streams.end() || it->second.closing) return serr(DDM_ERANGE, "unknown stream");
    *eng = it->second.eng;
    return DDM_OK;
}

extern "C" int ddm_server_drain(ddm_server *s) {
    if (!s) return DDM_EINVAL;
    if (s->batched) {
        std::unique_lock<std::mutex> lk(s->mu);
        s->draining += 1;  // no holding chunks back now
        s->work.notify_all();
        s->idle.wait(lk, [&] {
            if (any_queued(s)) return false;
            for (auto &kv : s->streams)
                if (kv.second.busy) return false;
            return true;
        });
        s->draining -= 1;
        lk.unlock();
        if (hipStreamSynchronize(s->tick_st) != hipSuccess) return serr(DDM_EDEVICE, "tick stream failed");
        if (!s->tick_err.empty()) return serr(DDM_EDEVICE, s->tick_err);
    }
    std::vector<std::pair<int, ddm_engine *>> engines;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto &kv : s->streams) {
            if (kv.second.closing) continue;
            kv.second.busy += 1;
            engines.emplace_back(kv.first, kv.second.eng);
        }
    }
    int rc = DDM_OK;
    for (auto &pe : engines) {
        if (rc == DDM_OK) {
            rc = ddm_sync(pe.second);
            if (rc) g_server_err = ddm_last_error(pe.second);
        }
        release_stream(s, pe.first);
    }
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &kv : s->streams) harvest(s, kv.second);
    return DDM_OK;
}

extern "C" int ddm_server_close(ddm_server *s, int stream_id) {
    if (!s) return DDM_EINVAL;
    Stream st;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->streams.find(stream_id);
        if (it == s->streams.end() || it->second.closing) return serr(DDM_ERANGE, "unknown stream");
        it->second.closing = true;  // no new user from here on
        it->second.urgent += 1;     // and no holding its chunks back
        s->work.notify_all();
        s->idle.wait(lk, [&] { return it->second.busy == 0 && it->second.queue.empty(); });
        st = std::move(it->second);
        s->streams.erase(it);
    }
    if (s->batched) {
        std::lock_guard<std::mutex> g(s->tick_mu);
        (void)hipStreamSynchronize(s->tick_st);  // its last tick
    }
    ddm_sync(st.eng);
    ddm_destroy(st.eng);
    std::lock_guard<std::mutex> lk(s->mu);
    for (Pending &p : st.pending) s->pool.push_back(p.ev);
    s->used -= st.footprint;
    return DDM_OK;
}

extern "C" int ddm_server_latencies(ddm_server *s, int stream_id, float *ms, int max) {
    if (!s || (!ms && max > 0)) return DDM_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    int n = 0;
    for (auto &kv : s->streams) {
        if (stream_id >= 0 && kv.first != stream_id) continue;
        harvest(s, kv.second);
        for (float v : kv.second.latency_ms) {
            if (n >= max) return n;
            ms[n++] = v;
        }
    }
    return n;
}

extern "C" int ddm_server_stats(ddm_server *s, ddm_server_info *info) {
    if (!s || !info) return DDM_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    info->budget_bytes = s->budget;
    info->used_bytes = s->used;
    info->streams = (int)s->streams.size();
    info->max_streams = s->max_streams;
    info->frames = s->frames_total;
    info->chunks = s->chunks_total;
    long long pending = 0;
    for (auto &kv : s->streams) pending += (long long)kv.second.pending.size() + (long long)kv.second.queue.size();
    info->chunks_in_flight = pending;
    info->ticks = s->ticks;
    info->graph_ticks = s->graph_ticks;
    info->tick_items = s->items;
    return DDM_OK;
}

extern "C" const char *ddm_server_last_error(const ddm_server *) { return g_server_err.c_str(); }
