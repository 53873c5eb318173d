// Serving loop: many DDM streams on one device.
//
// A stream is an engine (its own spectrum ring, accumulators and HIP stream)
// admitted when its device footprint fits in the server's budget.  Chunks
// of different streams are submitted as they arrive; each stream's work is
// queued on its own HIP stream, so chunks of different streams overlap on
// the device while those of one stream stay in order.  Every chunk's lag
// contributions are flushed right away, and the chunk-to-result latency
// (host submission -> chunk fully accumulated on the device) is recorded by
// a poller thread from an event recorded after the chunk.  A stream's
// results depend only on its own frames, so they are the same alone and in
// a mix.  The reference has no counterpart (it analyses one video per
// process, src/DDM.cu:268-896); its memory report (src/DDM.cu:493-634) is
// the seed of the footprint accounting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ddm_internal.h"
#include "ddm_kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
    hipEvent_t ev;
    Clock::time_point t_submit;
};

struct Stream {
    ddm_engine *eng = nullptr;
    size_t footprint = 0;
    std::deque<Pending> pending;
    std::vector<float> latency_ms;
    long long frames = 0;
    int busy = 0;          // calls using the engine outside the lock
    bool closing = false;  // ddm_server_close waits for busy == 0
};

}  // namespace

struct ddm_server {
    size_t budget = 0, used = 0;
    int max_streams = 0;
    std::mutex mu;
    std::condition_variable idle;  // a stream's busy count dropped
    std::map<int, Stream> streams;
    int next_id = 0;
    std::vector<hipEvent_t> pool;
    std::thread poller;
    std::atomic<bool> stop{false};
    long long frames_total = 0, chunks_total = 0;
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

extern "C" int ddm_server_create(size_t device_budget, int max_streams, ddm_server **out) {
    if (!out || max_streams <= 0) return DDM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_server_err = "no HIP device available";
        return DDM_EDEVICE;
    }
    if (device_budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return DDM_EDEVICE;
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
    s->poller = std::thread(poll_loop, s);
    *out = s;
    return DDM_OK;
}

extern "C" void ddm_server_destroy(ddm_server *s) {
    if (!s) return;
    s->stop.store(true);
    if (s->poller.joinable()) s->poller.join();
    for (auto &kv : s->streams) {
        ddm_sync(kv.second.eng);
        for (Pending &p : kv.second.pending) s->pool.push_back(p.ev);
        ddm_destroy(kv.second.eng);
    }
    for (hipEvent_t ev : s->pool) (void)hipEventDestroy(ev);
    delete s;
}

extern "C" int ddm_server_open(ddm_server *s, const ddm_desc *desc, int *stream_id) {
    if (!s || !desc || !stream_id) return DDM_EINVAL;
    ddm::Params p;
    std::string msg;
    int rc = ddm::check_desc(desc, &p, &msg);
    if (rc != DDM_OK) return serr(rc, msg);
    const size_t need = ddm::device_footprint(p);
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
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != DDM_OK) {
        s->used -= need;
        const std::string why = ddm_last_error(nullptr);
        // a refusal like any other over-budget request: the caller may try
        // again when streams have left
        if (rc == DDM_ENOMEM) return serr(DDM_EBUSY, "device memory exhausted below the budget (" + why + ")");
        return serr(rc, why);
    }
    const int id = s->next_id++;
    Stream &st = s->streams[id];
    st.eng = e;
    st.footprint = need;
    *stream_id = id;
    return DDM_OK;
}

extern "C" int ddm_server_submit(ddm_server *s, int stream_id, const void *frames, int n_frames) {
    if (!s) return DDM_EINVAL;
    ddm_engine *e = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        int rc = acquire(s, stream_id, &e);
        if (rc) return rc;
    }
    const Clock::time_point t0 = Clock::now();
    int rc = ddm_submit_host(e, frames, n_frames);
    if (rc == DDM_OK) rc = ddm_flush(e);
    hipEvent_t ev = nullptr;
    if (rc == DDM_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->pool.empty()) {
            ev = s->pool.back();
            s->pool.pop_back();
        }
    } else {
        g_server_err = ddm_last_error(e);
    }
    if (rc == DDM_OK && !ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        rc = serr(DDM_EDEVICE, "event creation failed");
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
        std::lock_guard<std::mutex> lk(s->mu);
        int rc = acquire(s, stream_id, &e);
        if (rc) return rc;
    }
    int rc = ddm_results(e, results);
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
        s->idle.wait(lk, [&] { return it->second.busy == 0; });
        st = std::move(it->second);
        s->streams.erase(it);
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
    for (auto &kv : s->streams) pending += (long long)kv.second.pending.size();
    info->chunks_in_flight = pending;
    return DDM_OK;
}

extern "C" const char *ddm_server_last_error(const ddm_server *) { return g_server_err.c_str(); }
