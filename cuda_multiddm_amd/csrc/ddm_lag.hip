// Lag engine of the DDM pipeline (processFFT / analyseChunk,
// src/DDM_kernel.cuh:85-101, src/DDM.cu:222-262): for a block of n new
// frames t0 .. t0+n-1 held in the spectrum ring and every lag tau,
//   acc[l][e] += sum_{j < n, t0+j-tau >= 0} |F_{t0+j}[e] - F_{t0+j-tau}[e]|^2
// (no 1/s^2 factor: it is a power of two, applied exactly in the average).
//
// Mapping: one workgroup per ring group (RING_GROUP = 64 consecutive
// spectrum elements, one per lane), 16 waves.  The block's new frames (and,
// for 256-frame blocks, the frames before it that the register window
// reaches) are staged once in LDS; the rest of the group's ring history is
// read with buffer loads.
// Three ways to form the pairs (new frame j, frame j - tau):
//  * the leading run of consecutive lags (1, 2, 3, ... in the usual lag
//    lists): a register window, one new frame and one entering partner per
//    step serving DENSE pairs;
//  * up to HIST_LAGS other lags, history-major: a block of consecutive
//    history frames in registers is paired with the LDS rows of every lag
//    that reaches it, so each history frame is read once per wave;
//  * more lags than that: each lag over ranges of j, whose partners are at
//    most two contiguous runs of ring slots (one wrap) or LDS rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ddm_batch.h"
#include "ddm_internal.h"
#include "ddm_kernels.h"

namespace ddm {

constexpr int DENSE_SEG = 32;                   // frames per segment of a lag's units
// Word 3 of a raw (untyped, stride 0) buffer resource on gfx9-family parts:
// a valid 32-bit data format, so dword loads return memory.
constexpr int BUFFER_RSRC_WORD3 = 0x00020000;
constexpr int DENSE_BATCH = 2;                  // register-window steps scheduled together
constexpr int ACC_PREFETCH = 4;                 // accumulators each wave requests ahead of its sums

// Per block size: J = 256 stages 128 KB of new frames, so one workgroup (4
// waves per SIMD) runs per CU and each wave has room for a wider window,
// more partial sums and longer load batches; smaller blocks run two
// workgroups per CU within 64 registers.
template <int J>
struct LagCfg {
    static constexpr bool WIDE = J >= 256;
    static constexpr int DENSE = WIDE ? 16 : 8;      // lags per register window
    static constexpr int SLOTS_MAX = WIDE ? 16 : 8;  // partial sums per wave
    static constexpr int UNROLL = WIDE ? 16 : 8;     // pairs per batch of loads
    static constexpr int WAVES_PER_EU = WIDE ? 4 : 8;
    // history-major path for the other lags: at most HIST_LAGS of them, one
    // register sum each, over blocks of HIST_BLOCK history frames
    static constexpr int HIST_LAGS = WIDE ? 32 : 0;
    static constexpr int HIST_BLOCK = 16;
    // frames before the block staged in LDS beside it (16 KB more: the
    // register window of lags up to PRE then finds every partner in LDS);
    // none for smaller blocks, whose two workgroups per CU fill the LDS
    static constexpr int PRE = WIDE ? 32 : 0;
};

// Workgroup barrier ordering LDS only.  Waves never exchange data through
// global memory here, so loads still in flight (the accumulator reads issued
// before the partial sums are exchanged) need not complete before it, as they
// would under __syncthreads().
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Relative cost of a pair in the register window and in the history-major
// path, for the split of the waves.
constexpr int COST_DENSE = 4, COST_HIST = 8;

// Host estimate of a history block's cost in VALU-instruction units: a block
// costs its load (a full memory latency, the same however few lags reach
// it), plus each lag that reaches it and each pair.
constexpr int PLAN_COST_BLOCK = 200, PLAN_COST_VISIT = 10, PLAN_COST_PAIR = 4;
// a window wave's own work in the same units: per step its window's pairs at
// COST_DENSE / COST_HIST of a history pair's cost, plus the step's loads
constexpr int PLAN_COST_STEP = 8;
constexpr int PLAN_DENSE_PCT = 70;  // scale of a window wave's starting load (measured best)
// A block's cost as charged to waves 0-3, 4-7, 8-11 and 12-15: the SIMD
// issues the oldest ready wave first, so the later waves of a SIMD progress
// more slowly and are given less (measured best on a 64-lag list)
constexpr int PLAN_WAVE_PCT[4] = {100, 100, 106, 116};

// The R ring slots of one ring group, read with buffer loads: the lane's
// offset in the vector operand, the wave-uniform slot in the scalar one.  The
// buffer's record count bounds every read to the group's own slots, and LDS
// and ring reads stay distinct instructions (no generic-address loads).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
// The LLVM intrinsic itself: this compiler's raw_buffer_load_b64 builtin
// lowers to a single-dword load.
__device__ f32x2_t raw_buffer_load_f32x2(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v2f32");
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ f32x4_t raw_buffer_load_f32x4(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v4f32");

struct GroupRing {
    __amdgpu_buffer_rsrc_t rs;
    int voff;  // lane * sizeof(float2)
    int hs;    // frames t0 - hs .. t0 - 1 are staged in LDS too (rows -hs .. -1)
    static constexpr int SLOT_BYTES = RING_GROUP * sizeof(float2);
    // slot + k, k a compile-time constant: k % 8 goes into the instruction's
    // 12-bit offset, the rest into the scalar offset, so runs of loads share
    // one or two scalar registers
    __device__ __forceinline__ float2 load(int slot, int k) const {
        const f32x2_t v = raw_buffer_load_f32x2(rs, voff + (k & 7) * SLOT_BYTES, (slot + (k & ~7)) * SLOT_BYTES, 0);
        return make_float2(v.x, v.y);
    }
    __device__ __forceinline__ float2 operator[](int slot) const { return load(slot, 0); }
};

// Row k of a run of consecutive frames: ring slots from `slot` on, or LDS
// rows RING_GROUP complex values apart.
struct RingRun {
    GroupRing g;
    int slot;
    __device__ __forceinline__ float2 operator[](int k) const { return g.load(slot, k); }
    __device__ __forceinline__ RingRun shift(int k) const { return RingRun{g, slot + k}; }
};
struct LdsRun {
    const float2 *p;
    __device__ __forceinline__ float2 operator[](int k) const { return p[k * RING_GROUP]; }
    __device__ __forceinline__ LdsRun shift(int k) const { return LdsRun{p + k * RING_GROUP}; }
};

// sum_{k < count} |a[k] - b[k]|^2 over two runs of frames.  Loads come
// in batches of U pairs; each pair costs two subtractions and two fused
// multiply-adds into one of four independent partial sums.
template <int U, class A>
__device__ __forceinline__ float run_sum(A a, LdsRun b, int count) {
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    int k = 0;
    for (; k + U <= count; k += U) {
        const A ak = a.shift(k);
        const LdsRun bk = b.shift(k);
        float2 va[U], vb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            va[u] = ak[u];
            vb[u] = bk[u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float dx = va[u].x - vb[u].x, dy = va[u].y - vb[u].y;
            s[u & 3] = fmaf(dx, dx, s[u & 3]);
            s[u & 3] = fmaf(dy, dy, s[u & 3]);
        }
    }
    for (; k < count; ++k) {
        const float2 va = a.shift(k)[0], vb = b.shift(k)[0];
        const float dx = va.x - vb.x, dy = va.y - vb.y;
        s[0] = fmaf(dx, dx, s[0]);
        s[0] = fmaf(dy, dy, s[0]);
    }
    return (s[0] + s[1]) + (s[2] + s[3]);
}

// Sum over j in [ja, jb) of the valid pairs (t0 + j, t0 + j - tau): the
// partner is valid when j - tau >= -dmax; it is in the ring history for
// j < tau and in LDS for j >= tau.
template <int U>
__device__ __forceinline__ float lag_range_sum(const GroupRing &g, const float2 *nl, int R, int slot0, int dmax,
                                               int tau, int ja, int jb) {
    float sum = 0.f;
    const int jh0 = max(tau - dmax, ja), jh1 = min(tau, jb);
    if (jh0 < jh1) {
        int jw = tau - slot0;  // j < jw: the partner slot wrapped
        jw = jw < jh0 ? jh0 : (jw > jh1 ? jh1 : jw);
        if (jh0 < jw)
            sum += run_sum<U>(RingRun{g, slot0 + jh0 - tau + R}, LdsRun{nl + jh0 * RING_GROUP}, jw - jh0);
        if (jw < jh1)
            sum += run_sum<U>(RingRun{g, slot0 + jw - tau}, LdsRun{nl + jw * RING_GROUP}, jh1 - jw);
    }
    const int ji = max(tau, ja);
    if (ji < jb) sum += run_sum<U>(LdsRun{nl + (ji - tau) * RING_GROUP}, LdsRun{nl + ji * RING_GROUP}, jb - ji);
    return sum;
}

// Frame t0 + r of the block's window: LDS for r >= -hs, ring history below.
// Frames of valid pairs are at most max lag <= R - n back, one wrap away;
// older ones (read only for discarded lags) are clamped to slot 0 so the
// load stays inside the ring.
__device__ __forceinline__ float2 window_frame(const GroupRing &g, const float2 *nl, int R, int slot0, int r) {
    float2 v;
    if (r >= -g.hs) {
        v = nl[r * RING_GROUP];
    } else {
        int s = slot0 + r;
        s += (s < 0) ? R : 0;
        s = s < 0 ? 0 : s;
        v = g[s];
    }
    return v;
}

// D steps of the register window: step u loads the entering partner pb[u]
// and the new frame at LDS row u of pa, and forms K pairs (the first K lags
// of the window).
template <int D, int K, int S, class A>
__device__ __forceinline__ void dense_block(A pb, const float2 *pa, float2 (&win)[D], float (&s)[S]) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
        win[u] = pb[u];
        const float2 a = pa[u * RING_GROUP];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const float2 b = win[(u - i) & (D - 1)];
            const float dx = a.x - b.x, dy = a.y - b.y;
            s[i] = fmaf(dx, dx, s[i]);
            s[i] = fmaf(dy, dy, s[i]);
        }
        // loads of at most DENSE_BATCH steps ahead: keeps the window within
        // the register budget
        if ((u + 1) % DENSE_BATCH == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// K <= D consecutive lags tau_a .. tau_a + K - 1 over new frames j in
// [j0, j1).  The partners of frame j are the frames j - tau_a - i, a sliding
// window of D registers: each step loads one new frame and one entering
// partner and forms K pairs.  Window slot (u - i) & (D - 1) holds the
// partner of step u for lag tau_a + i, so every register index is a
// compile-time constant.  Sums of lags beyond the caller's count (K rounds
// it up) come from in-bounds (possibly stale) slots and are discarded by the
// caller.
template <int D, int K, int S>
__device__ __forceinline__ void dense_sums(const GroupRing &g, const float2 *nl, int R, int slot0, int tau_a, int j0,
                                           int j1, float (&s)[S]) {
    static_assert(D <= S && (D & (D - 1)) == 0 && K <= D, "window must be a power of two within the partial sums");
    float2 win[D];
    // slots read before a step writes them: the partners of lags 1 .. K - 1
    // of the first steps
#pragma unroll
    for (int k = D - K + 1; k < D; ++k) win[k] = window_frame(g, nl, R, slot0, j0 - tau_a - (D - k));
#pragma unroll
    for (int i = 0; i < D; ++i) s[i] = 0.f;
    for (int jb = j0; jb < j1; jb += D) {
        const int cnt = j1 - jb, r0 = jb - tau_a;  // r0: entering partner of step 0
        // A full block of D steps whose entering partners all sit in LDS (the
        // staged frames before the block included), or
        // all in the ring history without a wrap, reads them from one base
        // with constant offsets; other blocks go frame by frame.
        if (cnt >= D && r0 >= -g.hs) {
            dense_block<D, K>(LdsRun{nl + r0 * RING_GROUP}, nl + jb * RING_GROUP, win, s);
            continue;
        }
        if (cnt >= D && r0 + D <= 0 && slot0 + r0 >= 0) {
            dense_block<D, K>(RingRun{g, slot0 + r0}, nl + jb * RING_GROUP, win, s);
            continue;
        }
#pragma unroll
        for (int u = 0; u < D; ++u) {
            if (u < cnt) {
                win[u] = window_frame(g, nl, R, slot0, jb + u - tau_a);
                const float2 a = nl[(jb + u) * RING_GROUP];
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const float2 b = win[(u - i) & (D - 1)];
                    const float dx = a.x - b.x, dy = a.y - b.y;
                    s[i] = fmaf(dx, dx, s[i]);
                    s[i] = fmaf(dy, dy, s[i]);
                }
            }
        }
    }
}

// The window for kc live lags: K = D / 2, 3D / 4, 7D / 8 or D, the smallest
// that covers kc (a wider window computes lags that are thrown away).
template <int D>
__host__ __device__ inline int window_width(int kc) {
    if (D < 16) return D;
    return kc <= D / 2 ? D / 2 : (kc <= 3 * D / 4 ? 3 * D / 4 : (kc <= 7 * D / 8 ? 7 * D / 8 : D));
}
template <int D, int S>
__device__ __forceinline__ void dense_sums_live(int kc, const GroupRing &g, const float2 *nl, int R, int slot0,
                                                int tau_a, int j0, int j1, float (&s)[S]) {
    if constexpr (D >= 16) {
        if (kc <= D / 2) {
            dense_sums<D, D / 2>(g, nl, R, slot0, tau_a, j0, j1, s);
            return;
        }
        if (kc <= 3 * D / 4) {
            dense_sums<D, 3 * D / 4>(g, nl, R, slot0, tau_a, j0, j1, s);
            return;
        }
        if (kc <= 7 * D / 8) {
            dense_sums<D, 7 * D / 8>(g, nl, R, slot0, tau_a, j0, j1, s);
            return;
        }
    }
    dense_sums<D, D>(g, nl, R, slot0, tau_a, j0, j1, s);
}

// History frames p0 .. p0 + B - 1 of the window, clamped to pmax (frames
// beyond it pair with no new frame).
template <int B>
__device__ __forceinline__ void hist_load(const GroupRing &g, const float2 *nl, int R, int slot0, int p0, int pmax,
                                          float2 (&fp)[B]) {
    const int pe = p0 + B - 1;  // last frame of the block
    if (pe <= pmax) {
        // whole block in LDS, or in the history as one run of slots
        int base = slot0 + p0;
        base += (base < 0) ? R : 0;
        if (p0 >= -g.hs) {
#pragma unroll
            for (int u = 0; u < B; ++u) fp[u] = nl[(p0 + u) * RING_GROUP];
            return;
        }
        if (pe < 0 && base + B <= R) {
#pragma unroll
            for (int u = 0; u < B; ++u) fp[u] = g.load(base, u);
            return;
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) fp[u] = window_frame(g, nl, R, slot0, min(p0 + u, pmax));
}

// Pairs of history frames p0 + u (in fp) with the new frames p0 + u + tau_l
// inside [0, n), for every lag that reaches the block: -p0 - B < tau < n - p0.
template <int B, int LSM>
__device__ __forceinline__ void hist_block(const float2 *nl, const float2 (&fp)[B], int p0, int n, int vtau, bool live,
                                           float (&s)[LSM]) {
    const int la = __builtin_popcountll(__ballot(live && vtau <= -p0 - B));
    const int lb = __builtin_popcountll(__ballot(live && vtau < n - p0));
    for (int l = la; l < lb; ++l) {
        const int j0 = p0 + __builtin_amdgcn_readlane(vtau, l);
        const float2 *a = nl + j0 * RING_GROUP;
        // two partial sums: chains of 16 FMAs, the fewest adds to fold them
        float t[2] = {0.f, 0.f};
        if (j0 >= 0 && j0 + B <= n) {
            // all B rows requested before the first pair: the reads overlap
            // (the compiler otherwise keeps a few in flight to save registers)
            float2 va[B];
#pragma unroll
            for (int u = 0; u < B; ++u) va[u] = a[u * RING_GROUP];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float2 v = va[u];
                const float dx = v.x - fp[u].x, dy = v.y - fp[u].y;
                t[u & 1] = fmaf(dx, dx, t[u & 1]);
                t[u & 1] = fmaf(dy, dy, t[u & 1]);
            }
        } else {
            const int ua = -j0, ub = n - j0;  // valid u: ua <= u < ub
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (u >= ua && u < ub) {
                    const float2 v = a[u * RING_GROUP];
                    const float dx = v.x - fp[u].x, dy = v.y - fp[u].y;
                    t[u & 1] = fmaf(dx, dx, t[u & 1]);
                    t[u & 1] = fmaf(dy, dy, t[u & 1]);
                }
            }
        }
        s[l] += t[0] + t[1];
    }
}

// Lags tau_0 < ... < tau_{ls-1} (lane l of `vtau` holds tau_l) by history
// frame: frame p of the window (ring history for p < 0, LDS for p >= 0)
// pairs with the new frames j = p + tau_l inside [0, n).  The frames
// [pmin, pmax] are cut into blocks of B; list k of the launch plan (k the
// wave's index among the history waves, or among all waves when the plan
// gives window waves blocks too) names the wave's blocks, cost-balanced on
// the host; without lists history wave k takes blocks k, k + ws, ...  Each block is loaded once into registers
// and paired with the LDS rows of every lag that reaches it, so a history
// frame is requested once per wave rather than once per lag.  s[l] collects
// lag l in a fixed block order.
template <int B, int LSM, class PL>
__device__ __forceinline__ void hist_sums(const GroupRing &g, const float2 *nl, int R, int slot0, int n, int vtau,
                                          int ls, const PL &pl, int k, int ws, float (&s)[LSM]) {
    const bool live = (int)(threadIdx.x % RING_GROUP) < ls;
    float2 fa[B];
    if (pl.listed) {
        const int i1 = pl.blk_start[k + 1];
        for (int i = pl.blk_start[k]; i < i1; ++i) {
            const int p0 = pl.pmin + (int)pl.blk_list[i] * B;
            hist_load<B>(g, nl, R, slot0, p0, pl.pmax, fa);
            hist_block<B>(nl, fa, p0, n, vtau, live, s);
        }
        return;
    }
    for (int p0 = pl.pmin + k * B; p0 <= pl.pmax; p0 += ws * B) {
        hist_load<B>(g, nl, R, slot0, p0, pl.pmax, fa);
        hist_block<B>(nl, fa, p0, n, vtau, live, s);
    }
}

// A lane's accumulators: acc[l][e] and, for compensated accumulation,
// comp[l][e], the rounding error of every addition into acc so far
// (acc + comp is the sum carried to twice the precision of FP32).  The
// pointers are the ring group's (wave-uniform, so they stay in scalar
// registers and each access is a scalar base plus the lane's 32-bit offset).
template <bool COMP>
struct AccRef {
    float *a;
    float *c;  // unused unless COMP
    int lane;
    __device__ __forceinline__ float *at(int l, long long E) const { return a + l * E + lane; }
    __device__ __forceinline__ float *cat(int l, long long E) const { return c + l * E + lane; }
};

// The accumulators of the first PF lags wave w adds (lags w, w + 16, ...),
// all requested when the wave's pair work is done, so that they arrive while
// the partial sums are exchanged in LDS (requested when the workgroup starts,
// they held registers through the pair work and the compiler spilled them,
// waiting for each load first).
template <int PF, bool COMP>
struct AccPre {
    float a[PF];
    float c[COMP ? PF : 1];
};

template <int PF, bool COMP>
__device__ __forceinline__ void load_acc(AccRef<COMP> ar, long long E, int w, int nred, AccPre<PF, COMP> &pre) {
#pragma unroll
    for (int m = 0; m < PF; ++m) {
        const int l = w + m * LAG_WAVES;
        pre.a[m] = l < nred ? *ar.at(l, E) : 0.f;
    }
}

// The compensation terms of the same lags, requested when the pair work is
// done (holding them through it would take registers the pair loops need).
template <int PF, bool COMP>
__device__ __forceinline__ void load_comp(AccRef<COMP> ar, long long E, int w, int nred, AccPre<PF, COMP> &pre) {
    if constexpr (COMP) {
#pragma unroll
        for (int m = 0; m < PF; ++m) {
            const int l = w + m * LAG_WAVES;
            pre.c[m] = l < nred ? *ar.cat(l, E) : 0.f;
        }
    }
}

// acc[l] += t for the m-th lag this wave adds (l = w + 16 m), from the read
// issued ahead when there is one (m < 0: none).  Compensated: the sum is rounded to
// FP32 and its exact rounding error (TwoSum, no multiplications, so no
// contraction can change it) is added to comp.
template <int PF, bool COMP>
__device__ __forceinline__ void add_acc(AccRef<COMP> ar, long long E, int l, int m, float t,
                                        const AccPre<PF, COMP> &pre) {
    float a = 0.f, c = 0.f;
    bool have = false;
#pragma unroll
    for (int i = 0; i < PF; ++i)
        if (i == m) {
            a = pre.a[i];
            if constexpr (COMP) c = pre.c[i];
            have = true;
        }
    float *p = ar.at(l, E);
    if (!have) a = *p;
    if constexpr (COMP) {
        float *q = ar.cat(l, E);
        if (!have) c = *q;
        const float s = a + t;
        const float bp = s - a;
        const float err = (a - (s - bp)) + (t - bp);
        *p = s;
        *q = c + err;
    } else {
        *p = a + t;
    }
}

// Partial sums of a launch with history-major lags meet in the staging buffer
// (the new frames are no longer read) in rounds of RL of those lags: in round
// r, history wave k writes its sums of lags RL*r + i to rows wd*SL + RL*k + i,
// and in round 0 dense wave w also writes its window sums to rows w*SL + i.
// Each lag is then added up by one wave in a fixed order (the same for any
// RL, so the results do not depend on it).
template <int LSM, int SL, int DN, int RL, int SM, int PF, bool COMP>
__device__ __forceinline__ void hist_reduce(float2 *nb, AccRef<COMP> accg, long long E, const float (&s)[SM],
                                            const float (&h)[LSM], int w, int wd, int hw0, int nhw, int sd,
                                            int dense, int ls, AccPre<PF, COMP> &pre) {
    static_assert(SL <= SM && LSM % RL == 0, "window sums exceed the partial registers");
    load_acc(accg, E, w, dense + ls, pre);
    load_comp(accg, E, w, dense + ls, pre);
    float *red = reinterpret_cast<float *>(nb);
    const int lane = threadIdx.x % RING_GROUP;
    const int hb = wd * SL;
#pragma unroll
    for (int r = 0; r < LSM / RL; ++r) {
        if (r * RL >= ls) break;
        lds_barrier();
        if (w < wd && r == 0) {
#pragma unroll
            for (int i = 0; i < SL; ++i) red[(w * SL + i) * RING_GROUP + lane] = s[i];
        }
        if (w >= hw0) {
            const int k = w - hw0;
#pragma unroll
            for (int i = 0; i < RL; ++i)
                if (r * RL + i < ls) red[(hb + k * RL + i) * RING_GROUP + lane] = h[r * RL + i];
        }
        lds_barrier();
        const int nd = r == 0 ? dense : 0, nh = min(RL, ls - r * RL);
        // wave w adds the round's lags l = w (mod 16): those are the lags
        // whose accumulators it requested ahead (w, w + 16, ...)
        const int lo = r == 0 ? 0 : dense + r * RL, hi = dense + r * RL + nh;
        for (int l = lo + (w - lo % LAG_WAVES + LAG_WAVES) % LAG_WAVES; l < hi; l += LAG_WAVES) {
            float t = 0.f;
            if (l < nd) {
                const int c = l / DN, i = l % DN;
                for (int seg = 0; seg < sd; ++seg) t += red[((c * sd + seg) * SL + i) * RING_GROUP + lane];
            } else {
                const int i = l - dense - r * RL;
                for (int k = 0; k < nhw; ++k) t += red[(hb + k * RL + i) * RING_GROUP + lane];
            }
            add_acc(accg, E, l, (l - w) / LAG_WAVES, t, pre);
        }
    }
}

// E: ring / accumulator row stride (padded), R: ring slots, slot0: slot of
// t0, dmax = min(max lag, t0): oldest history frame the block may pair with.
//
// Work split over the 16 waves (identical in every workgroup, from the
// host-made plan `pl`, see make_plan):
//  * dense part: the leading run of consecutive lags in chunks of DENSE
//    lags; each chunk is cut into sd segments of new frames, one wave per
//    (chunk, segment);
//  * the other ls lags on the remaining ws waves, history-major (blocks of
//    history frames from the plan's per-wave lists, which may give window
//    waves blocks too, after their window), or each lag cut into ss
//    segments ("units"; unit u goes to wave u % ws as its partial slot
//    u / ws), or, with more lags than partial slots, each lag owned by one
//    wave that adds its sum to the accumulator directly.
// Partial sums meet in LDS after the pair work and are added per lag in a
// fixed order, so results do not depend on timing.
// TAG makes each kernel's instantiation, and so its LDS stage, its own (an
// LDS array shared by two kernels is addressed less directly in both).
template <int J, bool COMP, int TAG, class PL>
__device__ __forceinline__ void lag_body(const float2 *__restrict__ ring, float *__restrict__ acc,
                                         float *__restrict__ comp, long long E, int R, int slot0, int n, int dmax,
                                         const int *__restrict__ lags, int L, const PL &pl, int group) {
    using C = LagCfg<J>;
    constexpr int DENSE = C::DENSE, U = C::UNROLL;
    constexpr int SLOTS = (J / 8) < C::SLOTS_MAX ? (J / 8) : C::SLOTS_MAX;  // partials per wave that fit in LDS
    static_assert(LAG_WAVES * SLOTS * RING_GROUP <= J * RING_GROUP * 2, "partials exceed the staging buffer");
    constexpr int PRE = C::PRE;
    __shared__ __attribute__((aligned(16))) float2 nb[(PRE + J) * RING_GROUP];
    const int tid = threadIdx.x;
    const int lane = tid % RING_GROUP;
    const int w = __builtin_amdgcn_readfirstlane(tid / RING_GROUP);
    GroupRing g;
    g.rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(ring + (long long)group * R * RING_GROUP), 0,
                                             R * (int)(RING_GROUP * sizeof(float2)), BUFFER_RSRC_WORD3);
    g.voff = lane * (int)sizeof(float2);
    g.hs = min(min(PRE, pl.pre), dmax);
    const int srows = g.hs + n, sfirst = slot0 - g.hs;  // staged rows -hs .. n - 1 and their first slot
    if (sfirst >= 0 && sfirst + srows <= R) {
        // the rows are one run of ring slots (no wrap): copied as 16-byte
        // pieces, all loads in flight before the LDS stores
        constexpr int PIECES_MAX = (PRE + J) * RING_GROUP / 2;
        constexpr int KB = (PIECES_MAX + LAG_NT - 1) / LAG_NT;
        const int pieces = srows * (RING_GROUP / 2);
        f32x4_t v[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int q = tid + k * LAG_NT;
            if (q < pieces) v[k] = raw_buffer_load_f32x4(g.rs, q * 16, sfirst * GroupRing::SLOT_BYTES, 0);
        }
        f32x4_t *dst = reinterpret_cast<f32x4_t *>(nb + (PRE - g.hs) * RING_GROUP);
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int q = tid + k * LAG_NT;
            if (q < pieces) dst[q] = v[k];
        }
    } else {
        // stage the new frames (rows 0 .. n - 1) and the hs frames before
        // them (rows -hs .. -1): all loads in flight before the LDS stores
        constexpr int K = (J + LAG_WAVES - 1) / LAG_WAVES;
        constexpr int KP = PRE > 0 ? (PRE + LAG_WAVES - 1) / LAG_WAVES : 1;
        float2 v[K], vp[KP];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = w + k * LAG_WAVES;
            if (j < n) {
                int s = slot0 + j;
                s -= (s >= R) ? R : 0;
                v[k] = g[s];
            }
        }
        if constexpr (PRE > 0) {
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int r = -1 - (w + k * LAG_WAVES);
                if (r >= -g.hs) {
                    int s = slot0 + r;
                    s += (s < 0) ? R : 0;
                    vp[k] = g[s];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = w + k * LAG_WAVES;
            if (j < n) nb[(PRE + j) * RING_GROUP + lane] = v[k];
        }
        if constexpr (PRE > 0) {
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int r = -1 - (w + k * LAG_WAVES);
                if (r >= -g.hs) nb[(PRE + r) * RING_GROUP + lane] = vp[k];
            }
        }
    }
    // ---- plan (host-computed, wave-uniform) ----
    const int dense_lags = pl.dense_lags;
    const int ls = L - dense_lags;
    const bool hist = C::HIST_LAGS && pl.hist;
    const int wd = pl.wd, sd = pl.sd, ss = pl.ss;
    const int ws = LAG_WAVES - wd;
    // the accumulators this wave adds to are all requested when its pair work
    // is done, before the partial sums are exchanged (held through the pair
    // work they would take registers its loops need)
    const long long group_off = (long long)group * RING_GROUP;
    AccRef<COMP> accg{acc + group_off, COMP ? comp + group_off : nullptr, lane};
    const int nred = hist ? L : (ss ? L : dense_lags);
    AccPre<ACC_PREFETCH, COMP> pre;
    lds_barrier();
    const float2 *nl = nb + PRE * RING_GROUP + lane;
    float s[C::SLOTS_MAX];
#pragma unroll
    for (int i = 0; i < C::SLOTS_MAX; ++i) s[i] = 0.f;
    constexpr int HL = C::HIST_LAGS > 0 ? C::HIST_LAGS : 1;
    float h[HL];  // history-major lag sums
#pragma unroll
    for (int i = 0; i < HL; ++i) h[i] = 0.f;
    if (w < wd) {
        const int c = w / sd, seg = w % sd;
        const int la = c * DENSE, kc = min(DENSE, dense_lags - la);
        const int tau_a = lags[la];
        const int len = (n + sd - 1) / sd;
        const int j0 = seg * len, j1 = min(j0 + len, n);
        if (j0 < j1) {
            // from jv on, every lag of the chunk has all its partners; near the
            // start of the stream the frames before jv pair only where a
            // partner exists, lag by lag
            const int jv = min(j1, max(j0, tau_a + kc - 1 - dmax));
            if (jv < j1) dense_sums_live<DENSE>(kc, g, nl, R, slot0, tau_a, jv, j1, s);
            if (j0 < jv) {
#pragma unroll
                for (int i = 0; i < DENSE; ++i)
                    if (i < kc) s[i] += lag_range_sum<U>(g, nl, R, slot0, dmax, tau_a + i, j0, jv);
            }
        }
    } else if (hist) {
        // history-major work below
    } else if (ss) {
        const int k = w - wd, units = ls * ss, len = (n + ss - 1) / ss;
#pragma unroll 1
        for (int m = 0; m < SLOTS; ++m) {
            const int u = k + m * ws;
            if (u < units) {
                const int l = dense_lags + u / ss, seg = u % ss;
                const int j0 = seg * len, j1 = min(j0 + len, n);
                if (j0 < j1) s[m] = lag_range_sum<U>(g, nl, R, slot0, dmax, lags[l], j0, j1);
            }
        }
    } else {
        for (int l = dense_lags + (w - wd); l < L; l += ws)
            add_acc(accg, E, l, -1, lag_range_sum<U>(g, nl, R, slot0, dmax, lags[l], 0, n), pre);
    }
    if constexpr (C::HIST_LAGS > 0) {
        if (hist) {
            // history waves, and window waves once their window is done when
            // the plan lists blocks for them too
            if (w >= wd || pl.all_hist) {
                const int vtau = lags[dense_lags + (lane < ls ? lane : 0)];
                hist_sums<C::HIST_BLOCK>(g, nl, R, slot0, n, vtau, ls, pl, pl.all_hist ? w : w - wd, ws, h);
            }
            // every wave's sums of all its lags in one round when their rows
            // fit in the staging buffer, else in rounds of 16 lags
            if (pl.all_hist && wd * SLOTS + LAG_WAVES * C::HIST_LAGS <= (PRE + J) * 2)
                hist_reduce<C::HIST_LAGS, SLOTS, DENSE, C::HIST_LAGS>(nb, accg, E, s, h, w, wd, 0, LAG_WAVES, sd,
                                                                      dense_lags, ls, pre);
            else if (pl.all_hist)
                hist_reduce<C::HIST_LAGS, SLOTS, DENSE, 16>(nb, accg, E, s, h, w, wd, 0, LAG_WAVES, sd, dense_lags,
                                                            ls, pre);
            else
                hist_reduce<C::HIST_LAGS, SLOTS, DENSE, 32>(nb, accg, E, s, h, w, wd, wd, ws, sd, dense_lags, ls,
                                                            pre);
            return;
        }
    }
    if (wd == 0 && ss == 0) return;
    // ---- combine the partials: [wave][slot][lane] in the staging buffer ----
    load_acc(accg, E, w, nred, pre);
    load_comp(accg, E, w, nred, pre);
    lds_barrier();
    float *red = reinterpret_cast<float *>(nb);
#pragma unroll
    for (int i = 0; i < SLOTS; ++i) red[(w * SLOTS + i) * RING_GROUP + lane] = s[i];
    lds_barrier();
    // one lag per wave and step, so every index below is wave-uniform
    for (int l = w; l < nred; l += LAG_WAVES) {
        float t = 0.f;
        if (l < dense_lags) {
            const int c = l / DENSE, i = l % DENSE;
            for (int seg = 0; seg < sd; ++seg) t += red[((c * sd + seg) * SLOTS + i) * RING_GROUP + lane];
        } else {
            const int u0 = (l - dense_lags) * ss;
            int q = u0 / ws, r = u0 - q * ws;  // unit u -> wave wd + u % ws, slot u / ws
            for (int seg = 0; seg < ss; ++seg) {
                t += red[((wd + r) * SLOTS + q) * RING_GROUP + lane];
                if (++r == ws) {
                    r = 0;
                    ++q;
                }
            }
        }
        add_acc(accg, E, l, (l - w) / LAG_WAVES, t, pre);
    }
}

template <int J, bool COMP>
__global__ __launch_bounds__(LAG_NT) __attribute__((amdgpu_waves_per_eu(LagCfg<J>::WAVES_PER_EU))) void lag_kernel(
    const float2 *__restrict__ ring, float *__restrict__ acc, float *__restrict__ comp, long long E, int R,
    int slot0, int n, int dmax, const int *__restrict__ lags, int L, const LagPlan pl) {
    lag_body<J, COMP, 0>(ring, acc, comp, E, R, slot0, n, dmax, lags, L, pl, blockIdx.x);
}

// One pass over many streams: a workgroup per (stream, ring group) item of
// the batch (job = stream, group within its spectrum).
template <int J, bool COMP>
__global__ __launch_bounds__(LAG_NT) __attribute__((amdgpu_waves_per_eu(LagCfg<J>::WAVES_PER_EU))) void
lag_kernel_batch(const BatchHdr *__restrict__ hdr, const LagJob *__restrict__ jobs,
                 const LagPlan *__restrict__ plans) {
    // the tables are read-only for the whole launch: read them through the
    // constant address space, so that their loads stay scalar beside the
    // accumulator stores (which the compiler cannot otherwise prove apart)
    typedef const __attribute__((address_space(4))) LagJob CJob;
    typedef const __attribute__((address_space(4))) LagPlan CPlan;
    CJob *cj = (CJob *)jobs;
    const long long b = blockIdx.x;
    CJob &jb = cj[batch_job(cj, hdr->njobs, b)];
    CPlan &pl = ((CPlan *)plans)[jb.plan];
    lag_body<J, COMP, 1>(jb.ring, jb.acc, jb.comp, jb.E, jb.R, jb.slot0, jb.n, jb.dmax, jb.lags, jb.L, pl,
                         (int)(b - jb.first));
}

// The launch plan (same rules for every workgroup):
//  * the leading run of consecutive lags (lags[i] = lags[0] + i) of DENSE/2
//    to 8 * DENSE lags (at most 64) goes on the register window, cut back to
//    whole windows when the last one would be short and history-major lags
//    take the rest;
//  * waves are split between window chunks and the other lags by relative
//    pair cost; the other lags go history-major when there are at most
//    HIST_LAGS of them, else each is cut into ss units (or owned by one wave
//    when even one unit per lag exceeds the partial slots);
//  * history-major blocks are handed out greedily, last block first, each to
//    the wave with the least estimated work, so the slowest wave (which ends
//    the pair phase) carries little more than the mean; with window waves
//    every wave takes blocks, a window wave starting from its window's
//    estimated work, so it turns to history blocks when its window is done
//    instead of idling while the history waves finish.
template <int J>
static void make_plan(const int *lags, int L, int n, int dmax, LagPlan &pl) {
    using C = LagCfg<J>;
    constexpr int DENSE = C::DENSE, B = C::HIST_BLOCK;
    constexpr int NSEG = J / DENSE_SEG;
    constexpr int SLOTS = (J / 8) < C::SLOTS_MAX ? (J / 8) : C::SLOTS_MAX;
    constexpr int MAX_DENSE = (8 * DENSE) < RING_GROUP ? (8 * DENSE) : RING_GROUP;
    pl = LagPlan{};
    int run = 1;
    while (run < L && run < RING_GROUP && lags[run] == lags[0] + run) ++run;
    int dense_lags = (SLOTS >= DENSE && run >= DENSE / 2) ? std::min(run, MAX_DENSE) : 0;
    if (C::HIST_LAGS && dense_lags > DENSE && dense_lags % DENSE < DENSE / 2) dense_lags -= dense_lags % DENSE;
    const int nchunks = (dense_lags + DENSE - 1) / DENSE;
    const int ls = L - dense_lags;
    const bool hist = C::HIST_LAGS && ls > 0 && ls <= C::HIST_LAGS;
    int wd = 0;
    if (nchunks) {
        // the lags' shares of the pair work in units of full lags: lag tau
        // pairs the frames j >= tau - dmax of the block (near the start of the
        // stream fewer than n, and a lag beyond n + dmax none)
        auto share = [&](int l) {
            const int first = std::max(0, lags[l] - dmax);
            return first >= n ? 0.0 : (double)(n - first) / n;
        };
        double dense_share = 0.0, other_share = 0.0;
        for (int l = 0; l < dense_lags; ++l) dense_share += share(l);
        for (int l = dense_lags; l < L; ++l) other_share += share(l);
        // split by these shares only when fewer than half of the other lags'
        // pairs exist (the first blocks of a stream: 4 -> 8 window waves,
        // 13 % faster on the 64-lag 2048^2 workload); otherwise by whole lags,
        // the split measured best once the history is there
        if (other_share >= 0.5 * ls) {
            dense_share = dense_lags;
            other_share = ls;
        }
        // the window computes window_width lags of a short last chunk
        const double width = (nchunks - 1) * DENSE + window_width<DENSE>(dense_lags - (nchunks - 1) * DENSE);
        const double wdense = (hist ? COST_DENSE : 3) * dense_share * width / dense_lags;
        const double wother = (hist ? COST_HIST : 5) * other_share;
        wd = (int)std::lround(LAG_WAVES * wdense / std::max(wdense + wother, 1e-9));
        wd = std::max(wd, nchunks);
        wd = std::min(wd, ls ? LAG_WAVES - 1 : LAG_WAVES);
        wd -= wd % nchunks;
    }
    const int ws = LAG_WAVES - wd;
    int ss = 0;
    if (!hist)
        for (int cand = NSEG; cand >= 1; cand /= 2)
            if (ls * cand <= SLOTS * ws) {
                ss = cand;
                break;
            }
    pl.dense_lags = dense_lags;
    // the window of the chunk from lag tau_a reads frames down to
    // tau_a + DENSE - 1 before the first new frame
    pl.pre = nchunks ? lags[(nchunks - 1) * DENSE] + DENSE - 1 : 0;
    pl.wd = wd;
    pl.sd = nchunks ? wd / nchunks : 1;
    pl.ss = ss;
    pl.hist = hist;
    if (!hist) return;
    const int tmin = lags[dense_lags], tmax = lags[L - 1];
    pl.pmin = std::max(-dmax, -tmax);
    pl.pmax = n - 1 - tmin;
    const int nb = pl.pmax >= pl.pmin ? (pl.pmax - pl.pmin) / B + 1 : 0;
    if (nb > PLAN_MAX_BLOCKS) return;  // round-robin blocks
    // every wave takes blocks when there are window waves: each window wave
    // starts with its window's work, in history-pair units
    const bool all = wd > 0;
    const int nw = all ? LAG_WAVES : ws;
    std::vector<long long> load(nw, 0);
    if (all) {
        const int len = (n + pl.sd - 1) / pl.sd;
        for (int w = 0; w < wd; ++w) {
            const int c = w / pl.sd, seg = w % pl.sd;
            const int kc = std::min(DENSE, dense_lags - c * DENSE);
            const int steps = std::max(0, std::min(len, n - seg * len));
            load[w] = (long long)steps *
                      (window_width<DENSE>(kc) * PLAN_COST_PAIR * COST_DENSE / COST_HIST + PLAN_COST_STEP) *
                      PLAN_DENSE_PCT / 100;
        }
    }
    std::vector<int> owner(nb);
    for (int b = nb - 1; b >= 0; --b) {
        const int p0 = pl.pmin + b * B;
        long long cost = PLAN_COST_BLOCK;
        for (int l = dense_lags; l < L; ++l) {
            const int j0 = p0 + lags[l];
            const int lo = std::max(0, -j0), hi = std::min(B, n - j0);
            if (hi > lo) cost += PLAN_COST_VISIT + (long long)PLAN_COST_PAIR * (hi - lo);
        }
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        owner[b] = k;
        load[k] += cost * PLAN_WAVE_PCT[(all ? k : k + wd) / 4] / 100;
    }
    int at = 0;
    for (int k = 0; k < nw; ++k) {
        pl.blk_start[k] = (unsigned char)at;
        for (int b = 0; b < nb; ++b)
            if (owner[b] == k) pl.blk_list[at++] = (unsigned char)b;
    }
    for (int k = nw; k <= LAG_WAVES; ++k) pl.blk_start[k] = (unsigned char)at;
    pl.listed = 1;
    pl.all_hist = all;
}

template <int J>
static void launch_lag_t(const LagArgs &a, hipStream_t st) {
    LagPlan pl;
    make_plan<J>(a.lags_host, a.lag_count, a.n_new, a.dmax, pl);
    const unsigned grid = (unsigned)(a.elems_per_frame / RING_GROUP);
    if (a.comp)
        hipLaunchKernelGGL((lag_kernel<J, true>), dim3(grid), dim3(LAG_NT), 0, st, (const float2 *)a.ring, a.acc,
                           a.comp, a.elems_per_frame, a.ring_slots, a.slot0, a.n_new, a.dmax, a.lags, a.lag_count,
                           pl);
    else
        hipLaunchKernelGGL((lag_kernel<J, false>), dim3(grid), dim3(LAG_NT), 0, st, (const float2 *)a.ring, a.acc,
                           nullptr, a.elems_per_frame, a.ring_slots, a.slot0, a.n_new, a.dmax, a.lags, a.lag_count,
                           pl);
}

void make_lag_plan(int J, const int *lags, int L, int n, int dmax, LagPlan *pl) {
    switch (J) {
    case 32: make_plan<32>(lags, L, n, dmax, *pl); break;
    case 64: make_plan<64>(lags, L, n, dmax, *pl); break;
    case 128: make_plan<128>(lags, L, n, dmax, *pl); break;
    default: make_plan<256>(lags, L, n, dmax, *pl); break;
    }
}

template <int J>
static hipError_t launch_lag_batch_t(bool comp, const BatchHdr *hdr, const LagJob *jobs, const LagPlan *plans,
                                     long long groups, hipStream_t st) {
    const dim3 grid((unsigned)groups);
    if (comp)
        hipLaunchKernelGGL((lag_kernel_batch<J, true>), grid, dim3(LAG_NT), 0, st, hdr, jobs, plans);
    else
        hipLaunchKernelGGL((lag_kernel_batch<J, false>), grid, dim3(LAG_NT), 0, st, hdr, jobs, plans);
    return hipGetLastError();
}

// Blocks of 32 .. 128 frames only: serving flushes every chunk through the
// lag engine, so its blocks are small; instantiating the 256-frame batch
// kernel as well made the compiler spill more in the 256-frame single
// kernel of the whole-video path.
hipError_t launch_lag_batch(int J, bool comp, const BatchHdr *hdr, const LagJob *jobs, const LagPlan *plans,
                            long long groups, hipStream_t st) {
    if (groups < 1 || groups > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (J) {
    case 32: return launch_lag_batch_t<32>(comp, hdr, jobs, plans, groups, st);
    case 64: return launch_lag_batch_t<64>(comp, hdr, jobs, plans, groups, st);
    case 128: return launch_lag_batch_t<128>(comp, hdr, jobs, plans, groups, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lag(const LagArgs &a, hipStream_t st) {
    if (a.elems_per_frame % RING_GROUP != 0 || a.n_new < 1 || a.n_new > a.lag_block || a.dmax < 0 ||
        a.slot0 < 0 || a.slot0 >= a.ring_slots || !a.lags_host || a.lag_count < 1)
        return hipErrorInvalidValue;
    switch (a.lag_block) {
    case 32: launch_lag_t<32>(a, st); break;
    case 64: launch_lag_t<64>(a, st); break;
    case 128: launch_lag_t<128>(a, st); break;
    case 256: launch_lag_t<256>(a, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ddm
