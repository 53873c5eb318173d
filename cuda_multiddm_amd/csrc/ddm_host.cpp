// Host-only parts of the DDM engine: descriptor validation, the q-ring
// tables, synthetic video, the I(lambda, tau) writer.  Compiled with g++
// -ffp-contract=off so the FP32 ring-membership test rounds every step the
// way the reference's host code does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

#include "ddm_internal.h"

namespace ddm {

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

static std::shared_ptr<const std::vector<int>> ring_members(int s, float lambda, float tol);

// The part of each s x s tile's half spectrum an engine's ring holds
// (LayoutTables).  The 256 .. 1024 tiles are kept column by column (their
// transform ends in a column pass), the others row by row.  With
// DDM_SPECTRUM_MASKED an 8 .. 1024 tile keeps, of every row (column), the
// prefix of columns (the band of low |ky| rows) reaching its last q-ring
// member, when that is less than 90 % of its elements (below that the
// saving in lag work outweighs the costlier stores); the 2 / 4 / 2048
// transforms write the full spectrum.  Tables are built once per (scale,
// tolerance, lambdas, mode) and shared.
static std::mutex g_lay_mu;
static std::map<std::tuple<int, int, uint32_t, std::vector<float>>, std::shared_ptr<const LayoutTables>> g_lays;

static std::shared_ptr<const LayoutTables> layout_tables(int s, const Params &p) {
    uint32_t tb;
    std::memcpy(&tb, &p.tol, 4);
    const bool masked = p.spectrum == DDM_SPECTRUM_MASKED && s >= 8 && s <= 1024;
    const auto key = std::make_tuple(s, masked ? 1 : 0, masked ? tb : 0u,
                                     masked ? p.lambdas : std::vector<float>());
    {
        std::lock_guard<std::mutex> lk(g_lay_mu);
        auto it = g_lays.find(key);
        if (it != g_lays.end()) return it->second;
    }
    const int H = s / 2, full = s * (H + 1);
    const bool cols = s >= 256 && s <= 1024;
    // per row: columns kept (prefix); per column: rows kept from the top
    // (ky < lo) and from the bottom (ky >= s - hi)
    std::vector<int> width(s, H + 1), lo(H + 1, s), hi(H + 1, 0);
    if (masked) {
        std::fill(width.begin(), width.end(), 0);
        std::fill(lo.begin(), lo.end(), 0);
        bool any = false;
        for (float lam : p.lambdas) {
            auto m = ring_members(s, lam, p.tol);
            if (!m) return nullptr;
            for (int li : *m) {
                const int ky = li / (H + 1), kx = li % (H + 1);
                any = true;
                width[ky] = std::max(width[ky], kx + 1);
                if (ky < H)
                    lo[kx] = std::max(lo[kx], ky + 1);
                else
                    hi[kx] = std::max(hi[kx], s - ky);
            }
        }
        // a scale without ring members keeps one (unused) element: DC
        if (!any) width[0] = lo[0] = 1;
        long long n = 0;
        if (cols)
            for (int kx = 0; kx <= H; ++kx) n += lo[kx] + hi[kx];
        else
            for (int w : width) n += w;
        if (n * 10 >= (long long)full * 9) {
            std::fill(width.begin(), width.end(), H + 1);
            std::fill(lo.begin(), lo.end(), s);
            std::fill(hi.begin(), hi.end(), 0);
        }
    }
    auto t = std::make_shared<LayoutTables>();
    t->pos.assign(full, -1);
    int n = 0;
    if (cols) {
        t->kind = LAY_COLS;
        t->start.assign(H + 2, 0);
        t->band.assign(H + 1, 0);
        int last = 0;
        for (int kx = 0; kx <= H; ++kx) {
            t->start[kx] = n;
            t->band[kx] = lo[kx];
            for (int r = 0; r < lo[kx] + hi[kx]; ++r) {
                const int ky = r < lo[kx] ? r : r + (s - lo[kx] - hi[kx]);
                t->pos[(size_t)ky * (H + 1) + kx] = n++;
            }
            if (kx < H && lo[kx] + hi[kx] > 0) last = kx;
        }
        t->start[H + 1] = n;
        t->pass_cols = std::min(H, (last / 16 + 1) * 16);
    } else {
        t->start.assign(s + 1, 0);
        int widest = 1;
        for (int ky = 0; ky < s; ++ky) {
            t->start[ky] = n;
            for (int kx = 0; kx < width[ky]; ++kx) t->pos[(size_t)ky * (H + 1) + kx] = n++;
            widest = std::max(widest, width[ky]);
        }
        t->row_lanes = 1;
        while (t->row_lanes < widest && t->row_lanes < 64) t->row_lanes *= 2;
        t->start[s] = n;
        t->kind = n == full ? LAY_FULL : LAY_ROWS;
        t->pass_cols = H;
        if (t->kind == LAY_FULL) t->start.clear();
    }
    t->elems = n;
    if (n < 1) return nullptr;
    std::lock_guard<std::mutex> lk(g_lay_mu);
    return g_lays.emplace(key, t).first->second;
}

int check_desc(const ddm_desc *d, Params *out, std::string *msg) {
    auto fail = [&](int code, const char *text) {
        if (msg) *msg = text;
        return code;
    };
    if (!d) return fail(DDM_EINVAL, "descriptor is NULL");
    Params p;
    if (d->width <= 0 || d->height <= 0) return fail(DDM_EINVAL, "frame size must be positive");
    if (d->bytes_per_pixel <= 0) return fail(DDM_EINVAL, "bytes per pixel must be positive");
    if (d->pixel_format < DDM_PIX_U8 || d->pixel_format > DDM_PIX_U16BE)
        return fail(DDM_EINVAL, "unknown pixel format");
    if (d->bytes_per_pixel < pixel_bytes_needed(d->pixel_format))
        return fail(DDM_EINVAL, "16-bit pixel formats need at least 2 bytes per pixel");
    if (d->scale_count <= 0 || !d->scales) return fail(DDM_EINVAL, "no scales given");
    if (d->lag_count <= 0 || !d->lags) return fail(DDM_EINVAL, "no tau values given");
    if (d->lambda_count <= 0 || !d->lambdas) return fail(DDM_EINVAL, "no lambda values given");
    if (d->chunk_frames <= 0) return fail(DDM_EINVAL, "chunk frame count must be positive");
    if (d->norm_mode != DDM_NORM_FRAMES && d->norm_mode != DDM_NORM_PAIRS)
        return fail(DDM_EINVAL, "unknown normalisation mode");
    if (d->accumulation != DDM_ACC_COMPENSATED && d->accumulation != DDM_ACC_PLAIN)
        return fail(DDM_EINVAL, "unknown accumulation mode");
    if (d->spectrum != DDM_SPECTRUM_MASKED && d->spectrum != DDM_SPECTRUM_FULL)
        return fail(DDM_EINVAL, "unknown spectrum mode");
    if (d->lag_block != 0 && d->lag_block != 32 && d->lag_block != 64 && d->lag_block != 128 &&
        d->lag_block != 256)
        return fail(DDM_EINVAL, "lag block must be 0 (automatic), 32, 64, 128 or 256");

    p.width = d->width;
    p.height = d->height;
    p.bpp = d->bytes_per_pixel;
    p.fmt = d->pixel_format;
    p.x_off = d->x_off;
    p.y_off = d->y_off;
    p.scales.assign(d->scales, d->scales + d->scale_count);
    p.lags.assign(d->lags, d->lags + d->lag_count);
    p.lambdas.assign(d->lambdas, d->lambdas + d->lambda_count);
    p.tol = d->mask_tolerance;
    p.chunk = d->chunk_frames;
    p.norm_mode = d->norm_mode;
    p.lag_block = d->lag_block ? d->lag_block : kDefaultLagBlock;
    p.acc_mode = d->accumulation;
    p.spectrum = d->spectrum;

    // src/DDM.cu:319-321
    std::sort(p.lags.begin(), p.lags.end());
    std::sort(p.lambdas.begin(), p.lambdas.end());
    std::sort(p.scales.begin(), p.scales.end(), std::greater<int>());

    // src/DDM.cu:327-344
    for (size_t i = 0; i < p.scales.size(); ++i) {
        if (!is_pow2(p.scales[i])) return fail(DDM_ESCALE, "scales must be powers of two (> 0)");
        if (p.scales[i] > kMaxScale) return fail(DDM_ESCALE, "scales above 2048 are not supported");
        if (i + 1 < p.scales.size() && !(p.scales[i] > p.scales[i + 1]))
            return fail(DDM_ESCALE, "scales should be descending order");
    }
    for (size_t i = 0; i < p.lags.size(); ++i) {
        if (p.lags[i] < 0) return fail(DDM_ELAG, "Tau values should be positive");
        if (p.lags[i] > kMaxLag) return fail(DDM_ELAG, "Tau value too large");
        if (i + 1 < p.lags.size() && !(p.lags[i] < p.lags[i + 1]))
            return fail(DDM_ELAG, "Tau vector should be ascending order");
    }
    for (size_t i = 0; i < p.lambdas.size(); ++i) {
        if (!(p.lambdas[i] >= 0.0f)) return fail(DDM_ELAMBDA, "q-vector values should be positive");
        if (i + 1 < p.lambdas.size() && !(p.lambdas[i] < p.lambdas[i + 1]))
            return fail(DDM_ELAMBDA, "q-vector vector should be ascending order");
    }
    // src/DDM.cu:346-347: a warning only
    if (!(p.tol < 10 && p.tol > 1.0f))
        p.warnings.push_back("mask_tolerance is likely undesired value, refer to README for more information");
    // src/DDM.cu:349-350 is a hard error in the reference; here an empty ring
    // yields 0, so it is reported as a warning.
    if (p.lambdas.back() > (float)p.scales.back())
        p.warnings.push_back("The largest q-vector should be smaller than the smallest scale.");

    p.M = p.scales.front();
    // src/DDM.cu:459-460
    if (p.x_off < 0 || p.y_off < 0 || (long long)p.M + p.x_off > p.width ||
        (long long)p.M + p.y_off > p.height)
        return fail(DDM_EROI, "the specified out dimensions must be smaller than actual image size");

    p.max_lag = p.lags.back();
    // The ring keeps every frame a lag can reach back to: the frames waiting
    // for the lag engine (less than one lag block plus one chunk) and max_lag
    // frames before them (the reference keeps 3 chunks and so requires
    // max_lag <= chunk, src/DDM.cu:462).
    long long slots = (long long)p.chunk + p.lag_block + p.max_lag;
    // the lag engine addresses one ring group's slots with 32-bit byte
    // offsets (slots * 64 elements * 8 bytes < 2^31)
    if (slots >= kMaxRingSlots) return fail(DDM_EINVAL, "chunk + lag block + largest tau is too large");
    p.ring_slots = (int)slots;

    long long off = 0, results = 0;
    for (int s : p.scales) {
        ScaleGeom g;
        g.s = s;
        g.tps = p.M / s;
        g.tiles = g.tps * g.tps;
        g.lay = layout_tables(s, p);
        if (!g.lay) return fail(DDM_EINVAL, "q-ring tables could not be built");
        g.tile_elems = g.lay->elems;
        g.elem_off = off;
        off += g.tile_elems * g.tiles;
        results += (long long)g.tiles * (long long)p.lambdas.size() * (long long)p.lags.size();
        p.geom.push_back(g);
    }
    p.elems_per_frame = off;
    p.elems_padded = pad_elems(off);
    p.result_count = results;
    if (out) *out = std::move(p);
    if (msg) msg->clear();
    return DDM_OK;
}

size_t device_footprint(const Params &p) {
    const size_t L = p.lags.size(), Q = p.lambdas.size();
    size_t bytes = (size_t)p.ring_slots * (size_t)p.elems_padded * 8;  // spectrum ring
    bytes += L * (size_t)p.elems_padded * 4;                           // accumulators
    if (p.acc_mode == DDM_ACC_COMPENSATED)                              // compensation terms and the FP64
        bytes += L * (size_t)p.elems_padded * (4 + 8);                  // accumulators they fold into
    bytes += UPLOAD_STAGES * (size_t)p.chunk * p.M * p.M * p.bpp;      // upload staging
    bytes += (size_t)p.result_count * 8 + L * 12;                      // results, lags, norms
    for (const ScaleGeom &g : p.geom) {
        bytes += Q * ((size_t)g.tile_elems * 4 + 8);  // ring tables (bound)
        bytes += (g.lay->start.size() + g.lay->band.size()) * 4;  // tile layout tables
    }
    // upload staging and the between-pass buffer of 256 .. 1024 tile FFTs
    // grow to one FFT batch of frames when a call hands over more than a
    // chunk: counted at that size, so an admitted stream never outgrows it
    const int batch = default_fft_batch(p.chunk);
    bytes += UPLOAD_STAGES * (size_t)(batch - p.chunk) * p.M * p.M * p.bpp;
    // (one region per 256 .. 1024 scale when the engine's default tiler mode,
    // set by the region size, is the multi-scale tiler)
    return bytes + (size_t)fft_pass_elems(p.M, p.scales, batch, p.M <= kTilerFusedMaxRegion) * 8;
}

// Member lists of one (scale, lambda, tolerance) ring, cached for the life of
// the process: a serving process opens many streams with the same rings.
static std::mutex g_ring_mu;
static std::map<std::tuple<int, uint32_t, uint32_t>, std::shared_ptr<const std::vector<int>>> g_rings;

static std::shared_ptr<const std::vector<int>> ring_members(int s, float lambda, float tol) {
    uint32_t lb, tb;
    std::memcpy(&lb, &lambda, 4);
    std::memcpy(&tb, &tol, 4);
    const auto key = std::make_tuple(s, lb, tb);
    {
        std::lock_guard<std::mutex> lk(g_ring_mu);
        auto it = g_rings.find(key);
        if (it != g_rings.end()) return it->second;
    }
    std::vector<uint8_t> m((size_t)s * (s / 2 + 1));
    if (ddm_ring_mask(s, lambda, tol, m.data()) < 0) return nullptr;
    auto v = std::make_shared<std::vector<int>>();
    for (size_t e = 0; e < m.size(); ++e)
        if (m[e]) v->push_back((int)e);
    std::lock_guard<std::mutex> lk(g_ring_mu);
    g_rings.emplace(key, v);
    return v;
}

int build_ring_tables(const Params &p, RingTables *out) {
    const int Q = (int)p.lambdas.size();
    out->idx.clear();
    out->off.assign(p.geom.size() * Q, 0);
    out->cnt.assign(p.geom.size() * Q, 0);
    for (size_t si = 0; si < p.geom.size(); ++si) {
        const int s = p.geom[si].s;
        const LayoutTables &lay = *p.geom[si].lay;
        for (int q = 0; q < Q; ++q) {
            auto members = ring_members(s, p.lambdas[q], p.tol);
            if (!members) return DDM_EINVAL;
            out->off[si * Q + q] = (int)out->idx.size();
            out->cnt[si * Q + q] = (int)members->size();
            // each member's place in the tile's stored layout, in member
            // order (the averages sum in list order, whatever the layout)
            for (int li : *members) {
                const int at = lay.pos[li];
                if (at < 0) return DDM_EINVAL;
                out->idx.push_back(at);
            }
        }
    }
    return DDM_OK;
}

void fill_twiddles(std::vector<float> *tw) {
    tw->assign(4 * (size_t)kMaxScale, 0.0f);
    for (int n = 1; n <= kMaxScale; n <<= 1) {
        for (int m = 0; m < n; ++m) {
            double a = -2.0 * M_PI * (double)m / (double)n;
            (*tw)[2 * ((size_t)n + m)] = (float)std::cos(a);
            (*tw)[2 * ((size_t)n + m) + 1] = (float)std::sin(a);
        }
    }
}

// ---------------------------------------------------------------------------
// Counter-based random numbers (splitmix64 finaliser over a counter).
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t counter) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) + (counter + 1) * 0x9E3779B97F4A7C15ULL);
}
static inline double u01(uint64_t x) { return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
static inline double gauss(uint64_t seed, uint64_t stream, uint64_t counter) {
    uint64_t a = rng_at(seed, stream, 2 * counter), b = rng_at(seed, stream, 2 * counter + 1);
    return std::sqrt(-2.0 * std::log(u01(a))) * std::cos(2.0 * M_PI * u01(b));
}

}  // namespace ddm

using namespace ddm;

extern "C" int ddm_ring_mask(int scale, float lambda, float tolerance, uint8_t *mask) {
    if (!is_pow2(scale) || !mask) return DDM_EINVAL;
    // src/DDM.cu:102 and src/azimuthal_average.cu:235-270, in FP32.
    volatile float q = static_cast<float>(scale) / lambda;
    volatile float q2 = q * q;
    volatile float tol2 = tolerance * tolerance;
    const int w = scale, h = scale, half_w = w / 2, half_h = h / 2;
    int count = 0;
    for (int x = 0; x < w / 2 + 1; ++x) {
        for (int y = 0; y < h; ++y) {
            int xs = (x + half_w) % w - half_w;
            int ys = (y + half_h) % h - half_h;
            volatile float r2 = (float)(xs * xs + ys * ys);
            volatile float ratio = r2 / q2;
            bool px = (1.0f <= ratio) && (ratio <= tol2);
            count += px;
            mask[(size_t)y * (w / 2 + 1) + x] = px;
        }
    }
    return count;
}

extern "C" int ddm_synth_uniform(void *out, int n_frames, int width, int height, int bits,
                                 uint64_t seed, long long first_frame) {
    if (!out || n_frames < 0 || width <= 0 || height <= 0 || bits < 1 || bits > 16) return DDM_EINVAL;
    const uint64_t npx = (uint64_t)width * height;
    const uint32_t maxv = (1u << bits) - 1;
    for (int f = 0; f < n_frames; ++f) {
        uint64_t fr = (uint64_t)(first_frame + f);
        for (uint64_t i = 0; i < npx; ++i) {
            uint32_t v = (uint32_t)(rng_at(seed, 7, fr * npx + i) >> 40);
            v = bits == 8 ? (v & 0xFF) : (uint32_t)(((uint64_t)v * (maxv + 1ULL)) >> 24);
            if (bits <= 8)
                ((uint8_t *)out)[(uint64_t)f * npx + i] = (uint8_t)v;
            else
                ((uint16_t *)out)[(uint64_t)f * npx + i] = (uint16_t)v;
        }
    }
    return DDM_OK;
}

extern "C" int ddm_synth_brownian(void *out, int n_frames, int width, int height, int bits,
                                  int n_spots, float sigma_px, float amplitude, float diffusion,
                                  float background, float noise_sigma, uint64_t seed, int nthreads) {
    if (!out || n_frames < 0 || width <= 0 || height <= 0 || bits < 1 || bits > 16 || n_spots < 0 ||
        !(sigma_px > 0) || diffusion < 0)
        return DDM_EINVAL;
    const double step_sd = std::sqrt(2.0 * diffusion);
    // trajectories, sequential in time
    std::vector<float> pos((size_t)n_frames * n_spots * 2);
    for (int k = 0; k < n_spots; ++k) {
        double x = u01(rng_at(seed, 1, 2 * (uint64_t)k)) * width;
        double y = u01(rng_at(seed, 1, 2 * (uint64_t)k + 1)) * height;
        for (int f = 0; f < n_frames; ++f) {
            if (f > 0) {
                uint64_t c = ((uint64_t)f * n_spots + k) * 2;
                x += step_sd * gauss(seed, 2, c);
                y += step_sd * gauss(seed, 2, c + 1);
                x -= std::floor(x / width) * width;
                y -= std::floor(y / height) * height;
            }
            pos[((size_t)f * n_spots + k) * 2] = (float)x;
            pos[((size_t)f * n_spots + k) * 2 + 1] = (float)y;
        }
    }
    const int rad = (int)std::ceil(4.0 * sigma_px);
    const double inv2s2 = 1.0 / (2.0 * (double)sigma_px * sigma_px);
    const double maxv = (double)((1u << bits) - 1);
    const uint64_t npx = (uint64_t)width * height;
    auto render = [&](int f0, int f1) {
        std::vector<float> img(npx);
        for (int f = f0; f < f1; ++f) {
            std::fill(img.begin(), img.end(), 0.0f);
            for (int k = 0; k < n_spots; ++k) {
                double cx = pos[((size_t)f * n_spots + k) * 2];
                double cy = pos[((size_t)f * n_spots + k) * 2 + 1];
                int ix = (int)std::floor(cx), iy = (int)std::floor(cy);
                for (int dy = -rad; dy <= rad + 1; ++dy) {
                    int py = iy + dy;
                    double ddy = (double)py - cy;
                    int wy = ((py % height) + height) % height;
                    for (int dx = -rad; dx <= rad + 1; ++dx) {
                        int px = ix + dx;
                        double ddx = (double)px - cx;
                        int wx = ((px % width) + width) % width;
                        img[(size_t)wy * width + wx] +=
                            (float)(amplitude * std::exp(-(ddx * ddx + ddy * ddy) * inv2s2));
                    }
                }
            }
            for (uint64_t i = 0; i < npx; ++i) {
                double v = background + img[i];
                if (noise_sigma > 0) v += noise_sigma * gauss(seed, 3, (uint64_t)f * npx + i);
                v = std::nearbyint(v);
                v = v < 0 ? 0 : (v > maxv ? maxv : v);
                if (bits <= 8)
                    ((uint8_t *)out)[(uint64_t)f * npx + i] = (uint8_t)v;
                else
                    ((uint16_t *)out)[(uint64_t)f * npx + i] = (uint16_t)v;
            }
        }
    };
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min(nthreads, std::max(1, n_frames));
    std::vector<std::thread> th;
    int per = (n_frames + nthreads - 1) / std::max(1, nthreads);
    for (int t = 0; t < nthreads; ++t) {
        int f0 = t * per, f1 = std::min(n_frames, f0 + per);
        if (f0 < f1) th.emplace_back(render, f0, f1);
    }
    for (auto &t : th) t.join();
    return DDM_OK;
}

extern "C" int ddm_write_iqt_file(const char *path, const double *iqt, const float *lambdas,
                                  int lambda_count, const int *lags, int lag_count, float fps) {
    if (!path || !iqt || !lambdas || !lags || lambda_count <= 0 || lag_count <= 0) return DDM_EINVAL;
    std::ofstream f(path);
    if (!f.is_open()) return DDM_EINVAL;
    // src/azimuthal_average.cu:120-141: default ostream formatting of floats
    for (int i = 0; i < lambda_count; ++i) f << lambdas[i] << " ";
    f << "\n";
    for (int t = 0; t < lag_count; ++t) f << static_cast<float>(lags[t]) / fps << " ";
    f << "\n";
    for (int q = 0; q < lambda_count; ++q) {
        for (int t = 0; t < lag_count; ++t) f << static_cast<float>(iqt[(size_t)q * lag_count + t]) << " ";
        f << "\n";
    }
    return f.good() ? DDM_OK : DDM_EINVAL;
}

extern "C" const char *ddm_status_string(int code) {
    switch (code) {
    case DDM_OK: return "ok";
    case DDM_EINVAL: return "invalid argument";
    case DDM_ENOMEM: return "out of memory";
    case DDM_EDEVICE: return "device error";
    case DDM_EROI: return "ROI outside the frame";
    case DDM_ESCALE: return "invalid scale";
    case DDM_ELAG: return "invalid tau";
    case DDM_ELAMBDA: return "invalid lambda";
    case DDM_ESTATE: return "invalid engine state";
    case DDM_ERANGE: return "out of range";
    case DDM_EBUSY: return "busy: no slot or device budget left";
    default: return "unknown status";
    }
}

extern "C" int ddm_validate(const ddm_desc *desc, ddm_info *info, char *msg, size_t msg_len) {
    Params p;
    std::string m;
    int rc = check_desc(desc, &p, &m);
    if (msg && msg_len) {
        if (rc == DDM_OK && !p.warnings.empty()) m = p.warnings.front();
        std::snprintf(msg, msg_len, "%s", m.c_str());
    }
    if (rc == DDM_OK && info) {
        std::memset(info, 0, sizeof(*info));
        info->largest_scale = p.M;
        info->scale_count = (int)p.scales.size();
        info->lag_count = (int)p.lags.size();
        info->lambda_count = (int)p.lambdas.size();
        info->ring_slots = p.ring_slots;
        info->elems_per_frame = p.elems_per_frame;
        info->result_count = p.result_count;
        info->elems_per_frame_padded = p.elems_padded;
        info->lag_block = p.lag_block;
        info->device_bytes = device_footprint(p);
    }
    return rc;
}
