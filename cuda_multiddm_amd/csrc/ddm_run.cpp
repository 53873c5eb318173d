// runDDM -- the reference's analysis driver (defined src/DDM.cu:268-896,
// declared src/main.cpp:40-64) with the same signature, on top of the C ABI
// of the MI355X engine (include/ddm_amd.h): video setup, every parameter
// check before the device is touched, the chunk loop (host frames ->
// ddm_submit_host), -G windows, -K / -A dumps, the output files
// <prefix><scale>-<tile> (writeIqtToFile, src/azimuthal_average.cu:112-150)
// and the [Time elapsed] block (src/DDM.cu:881-895).  ./ddm's main() calls
// it after parsing the flags, as the reference's main does
// (src/main.cpp:315-339); a build that keeps the reference's main.cpp links
// this file in place of src/DDM.cu.
//
// Differences from the reference, all deliberate: the caller's tau / lambda
// / scale arrays are copied and sorted inside the engine, never in place
// (src/DDM.cu:319-321); tau may exceed the chunk length and lambda the
// smallest scale (hard errors there, src/DDM.cu:349, 462); OpenCV decoding
// and web cameras are not part of this build; `multistream` and
// `webcam_idx` are accepted and unused (uploads always go through three
// staging buffers in turn).
#include "ddm_run.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "ddm_amd.h"
#include "movie_reader.h"

namespace {

bool g_verbose = false;

void verbose(const char *fmt, ...) {
    if (!g_verbose) return;
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
}

using ddm_run::condition_assert;

void write_iqt(const std::string &path, const double *iqt, const std::vector<float> &lambdas,
               const std::vector<int> &lags, float fps) {
    int rc = ddm_write_iqt_file(path.c_str(), iqt, lambdas.data(), (int)lambdas.size(), lags.data(),
                                (int)lags.size(), fps);
    if (rc != DDM_OK) {
        std::fprintf(stderr, "[Out Error] Unable to open %s.\n", path.c_str());
        std::exit(EXIT_FAILURE);
    }
    verbose("I(lambda, tau) written to %s\n", path.c_str());
}

struct Analysis {
    ddm_engine *eng = nullptr;
    ddm_info info{};
    std::vector<int> scales, lags;
    std::vector<float> lambdas;
    float fps = 1.0f;
};

void engine_check(int rc, ddm_engine *eng, const char *what) {
    if (rc == DDM_OK) return;
    std::fprintf(stderr, "[Error] %s: %s (%s)\n", what, ddm_last_error(eng), ddm_status_string(rc));
    std::exit(EXIT_FAILURE);
}

// Averages and writes every (scale, tile) file for the accumulated window.
void emit(Analysis &a, const std::string &prefix, bool non_ave) {
    std::vector<double> res((size_t)a.info.result_count);
    engine_check(ddm_results(a.eng, res.data()), a.eng, "azimuthal average");
    const size_t Q = a.lambdas.size(), L = a.lags.size();
    size_t off = 0;
    const int M = a.info.largest_scale;
    for (size_t si = 0; si < a.scales.size(); ++si) {
        const int s = a.scales[si];
        const int T = (M / s) * (M / s);
        std::vector<float> acc;
        const size_t tile_elems = (size_t)s * (s / 2 + 1);
        if (non_ave) {
            acc.resize(L * T * tile_elems);
            engine_check(ddm_get_accumulators(a.eng, (int)si, acc.data()), a.eng, "accumulator dump");
        }
        for (int t = 0; t < T; ++t) {
            const std::string name = prefix + std::to_string(s) + "-" + std::to_string(t);
            if (non_ave) {
                // intended layout of writeNonAveIqtToFile (src/azimuthal_average.cu:65-106):
                // lambda line, tau line, then one row of tile elements per tau
                const std::string nn = name + "-nonAve";
                std::ofstream f(nn);
                if (!f.is_open()) {
                    std::fprintf(stderr, "[Out Error] Unable to open %s.\n", nn.c_str());
                    std::exit(EXIT_FAILURE);
                }
                for (float lam : a.lambdas) f << lam << " ";
                f << "\n";
                for (int tau : a.lags) f << static_cast<float>(tau) / a.fps << " ";
                f << "\n";
                for (size_t l = 0; l < L; ++l) {
                    const float *row = acc.data() + (l * T + t) * tile_elems;
                    for (size_t e = 0; e < tile_elems; ++e) f << row[e] << " ";
                    f << "\n";
                }
                verbose("Non-averaged ISF written to %s\n", nn.c_str());
            }
            write_iqt(name, res.data() + off + (size_t)t * Q * L, a.lambdas, a.lags, a.fps);
        }
        off += (size_t)T * Q * L;
    }
}

// -A: |F / s^2|^2 of every tile of every frame of the chunk just submitted
// (calculateAmplitude + writeFFTToFile, src/DDM_kernel.cuh:107-122,
// src/azimuthal_average.cu:31-59), one row per (frame, tile).
void dump_fft(Analysis &a, const std::string &prefix, int chunk_idx, long long first, int n) {
    const int M = a.info.largest_scale;
    for (size_t si = 0; si < a.scales.size(); ++si) {
        const int s = a.scales[si];
        const int T = (M / s) * (M / s);
        const size_t tile_elems = (size_t)s * (s / 2 + 1);
        std::vector<float> spec(T * tile_elems * 2);
        const std::string name = prefix + "_fft_chunk_" + std::to_string(chunk_idx) + "-" + std::to_string(s);
        std::ofstream f(name);
        if (!f.is_open()) {
            std::fprintf(stderr, "[Out Error] Unable to open %s.\n", name.c_str());
            std::exit(EXIT_FAILURE);
        }
        const float norm = 1.0f / (float)(s * s);
        for (int fr = 0; fr < n; ++fr) {
            engine_check(ddm_get_spectra(a.eng, first + fr, (int)si, spec.data()), a.eng, "spectrum dump");
            for (int t = 0; t < T; ++t) {
                for (size_t e = 0; e < tile_elems; ++e) {
                    const float re = norm * spec[2 * (t * tile_elems + e)];
                    const float im = norm * spec[2 * (t * tile_elems + e) + 1];
                    f << re * re + im * im << " ";
                }
                f << "\n";
            }
        }
        verbose("Primary FFT is written to %s\n", name.c_str());
    }
}

}  // namespace

namespace ddm_run {

void set_verbose(bool on) { g_verbose = on; }

// src/debug.cpp:33-45
void condition_assert(bool cond, const std::string &text, bool abort_run) {
    if (cond) return;
    if (abort_run) {
        std::fprintf(stderr, "[Error] %s\n", text.c_str());
        std::exit(EXIT_FAILURE);
    }
    std::fprintf(stderr, "[Warning] %s\n", text.c_str());
}

}  // namespace ddm_run

void runDDM(std::string file_in, std::string file_out, int *tau_arr, int tau_count, float *lambda_arr,
            int lambda_count, int *scale_arr, int scale_count, int x_off, int y_off, int frame_count,
            int frame_offset, int chunk_frame_count, bool /*multistream*/, bool use_webcam, int /*webcam_idx*/,
            float q_tolerance, bool use_movie_file, bool use_index_fps, bool use_explicit_fps, float explicit_fps,
            int dump_accum_after, bool benchmark_mode, bool nonAveOutput, bool FFTOutput) {
    condition_assert(tau_arr && tau_count > 0, "tau-file holds no values.", true);
    condition_assert(lambda_arr && lambda_count > 0, "lambda-file holds no values.", true);
    condition_assert(scale_arr && scale_count > 0, "scales-file holds no values.", true);
    condition_assert(frame_count > 0, "Number of frames to analyse (-N) must be positive.", true);
    condition_assert(chunk_frame_count > 0, "Chunk frame count (-C) must be positive.", true);
    condition_assert(frame_offset >= 0, "Frame offset (-s) must not be negative.", true);
    condition_assert(dump_accum_after >= 0, "Purge interval (-G) must not be negative.", true);
    const std::vector<int> lags(tau_arr, tau_arr + tau_count);
    const std::vector<float> lambdas(lambda_arr, lambda_arr + lambda_count);
    const std::vector<int> scales(scale_arr, scale_arr + scale_count);
    auto start = std::chrono::high_resolution_clock::now();
    verbose("[multiDDM Begin]\n");

    // video setup (src/DDM.cu:390-452)
    int w = 0, h = 0, bpp = 1, fmt = DDM_PIX_U8;
    float fps = 1.0f;
    ddm::MovieReader movie;
    int max_scale = 0;
    for (int s : scales) max_scale = std::max(max_scale, s);
    if (benchmark_mode) {
        w = h = max_scale;
    } else if (use_webcam) {
        condition_assert(false, "web-camera input is not available in this build", true);
    } else if (use_movie_file) {
        if (!movie.open(file_in, frame_offset)) condition_assert(false, movie.error(), true);
        w = movie.info().width;
        h = movie.info().height;
        bpp = movie.info().bpp;
        fmt = movie.info().pixel_format;
        fps = movie.info().fps;
    } else {
        condition_assert(false, "error opening video file with openCV (only .movie input, -M, is available in this build)",
                         true);
    }
    if (use_index_fps) fps = 1.0f;
    if (use_explicit_fps) fps = explicit_fps;

    ddm_desc d{};
    d.width = w;
    d.height = h;
    d.bytes_per_pixel = bpp;
    d.pixel_format = fmt;
    d.x_off = x_off;
    d.y_off = y_off;
    d.scales = scales.data();
    d.scale_count = (int)scales.size();
    d.lags = lags.data();
    d.lag_count = (int)lags.size();
    d.lambdas = lambdas.data();
    d.lambda_count = (int)lambdas.size();
    d.mask_tolerance = q_tolerance;
    d.chunk_frames = chunk_frame_count;
    d.norm_mode = DDM_NORM_FRAMES;
    d.lag_block = 0;
    // per-element output (-K) and the spectrum dump (-A) read every element;
    // the averages alone need only the q-ring band
    d.spectrum = (nonAveOutput || FFTOutput) ? DDM_SPECTRUM_FULL : DDM_SPECTRUM_MASKED;

    // every parameter check happens before the device is touched
    char msg[512];
    ddm_info vinfo{};
    int rc = ddm_validate(&d, &vinfo, msg, sizeof msg);
    if (rc != DDM_OK) condition_assert(false, msg, true);
    if (msg[0]) condition_assert(false, msg, false);
    verbose("%d tau-values.\n", (int)lags.size());
    verbose("%d q-vector values.\n", (int)lambdas.size());
    verbose("Parameter Check Done.\n");
    verbose("[Video info - (%d x %d), %d Frames (offset %d), %.4f FPS, %d bytes per pixel]\n", w, h,
            frame_count, frame_offset, fps, bpp);

    Analysis a;
    a.fps = fps;
    rc = ddm_create(&d, &a.eng);
    if (rc != DDM_OK) condition_assert(false, ddm_last_error(nullptr), true);
    engine_check(ddm_get_info(a.eng, &a.info), a.eng, "engine info");
    a.scales.resize(a.info.scale_count);
    a.lags.resize(a.info.lag_count);
    a.lambdas.resize(a.info.lambda_count);
    engine_check(ddm_get_params(a.eng, a.scales.data(), a.lags.data(), a.lambdas.data()), a.eng, "params");
    verbose("Memory allocation finished.\n\tDevice:\t %f GB\n\tHost (pinned):\t %f GB\n",
            a.info.device_bytes / 1073741824.0, a.info.pinned_bytes / 1073741824.0);

    const size_t frame_bytes = (size_t)w * h * bpp;
    std::vector<uint8_t> host((size_t)chunk_frame_count * frame_bytes);
    const int total_chunks = (frame_count + chunk_frame_count - 1) / chunk_frame_count;
    long long done = 0;
    int window = 0;
    for (int ci = 0; ci < total_chunks; ++ci) {
        const int n = std::min<long long>(chunk_frame_count, frame_count - done);
        if (benchmark_mode) {
            engine_check(ddm_synth_uniform(host.data(), n, w, h, 8, 0x5EEDULL, frame_offset + done), nullptr,
                         "synthetic frames");
        } else if (!movie.read(host.data(), n)) {
            condition_assert(false, movie.error(), true);
        }
        engine_check(ddm_submit_host(a.eng, host.data(), n), a.eng, "chunk submission");
        if (FFTOutput) dump_fft(a, file_out, ci, done, n);
        done += n;
        verbose("[Chunk complete (%d \\ %d)]\n", ci + 1, total_chunks);
        // -G: output and purge every purge_every chunks (src/DDM.cu:800-826)
        if (dump_accum_after > 0 && (ci + 1) % dump_accum_after == 0 && ci + 1 < total_chunks) {
            verbose("[Parsing Accumulator]\n");
            emit(a, file_out + "_t" + std::to_string(window) + "_", nonAveOutput);
            engine_check(ddm_reset(a.eng), a.eng, "purge");
            verbose("[Purging Accumulator]\n");
            ++window;
        }
    }
    engine_check(ddm_sync(a.eng), a.eng, "main loop");
    auto end_main = std::chrono::high_resolution_clock::now();
    verbose("Analysis.\n");
    emit(a, file_out, nonAveOutput);
    ddm_destroy(a.eng);
    auto end_out = std::chrono::high_resolution_clock::now();

    // src/DDM.cu:881-895
    const double d1 = std::chrono::duration_cast<std::chrono::microseconds>(end_main - start).count();
    const double d2 = std::chrono::duration_cast<std::chrono::microseconds>(end_out - end_main).count();
    std::printf("[Time elapsed] "
                "\n\tMain:\t\t%f s, "
                "\n\tRadial Average:\t%f s,"
                "\n\tTotal\t\t%f s,\t(%f frame / second)\n",
                (float)d1 / 1e6, (float)d2 / 1e6, ((float)d1 + (float)d2) / 1e6,
                (float)(frame_count * 1e6) / ((float)d1 + (float)d2));
}
