// ./ddm -- command-line front end of the MI355X DDM engine, flag- and
// file-compatible with the reference CLI (src/main.cpp:104-342) and its
// runDDM driver (src/DDM.cu:268-896):
//   -o PATH -N INT -Q PATH -T PATH -S PATH   required
//   -f PATH [-M] | -W[INT] | -B               input (.movie via -M; -B synthetic)
//   -s -x -y -I -v -Z -t -C -G -F -K -A -h    options
// Output: one file <prefix><scale>-<tile> per (scale, tile) in the layout of
// writeIqtToFile (src/azimuthal_average.cu:112-150); -G windows write
// <prefix>_t<i>_<scale>-<tile>.  Differences from the reference, all
// deliberate: -B and -K take no argument (the reference's optstring makes
// them swallow one, src/main.cpp:115); -F sets the frame rate (a no-op in
// the reference, src/main.cpp:214-218); tau may exceed the chunk length and
// lambda the smallest scale (hard errors there, src/DDM.cu:349, 462);
// OpenCV decoding and web cameras are not part of this build.
#include <getopt.h>
#include <unistd.h>

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

// src/debug.cpp:33-45
void condition_assert(bool cond, const std::string &text, bool abort_run = false) {
    if (cond) return;
    if (abort_run) {
        std::fprintf(stderr, "[Error] %s\n", text.c_str());
        std::exit(EXIT_FAILURE);
    }
    std::fprintf(stderr, "[Warning] %s\n", text.c_str());
}

void print_help() {
    std::fprintf(stderr,
                 "\n ~~ multiscale DDM - MI355X - HELP ~~ \n"
                 "\n"
                 "  Usage ./ddm [OPTION]..\n"
                 "  -h           Print out this help.\n"
                 "   REQUIRED ARGS\n"
                 "  -o PATH      Output file-path.\n"
                 "  -N INT       Number of frames to analyse.\n"
                 "  -Q PATH      Specify path to lambda-value file (line separated).\n"
                 "  -T PATH      Specify path to tau-value file (line separated). \n"
                 "  -S PATH      Specify path to scale-value file (line separated). \n"
                 "   INPUT ARGS\n"
                 "  -f PATH      Specify path to input video (with -M: .movie file).\n"
                 "  -W INT       Use web-camera as input video (not available in this build).\n"
                 "  -B           Benchmark mode, will perform analysis on random data.\n"
                 "   OPTIONAL ARGS\n"
                 "  -s OFFSET    Set first frame offset (default 0).\n"
                 "  -x OFFSET    Set x-offset (default 0).\n"
                 "  -y OFFSET    Set y-offset (default 0).\n"
                 "  -I           Use frame indices for tau-labels not real time.\n"
                 "  -v           Verbose mode on.\n"
                 "  -Z           Single-stream mode (accepted; the engine always double-buffers uploads).\n"
                 "  -t INT       Set the q-vector mask tolerance - percent (integer only) (default 20 i.e. radial mask (1 - 1.2) * q).\n"
                 "  -C INT       Set main chunk frame count (default 100 frames).\n"
                 "  -G SIZE      Sub-divide analysis, buffer will be output and purged every SIZE chunks\n"
                 "  -M           Set if using movie-file format.\n"
                 "  -F FPS       Force the analysis to assume a specific frame-rate, over-rides other options.\n"
                 "  -K           Output non-averaged ISF results.\n"
                 "  -A           Output primary FFT results.\n");
}

struct Options {
    std::string file_in, file_out, q_file, t_file, s_file;
    int frame_count = 0, frame_offset = 0, x_off = 0, y_off = 0;
    int chunk = 100;         // src/main.cpp:23
    int purge_every = 0;
    bool use_webcam = false, use_movie = false, use_index_fps = false, use_explicit_fps = false;
    float explicit_fps = 1.0f;
    bool multi_stream = true;
    float tolerance = 1.2f;  // src/main.cpp:33
    bool benchmark = false, non_ave = false, fft_out = false;
};

template <class T>
std::vector<T> read_values(const std::string &path, const char *what) {
    std::ifstream f(path);
    condition_assert(f.is_open(), std::string("cannot open ") + what, true);
    std::vector<T> v;
    T x;
    while (f >> x) v.push_back(x);
    return v;
}

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

int main(int argc, char **argv) {
    std::printf("DDM Start\n");
    std::fflush(stdout);
    Options o;
    bool input_specified = false;
    for (;;) {
        const int c = getopt(argc, argv, "ho:N:s:x:y:Q:T:S:If:W::vZt:C:MG:F:BKA");
        if (c == -1) break;
        switch (c) {
        case '?':
        case 'h': print_help(); return -1;
        case 'o': o.file_out = optarg; break;
        case 'N': o.frame_count = std::atoi(optarg); break;
        case 'x': o.x_off = std::atoi(optarg); break;
        case 'y': o.y_off = std::atoi(optarg); break;
        case 'Q': o.q_file = optarg; break;
        case 'T': o.t_file = optarg; break;
        case 'S': o.s_file = optarg; break;
        case 'I': o.use_index_fps = true; break;
        case 's': o.frame_offset = std::atoi(optarg); break;
        case 'f':
            condition_assert(!input_specified, "Cannot use both in-filepath and web-cam option at same time.", true);
            o.file_in = optarg;
            input_specified = true;
            break;
        case 'W':
            condition_assert(!input_specified, "Cannot use both in-filepath and web-cam option at same time.", true);
            o.use_webcam = true;
            input_specified = true;
            break;
        case 'B':
            o.benchmark = true;
            input_specified = true;
            break;
        case 'v': g_verbose = true; break;
        case 'Z': o.multi_stream = false; break;
        case 't': o.tolerance = (float)(1.0 + (std::atoi(optarg) / 100.0)); break;  // src/main.cpp:197-198
        case 'C': o.chunk = std::atoi(optarg); break;
        case 'M': o.use_movie = true; break;
        case 'G': o.purge_every = std::atoi(optarg); break;
        case 'F':
            o.use_explicit_fps = true;
            o.explicit_fps = std::strtof(optarg, nullptr);
            break;
        case 'K': o.non_ave = true; break;
        case 'A': o.fft_out = true; break;
        }
    }
    if (optind != argc) {
        print_help();
        condition_assert(false, "An unexpected option was found.", true);
    }
    condition_assert(input_specified, "Must specify input.", true);

    // src/main.cpp:246-313
    std::ifstream qf(o.q_file), tf(o.t_file), sf(o.s_file);
    condition_assert(qf.is_open(), "cannot open lambda-file.", true);
    condition_assert(tf.is_open(), "cannot open tau-file", true);
    condition_assert(sf.is_open(), "cannot open scales-file.", true);
    qf.close();
    tf.close();
    sf.close();
    std::vector<float> lambdas = read_values<float>(o.q_file, "lambda-file.");
    std::vector<int> lags = read_values<int>(o.t_file, "tau-file");
    std::vector<int> scales = read_values<int>(o.s_file, "scales-file.");
    condition_assert(!lambdas.empty(), "lambda-file holds no values.", true);
    condition_assert(!lags.empty(), "tau-file holds no values.", true);
    condition_assert(!scales.empty(), "scales-file holds no values.", true);
    condition_assert(o.frame_count > 0, "Number of frames to analyse (-N) must be positive.", true);
    condition_assert(o.chunk > 0, "Chunk frame count (-C) must be positive.", true);
    condition_assert(o.frame_offset >= 0, "Frame offset (-s) must not be negative.", true);
    condition_assert(o.purge_every >= 0, "Purge interval (-G) must not be negative.", true);

    auto start = std::chrono::high_resolution_clock::now();
    verbose("[multiDDM Begin]\n");

    // video setup (src/DDM.cu:390-452)
    int w = 0, h = 0, bpp = 1, fmt = DDM_PIX_U8;
    float fps = 1.0f;
    ddm::MovieReader movie;
    int max_scale = 0;
    for (int s : scales) max_scale = std::max(max_scale, s);
    if (o.benchmark) {
        w = h = max_scale;
    } else if (o.use_webcam) {
        condition_assert(false, "web-camera input is not available in this build", true);
    } else if (o.use_movie) {
        if (!movie.open(o.file_in, o.frame_offset)) condition_assert(false, movie.error(), true);
        w = movie.info().width;
        h = movie.info().height;
        bpp = movie.info().bpp;
        fmt = movie.info().pixel_format;
        fps = movie.info().fps;
    } else {
        condition_assert(false, "error opening video file with openCV (only .movie input, -M, is available in this build)",
                         true);
    }
    if (o.use_index_fps) fps = 1.0f;
    if (o.use_explicit_fps) fps = o.explicit_fps;

    ddm_desc d{};
    d.width = w;
    d.height = h;
    d.bytes_per_pixel = bpp;
    d.pixel_format = fmt;
    d.x_off = o.x_off;
    d.y_off = o.y_off;
    d.scales = scales.data();
    d.scale_count = (int)scales.size();
    d.lags = lags.data();
    d.lag_count = (int)lags.size();
    d.lambdas = lambdas.data();
    d.lambda_count = (int)lambdas.size();
    d.mask_tolerance = o.tolerance;
    d.chunk_frames = o.chunk;
    d.norm_mode = DDM_NORM_FRAMES;
    d.lag_block = 0;

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
            o.frame_count, o.frame_offset, fps, bpp);

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
    std::vector<uint8_t> host((size_t)o.chunk * frame_bytes);
    const int total_chunks = (o.frame_count + o.chunk - 1) / o.chunk;
    long long done = 0;
    int window = 0;
    for (int ci = 0; ci < total_chunks; ++ci) {
        const int n = std::min<long long>(o.chunk, o.frame_count - done);
        if (o.benchmark) {
            engine_check(ddm_synth_uniform(host.data(), n, w, h, 8, 0x5EEDULL, o.frame_offset + done), nullptr,
                         "synthetic frames");
        } else if (!movie.read(host.data(), n)) {
            condition_assert(false, movie.error(), true);
        }
        engine_check(ddm_submit_host(a.eng, host.data(), n), a.eng, "chunk submission");
        if (o.fft_out) dump_fft(a, o.file_out, ci, done, n);
        done += n;
        verbose("[Chunk complete (%d \\ %d)]\n", ci + 1, total_chunks);
        // -G: output and purge every purge_every chunks (src/DDM.cu:800-826)
        if (o.purge_every > 0 && (ci + 1) % o.purge_every == 0 && ci + 1 < total_chunks) {
            verbose("[Parsing Accumulator]\n");
            emit(a, o.file_out + "_t" + std::to_string(window) + "_", o.non_ave);
            engine_check(ddm_reset(a.eng), a.eng, "purge");
            verbose("[Purging Accumulator]\n");
            ++window;
        }
    }
    engine_check(ddm_sync(a.eng), a.eng, "main loop");
    auto end_main = std::chrono::high_resolution_clock::now();
    verbose("Analysis.\n");
    emit(a, o.file_out, o.non_ave);
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
                (float)(o.frame_count * 1e6) / ((float)d1 + (float)d2));
    std::printf("DDM End\n");
    return 0;
}
