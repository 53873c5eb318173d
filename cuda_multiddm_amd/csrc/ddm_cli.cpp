// ./ddm -- command-line front end of the MI355X DDM engine, flag- and
// file-compatible with the reference CLI (src/main.cpp:104-342):
//   -o PATH -N INT -Q PATH -T PATH -S PATH   required
//   -f PATH [-M] | -W[INT] | -B               input (.movie via -M; -B synthetic)
//   -s -x -y -I -v -Z -t -C -G -F -K -A -h    options
// It parses the flags and parameter files as the reference's main does and
// calls runDDM (ddm_run.cpp, the reference's signature), which writes one
// file <prefix><scale>-<tile> per (scale, tile).  Differences from the
// reference, all deliberate: -B and -K take no argument (the reference's
// optstring makes them swallow one, src/main.cpp:115); -F sets the frame
// rate (a no-op in the reference, src/main.cpp:214-218).
#include <getopt.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "ddm_amd.h"
#include "ddm_run.hpp"

namespace {

bool g_verbose = false;
using ddm_run::condition_assert;

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

    ddm_run::set_verbose(g_verbose);
    runDDM(o.file_in, o.file_out, lags.data(), (int)lags.size(), lambdas.data(), (int)lambdas.size(), scales.data(),
           (int)scales.size(), o.x_off, o.y_off, o.frame_count, o.frame_offset, o.chunk, o.multi_stream,
           o.use_webcam, 0, o.tolerance, o.use_movie, o.use_index_fps, o.use_explicit_fps, o.explicit_fps,
           o.purge_every, o.benchmark, o.non_ave, o.fft_out);
    std::printf("DDM End\n");
    return 0;
}
