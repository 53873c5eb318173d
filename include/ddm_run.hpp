/*
 * ddm_run.hpp -- C++ drop-in for the reference's analysis driver.
 *
 * runDDM has the signature of the reference's runDDM (declared
 * src/main.cpp:40-64, defined src/DDM.cu:268-292) and runs a whole video
 * through the MI355X engine (include/ddm_amd.h): a C++ build that keeps the
 * reference's main.cpp links cuda_multiddm_amd/csrc/ddm_run.cpp and
 * movie_reader.cpp in place of src/DDM.cu, plus libddm_amd.so.  Errors print
 * "[Error] ..." and exit(1), as the reference's conditionAssert does
 * (src/debug.cpp:33-45).
 */
#ifndef DDM_RUN_HPP
#define DDM_RUN_HPP

#include <string>

void runDDM(std::string file_in, std::string file_out, int *tau_arr, int tau_count, float *lambda_arr,
            int lambda_count, int *scale_arr, int scale_count, int x_off, int y_off, int frame_count,
            int frame_offset, int chunk_frame_count, bool multistream, bool use_webcam, int webcam_idx,
            float q_tolerance, bool use_movie_file, bool use_index_fps, bool use_explicit_fps, float explicit_fps,
            int dump_accum_after, bool benchmark_mode, bool nonAveOutput, bool FFTOutput);

namespace ddm_run {
/* -v: progress messages on stdout (the reference's global Verbose). */
void set_verbose(bool on);
/* src/debug.cpp:33-45: "[Error] text" and exit(1) when abort_run, else
 * "[Warning] text". */
void condition_assert(bool cond, const std::string &text, bool abort_run = false);
}  // namespace ddm_run

#endif /* DDM_RUN_HPP */
