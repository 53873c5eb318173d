// Reader for the camera ".movie" container the reference's -M option reads
// (src/video_reader.hpp:13-214, src/video_reader.cpp:43-290): a sequence of
// frames, each a little-endian header (IIDC 172 B, Andor 128 B, Ximea 240 B,
// all starting with magic 0x496d6554 'TemI', version 1) followed by
// length_data bytes of MONO8 / MONO16BE / MONO16LE pixels.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

namespace ddm {

struct MovieInfo {
    int width = 0, height = 0;
    int bpp = 1;              // bytes per pixel
    int pixel_format = 0;     // DDM_PIX_*
    float fps = 1.0f;
    uint32_t camera_type = 0; // 1 IIDC, 2 Andor, 3 Ximea
    uint32_t header_len = 0;
    uint32_t data_len = 0;
    long first_offset = 0;    // file offset of the first frame to read
};

class MovieReader {
  public:
    ~MovieReader();
    // Opens the file, parses the first two frame headers (size, pixel mode,
    // frame rate) and seeks to frame `frame_offset`.  Returns false with a
    // message in error() on failure.
    bool open(const std::string &path, int frame_offset);
    // Reads n frames of width*height*bpp bytes into out.  False on a short
    // or corrupted frame.
    bool read(uint8_t *out, int n);
    const MovieInfo &info() const { return info_; }
    const std::string &error() const { return err_; }

  private:
    FILE *f_ = nullptr;
    MovieInfo info_;
    std::string err_;
};

}  // namespace ddm
