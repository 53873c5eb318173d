// .movie container reader; format notes in movie_reader.h.
#include "movie_reader.h"

#include <cstring>
#include <vector>

#include "ddm_amd.h"

namespace ddm {

namespace {

constexpr uint32_t kMagic = 0x496d6554;  // 'TemI'
constexpr uint32_t kVersion = 1;
constexpr uint32_t kTypeIIDC = 1, kTypeAndor = 2, kTypeXimea = 3;
constexpr uint32_t kMono8 = 1, kMono16BE = 2, kMono16LE = 3;

uint32_t rd32(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
uint64_t rd64(const uint8_t *p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

uint32_t header_length(uint32_t type) {
    switch (type) {
    case kTypeIIDC: return 172;
    case kTypeAndor: return 128;
    case kTypeXimea: return 240;
    default: return 0;
    }
}

// Frame size and timestamp (microseconds) from one camera header
// (field offsets of the packed structs in src/video_reader.hpp:48-214).
bool parse_header(uint32_t type, const uint8_t *h, uint32_t *w, uint32_t *hgt, double *t_us) {
    switch (type) {
    case kTypeIIDC:
        *w = rd32(h + 64);
        *hgt = rd32(h + 68);
        *t_us = (double)rd64(h + 48);
        return true;
    case kTypeAndor: {
        int32_t xs = (int32_t)rd32(h + 48), xe = (int32_t)rd32(h + 52);
        int32_t ys = (int32_t)rd32(h + 56), ye = (int32_t)rd32(h + 60);
        int32_t xb = (int32_t)rd32(h + 64), yb = (int32_t)rd32(h + 68);
        if (xb <= 0 || yb <= 0) return false;
        *w = (uint32_t)((xe - xs + 1) / xb);
        *hgt = (uint32_t)((ye - ys + 1) / yb);
        // seconds + nanoseconds (the reference uses whole seconds only and
        // divides by zero for frame rates above 1 Hz, src/video_reader.cpp:158)
        *t_us = (double)rd64(h + 24) * 1e6 + (double)rd64(h + 32) * 1e-3;
        return true;
    }
    case kTypeXimea:
        *w = rd32(h + 152);
        *hgt = rd32(h + 156);
        *t_us = (double)rd64(h + 128) * 1e6 + (double)rd64(h + 136) * 1e-3;
        return true;
    default: return false;
    }
}

}  // namespace

MovieReader::~MovieReader() {
    if (f_) std::fclose(f_);
}

bool MovieReader::open(const std::string &path, int frame_offset) {
    f_ = std::fopen(path.c_str(), "rb");
    if (!f_) {
        err_ = "couldn't open .movie file";
        return false;
    }
    // locate the first header by its magic value, byte by byte (src/video_reader.cpp:66-79)
    long magic_off = 0;
    bool found = false;
    uint8_t b4[4];
    while (std::fseek(f_, magic_off, SEEK_SET) == 0 && std::fread(b4, 4, 1, f_) == 1) {
        if (rd32(b4) == kMagic) {
            found = true;
            break;
        }
        ++magic_off;
    }
    if (!found) {
        err_ = "[.moviefile Initialise Error] Wrong magic at offset " + std::to_string(magic_off);
        return false;
    }
    uint8_t common[24];
    std::fseek(f_, magic_off, SEEK_SET);
    if (std::fread(common, sizeof(common), 1, f_) != 1) {
        err_ = "[.moviefile Initialise Error] Corrupted header at offset " + std::to_string(magic_off);
        return false;
    }
    if (rd32(common + 4) != kVersion) {
        err_ = "[.moviefile Initialise Error] Unsupported movie version " + std::to_string(rd32(common + 4));
        return false;
    }
    const uint32_t type = rd32(common + 8), pixmode = rd32(common + 12), data_len = rd32(common + 20);
    const uint32_t hl = header_length(type);
    if (!hl) {
        err_ = "[Movie-file Init. Error] Unsupported camera type.";
        return false;
    }
    std::vector<uint8_t> h0(hl), h1(hl);
    std::fseek(f_, magic_off, SEEK_SET);
    uint32_t w = 0, hh = 0, w1 = 0, hh1 = 0;
    double t0 = 0, t1 = 0;
    if (std::fread(h0.data(), hl, 1, f_) != 1 || !parse_header(type, h0.data(), &w, &hh, &t0)) {
        err_ = "[Movie-file Init. Error] Corrupted header at offset " + std::to_string(std::ftell(f_));
        return false;
    }
    // second frame's header for the frame rate (src/video_reader.cpp:128-137)
    std::fseek(f_, (long)data_len, SEEK_CUR);
    bool have_second = std::fread(h1.data(), hl, 1, f_) == 1 && parse_header(type, h1.data(), &w1, &hh1, &t1);
    info_.width = (int)w;
    info_.height = (int)hh;
    info_.camera_type = type;
    info_.header_len = hl;
    info_.data_len = data_len;
    info_.bpp = (pixmode == kMono8) ? 1 : 2;
    info_.pixel_format = pixmode == kMono8 ? DDM_PIX_U8 : (pixmode == kMono16BE ? DDM_PIX_U16BE : DDM_PIX_U16LE);
    info_.fps = (have_second && t1 > t0) ? (float)(1.0e6 / (t1 - t0)) : 1.0f;
    if ((uint64_t)w * hh * info_.bpp > data_len || w == 0 || hh == 0) {
        err_ = "[Movie-file Init. Error] frame size does not fit the frame data length";
        return false;
    }
    info_.first_offset = magic_off + (long)frame_offset * (long)(hl + data_len);
    if (std::fseek(f_, info_.first_offset, SEEK_SET) != 0) {
        err_ = "[Movie-file Init. Error] cannot seek to the first frame";
        return false;
    }
    return true;
}

bool MovieReader::read(uint8_t *out, int n) {
    std::vector<uint8_t> hdr(info_.header_len), data(info_.data_len);
    const size_t frame_bytes = (size_t)info_.width * info_.height * info_.bpp;
    for (int i = 0; i < n; ++i) {
        if (std::fread(hdr.data(), info_.header_len, 1, f_) != 1 || rd32(hdr.data()) != kMagic) {
            err_ = "[.moviefile Read Error] Corrupted header at offset " + std::to_string(std::ftell(f_));
            return false;
        }
        if (std::fread(data.data(), info_.data_len, 1, f_) != 1) {
            err_ = "[.moviefile Read Error] Corrupted data at offset " + std::to_string(std::ftell(f_));
            return false;
        }
        std::memcpy(out + (size_t)i * frame_bytes, data.data(), frame_bytes);
    }
    return true;
}

}  // namespace ddm
