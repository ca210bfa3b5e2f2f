// Minimal FITS primary-HDU image I/O (cfitsio is not part of this build).
//
// Writing follows what the reference asks of cfitsio for its dirty image (write_fits_2d,
// SVDIntegratedImager.cu:153-176; write_fits_image, archive/pipe_interface.cpp:98-137):
// a primary HDU with BITPIX = -32, NAXIS = 2, NAXIS1 = NAXIS2 = N, big-endian float data in
// memory order, padded to 2880-byte records, replacing an existing file.
// Reading covers what read_fits_image (archive/pipe_interface.cpp:12-96) accepts: 1-D or 2-D
// primary images of any standard BITPIX, converted to float with BSCALE / BZERO applied.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "svdimager.h"

namespace {

constexpr size_t kBlock = 2880;

std::string card(const std::string& key, const std::string& value, const std::string& comment) {
    char buf[81];
    if (comment.empty()) std::snprintf(buf, sizeof buf, "%-8s= %20s", key.c_str(), value.c_str());
    else std::snprintf(buf, sizeof buf, "%-8s= %20s / %s", key.c_str(), value.c_str(), comment.c_str());
    std::string s(buf);
    s.resize(80, ' ');
    return s;
}

std::string raw_card(const std::string& text) {
    std::string s = text;
    s.resize(80, ' ');
    return s;
}

bool little_endian() {
    const uint16_t x = 1;
    return *reinterpret_cast<const uint8_t*>(&x) == 1;
}

void swap_bytes(uint8_t* p, size_t width, size_t count) {
    if (!little_endian()) return;
    for (size_t i = 0; i < count; ++i) {
        uint8_t* e = p + i * width;
        for (size_t a = 0, b = width - 1; a < b; ++a, --b) {
            const uint8_t t = e[a];
            e[a] = e[b];
            e[b] = t;
        }
    }
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && s[a] == ' ') ++a;
    while (b > a && s[b - 1] == ' ') --b;
    return s.substr(a, b - a);
}

}  // namespace

extern "C" {

int svdi_fits_write(const char* path, const float* data, long naxis1, long naxis2) {
    if (!path || !data || naxis1 <= 0 || naxis2 <= 0) return SVDI_ERR_INVALID_ARGUMENT;
    std::string header;
    header += card("SIMPLE", "T", "file does conform to FITS standard");
    header += card("BITPIX", "-32", "number of bits per data pixel");
    header += card("NAXIS", "2", "number of data axes");
    header += card("NAXIS1", std::to_string(naxis1), "length of data axis 1");
    header += card("NAXIS2", std::to_string(naxis2), "length of data axis 2");
    header += card("EXTEND", "T", "FITS dataset may contain extensions");
    header += raw_card("COMMENT   FITS (Flexible Image Transport System) format is defined in 'Astronomy");
    header += raw_card("COMMENT   and Astrophysics', volume 376, page 359; bibcode: 2001A&A...376..359H");
    header += raw_card("END");
    header.resize((header.size() + kBlock - 1) / kBlock * kBlock, ' ');
    const size_t n = static_cast<size_t>(naxis1) * static_cast<size_t>(naxis2);
    std::vector<uint8_t> body(n * 4);
    std::memcpy(body.data(), data, n * 4);
    swap_bytes(body.data(), 4, n);
    body.resize((body.size() + kBlock - 1) / kBlock * kBlock, 0);
    std::remove(path);                                  // overwrite, as cfitsio's "!" prefix does
    FILE* f = std::fopen(path, "wb");
    if (!f) return SVDI_ERR_IO;
    const bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size() &&
                    std::fwrite(body.data(), 1, body.size(), f) == body.size();
    const bool closed = std::fclose(f) == 0;
    return (ok && closed) ? SVDI_OK : SVDI_ERR_IO;
}

int svdi_fits_read(const char* path, float** out, long naxes[2]) {
    if (!path || !out || !naxes) return SVDI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SVDI_ERR_IO;
    long bitpix = 0, naxis = -1, n1 = 0, n2 = 1;
    double bscale = 1.0, bzero = 0.0;
    bool simple = false, end = false;
    size_t header_bytes = 0;
    char rec[kBlock];
    while (!end) {
        if (std::fread(rec, 1, kBlock, f) != kBlock) {
            std::fclose(f);
            return SVDI_ERR_IO;
        }
        header_bytes += kBlock;
        for (size_t c = 0; c < kBlock / 80 && !end; ++c) {
            const std::string cardtxt(rec + c * 80, 80);
            const std::string key = trim(cardtxt.substr(0, 8));
            if (key == "END") {
                end = true;
                break;
            }
            if (cardtxt.size() < 10 || cardtxt[8] != '=') continue;
            std::string val = cardtxt.substr(10);
            const size_t slash = val.find('/');
            if (slash != std::string::npos && val.find('\'') == std::string::npos) val = val.substr(0, slash);
            val = trim(val);
            if (key == "SIMPLE") simple = (val == "T");
            else if (key == "BITPIX") bitpix = std::strtol(val.c_str(), nullptr, 10);
            else if (key == "NAXIS") naxis = std::strtol(val.c_str(), nullptr, 10);
            else if (key == "NAXIS1") n1 = std::strtol(val.c_str(), nullptr, 10);
            else if (key == "NAXIS2") n2 = std::strtol(val.c_str(), nullptr, 10);
            else if (key == "BSCALE") bscale = std::strtod(val.c_str(), nullptr);
            else if (key == "BZERO") bzero = std::strtod(val.c_str(), nullptr);
        }
        if (header_bytes > 1000 * kBlock) break;
    }
    if (!simple || !end || (naxis != 1 && naxis != 2) || n1 <= 0 || n2 <= 0) {
        std::fclose(f);
        return SVDI_ERR_IO;
    }
    if (naxis == 1) n2 = 1;
    const size_t count = static_cast<size_t>(n1) * static_cast<size_t>(n2);
    size_t width;
    switch (bitpix) {
        case 8: width = 1; break;
        case 16: width = 2; break;
        case 32: case -32: width = 4; break;
        case 64: case -64: width = 8; break;
        default:
            std::fclose(f);
            return SVDI_ERR_IO;
    }
    std::vector<uint8_t> raw(count * width);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) {
        std::fclose(f);
        return SVDI_ERR_IO;
    }
    std::fclose(f);
    swap_bytes(raw.data(), width, count);
    float* data = static_cast<float*>(std::malloc(count * sizeof(float)));
    if (!data) return SVDI_ERR_OUT_OF_MEMORY;
    for (size_t i = 0; i < count; ++i) {
        double v;
        const uint8_t* p = raw.data() + i * width;
        switch (bitpix) {
            case 8: v = *p; break;
            case 16: { int16_t t; std::memcpy(&t, p, 2); v = t; break; }
            case 32: { int32_t t; std::memcpy(&t, p, 4); v = t; break; }
            case 64: { int64_t t; std::memcpy(&t, p, 8); v = static_cast<double>(t); break; }
            case -32: { float t; std::memcpy(&t, p, 4); v = t; break; }
            default: { double t; std::memcpy(&t, p, 8); v = t; break; }
        }
        data[i] = static_cast<float>(bzero + bscale * v);
    }
    *out = data;
    naxes[0] = n1;
    naxes[1] = n2;
    return SVDI_OK;
}

}  // extern "C"
