// png.cpp — PNG output of a rendered frame: what the reference's viewer does
// with RayTracer::output on screen (main.cpp:207-208 uploads it as an RGB
// texture, row 0 at the bottom), written to a file instead.
//
// The file is an 8-bit RGB PNG (ISO/IEC 15948): IHDR, one IDAT holding a zlib
// stream (RFC 1950) of stored deflate blocks (RFC 1951 block type 0, no
// compression), IEND.  Every scanline uses filter type 0.  CRC-32 (the PNG
// polynomial 0xEDB88320, reflected) covers each chunk's type and data;
// Adler-32 covers the uncompressed stream.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rtg_host.h"

namespace {

uint32_t crc_table[256];
bool crc_ready = false;

void crc_init() {
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
    crc_ready = true;
}

uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24)), v.push_back((uint8_t)(x >> 16)), v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t> &out, const char type[4], const std::vector<uint8_t> &data) {
    put_be32(out, (uint32_t)data.size());
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32(out.data() + at, 4 + data.size()) ^ 0xFFFFFFFFu);
}

}  // namespace

extern "C" int rtg_host_write_png(const char *path, const uint8_t *rgb, int width, int height, int stride,
                                  int bottom_up) {
    if (path == nullptr || rgb == nullptr || width <= 0 || height <= 0 || width > 65535 || height > 65535)
        return -1;
    if (stride <= 0) stride = 3 * width;
    if (stride < 3 * width) return -1;
    if (!crc_ready) crc_init();
    // raw scanlines: filter byte 0 + RGB
    const size_t row = 1 + 3 * (size_t)width;
    std::vector<uint8_t> raw(row * height);
    for (int y = 0; y < height; ++y) {
        const int src = bottom_up ? height - 1 - y : y;
        raw[row * y] = 0;
        std::memcpy(&raw[row * y + 1], rgb + (size_t)src * stride, 3 * (size_t)width);
    }
    // zlib: CMF/FLG (deflate, 32K window, no dictionary, check bits), stored blocks, Adler-32
    std::vector<uint8_t> z;
    z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
    z.push_back(0x78), z.push_back(0x01);
    for (size_t off = 0;;) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)len), z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len), z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        off += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (const uint8_t c : raw) a = (a + c) % 65521u, b = (b + a) % 65521u;
    put_be32(z, (b << 16) | a);

    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width), put_be32(ihdr, (uint32_t)height);
    ihdr.push_back(8), ihdr.push_back(2), ihdr.push_back(0), ihdr.push_back(0), ihdr.push_back(0);
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE *f = std::fopen(path, "wb");
    if (f == nullptr) return -2;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return std::fclose(f) == 0 && ok ? 0 : -2;
}
