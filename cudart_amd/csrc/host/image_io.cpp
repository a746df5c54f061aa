// Image output: RGBA8 PNG (Launcher::saveToImage, src/context/launcher.cpp:13-21,
// which calls stbi_write_png) and float PFM for radiance. The PNG is written with
// stored (uncompressed) deflate blocks: CRC-32 (ISO 3309 polynomial 0xEDB88320)
// and Adler-32 (RFC 1950) are computed here, so no zlib is needed.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace prt {
namespace {

uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
    return ~crc;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, uint32_t(data.size()));
    size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, crc32_update(0, out.data() + start, out.size() - start));
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_write_png(const char* path, uint32_t w, uint32_t h, const uint8_t* rgba) {
    clear_error();
    if (!path || !rgba || w == 0 || h == 0) return fail(PRT_ERR_INVALID_ARG, "bad argument");
    std::vector<uint8_t> raw;
    raw.reserve((size_t(w) * 4 + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgba + size_t(y) * w * 4, rgba + size_t(y + 1) * w * 4);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    size_t pos = 0;
    do {
        size_t len = std::min<size_t>(65535, raw.size() - pos);
        bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(uint8_t(len));
        z.push_back(uint8_t(len >> 8));
        z.push_back(uint8_t(~len));
        z.push_back(uint8_t(~len >> 8));
        z.insert(z.end(), raw.begin() + long(pos), raw.begin() + long(pos + len));
        pos += len;
    } while (pos < raw.size());
    put32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(PRT_ERR_IO, std::string("cannot write ") + path);
    size_t wr = std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    if (wr != out.size()) return fail(PRT_ERR_IO, std::string("short write to ") + path);
    return PRT_OK;
}

// Portable float map, little-endian ("PF", scale -1), rows bottom to top.
extern "C" int prt_write_pfm(const char* path, uint32_t w, uint32_t h, const float* rgb) {
    clear_error();
    if (!path || !rgb || w == 0 || h == 0) return fail(PRT_ERR_INVALID_ARG, "bad argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(PRT_ERR_IO, std::string("cannot write ") + path);
    std::fprintf(f, "PF\n%u %u\n-1.0\n", w, h);
    bool ok = true;
    for (uint32_t y = h; y-- > 0;) ok = ok && std::fwrite(rgb + size_t(y) * w * 3, sizeof(float), size_t(w) * 3, f) == size_t(w) * 3;
    std::fclose(f);
    if (!ok) return fail(PRT_ERR_IO, std::string("short write to ") + path);
    return PRT_OK;
}
