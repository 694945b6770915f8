// Image output in place of present.comp and the swapchain
// (shaders/glsl/WaveFrontPathTracer/present.comp:19-31, which stores the
// tone-mapped float4 pixels into an rgba8 image): PNG with the same UNORM
// conversion (clamp to [0, 1], x * 255 rounded to nearest), and PFM for the
// linear frame estimate.  The PNG writer uses stored (uncompressed) deflate
// blocks, so it needs no compression library.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "host.h"

namespace wfpt {
namespace {

uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put_be32(out, crc32_update(0u, td.data(), td.size()));
}

// UNORM8 conversion of imageStore into an rgba8 image.
uint8_t unorm8(float x) {
    if (!(x > 0.0f)) return 0;  // NaN and negatives
    if (x >= 1.0f) return 255;
    return (uint8_t)std::lrint(x * 255.0f);
}

int write_file(const char* path, const std::vector<uint8_t>& bytes) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(WF_ERR_IO, std::string("cannot open ") + path);
    const size_t w = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const int c = std::fclose(f);
    if (w != bytes.size() || c != 0) return fail(WF_ERR_IO, std::string("cannot write ") + path);
    return WF_OK;
}

}  // namespace

int encode_png_rgb8(const float* rgba, int width, int height, std::vector<uint8_t>& out) {
    // raw scanlines: filter byte 0, then RGB
    std::vector<uint8_t> raw;
    raw.reserve((size_t)height * (3 * (size_t)width + 1));
    for (int y = 0; y < height; ++y) {
        raw.push_back(0);
        for (int x = 0; x < width; ++x) {
            const float* p = rgba + 4 * ((size_t)y * width + x);
            raw.push_back(unorm8(p[0]));
            raw.push_back(unorm8(p[1]));
            raw.push_back(unorm8(p[2]));
        }
    }
    // zlib stream with stored deflate blocks and the Adler-32 checksum
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    out.assign({0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'});
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, truecolour, deflate, no filter, no interlace
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    return WF_OK;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_write_png(const char* path, const float* host_rgba, int width, int height) {
    WF_CHECK_ARG(path && host_rgba, "wf_write_png: null argument");
    WF_CHECK_ARG(width > 0 && height > 0, "wf_write_png: empty image");
    std::vector<uint8_t> png;
    WF_TRY(encode_png_rgb8(host_rgba, width, height, png));
    return write_file(path, png);
}

extern "C" int wf_write_pfm(const char* path, const float* host_rgba, int width, int height) {
    WF_CHECK_ARG(path && host_rgba, "wf_write_pfm: null argument");
    WF_CHECK_ARG(width > 0 && height > 0, "wf_write_pfm: empty image");
    const std::string head = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1.0\n";
    std::vector<uint8_t> out(head.begin(), head.end());
    // PFM rows run bottom to top, little-endian floats
    for (int y = height - 1; y >= 0; --y)
        for (int x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) {
                const float v = host_rgba[4 * ((size_t)y * width + x) + c];
                uint8_t b[4];
                std::memcpy(b, &v, 4);
                out.insert(out.end(), b, b + 4);
            }
    return write_file(path, out);
}
