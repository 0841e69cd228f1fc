// image_io.cpp — headless image output standing in for the GL read-back +
// stb_image_write of RendererCore::saveImage (src/RendererCore.cpp:608-646).
//   .hdr  Radiance RGBE, flat scanlines (a valid encoding of the format)
//   .png  8-bit RGB; deflate "stored" blocks, CRC-32 and Adler-32 per RFC 1950/1951/PNG
// Both take rows top-down (the caller applies the reference's vertical flip).
#include "yn_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace yune {

int write_hdr(const std::string& path, int w, int h, const float* rgb)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return -1;
    std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    std::vector<unsigned char> row((size_t)w * 4);
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const float* p = rgb + ((size_t)y * w + x) * 3;
            float r = p[0] > 0 ? p[0] : 0, g = p[1] > 0 ? p[1] : 0, b = p[2] > 0 ? p[2] : 0;
            float m = std::fmax(r, std::fmax(g, b));
            unsigned char* o = &row[(size_t)x * 4];
            if (!(m >= 1e-32f) || !std::isfinite(m)) {
                if (std::isfinite(m)) { o[0] = o[1] = o[2] = o[3] = 0; continue; }
                o[0] = o[1] = o[2] = 255; o[3] = 255; continue;   // non-finite: brightest code
            }
            int e;
            float s = std::frexp(m, &e) * 256.0f / m;
            o[0] = (unsigned char)(r * s); o[1] = (unsigned char)(g * s); o[2] = (unsigned char)(b * s);
            o[3] = (unsigned char)(e + 128);
        }
        if (std::fwrite(row.data(), 1, row.size(), f) != row.size()) { std::fclose(f); return -1; }
    }
    return std::fclose(f) == 0 ? 0 : -1;
}

namespace {

uint32_t crc_table[256];
bool crc_init = false;

uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xffffffffu)
{
    if (!crc_init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t k = i;
            for (int j = 0; j < 8; ++j) k = (k & 1) ? 0xedb88320u ^ (k >> 1) : k >> 1;
            crc_table[i] = k;
        }
        crc_init = true;
    }
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void be32(std::vector<unsigned char>& v, uint32_t x)
{
    v.push_back((unsigned char)(x >> 24)); v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8)); v.push_back((unsigned char)x);
}

void chunk(std::vector<unsigned char>& out, const char* type, const std::vector<unsigned char>& data)
{
    be32(out, (uint32_t)data.size());
    size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    be32(out, crc32(&out[start], out.size() - start) ^ 0xffffffffu);
}

}  // namespace

int write_png(const std::string& path, int w, int h, const uint8_t* rgb)
{
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (w * 3 + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);   // filter: none
        raw.insert(raw.end(), rgb + (size_t)y * w * 3, rgb + (size_t)(y + 1) * w * 3);
    }
    std::vector<unsigned char> z;
    z.push_back(0x78); z.push_back(0x01);
    size_t pos = 0;
    do {
        size_t n = std::min<size_t>(65535, raw.size() - pos);
        bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char)(n & 0xff)); z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 0xff)); z.push_back((unsigned char)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char ch : raw) { a = (a + ch) % 65521u; b = (b + a) % 65521u; }
    be32(z, (b << 16) | a);
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    be32(ihdr, (uint32_t)w); be32(ihdr, (uint32_t)h);
    ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return -1;
    bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

}  // namespace yune
