// Image output: linear output-space RGB -> sRGB 8-bit RGBA -> PNG.
//
// Replaces Film::write_to_png (base/film.cu:142-181) with SRGBColorEncoding::from_linear
// (spectrum_util/color_encoding.h:24-60): the transfer curve is the rational
// minimax fit of enoki's color.h that the reference evaluates, so 8-bit values
// match it exactly. PNG bytes come from zlib rather than lodepng; the decoded
// pixels are what is specified.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <zlib.h>

#include "../../include/pbrtm.h"

namespace {

float horner(float t, const float *c, int n) {  // evaluate_polynomial (basic_math.h:110-117)
    float acc = c[n - 1];
    for (int i = n - 2; i >= 0; --i) acc = std::fma(t, acc, c[i]);
    return acc;
}

float linear_to_srgb(float v) {
    if (v <= 0.0031308f) return 12.92f * v;
    static const float P[6] = {-0.0016829072605308378f, 0.03453868659826638f, 0.7642611304733891f,
                               2.0041169284241644f,     0.7551545191665577f,  -0.016202083165206348f};
    static const float Q[6] = {4.178892964897981e-7f, -0.00004375359692957097f, 0.03467195408529984f,
                               0.6085338522168684f,   1.8970238036421054f,      1.f};
    const float s = std::sqrt(v > 0.0f ? v : 0.0f);
    return horner(s, P, 6) / horner(s, Q, 6) * v;
}

uint8_t encode(float v) {
    if (v <= 0) return 0;
    if (v >= 1) return 255;
    if (std::isnan(v)) return 0;
    const double r = std::round(255.0 * double(linear_to_srgb(v)));
    return uint8_t(r < 0 ? 0 : (r > 255 ? 255 : r));
}

void put_be32(std::vector<uint8_t> &o, uint32_t v) {
    o.push_back(uint8_t(v >> 24));
    o.push_back(uint8_t(v >> 16));
    o.push_back(uint8_t(v >> 8));
    o.push_back(uint8_t(v));
}

void chunk(std::vector<uint8_t> &o, const char *type, const uint8_t *data, size_t n) {
    put_be32(o, uint32_t(n));
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data, data + n);
    put_be32(o, uint32_t(crc32(0L, o.data() + start, uInt(n + 4))));
}

}  // namespace

extern "C" {

int pbrtm_srgb8(const float *rgb, uint64_t n, uint8_t *rgba_out) {
    if (!rgb || !rgba_out) return PBRTM_ERR_INVALID;
    for (uint64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) rgba_out[4 * i + c] = encode(rgb[3 * i + c]);
        rgba_out[4 * i + 3] = 255;
    }
    return PBRTM_OK;
}

int pbrtm_write_png(const float *rgb, uint32_t width, uint32_t height, const char *path) {
    if (!rgb || !path || width == 0 || height == 0) return PBRTM_ERR_INVALID;
    const size_t row = size_t(width) * 4 + 1;
    std::vector<uint8_t> raw(row * height);
    for (uint32_t y = 0; y < height; ++y) {
        raw[y * row] = 0;  // filter type None
        pbrtm_srgb8(rgb + size_t(y) * width * 3, width, raw.data() + y * row + 1);
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) return PBRTM_ERR_NOMEM;
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    const uint8_t rest[5] = {8, 6, 0, 0, 0};  // 8-bit RGBA, deflate, adaptive filtering, no interlace
    ihdr.insert(ihdr.end(), rest, rest + 5);
    chunk(png, "IHDR", ihdr.data(), ihdr.size());
    chunk(png, "IDAT", z.data(), zlen);
    chunk(png, "IEND", nullptr, 0);
    FILE *f = std::fopen(path, "wb");
    if (!f) return PBRTM_ERR_INVALID;
    const size_t w = std::fwrite(png.data(), 1, png.size(), f);
    const int rc = std::fclose(f);
    return (w == png.size() && rc == 0) ? PBRTM_OK : PBRTM_ERR_INVALID;
}

int pbrtm_write_pfm(const float *rgb, uint32_t width, uint32_t height, const char *path) {
    if (!rgb || !path || width == 0 || height == 0) return PBRTM_ERR_INVALID;
    FILE *f = std::fopen(path, "wb");
    if (!f) return PBRTM_ERR_INVALID;
    // "PF": three channels; a negative scale marks little-endian floats; rows bottom-up
    bool ok = std::fprintf(f, "PF\n%u %u\n-1.0\n", width, height) > 0;
    for (uint32_t y = height; ok && y-- > 0;)
        ok = std::fwrite(rgb + size_t(y) * width * 3, sizeof(float), size_t(width) * 3, f) == size_t(width) * 3;
    const int rc = std::fclose(f);
    return (ok && rc == 0) ? PBRTM_OK : PBRTM_ERR_INVALID;
}

}  // extern "C"
