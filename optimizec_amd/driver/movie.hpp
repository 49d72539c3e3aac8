// movie.hpp -- MOV movie frames of the reference driver (color_cuda.cu:19-597,
// mem_cuda.cu:245-332, 1834-1910): the physical vorticity of every diagnostic step that is a
// multiple of UPDATE, sampled every F-th point, mapped to 8-bit grey as in the PGM dumps,
// stretched to mean +- 2.5 sigma of those grey levels, coloured with a five-stop inferno
// ramp and written as an RGB PNG (stored deflate blocks, no compression).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace movie {

// dnsCudaMovieScaleF (color_cuda.cu:486-493): frames at most ~3840 pixels wide
inline int scale_f(int N)
{
    if (N == 9216) return 8;
    if (N == 16384) return 12;
    if (N == 29400) return 14;
    return std::max(1, (3 * N / 2 + 3839) / 3840);
}

// 256-entry RGB ramp through the five stops of the reference (color_cuda.cu:40-70):
// linear in between, truncated to bytes
inline std::vector<unsigned char> inferno_lut()
{
    struct Stop {
        float pos;
        unsigned char r, g, b;
    };
    static const Stop stops[5] = {
        {0.00f, 0, 0, 4}, {0.25f, 87, 15, 109}, {0.50f, 187, 55, 84}, {0.75f, 249, 142, 8}, {1.00f, 252, 255, 164}};
    std::vector<unsigned char> lut(256 * 3);
    for (int i = 0; i < 256; ++i) {
        const float x = (float)i / 255.0f;
        int s = 0;
        while (s + 1 < 5 && x > stops[s + 1].pos) ++s;
        const Stop &a = stops[s], &b = stops[std::min(s + 1, 4)];
        const float span = std::max(1.0e-12f, b.pos - a.pos);
        const float t = std::max(0.0f, std::min(1.0f, (x - a.pos) / span));
        lut[3 * i + 0] = (unsigned char)((1.0f - t) * a.r + t * b.r);
        lut[3 * i + 1] = (unsigned char)((1.0f - t) * a.g + t * b.g);
        lut[3 * i + 2] = (unsigned char)((1.0f - t) * a.b + t * b.b);
    }
    return lut;
}

// grey level of one value: floor(1 + 254 (v - min) / range) clamped to [1, 255]; 128 for a
// constant field or a non-finite value
inline int grey(double v, double mn, double inv_range, bool constant)
{
    if (constant || !std::isfinite(v)) return 128;
    double p = 1.0 + (v - mn) * inv_range * 254.0;
    p = std::min(255.0, std::max(1.0, p));
    return (int)p;
}

// RGB frame (w x h x 3) of an M x M plane [z][x]; sigma_out receives the grey-level sigma
inline std::vector<unsigned char> render(const std::vector<double> &f, int M, int F, int *w_out, int *h_out,
                                         double *sigma_out)
{
    const int w = (M + F - 1) / F, h = (M + F - 1) / F;
    double mn = 1.0e300, mx = -1.0e300;
    for (double v : f)
        if (std::isfinite(v)) {
            mn = std::min(mn, v);
            mx = std::max(mx, v);
        }
    if (!(mx >= mn)) mn = mx = 0.0;
    const double range = mx - mn;
    const bool constant = !(range > 1.0e-12);
    const double inv = constant ? 1.0 : 1.0 / range;
    double sum = 0.0, sum2 = 0.0;
    for (double v : f) {
        const double p = grey(v, mn, inv, constant);
        sum += p;
        sum2 += p * p;
    }
    const double n = (double)f.size();
    const double mean = sum / n, sigma = std::sqrt(std::max(0.0, sum2 / n - mean * mean));
    const double lo = mean - 2.5 * sigma, hi = mean + 2.5 * sigma;
    static const std::vector<unsigned char> lut = inferno_lut();
    std::vector<unsigned char> rgb((size_t)w * h * 3);
    for (int oy = 0; oy < h; ++oy)
        for (int ox = 0; ox < w; ++ox) {
            const int x = std::min(ox * F, M - 1), y = std::min(oy * F, M - 1);
            const int pix = grey(f[(size_t)y * M + x], mn, inv, constant);
            int c = pix;
            if (sigma >= 1.0 && hi > lo) c = std::min(255, std::max(0, (int)std::floor((pix - lo) * 255.0 / (hi - lo) + 0.5)));
            unsigned char *d = &rgb[3 * ((size_t)oy * w + ox)];
            d[0] = lut[3 * c];
            d[1] = lut[3 * c + 1];
            d[2] = lut[3 * c + 2];
        }
    *w_out = w;
    *h_out = h;
    *sigma_out = sigma;
    return rgb;
}

inline uint32_t crc32(uint32_t crc, const unsigned char *p, size_t n)
{
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) {
        crc ^= p[i];
        for (int k = 0; k < 8; ++k) crc = (crc & 1u) ? (0xedb88320u ^ (crc >> 1)) : (crc >> 1);
    }
    return ~crc;
}

inline void put_be32(std::vector<unsigned char> &v, uint32_t x)
{
    for (int s = 24; s >= 0; s -= 8) v.push_back((unsigned char)(x >> s));
}

// PNG, 8-bit RGB, filter 0 on every row, zlib stream of stored blocks (<= 65535 bytes each)
inline bool write_png(const std::string &path, int w, int h, const std::vector<unsigned char> &rgb)
{
    std::vector<unsigned char> raw;
    raw.reserve(((size_t)w * 3 + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + (size_t)y * w * 3, rgb.begin() + (size_t)(y + 1) * w * 3);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0;) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char)(n & 255));
        z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 255));
        z.push_back((unsigned char)((~n >> 8) & 255));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        off += n;
        if (last) break;
    }
    uint32_t a = 1, b = 0;  // Adler-32 of the uncompressed data
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        std::perror("writeRgbPng: fopen");
        return false;
    }
    auto chunk = [&](const char *type, const unsigned char *data, size_t n) {
        std::vector<unsigned char> c;
        put_be32(c, (uint32_t)n);
        c.insert(c.end(), type, type + 4);
        if (n) c.insert(c.end(), data, data + n);
        put_be32(c, crc32(0, c.data() + 4, n + 4));
        return std::fwrite(c.data(), 1, c.size(), fp) == c.size();
    };
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    bool ok = std::fwrite(sig, 1, 8, fp) == 8 && chunk("IHDR", ihdr.data(), ihdr.size()) &&
              chunk("IDAT", z.data(), z.size()) && chunk("IEND", nullptr, 0);
    if (std::fclose(fp) != 0) ok = false;
    return ok;
}

}  // namespace movie
