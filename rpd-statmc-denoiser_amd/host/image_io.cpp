#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace host {

bool tonemap_from_string(const std::string& name, ToneMapping& out) {
    std::string n = name;
    for (auto& c : n) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (n == "aces") { out = ToneMapping::Aces; return true; }
    if (n == "agx" || n == "agx-default" || n == "agx_default") { out = ToneMapping::AgxDefault; return true; }
    if (n == "agx-golden" || n == "agx_golden") { out = ToneMapping::AgxGolden; return true; }
    if (n == "agx-punchy" || n == "agx_punchy") { out = ToneMapping::AgxPunchy; return true; }
    return false;
}

// ---- tone mapping --------------------------------------------------------------------
// ACES filmic fit (K. Narkowicz 2016) with 0.6 exposure; AgX with the
// minimal-AgX inset/outset matrices and 6th-order contrast polynomial
// (iolite-engine.com "minimal AgX"), as in shared/image.h:29-100.
static float clamp01(float v) { return std::min(1.0f, std::max(0.0f, v)); }

void tonemap_pixel(const float in[3], ToneMapping tm, float out[3]) {
    if (tm == ToneMapping::Aces) {
        for (int k = 0; k < 3; ++k) {
            const float c = 0.6f * in[k];
            out[k] = clamp01((c * (2.51f * c + 0.03f)) / (c * (2.43f * c + 0.59f) + 0.14f));
        }
        return;
    }
    static const float inset[3][3] = {{0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f},
                                      {0.0784335999999992f, 0.878468636469772f, 0.0784336f},
                                      {0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f}};
    static const float outset[3][3] = {{1.19687900512017f, -0.0528968517574562f, -0.0529716355144438f},
                                       {-0.0980208811401368f, 1.15190312990417f, -0.0980434501171241f},
                                       {-0.0990297440797205f, -0.0989611768448433f, 1.15107367264116f}};
    const float min_ev = -12.47393f, max_ev = 4.026069f;
    float col[3];
    for (int r = 0; r < 3; ++r) {
        float v = inset[r][0] * in[0] + (inset[r][1] * in[1] + inset[r][2] * in[2]);
        v = std::max(v, 1e-8f);
        float e = std::log2(std::max(v, 1e-8f));
        e = std::min(max_ev, std::max(min_ev, e));
        col[r] = (e - min_ev) / (max_ev - min_ev);
    }
    for (int k = 0; k < 3; ++k) {
        const float x = col[k], x2 = x * x, x4 = x2 * x2, x6 = x4 * x2;
        col[k] = -17.86f * (x6 * x) + 78.01f * x6 - 126.7f * (x4 * x) + 92.06f * x4 - 28.72f * (x2 * x) +
                 4.361f * x2 - 0.1718f * x + 0.002857f;
    }
    float slope[3] = {1, 1, 1}, power = 1.0f, sat = 1.0f;
    if (tm == ToneMapping::AgxGolden) { slope[0] = 1.0f; slope[1] = 0.9f; slope[2] = 0.5f; power = 0.8f; sat = 0.8f; }
    if (tm == ToneMapping::AgxPunchy) { power = 1.35f; sat = 1.4f; }
    float val[3];
    for (int k = 0; k < 3; ++k) val[k] = std::pow(col[k] * slope[k], power);
    const float luma = 0.2126f * val[0] + (0.7152f * val[1] + 0.0722f * val[2]);
    for (int k = 0; k < 3; ++k) val[k] = luma + sat * (val[k] - luma);
    for (int r = 0; r < 3; ++r)
        out[r] = clamp01(outset[r][0] * val[0] + (outset[r][1] * val[1] + outset[r][2] * val[2]));
}

// ---- Radiance RGBE ---------------------------------------------------------------------
static void to_rgbe(const float* c, unsigned char* e) {
    const float m = std::max(c[0], std::max(c[1], c[2]));
    if (!(m >= 1e-32f)) { e[0] = e[1] = e[2] = e[3] = 0; return; }
    int ex;
    const float f = std::frexp(m, &ex) * 256.0f / m;
    for (int k = 0; k < 3; ++k) e[k] = (unsigned char)std::max(0.0f, std::min(255.0f, c[k] * f));
    e[3] = (unsigned char)(ex + 128);
}

// new-style run-length encoding of one component of a scanline
static void rle_component(FILE* f, const unsigned char* v, int n) {
    int i = 0;
    while (i < n) {
        int run = 1;
        while (i + run < n && run < 127 && v[i + run] == v[i]) ++run;
        if (run >= 3) {
            const unsigned char hdr = (unsigned char)(128 + run);
            fputc(hdr, f);
            fputc(v[i], f);
            i += run;
            continue;
        }
        int lit = 0;  // literal span up to the next run of 3
        while (i + lit < n && lit < 128) {
            if (i + lit + 2 < n && v[i + lit] == v[i + lit + 1] && v[i + lit] == v[i + lit + 2]) break;
            ++lit;
        }
        fputc((unsigned char)lit, f);
        fwrite(v + i, 1, size_t(lit), f);
        i += lit;
    }
}

bool write_hdr(const std::string& path, int w, int h, const float* rgb) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    std::vector<unsigned char> line(size_t(w) * 4);
    std::vector<unsigned char> comp(static_cast<size_t>(w));
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) to_rgbe(rgb + (size_t(y) * w + x) * 3, &line[size_t(x) * 4]);
        if (w < 8 || w > 0x7fff) {
            fwrite(line.data(), 1, line.size(), f);
            continue;
        }
        const unsigned char head[4] = {2, 2, (unsigned char)(w >> 8), (unsigned char)(w & 255)};
        fwrite(head, 1, 4, f);
        for (int c = 0; c < 4; ++c) {
            for (int x = 0; x < w; ++x) comp[size_t(x)] = line[size_t(x) * 4 + c];
            rle_component(f, comp.data(), w);
        }
    }
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0) && ok;
}

bool write_hdr_scalar(const std::string& path, int w, int h, const float* v) {
    std::vector<float> rgb(size_t(w) * h * 3);
    for (size_t i = 0; i < size_t(w) * h; ++i) rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = v[i];
    return write_hdr(path, w, h, rgb.data());
}

// ---- PNG (8-bit RGB, filter 0, zlib) ----------------------------------------------------
static void put_be32(std::vector<unsigned char>& b, uint32_t v) {
    b.push_back((unsigned char)(v >> 24)); b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8)); b.push_back((unsigned char)v);
}
static bool chunk(FILE* f, const char* type, const std::vector<unsigned char>& data) {
    std::vector<unsigned char> b;
    put_be32(b, uint32_t(data.size()));
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    const uLong crc = crc32(crc32(0L, Z_NULL, 0), b.data() + 4, uInt(b.size() - 4));
    put_be32(b, uint32_t(crc));
    return fwrite(b.data(), 1, b.size(), f) == b.size();
}

bool write_png(const std::string& path, int w, int h, const unsigned char* rgb8) {
    std::vector<unsigned char> raw(size_t(h) * (size_t(w) * 3 + 1));
    for (int y = 0; y < h; ++y) {
        raw[size_t(y) * (w * 3 + 1)] = 0;
        std::memcpy(&raw[size_t(y) * (w * 3 + 1) + 1], rgb8 + size_t(y) * w * 3, size_t(w) * 3);
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) return false;
    z.resize(zlen);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    bool ok = fwrite(sig, 1, 8, f) == 8;
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, uint32_t(w));
    put_be32(ihdr, uint32_t(h));
    ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    ok = ok && chunk(f, "IHDR", ihdr) && chunk(f, "IDAT", z) && chunk(f, "IEND", {});
    return (std::fclose(f) == 0) && ok;
}

bool write_png_tonemapped(const std::string& path, int w, int h, const float* rgb, ToneMapping tm) {
    std::vector<unsigned char> px(size_t(w) * h * 3);
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        float o[3];
        tonemap_pixel(rgb + 3 * i, tm, o);
        for (int k = 0; k < 3; ++k) px[3 * i + k] = (unsigned char)(255.0f * std::max(0.0f, std::min(1.0f, o[k])));
    }
    return write_png(path, w, h, px.data());
}

}  // namespace host
