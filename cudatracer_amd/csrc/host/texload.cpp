// TriangleList::loadTexture / core::Surface (TriangleList.cpp:377-391,
// src/Utils/Surface.cpp:33-100) without FreeImage: TGA (the format of the
// reference's Models/sponza textures), binary PPM (P6) and PFM.
//
// The reference's Surface keeps a float RGBA copy of the image in FreeImage's
// scan-line order, whose row 0 is the bottom of the picture: 8-bit images
// become colour / 255 with alpha 1 (after FreeImage_ConvertTo32Bits), float
// images keep their values.  PPM stores rows top-down, so they are flipped;
// PFM stores them bottom-up already.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

#include "scene.h"

namespace ct {
namespace {

// Next header token of a PNM/PFM file (whitespace separated, '#' comments).
bool next_token(std::istream& in, std::string& tok) {
    tok.clear();
    int c;
    while ((c = in.get()) != EOF) {
        if (c == '#') {
            while ((c = in.get()) != EOF && c != '\n') {
            }
            continue;
        }
        if (!std::isspace(c)) {
            tok.push_back((char)c);
            break;
        }
    }
    while ((c = in.peek()) != EOF && !std::isspace(c)) tok.push_back((char)in.get());
    return !tok.empty();
}

bool to_uint(const std::string& s, uint32_t& v) {
    if (s.empty() || s.size() > 9) return false;
    for (char ch : s)
        if (ch < '0' || ch > '9') return false;
    v = (uint32_t)std::stoul(s);
    return true;
}

bool load_ppm(std::istream& in, Texture& t, std::string& why) {
    std::string tok;
    uint32_t w = 0, h = 0, maxval = 0;
    if (!next_token(in, tok) || !to_uint(tok, w) || !next_token(in, tok) || !to_uint(tok, h) ||
        !next_token(in, tok) || !to_uint(tok, maxval)) {
        why = "bad PPM header";
        return false;
    }
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28)) {
        why = "bad PPM size";
        return false;
    }
    if (maxval == 0 || maxval > 255) {
        why = "PPM maxval must be 1..255";
        return false;
    }
    in.get();  // the single whitespace after maxval
    std::vector<unsigned char> px((size_t)w * h * 3);
    in.read(reinterpret_cast<char*>(px.data()), (std::streamsize)px.size());
    if ((size_t)in.gcount() != px.size()) {
        why = "truncated PPM data";
        return false;
    }
    t.width = w;
    t.height = h;
    t.rgba.resize((size_t)w * h * 4);
    for (uint32_t y = 0; y < h; ++y) {
        const unsigned char* row = px.data() + (size_t)(h - 1 - y) * w * 3;  // file rows run top-down
        for (uint32_t x = 0; x < w; ++x)
            for (int k = 0; k < 4; ++k) {
                float v = 1.0f;
                if (k < 3) {
                    uint32_t c = row[3 * x + k];
                    if (maxval != 255) c = (c * 255 + maxval / 2) / maxval;  // to 8 bits
                    v = (float)c / 255.0f;
                }
                t.rgba[4 * ((size_t)x + (size_t)y * w) + k] = v;
            }
    }
    return true;
}

bool load_pfm(std::istream& in, bool colour, Texture& t, std::string& why) {
    std::string tok;
    uint32_t w = 0, h = 0;
    if (!next_token(in, tok) || !to_uint(tok, w) || !next_token(in, tok) || !to_uint(tok, h) || !next_token(in, tok)) {
        why = "bad PFM header";
        return false;
    }
    const double scale = std::atof(tok.c_str());
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28) || scale == 0.0) {
        why = "bad PFM size or scale";
        return false;
    }
    in.get();
    const int nc = colour ? 3 : 1;
    std::vector<unsigned char> raw((size_t)w * h * nc * 4);
    in.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
    if ((size_t)in.gcount() != raw.size()) {
        why = "truncated PFM data";
        return false;
    }
    const bool little = scale < 0.0;  // negative scale: little-endian samples
    t.width = w;
    t.height = h;
    t.rgba.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; ++i) {  // PFM rows run bottom-up, like the texel rows
        float c[3];
        for (int k = 0; k < nc; ++k) {
            const unsigned char* b = raw.data() + (i * nc + k) * 4;
            const uint32_t u = little ? (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24
                                      : (uint32_t)b[3] | (uint32_t)b[2] << 8 | (uint32_t)b[1] << 16 | (uint32_t)b[0] << 24;
            std::memcpy(&c[k], &u, 4);
        }
        if (!colour) c[1] = c[2] = c[0];
        for (int k = 0; k < 3; ++k) t.rgba[4 * i + k] = c[k];
        t.rgba[4 * i + 3] = 1.0f;
    }
    return true;
}

// Truevision TGA: types 2 (true colour), 3 (grey) and their run-length
// encoded forms 10 and 11; 8, 24 or 32 bits per pixel stored B, G, R(, A).
// Rows run bottom-up unless bit 5 of the descriptor (top-left origin) is
// set.  Colour becomes c / 255 with alpha 1, the floats Surface keeps for
// 8-bit images.
bool load_tga(const std::vector<unsigned char>& f, Texture& t, std::string& why) {
    if (f.size() < 18) {
        why = "truncated TGA header";
        return false;
    }
    const uint32_t idlen = f[0], cmtype = f[1], type = f[2];
    const uint32_t w = f[12] | (uint32_t)f[13] << 8, h = f[14] | (uint32_t)f[15] << 8;
    const uint32_t bpp = f[16], desc = f[17];
    const bool rle = type == 10 || type == 11, grey = type == 3 || type == 11;
    if (cmtype != 0 || !(type == 2 || type == 3 || type == 10 || type == 11)) {
        why = "unsupported TGA type (true colour or grey, no colour map)";
        return false;
    }
    if ((grey && bpp != 8) || (!grey && bpp != 24 && bpp != 32) || (desc & 0x10)) {
        why = "unsupported TGA pixel format";
        return false;
    }
    if (w == 0 || h == 0) {
        why = "bad TGA size";
        return false;
    }
    const size_t px = bpp / 8, n = (size_t)w * h;
    std::vector<unsigned char> raw(n * px);
    size_t pos = 18 + idlen;
    if (!rle) {
        if (f.size() < pos + raw.size()) {
            why = "truncated TGA data";
            return false;
        }
        std::memcpy(raw.data(), f.data() + pos, raw.size());
    } else {
        size_t out = 0;
        while (out < n) {
            if (pos >= f.size()) {
                why = "truncated TGA data";
                return false;
            }
            const uint32_t hdr = f[pos++];
            const size_t cnt = (hdr & 0x7F) + 1;
            if (out + cnt > n) {
                why = "bad TGA run";
                return false;
            }
            if (hdr & 0x80) {  // run of one pixel value
                if (pos + px > f.size()) {
                    why = "truncated TGA data";
                    return false;
                }
                for (size_t i = 0; i < cnt; ++i) std::memcpy(&raw[(out + i) * px], &f[pos], px);
                pos += px;
            } else {  // raw packet
                if (pos + cnt * px > f.size()) {
                    why = "truncated TGA data";
                    return false;
                }
                std::memcpy(&raw[out * px], &f[pos], cnt * px);
                pos += cnt * px;
            }
            out += cnt;
        }
    }
    const bool top_down = (desc & 0x20) != 0;
    t.width = w;
    t.height = h;
    t.rgba.resize(n * 4);
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t src_row = top_down ? h - 1 - y : y;
        for (uint32_t x = 0; x < w; ++x) {
            const unsigned char* p = &raw[((size_t)src_row * w + x) * px];
            float* q = &t.rgba[4 * ((size_t)x + (size_t)y * w)];
            if (grey) {
                q[0] = q[1] = q[2] = (float)p[0] / 255.0f;
            } else {
                q[0] = (float)p[2] / 255.0f;
                q[1] = (float)p[1] / 255.0f;
                q[2] = (float)p[0] / 255.0f;
            }
            q[3] = 1.0f;
        }
    }
    return true;
}

bool has_suffix(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

}  // namespace

int TriangleList::addTexture(uint32_t width, uint32_t height, const float* rgba) {
    if (width == 0 || height == 0 || !rgba || (uint64_t)width * height > (1ull << 28)) return -1;
    Texture t;
    t.width = width;
    t.height = height;
    t.rgba.assign(rgba, rgba + (size_t)width * height * 4);
    m_Textures.push_back(std::move(t));
    return (int)m_Textures.size() - 1;
}

int TriangleList::loadTexture(const std::string& path, std::string* error) {
    auto it = m_LoadedTextures.find(path);
    if (it != m_LoadedTextures.end()) return it->second;  // already in memory
    std::string why;
    std::ifstream in(path, std::ios::binary);
    Texture t;
    bool ok = false;
    if (!in) {
        why = "cannot open " + path;
    } else if (has_suffix(path, ".tga")) {  // TGA has no magic number
        std::vector<unsigned char> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        ok = load_tga(f, t, why);
    } else {
        std::string magic;
        next_token(in, magic);
        if (magic == "P6") ok = load_ppm(in, t, why);
        else if (magic == "PF" || magic == "Pf") ok = load_pfm(in, magic == "PF", t, why);
        else why = "unsupported image format (TGA, binary PPM P6 or PFM): " + path;
    }
    if (!ok) {
        if (error) *error = why;
        return -1;
    }
    m_Textures.push_back(std::move(t));
    const int idx = (int)m_Textures.size() - 1;
    m_LoadedTextures[path] = idx;
    return idx;
}

}  // namespace ct
