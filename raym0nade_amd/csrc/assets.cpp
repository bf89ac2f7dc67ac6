// Host asset path: what Model::Model does with assimp and the embedded
// Python image scripts (src/model.cpp:88-215, src/material.cpp:150-298,
// src/component.cpp:69-114), without either:
//   - Wavefront OBJ + MTL import into an rm_scene_desc;
//   - a PNG decoder (zlib inflate + the five PNG row filters), with libpng's
//     conversions as the reference requests them (16-bit stripped to the
//     high byte, grey and palette expanded to RGB, alpha dropped);
//   - a DDS decoder (uncompressed, BC1-BC3) for textures;
//   - a Radiance HDR (RGBE) reader for the sky, flat and run-length scanlines.
// assimp, libpng, FreeImage and imageio are absent from this image; the
// formats are implemented from their published specifications.
#include <zlib.h>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>
#include "assets.h"
#include "rm_internal.h"

namespace rm {
namespace {

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + k);
    fclose(f);
    return true;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// PNG (ISO/IEC 15948): RGB bytes [h][w][3] as the reference's loader leaves
// them (src/material.cpp:231-260).
int32_t decode_png(const std::vector<uint8_t> &f, int &W, int &H, std::vector<uint8_t> &rgb) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || memcmp(f.data(), sig, 8) != 0) return fail(RM_ERR_INVALID, "PNG: bad signature");
    size_t p = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    W = H = 0;
    while (p + 12 <= f.size()) {
        uint32_t len = be32(&f[p]);
        if (p + 12 + (size_t)len > f.size()) return fail(RM_ERR_INVALID, "PNG: truncated chunk");
        const uint8_t *type = &f[p + 4], *d = &f[p + 8];
        if (!memcmp(type, "IHDR", 4)) {
            if (len < 13) return fail(RM_ERR_INVALID, "PNG: short IHDR");
            W = (int)be32(d);
            H = (int)be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0) return fail(RM_ERR_INVALID, "PNG: unknown compression or filter method");
        } else if (!memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        p += 12 + (size_t)len;
    }
    if (W <= 0 || H <= 0 || W > (1 << 15) || H > (1 << 15)) return fail(RM_ERR_INVALID, "PNG: bad size %dx%d", W, H);
    if (interlace) return fail(RM_ERR_UNSUPPORTED, "PNG: interlaced images are not supported");
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;  // grey
        case 2: channels = 3; break;  // RGB
        case 3: channels = 1; break;  // palette index
        case 4: channels = 2; break;  // grey + alpha
        case 6: channels = 4; break;  // RGBA
        default: return fail(RM_ERR_INVALID, "PNG: colour type %d", ctype);
    }
    const bool ok_depth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!ok_depth) return fail(RM_ERR_INVALID, "PNG: bit depth %d with colour type %d", depth, ctype);
    if (ctype == 3 && plte.size() < 3) return fail(RM_ERR_INVALID, "PNG: palette image without PLTE");
    const size_t bits_px = (size_t)channels * depth;
    const size_t stride = ((size_t)W * bits_px + 7) / 8;
    const size_t bpp = bits_px >= 8 ? bits_px / 8 : 1;  // filter unit
    std::vector<uint8_t> raw((stride + 1) * H);
    uLongf raw_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK || raw_len != raw.size())
        return fail(RM_ERR_INVALID, "PNG: image data does not inflate to %zu bytes", raw.size());
    // undo the row filters in place; row y's bytes at raw[y*(stride+1)+1 ...]
    std::vector<uint8_t> img(stride * H);
    for (int y = 0; y < H; ++y) {
        const uint8_t ft = raw[(size_t)y * (stride + 1)];
        const uint8_t *src = &raw[(size_t)y * (stride + 1) + 1];
        uint8_t *cur = &img[(size_t)y * stride];
        const uint8_t *up = y ? &img[(size_t)(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= bpp ? cur[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: {
                    int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: return fail(RM_ERR_INVALID, "PNG: filter type %d", ft);
            }
            cur[i] = (uint8_t)v;
        }
    }
    // sample k of row y, as 8 bits (16 -> high byte; 1, 2, 4 -> scaled to 0..255 for grey)
    auto sample = [&](int y, size_t k) -> int {
        const uint8_t *row = &img[(size_t)y * stride];
        if (depth == 8) return row[k];
        if (depth == 16) return row[2 * k];
        const size_t bit = k * depth;
        const int v = (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
        return ctype == 3 ? v : v * 255 / ((1 << depth) - 1);
    };
    rgb.resize((size_t)W * H * 3);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint8_t *o = &rgb[((size_t)y * W + x) * 3];
            const size_t k = (size_t)x * channels;
            if (ctype == 3) {
                const size_t e = (size_t)sample(y, k) * 3;
                if (e + 2 >= plte.size()) return fail(RM_ERR_INVALID, "PNG: palette index out of range");
                o[0] = plte[e], o[1] = plte[e + 1], o[2] = plte[e + 2];
            } else if (ctype == 0 || ctype == 4) {
                o[0] = o[1] = o[2] = (uint8_t)sample(y, k);
            } else {
                o[0] = (uint8_t)sample(y, k), o[1] = (uint8_t)sample(y, k + 1), o[2] = (uint8_t)sample(y, k + 2);
            }
        }
    return RM_OK;
}

// Radiance RGBE (Ward, "Real Pixels", Graphics Gems II): linear RGB [h][w][3],
// top row first.  A texel (r, g, b, e) is (r, g, b) * 2^(e - 136), zero when
// e = 0.
int32_t decode_hdr(const std::vector<uint8_t> &f, int &W, int &H, std::vector<float> &rgb) {
    size_t p = 0;
    auto line = [&](std::string &s) -> bool {
        s.clear();
        while (p < f.size() && f[p] != '\n') s.push_back((char)f[p++]);
        if (p >= f.size()) return false;
        ++p;
        return true;
    };
    std::string s;
    if (!line(s) || s.rfind("#?", 0) != 0) return fail(RM_ERR_INVALID, "HDR: missing #? header");
    bool rgbe = true;
    while (line(s) && !s.empty())
        if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe") rgbe = false;
    if (!rgbe) return fail(RM_ERR_UNSUPPORTED, "HDR: only FORMAT=32-bit_rle_rgbe");
    if (!line(s)) return fail(RM_ERR_INVALID, "HDR: missing resolution line");
    char ya[3] = {0}, xa[3] = {0};
    if (sscanf(s.c_str(), "%2s %d %2s %d", ya, &H, xa, &W) != 4 || strcmp(ya, "-Y") || strcmp(xa, "+X"))
        return fail(RM_ERR_UNSUPPORTED, "HDR: orientation '%s' (only -Y h +X w)", s.c_str());
    if (W <= 0 || H <= 0 || W > (1 << 16) || H > (1 << 16)) return fail(RM_ERR_INVALID, "HDR: bad size");
    std::vector<uint8_t> scan((size_t)W * 4);
    rgb.resize((size_t)W * H * 3);
    auto need = [&](size_t k) { return p + k <= f.size(); };
    for (int y = 0; y < H; ++y) {
        if (W >= 8 && W < 32768 && need(4) && f[p] == 2 && f[p + 1] == 2 && ((f[p + 2] << 8) | f[p + 3]) == W) {
            p += 4;  // per-component run-length scanline
            for (int c = 0; c < 4; ++c)
                for (int x = 0; x < W;) {
                    if (!need(1)) return fail(RM_ERR_INVALID, "HDR: truncated");
                    int n = f[p++];
                    if (n > 128) {
                        n -= 128;
                        if (x + n > W || !need(1)) return fail(RM_ERR_INVALID, "HDR: bad run");
                        const uint8_t v = f[p++];
                        for (int k = 0; k < n; ++k) scan[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (n == 0 || x + n > W || !need(n)) return fail(RM_ERR_INVALID, "HDR: bad literal run");
                        for (int k = 0; k < n; ++k) scan[(size_t)(x++) * 4 + c] = f[p++];
                    }
                }
        } else {
            // flat texels, with the original (1, 1, 1, n) repeat of the previous texel
            int shift = 0;
            for (int x = 0; x < W;) {
                if (!need(4)) return fail(RM_ERR_INVALID, "HDR: truncated");
                const uint8_t *t = &f[p];
                p += 4;
                if (t[0] == 1 && t[1] == 1 && t[2] == 1) {
                    if (x == 0) return fail(RM_ERR_INVALID, "HDR: repeat with no previous texel");
                    const int n = t[3] << shift;
                    if (x + n > W) return fail(RM_ERR_INVALID, "HDR: bad repeat");
                    for (int k = 0; k < n; ++k, ++x) memcpy(&scan[(size_t)x * 4], &scan[(size_t)(x - 1) * 4], 4);
                    shift += 8;
                } else {
                    memcpy(&scan[(size_t)(x++) * 4], t, 4);
                    shift = 0;
                }
            }
        }
        for (int x = 0; x < W; ++x) {
            const uint8_t *t = &scan[(size_t)x * 4];
            float *o = &rgb[((size_t)y * W + x) * 3];
            if (t[3] == 0) {
                o[0] = o[1] = o[2] = 0.0f;
            } else {
                const float m = (float)std::ldexp(1.0, (int)t[3] - 136);
                o[0] = t[0] * m, o[1] = t[1] * m, o[2] = t[2] * m;
            }
        }
    }
    return RM_OK;
}

// DDS (DirectDraw Surface) top level as RGB bytes: uncompressed 24/32-bit
// RGB(A) with channel masks, and the S3TC block formats BC1 (DXT1), BC2
// (DXT3) and BC3 (DXT5), by FourCC or DX10 header -- the formats the
// reference's imageio/FreeImage reader decodes to 8-bit RGBA.  Alpha is
// dropped (the scene format has RGB albedo only).
// alpha receives one byte per pixel: 255 for formats without alpha; BC1's
// punch-through black is 0; BC2 expands its 4-bit values (a * 17); BC3
// interpolates its two end points with integer division, as for the colours.
int32_t decode_dds(const std::vector<uint8_t> &f, int &W, int &H, std::vector<uint8_t> &rgb,
                   std::vector<uint8_t> &alpha, bool &has_alpha) {
    auto le32 = [&](size_t o) { return (uint32_t)f[o] | (uint32_t)f[o + 1] << 8 | (uint32_t)f[o + 2] << 16 | (uint32_t)f[o + 3] << 24; };
    if (f.size() < 128 || memcmp(f.data(), "DDS ", 4) != 0 || le32(4) != 124) return fail(RM_ERR_INVALID, "DDS: bad header");
    H = (int)le32(12);
    W = (int)le32(16);
    if (W <= 0 || H <= 0 || W > (1 << 15) || H > (1 << 15)) return fail(RM_ERR_INVALID, "DDS: bad size");
    const uint32_t pf_flags = le32(80), fourcc = le32(84), bits = le32(88);
    size_t data = 128;
    int bc = 0;  // 1, 2, 3 = BC1..BC3; 0 = uncompressed
    if (pf_flags & 0x4) {  // DDPF_FOURCC
        if (fourcc == 0x31545844u) bc = 1;         // "DXT1"
        else if (fourcc == 0x33545844u) bc = 2;    // "DXT3"
        else if (fourcc == 0x35545844u) bc = 3;    // "DXT5"
        else if (fourcc == 0x30315844u) {          // "DX10": DXGI format follows
            if (f.size() < 148) return fail(RM_ERR_INVALID, "DDS: short DX10 header");
            const uint32_t dxgi = le32(128);
            data = 148;
            if (dxgi == 71 || dxgi == 72) bc = 1;       // BC1_UNORM(_SRGB)
            else if (dxgi == 74 || dxgi == 75) bc = 2;  // BC2
            else if (dxgi == 77 || dxgi == 78) bc = 3;  // BC3
            else return fail(RM_ERR_UNSUPPORTED, "DDS: DXGI format %u", dxgi);
        } else {
            return fail(RM_ERR_UNSUPPORTED, "DDS: FourCC %08x", fourcc);
        }
    } else if (!(pf_flags & 0x40) || (bits != 24 && bits != 32)) {  // DDPF_RGB
        return fail(RM_ERR_UNSUPPORTED, "DDS: pixel format flags %x, %u bits", pf_flags, bits);
    }
    rgb.assign((size_t)W * H * 3, 0);
    alpha.assign((size_t)W * H, 255);
    has_alpha = bc != 0 || (pf_flags & 0x1);  // DDPF_ALPHAPIXELS
    if (bc == 0) {
        const uint32_t mask[3] = {le32(92), le32(96), le32(100)}, amask = (pf_flags & 0x1) ? le32(104) : 0u;
        const size_t bpp = bits / 8, row = (size_t)W * bpp;
        if (f.size() < data + row * H) return fail(RM_ERR_INVALID, "DDS: truncated");
        for (int c = 0; c < 3; ++c)
            if (!mask[c]) return fail(RM_ERR_UNSUPPORTED, "DDS: missing channel mask");
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const uint8_t *p = &f[data + (size_t)y * row + (size_t)x * bpp];
                const uint32_t v = (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (bpp == 4 ? (uint32_t)p[3] << 24 : 0u);
                for (int c = 0; c < 3; ++c) {
                    const int sh = __builtin_ctz(mask[c]);
                    const uint32_t m = mask[c] >> sh, q = (v & mask[c]) >> sh;
                    rgb[((size_t)y * W + x) * 3 + c] = (uint8_t)(m == 255 ? q : (q * 255 + m / 2) / m);
                }
                if (amask) {
                    const int sh = __builtin_ctz(amask);
                    const uint32_t m = amask >> sh, q = (v & amask) >> sh;
                    alpha[(size_t)y * W + x] = (uint8_t)(m == 255 ? q : (q * 255 + m / 2) / m);
                }
            }
        return RM_OK;
    }
    const int bw = (W + 3) / 4, bh = (H + 3) / 4;
    const size_t bsz = bc == 1 ? 8 : 16;
    if (f.size() < data + (size_t)bw * bh * bsz) return fail(RM_ERR_INVALID, "DDS: truncated block data");
    for (int by = 0; by < bh; ++by)
        for (int bx = 0; bx < bw; ++bx) {
            const uint8_t *b = &f[data + ((size_t)by * bw + bx) * bsz];
            const uint8_t *cb = bc == 1 ? b : b + 8;  // the colour half
            const uint32_t c0 = cb[0] | cb[1] << 8, c1 = cb[2] | cb[3] << 8;
            int pal[4][3];
            auto expand = [](uint32_t c, int *o) {
                const uint32_t r = (c >> 11) & 31, g = (c >> 5) & 63, bl = c & 31;
                o[0] = (int)(r << 3 | r >> 2), o[1] = (int)(g << 2 | g >> 4), o[2] = (int)(bl << 3 | bl >> 2);
            };
            expand(c0, pal[0]);
            expand(c1, pal[1]);
            for (int k = 0; k < 3; ++k) {
                if (bc != 1 || c0 > c1) {  // four colours (always for BC2/BC3)
                    pal[2][k] = (2 * pal[0][k] + pal[1][k]) / 3;
                    pal[3][k] = (pal[0][k] + 2 * pal[1][k]) / 3;
                } else {  // three colours and black (transparent in BC1)
                    pal[2][k] = (pal[0][k] + pal[1][k]) / 2;
                    pal[3][k] = 0;
                }
            }
            const uint32_t idx = cb[4] | cb[5] << 8 | cb[6] << 16 | (uint32_t)cb[7] << 24;
            int apal[8] = {b[0], b[1], 0, 0, 0, 0, 0, 0};  // BC3 alpha end points and steps
            if (bc == 3) {
                if (apal[0] > apal[1]) {
                    for (int i = 2; i < 8; ++i) apal[i] = ((8 - i) * apal[0] + (i - 1) * apal[1]) / 7;
                } else {
                    for (int i = 2; i < 6; ++i) apal[i] = ((6 - i) * apal[0] + (i - 1) * apal[1]) / 5;
                    apal[6] = 0;
                    apal[7] = 255;
                }
            }
            uint64_t abits = 0;
            for (int k = 0; k < 6; ++k) abits |= (uint64_t)b[2 + k] << (8 * k);
            for (int t = 0; t < 16; ++t) {
                const int x = bx * 4 + (t & 3), y = by * 4 + (t >> 2);
                if (x >= W || y >= H) continue;
                const int sel = (idx >> (2 * t)) & 3;
                const int* c = pal[sel];
                uint8_t *o = &rgb[((size_t)y * W + x) * 3];
                o[0] = (uint8_t)c[0], o[1] = (uint8_t)c[1], o[2] = (uint8_t)c[2];
                uint8_t &a = alpha[(size_t)y * W + x];
                if (bc == 1) a = (c0 <= c1 && sel == 3) ? 0 : 255;
                else if (bc == 2) a = (uint8_t)(((b[t >> 1] >> (4 * (t & 1))) & 15) * 17);
                else a = (uint8_t)apal[(abits >> (3 * t)) & 7];
            }
        }
    return RM_OK;
}

std::string ext_of(const std::string &path) {
    size_t d = path.find_last_of('.');
    std::string e = d == std::string::npos ? "" : path.substr(d + 1);
    for (auto &c : e) c = (char)tolower(c);
    return e;
}

// An image file as linear floats: PNG as c / 255, HDR as stored.
// alpha (optional) receives a / 255 per pixel for a DDS format with an alpha
// channel and is left empty otherwise (the reference strips PNG alpha,
// src/material.cpp:264-266).
int32_t load_image(const std::string &path, int &W, int &H, std::vector<float> &out, bool &is_hdr,
                   std::vector<float> *alpha = nullptr) {
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return fail(RM_ERR_INVALID, "cannot read %s", path.c_str());
    const std::string e = ext_of(path);
    if (e == "png") {
        std::vector<uint8_t> rgb;
        int32_t rc = decode_png(f, W, H, rgb);
        if (rc) return rc;
        out.resize(rgb.size());
        for (size_t i = 0; i < rgb.size(); ++i) out[i] = (float)rgb[i] / 255.0f;
        is_hdr = false;
        return RM_OK;
    }
    if (e == "dds") {
        std::vector<uint8_t> rgb, a8;
        bool has_alpha = false;
        int32_t rc = decode_dds(f, W, H, rgb, a8, has_alpha);
        if (rc) return rc;
        out.resize(rgb.size());
        for (size_t i = 0; i < rgb.size(); ++i) out[i] = (float)rgb[i] / 255.0f;
        if (alpha && has_alpha) {
            alpha->resize(a8.size());
            for (size_t i = 0; i < a8.size(); ++i) (*alpha)[i] = (float)a8[i] / 255.0f;
        }
        is_hdr = false;
        return RM_OK;
    }
    if (e == "hdr") {
        is_hdr = true;
        return decode_hdr(f, W, H, out);
    }
    return fail(RM_ERR_UNSUPPORTED, "%s: only .png, .dds and .hdr images are read by this build", path.c_str());
}

// An 8-bit texture as the reference's loaders store it (ImageData::data[0]):
// PNG as 3 bytes per pixel after libpng's conversions
// (src/material.cpp:214-267); DDS as FreeImage decodes it for the imageio
// script (scripts/dds_to_array.py): 4 bytes per pixel (RGBA) for
// block-compressed and 32-bit formats, 3 for 24-bit RGB.
int32_t load_image8(const std::string &path, int &W, int &H, int &channels, std::vector<uint8_t> &bytes) {
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return fail(RM_ERR_INVALID, "cannot read %s", path.c_str());
    const std::string e = ext_of(path);
    if (e == "png") {
        channels = 3;
        return decode_png(f, W, H, bytes);
    }
    if (e == "dds") {
        std::vector<uint8_t> rgb, a8;
        bool has_alpha = false;
        int32_t rc = decode_dds(f, W, H, rgb, a8, has_alpha);
        if (rc) return rc;
        const uint32_t pf_flags = f[80] | f[81] << 8 | f[82] << 16 | (uint32_t)f[83] << 24;
        const uint32_t bits = f[88] | f[89] << 8 | f[90] << 16 | (uint32_t)f[91] << 24;
        channels = ((pf_flags & 0x4) || bits == 32) ? 4 : 3;
        const size_t np = (size_t)W * H;
        bytes.resize(np * channels);
        for (size_t i = 0; i < np; ++i) {
            for (int c = 0; c < 3; ++c) bytes[i * channels + c] = rgb[i * 3 + c];
            if (channels == 4) bytes[i * 4 + 3] = a8[i];
        }
        return RM_OK;
    }
    return fail(RM_ERR_UNSUPPORTED, "%s: 8-bit texture maps are .png or .dds", path.c_str());
}

// ------------------------------------------------------------ OBJ + MTL

// Texture paths are URL-decoded (urlDecode, src/material.cpp:9-24): each
// "%XX" becomes the byte of hex value XX.
std::string url_decode(const std::string &s) {
    std::string out;
    for (size_t i = 0; i < s.size(); ++i) {
        auto hex = [](char c) { return std::isxdigit((unsigned char)c) != 0; };
        if (s[i] == '%' && i + 2 < s.size() && hex(s[i + 1]) && hex(s[i + 2])) {
            out += (char)std::stoi(s.substr(i + 1, 2), nullptr, 16);
            i += 2;
        } else {
            out += s[i];
        }
    }
    return out;
}

}  // namespace

rm_material default_material() {
    rm_material m{};
    for (int k = 0; k < 3; ++k) m.base_color[k] = 1.0f;  // white without a diffuse map (src/material.cpp:349-351)
    m.roughness = 0.8f;                                   // Material() (src/material.cpp:109-111)
    m.opacity = 1.0f;
    m.ior = 1.0f;
    m.albedo_tex = m.normal_tex = m.emissive_tex = m.specular_tex = -1;
    return m;
}

// Material::loadMaterialProperties (src/material.cpp:300-333): an opacity
// below 0.99 makes a glass (opacity 0, ior 1.25, roughness 5e-3, colours by
// name).
void glass_by_name(MtlDef &d) {
    rm_material &m = d.m;
    m.opacity = 0.0f;
    m.ior = 1.25f;
    m.roughness = 5e-3f;
    const std::string &n = d.name;
    auto set = [&](float r, float g, float b) { m.transmitting_color[0] = r, m.transmitting_color[1] = g, m.transmitting_color[2] = b; };
    if (n == "Ice") m.roughness = 0.5f;
    if (n == "TransparentGlassWine") set(0.2f, 0.08f, 0.07f);
    if (n == "TransparentGlass" || n == "Water" || n == "Ice") set(1.0f, 1.0f, 1.0f);
    if (n == "Beer") set(0.8f, 0.7f, 0.55f);
    if (n == "Red_Wine") set(0.24f, 0.09f, 0.07f);
    if (n == "White_Wine") set(0.85f, 0.78f, 0.6f);
}

namespace {

// Texture maps of a material under the keywords assimp's OBJ importer maps
// to the reference's texture types: map_Kd diffuse, norm / map_Kn normals,
// map_Ke / map_emissive emissive, map_Ks specular.
int32_t parse_mtl(const std::string &path, std::vector<MtlDef> &mats) {
    std::ifstream in(path);
    if (!in) return fail(RM_ERR_INVALID, "cannot read material library %s", path.c_str());
    std::string ln;
    MtlDef *cur = nullptr;
    std::vector<float> opacity;  // per material, 1 unless d / Tr says otherwise
    while (std::getline(in, ln)) {
        std::istringstream ss(ln);
        std::string k;
        if (!(ss >> k) || k[0] == '#') continue;
        if (k == "newmtl") {
            mats.emplace_back();
            cur = &mats.back();
            std::getline(ss >> std::ws, cur->name);
            while (!cur->name.empty() && (cur->name.back() == '\r' || cur->name.back() == ' ')) cur->name.pop_back();
            cur->m = default_material();
            opacity.push_back(1.0f);
        } else if (!cur) {
            continue;
        } else if (k == "Kd") {  // EXT: a constant albedo (the reference uses only map_Kd)
            ss >> cur->m.base_color[0] >> cur->m.base_color[1] >> cur->m.base_color[2];
        } else if (k == "Ke") {  // EXT: a constant emission (the reference uses only map_Ke)
            ss >> cur->m.emission[0] >> cur->m.emission[1] >> cur->m.emission[2];
        } else if (k == "Pr") {  // EXT: constant roughness (PBR extension of MTL)
            ss >> cur->m.roughness;
        } else if (k == "Pm") {  // EXT: constant metallic
            ss >> cur->m.metallic;
        } else if (k == "d") {
            ss >> opacity.back();
        } else if (k == "Tr") {
            float t = 0;
            ss >> t;
            opacity.back() = 1.0f - t;
        } else if (k == "map_Kd" || k == "norm" || k == "map_Kn" || k == "map_Ke" || k == "map_emissive" ||
                   k == "map_Ks") {
            std::string rest;
            std::getline(ss >> std::ws, rest);
            while (!rest.empty() && (rest.back() == '\r' || rest.back() == ' ')) rest.pop_back();
            size_t sp = rest.find_last_of(' ');  // options before the file name are ignored
            std::string file = sp == std::string::npos ? rest : rest.substr(sp + 1);
            if (k == "map_Kd") cur->map_kd = file;
            else if (k == "norm" || k == "map_Kn") cur->map_norm = file;
            else if (k == "map_Ks") cur->map_ks = file;
            else cur->map_ke = file;
        }
    }
    for (size_t i = 0; i < mats.size(); ++i)
        if (opacity[i] < 0.99f) glass_by_name(mats[i]);
    return RM_OK;
}

}  // namespace

int32_t parse_obj(const std::string &folder, const std::string &name, LoadedModel &S) {
    std::ifstream in(folder + name);
    if (!in) return fail(RM_ERR_INVALID, "cannot read model %s%s", folder.c_str(), name.c_str());
    std::vector<float> P, T, N;
    std::vector<MtlDef> defs;
    std::map<std::string, int> mat_index;
    int cur_mat = -1, default_mat = -1;
    std::string cur_group;
    std::map<std::pair<std::string, int>, int> mesh_of;  // (group, material) -> mesh id in order of appearance
    struct Tri { int mesh; int v[3], t[3], n[3]; int mat; };
    std::vector<Tri> tris;
    std::string ln;
    auto fix = [](long i, size_t count) -> long { return i < 0 ? (long)count + i : i - 1; };
    while (std::getline(in, ln)) {
        std::istringstream ss(ln);
        std::string k;
        if (!(ss >> k) || k[0] == '#') continue;
        if (k == "v") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            P.insert(P.end(), {x, y, z});
        } else if (k == "vt") {
            float u = 0, v = 0;
            ss >> u >> v;
            T.insert(T.end(), {u, v});
        } else if (k == "vn") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            N.insert(N.end(), {x, y, z});
        } else if (k == "mtllib") {
            std::string lib;
            std::getline(ss >> std::ws, lib);
            while (!lib.empty() && (lib.back() == '\r' || lib.back() == ' ')) lib.pop_back();
            const size_t before = defs.size();
            int32_t rc = parse_mtl(folder + lib, defs);
            if (rc) return rc;
            for (size_t i = before; i < defs.size(); ++i) mat_index[defs[i].name] = (int)i;
        } else if (k == "usemtl") {
            std::string m;
            std::getline(ss >> std::ws, m);
            while (!m.empty() && (m.back() == '\r' || m.back() == ' ')) m.pop_back();
            auto it = mat_index.find(m);
            if (it == mat_index.end()) return fail(RM_ERR_INVALID, "OBJ: unknown material '%s'", m.c_str());
            cur_mat = it->second;
        } else if (k == "o" || k == "g") {
            std::getline(ss >> std::ws, cur_group);
        } else if (k == "f") {
            std::vector<int> vi, ti, ni;
            std::string tok;
            while (ss >> tok) {
                long a = 0, b = 0, c = 0;
                int have_t = 0, have_n = 0;
                const char *s = tok.c_str();
                char *e;
                a = strtol(s, &e, 10);
                if (*e == '/') {
                    s = e + 1;
                    if (*s != '/') { b = strtol(s, &e, 10); have_t = 1; s = e; }
                    if (*s == '/') { c = strtol(s + 1, &e, 10); have_n = 1; }
                }
                long va = fix(a, P.size() / 3), tb = have_t ? fix(b, T.size() / 2) : -1, nc = have_n ? fix(c, N.size() / 3) : -1;
                if (va < 0 || va >= (long)(P.size() / 3) || (have_t && (tb < 0 || tb >= (long)(T.size() / 2))) ||
                    (have_n && (nc < 0 || nc >= (long)(N.size() / 3))))
                    return fail(RM_ERR_INVALID, "OBJ: face index out of range in '%s'", ln.c_str());
                vi.push_back((int)va);
                ti.push_back((int)tb);
                ni.push_back((int)nc);
            }
            if (vi.size() < 3) continue;
            int mat = cur_mat;
            if (mat < 0) {
                if (default_mat < 0) {
                    MtlDef d;
                    d.name = "DefaultMaterial";
                    d.m = default_material();
                    defs.push_back(d);
                    default_mat = (int)defs.size() - 1;
                }
                mat = default_mat;
            }
            auto key = std::make_pair(cur_group, mat);
            auto it = mesh_of.find(key);
            int mesh = it == mesh_of.end() ? (mesh_of[key] = (int)mesh_of.size()) : it->second;
            for (size_t j = 1; j + 1 < vi.size(); ++j) {  // a fan, as for a convex polygon
                Tri t{mesh, {vi[0], vi[j], vi[j + 1]}, {ti[0], ti[j], ti[j + 1]}, {ni[0], ni[j], ni[j + 1]}, mat};
                tris.push_back(t);
            }
        }
    }
    if (tris.empty()) return fail(RM_ERR_EMPTY, "OBJ: no faces in %s", name.c_str());
    // meshes contiguous, in order of first appearance; faces keep file order inside a mesh
    std::vector<std::vector<int>> by_mesh(mesh_of.size());
    for (size_t i = 0; i < tris.size(); ++i) by_mesh[tris[i].mesh].push_back((int)i);
    for (const auto &list : by_mesh)
        for (int i : list) {
            const Tri &t = tris[i];
            for (int c = 0; c < 3; ++c) {
                for (int a = 0; a < 3; ++a) S.verts.push_back(P[(size_t)t.v[c] * 3 + a]);
                S.uvs.push_back(t.t[c] >= 0 ? T[(size_t)t.t[c] * 2] : 0.0f);
                S.uvs.push_back(t.t[c] >= 0 ? T[(size_t)t.t[c] * 2 + 1] : 0.0f);
                if (t.n[c] < 0) S.all_normals = false;
                for (int a = 0; a < 3; ++a) S.normals.push_back(t.n[c] >= 0 ? N[(size_t)t.n[c] * 3 + a] : 0.0f);
            }
            S.material.push_back(t.mat);
            S.mesh.push_back(t.mesh);
        }
    return load_material_maps(folder, defs, S);
}

// Materials and their texture maps (Material::loadImageFromFile,
// src/material.cpp:290-298, as Model::processMaterial passes them,
// src/model.cpp:146-152: model folder + path, backslashes made slashes, then
// URL-decoded): PNG and DDS maps as the reference's 8-bit texels; an HDR
// diffuse map (extension) as linear fp32 texels.  Each file is loaded once.
int32_t load_material_maps(const std::string &folder, std::vector<MtlDef> &defs, LoadedModel &S) {
    auto texture = [&](const std::string &file, bool allow_hdr, int32_t &index) -> int32_t {
        index = -1;
        if (file.empty()) return RM_OK;
        std::string path = folder + file;
        for (auto &c : path)
            if (c == '\\') c = '/';
        path = url_decode(path);
        for (size_t i = 0; i < S.tex_paths.size(); ++i)
            if (S.tex_paths[i] == path) {
                index = (int)i;
                return RM_OK;
            }
        // Material::loadImageFromFile reads "png", "PNG", "dds" and "DDS"
        // files and leaves the map empty for any other extension
        const size_t dot = path.find_last_of('.');
        const std::string e = dot == std::string::npos ? path : path.substr(dot + 1);
        const bool hdr_map = allow_hdr && ext_of(path) == "hdr";
        if (!hdr_map && e != "png" && e != "PNG" && e != "dds" && e != "DDS") return RM_OK;
        int w = 0, h = 0;
        if (hdr_map) {
            bool hdr = false;
            std::vector<float> px;
            int32_t rc = load_image(path, w, h, px, hdr);
            if (rc) return rc;
            S.tex_data.push_back(std::move(px));
            S.tex_bytes.emplace_back();
            S.texs.push_back(rm_texture{w, h, nullptr, nullptr, nullptr, 0});
        } else {
            int ch = 0;
            std::vector<uint8_t> bytes;
            int32_t rc = load_image8(path, w, h, ch, bytes);
            if (rc) return rc;
            S.tex_data.emplace_back();
            S.tex_bytes.push_back(std::move(bytes));
            S.texs.push_back(rm_texture{w, h, nullptr, nullptr, nullptr, ch});
        }
        S.tex_paths.push_back(path);
        index = (int)S.texs.size() - 1;
        return RM_OK;
    };
    for (auto &d : defs) {
        int32_t rc;
        if ((rc = texture(d.map_kd, true, d.m.albedo_tex)) || (rc = texture(d.map_norm, false, d.m.normal_tex)) ||
            (rc = texture(d.map_ke, false, d.m.emissive_tex)) || (rc = texture(d.map_ks, false, d.m.specular_tex)))
            return rc;
        S.mats.push_back(d.m);
        S.mat_names.push_back(d.name);
    }
    return RM_OK;
}

void fill_desc(LoadedModel &S) {
    for (size_t i = 0; i < S.texs.size(); ++i) {
        if (S.texs[i].channels) {
            S.texs[i].bytes = S.tex_bytes[i].data();
        } else {
            S.texs[i].rgb = S.tex_data[i].data();
        }
    }
    rm_scene_desc &d = S.desc;
    d = rm_scene_desc{};
    d.n_tris = (int64_t)S.material.size();
    d.verts = S.verts.data();
    d.uvs = S.uvs.data();
    d.normals = S.all_normals ? S.normals.data() : nullptr;
    d.material = S.material.data();
    d.mesh = S.mesh.data();
    d.n_materials = (int32_t)S.mats.size();
    d.materials = S.mats.data();
    d.n_textures = (int32_t)S.texs.size();
    d.textures = S.texs.empty() ? nullptr : S.texs.data();
    d.sky_width = S.sky_w;
    d.sky_height = S.sky_h;
    d.sky_rgb = S.sky.empty() ? nullptr : S.sky.data();
}

}  // namespace rm

using namespace rm;

extern "C" int32_t rm_image_load(const char *path, int32_t *width, int32_t *height, float **rgb) {
    if (!path || !width || !height || !rgb) return fail(RM_ERR_INVALID, "rm_image_load: NULL argument");
    *rgb = nullptr;
    int w = 0, h = 0;
    bool hdr = false;
    std::vector<float> px;
    int32_t rc = load_image(path, w, h, px, hdr);
    if (rc) return rc;
    float *out = (float *)malloc(px.size() * sizeof(float));
    if (!out) return fail(RM_ERR_OOM, "rm_image_load: host allocation");
    memcpy(out, px.data(), px.size() * sizeof(float));
    *width = w;
    *height = h;
    *rgb = out;
    return RM_OK;
}

extern "C" void rm_image_free(float *rgb) { free(rgb); }

struct rm_model {
    rm::LoadedModel m;
};

extern "C" int32_t rm_model_read(const char *model_folder, const char *model_name, const char *sky_name,
                                 rm_scene_desc *desc, rm_model **out) {
    if (!model_folder || !model_name || !desc || !out) return fail(RM_ERR_INVALID, "rm_model_read: NULL argument");
    *out = nullptr;
    const std::string folder = model_folder, name = model_name, ext = ext_of(name);
    if (ext != "obj" && ext != "fbx")
        return fail(RM_ERR_UNSUPPORTED, "rm_model_read: %s: only Wavefront OBJ and binary FBX models are read",
                    model_name);
    rm_model *M = new (std::nothrow) rm_model;
    if (!M) return fail(RM_ERR_OOM, "rm_model_read: host allocation");
    LoadedModel &S = M->m;
    int32_t rc = ext == "obj" ? parse_obj(folder, name, S) : parse_fbx(folder, name, S);
    if (!rc && sky_name && strcmp(sky_name, "null") != 0) {
        bool hdr = false;
        rc = load_image(folder + sky_name, S.sky_w, S.sky_h, S.sky, hdr);
    }
    if (rc) {
        delete M;
        return rc;
    }
    fill_desc(S);
    *desc = S.desc;
    *out = M;
    return RM_OK;
}

extern "C" int32_t rm_model_destroy(rm_model *model) {
    delete model;
    return RM_OK;
}

extern "C" int32_t rm_scene_load(const char *model_folder, const char *model_name, const char *sky_name, rm_scene **out) {
    if (!model_folder || !model_name || !out) return fail(RM_ERR_INVALID, "rm_scene_load: NULL argument");
    *out = nullptr;
    rm_scene_desc d{};
    rm_model *M = nullptr;
    int32_t rc = rm_model_read(model_folder, model_name, sky_name, &d, &M);
    if (rc) return rc;
    rc = rm_scene_create(&d, out);
    rm_model_destroy(M);
    return rc;
}
