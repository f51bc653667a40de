// imageio.cpp -- image input for textures and environment maps.
//
// ReadImage (src/renderer/imageio.cpp:45-57) loads through stbi_loadf with a
// vertical flip and three channels.  Restated here for the formats the scenes
// use: PNG (zlib inflate of the IDAT stream, the five scanline filters,
// grey / RGB / palette / alpha colour types at 1-16 bits), Radiance RGBE .hdr
// (flat and new-style run-length scanlines) and binary PNM (P5 / P6).  What
// stbi_loadf makes of them: 8-bit samples (16-bit PNG samples reduced to their
// high byte, low-depth grey scaled to 0..255) become (c / 255)^2.2; RGBE pixels
// become m * 2^(e - 136); alpha is dropped, grey is replicated, and the rows
// are returned bottom row first.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gpr_internal.h"

namespace gpr {
namespace {

struct Image {
    uint32_t w = 0, h = 0;
    std::vector<float> rgb;  // top row first until flipped
};

// stbi__ldr_to_hdr: pow(c / 255, 2.2) per colour channel
float ldr_to_linear(unsigned c) { return (float)std::pow((double)((float)c / 255.0f), (double)2.2f); }

void flip_rows(Image& im) {
    const size_t row = 3ull * im.w;
    for (uint32_t y = 0; y < im.h / 2; ++y)
        std::swap_ranges(im.rgb.begin() + y * row, im.rgb.begin() + (y + 1) * row, im.rgb.begin() + (im.h - 1 - y) * row);
}

uint32_t be32(const unsigned char* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// ------------------------------------------------------------------ PNG --
int decode_png(const std::string& path, const std::string& buf, Image* out) {
    const unsigned char* b = reinterpret_cast<const unsigned char*>(buf.data());
    const size_t n = buf.size();
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(b, sig, 8) != 0) return fail(GPR_ERR_PARSE, path + ": not a PNG file");
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, pal;
    bool ended = false;
    for (size_t p = 8; p + 12 <= n && !ended;) {
        const uint32_t len = be32(b + p);
        if (len > n - p - 12) return fail(GPR_ERR_PARSE, path + ": truncated PNG chunk");
        const std::string type(reinterpret_cast<const char*>(b + p + 4), 4);
        const unsigned char* d = b + p + 8;
        if (type == "IHDR") {
            if (len < 13) return fail(GPR_ERR_PARSE, path + ": bad IHDR");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0) return fail(GPR_ERR_PARSE, path + ": unknown PNG compression / filter method");
        } else if (type == "PLTE") {
            pal.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            ended = true;
        }
        p += 12ull + len;
    }
    if (ctype < 0 || w == 0 || h == 0) return fail(GPR_ERR_PARSE, path + ": PNG without IHDR");
    if (interlace != 0) return fail(GPR_ERR_UNSUPPORTED, path + ": interlaced PNG");
    if ((uint64_t)w * h > (1ull << 28)) return fail(GPR_ERR_UNSUPPORTED, path + ": PNG larger than 2^28 pixels");
    int chans;
    switch (ctype) {
        case 0: chans = 1; break;
        case 2: chans = 3; break;
        case 3: chans = 1; break;
        case 4: chans = 2; break;
        case 6: chans = 4; break;
        default: return fail(GPR_ERR_PARSE, path + ": bad PNG colour type");
    }
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) return fail(GPR_ERR_PARSE, path + ": bad PNG bit depth");
    if (ctype == 3 && (pal.empty() || pal.size() % 3 != 0)) return fail(GPR_ERR_PARSE, path + ": palette PNG without PLTE");
    const size_t stride = ((size_t)w * chans * depth + 7) / 8;
    const size_t bpp = std::max<size_t>(1, (size_t)chans * depth / 8);
    std::vector<unsigned char> raw((stride + 1) * h);
    uLongf got = (uLongf)raw.size();
    if (uncompress(raw.data(), &got, idat.data(), (uLong)idat.size()) != Z_OK || got != raw.size())
        return fail(GPR_ERR_PARSE, path + ": corrupt PNG image data");
    // scanline filters (PNG specification, section 9)
    std::vector<unsigned char> cur(stride), prev(stride, 0), img(stride * h);
    for (uint32_t y = 0; y < h; ++y) {
        const unsigned char* r = raw.data() + y * (stride + 1);
        const int f = r[0];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? cur[x - bpp] : 0, up = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
            int v = r[1 + x];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) / 2; break;
                case 4: {
                    const int pp = a + up - c, pa = std::abs(pp - a), pb = std::abs(pp - up), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
                    break;
                }
                default: return fail(GPR_ERR_PARSE, path + ": bad PNG filter type");
            }
            cur[x] = (unsigned char)v;
        }
        std::memcpy(img.data() + y * stride, cur.data(), stride);
        std::swap(cur, prev);
    }
    out->w = w;
    out->h = h;
    out->rgb.resize(3ull * w * h);
    auto sample8 = [&](const unsigned char* row, uint32_t x, int c) -> unsigned {
        if (depth == 16) return row[2 * ((size_t)x * chans + c)];  // high byte
        if (depth == 8) return row[(size_t)x * chans + c];
        const size_t bit = (size_t)x * depth;  // 1/2/4-bit, single channel
        const unsigned v = (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
        return ctype == 3 ? v : v * (255u / ((1u << depth) - 1));
    };
    for (uint32_t y = 0; y < h; ++y) {
        const unsigned char* row = img.data() + y * stride;
        for (uint32_t x = 0; x < w; ++x) {
            unsigned c8[3];
            if (ctype == 3) {
                const unsigned i = sample8(row, x, 0);
                if (3 * i + 2 >= pal.size()) return fail(GPR_ERR_PARSE, path + ": palette index out of range");
                for (int c = 0; c < 3; ++c) c8[c] = pal[3 * i + c];
            } else if (chans <= 2) {
                c8[0] = c8[1] = c8[2] = sample8(row, x, 0);
            } else {
                for (int c = 0; c < 3; ++c) c8[c] = sample8(row, x, c);
            }
            for (int c = 0; c < 3; ++c) out->rgb[3 * ((size_t)y * w + x) + c] = ldr_to_linear(c8[c]);
        }
    }
    return GPR_OK;
}

// --------------------------------------------------------- Radiance HDR --
void rgbe_to_float(const unsigned char* e, float* o) {  // stbi__hdr_convert
    if (e[3] != 0) {
        const float f = (float)std::ldexp(1.0f, (int)e[3] - (int)(128 + 8));
        for (int c = 0; c < 3; ++c) o[c] = (float)e[c] * f;
    } else {
        o[0] = o[1] = o[2] = 0.f;
    }
}

int decode_hdr(const std::string& path, const std::string& buf, Image* out) {
    size_t p = 0;
    auto line = [&](std::string* l) -> bool {
        if (p >= buf.size()) return false;
        const size_t e = buf.find('\n', p);
        *l = buf.substr(p, e == std::string::npos ? std::string::npos : e - p);
        p = e == std::string::npos ? buf.size() : e + 1;
        return true;
    };
    std::string l;
    if (!line(&l) || (l != "#?RADIANCE" && l != "#?RGBE")) return fail(GPR_ERR_PARSE, path + ": not a Radiance file");
    bool fmt = false;
    while (true) {
        if (!line(&l)) return fail(GPR_ERR_PARSE, path + ": truncated HDR header");
        if (l.empty()) break;
        if (l == "FORMAT=32-bit_rle_rgbe") fmt = true;
    }
    if (!fmt) return fail(GPR_ERR_UNSUPPORTED, path + ": HDR format other than 32-bit_rle_rgbe");
    int h = 0, w = 0;
    char ya[3] = {0}, xa[3] = {0};
    if (!line(&l) || std::sscanf(l.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::strcmp(ya, "-Y") != 0 ||
        std::strcmp(xa, "+X") != 0 || w <= 0 || h <= 0)
        return fail(GPR_ERR_UNSUPPORTED, path + ": HDR orientation other than -Y h +X w");
    if ((uint64_t)w * h > (1ull << 28)) return fail(GPR_ERR_UNSUPPORTED, path + ": HDR larger than 2^28 pixels");
    const unsigned char* b = reinterpret_cast<const unsigned char*>(buf.data());
    const size_t n = buf.size();
    out->w = (uint32_t)w;
    out->h = (uint32_t)h;
    out->rgb.resize(3ull * w * h);
    auto flat = [&](size_t first_pixel) -> int {  // uncompressed RGBE from pixel first_pixel on
        const size_t need = 4ull * w * h - 4ull * first_pixel;
        if (n - p < need) return fail(GPR_ERR_PARSE, path + ": truncated HDR pixels");
        for (size_t i = first_pixel; i < (size_t)w * h; ++i, p += 4) rgbe_to_float(b + p, &out->rgb[3 * i]);
        return GPR_OK;
    };
    if (w < 8 || w >= 32768) return flat(0);
    std::vector<unsigned char> line_buf(4ull * w);
    for (int y = 0; y < h; ++y) {
        if (n - p < 4) return fail(GPR_ERR_PARSE, path + ": truncated HDR scanline");
        if (b[p] != 2 || b[p + 1] != 2 || (b[p + 2] & 0x80)) {
            if (y != 0) return fail(GPR_ERR_PARSE, path + ": mixed flat and RLE HDR scanlines");
            return flat(0);  // not run-length encoded at all
        }
        if (((int)b[p + 2] << 8 | b[p + 3]) != w) return fail(GPR_ERR_PARSE, path + ": HDR scanline width mismatch");
        p += 4;
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < w) {
                if (p >= n) return fail(GPR_ERR_PARSE, path + ": truncated HDR run");
                int cnt = b[p++];
                if (cnt > 128) {
                    cnt -= 128;
                    if (cnt > w - x || p >= n) return fail(GPR_ERR_PARSE, path + ": bad HDR run");
                    const unsigned char v = b[p++];
                    for (int k = 0; k < cnt; ++k) line_buf[4 * (x++) + c] = v;
                } else {
                    if (cnt == 0 || cnt > w - x || n - p < (size_t)cnt) return fail(GPR_ERR_PARSE, path + ": bad HDR dump");
                    for (int k = 0; k < cnt; ++k) line_buf[4 * (x++) + c] = b[p++];
                }
            }
        }
        for (int x = 0; x < w; ++x) rgbe_to_float(&line_buf[4 * x], &out->rgb[3 * ((size_t)y * w + x)]);
    }
    return GPR_OK;
}

// ------------------------------------------------------------ PNM P5/P6 --
int decode_pnm(const std::string& path, const std::string& buf, Image* out) {
    size_t p = 2;
    const int chans = buf[1] == '6' ? 3 : 1;
    auto num = [&](long* v) -> bool {
        while (p < buf.size()) {
            if (buf[p] == '#') {
                while (p < buf.size() && buf[p] != '\n') ++p;
            } else if (std::isspace((unsigned char)buf[p])) {
                ++p;
            } else {
                break;
            }
        }
        if (p >= buf.size() || !std::isdigit((unsigned char)buf[p])) return false;
        *v = 0;
        while (p < buf.size() && std::isdigit((unsigned char)buf[p]) && *v < (1L << 30)) *v = *v * 10 + (buf[p++] - '0');
        return true;
    };
    long w = 0, h = 0, mx = 0;
    if (!num(&w) || !num(&h) || !num(&mx) || w <= 0 || h <= 0 || mx <= 0)
        return fail(GPR_ERR_PARSE, path + ": bad PNM header");
    if (mx > 255) return fail(GPR_ERR_UNSUPPORTED, path + ": 16-bit PNM");
    if (p >= buf.size() || !std::isspace((unsigned char)buf[p])) return fail(GPR_ERR_PARSE, path + ": bad PNM header");
    ++p;
    if ((uint64_t)w * h > (1ull << 28) || buf.size() - p < (size_t)w * h * chans)
        return fail(GPR_ERR_PARSE, path + ": truncated PNM pixels");
    out->w = (uint32_t)w;
    out->h = (uint32_t)h;
    out->rgb.resize(3ull * w * h);
    const unsigned char* b = reinterpret_cast<const unsigned char*>(buf.data()) + p;
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) out->rgb[3 * i + c] = ldr_to_linear(b[i * chans + (chans == 3 ? c : 0)]);
    return GPR_OK;
}

}  // namespace

int read_image(const std::string& path, uint32_t* w, uint32_t* h, std::vector<float>* rgb) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(GPR_ERR_IO, "cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string buf = ss.str();
    Image im;
    int rc;
    if (buf.size() >= 8 && static_cast<unsigned char>(buf[0]) == 137 && buf.compare(1, 3, "PNG") == 0)
        rc = decode_png(path, buf, &im);
    else if (buf.compare(0, 2, "#?") == 0)
        rc = decode_hdr(path, buf, &im);
    else if (buf.size() >= 2 && buf[0] == 'P' && (buf[1] == '5' || buf[1] == '6'))
        rc = decode_pnm(path, buf, &im);
    else
        rc = fail(GPR_ERR_UNSUPPORTED, path + ": image format not recognised (PNG, Radiance HDR, binary PNM)");
    if (rc != GPR_OK) return rc;
    flip_rows(im);
    *w = im.w;
    *h = im.h;
    *rgb = std::move(im.rgb);
    return GPR_OK;
}

}  // namespace gpr

extern "C" int gpr_image_read(const char* path, float** rgb, uint32_t* w, uint32_t* h) {
    if (!path || !rgb || !w || !h) return gpr::fail(GPR_ERR_INVALID, "gpr_image_read: null argument");
    std::vector<float> v;
    uint32_t iw = 0, ih = 0;
    GPR_TRY(gpr::guarded("gpr_image_read", [&] { return gpr::read_image(path, &iw, &ih, &v); }));
    float* p = static_cast<float*>(std::malloc(v.size() * sizeof(float)));
    if (!p) return gpr::fail(GPR_ERR_OOM, "gpr_image_read: host allocation failed");
    std::memcpy(p, v.data(), v.size() * sizeof(float));
    *rgb = p;
    *w = iw;
    *h = ih;
    return GPR_OK;
}

extern "C" void gpr_image_free(float* rgb) { std::free(rgb); }
