// Film encoders (include/psy_image.h): PPM, PNG and EXR files and the
// base64 bucket payload of the reference's Blender protocol, from the XYZ
// film of a render.  Host code: the pixel conversion is split over host
// threads by rows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/psy_image.h"
#include "color_host.h"

namespace psy {
namespace api {
int fail(int code, const std::string& msg);  // api.hip: records psy_last_error()
}  // namespace api
}  // namespace psy

namespace {

using psy::color::apply;
using psy::color::matrices;

// srgb_gamma (src/image.rs:282-288) with the host C library's powf, as the
// reference's f32::powf.
float srgb_gamma(float n) {
    if (n < 0.0031308f) return n * 12.92f;
    return (1.055f * std::pow(n, 1.0f / 2.4f)) - 0.055f;
}

// quantize (image.rs:306-309): clamp to [0, 1] ignoring NaN, scale by 255
// and truncate.
uint8_t quantize(float n) {
    float c = 0.0f;
    if (n > 0.0f) c = n;       // 0.0.max(n): NaN and negatives give 0
    if (!(c < 1.0f)) c = 1.0f;  // 1.0.min(c)
    return (uint8_t)(c * 255.0f);
}

void srgb8(const float* xyz, uint8_t* rgb) {
    float r[3];
    apply(matrices().xyz_to_rec709_e, xyz[0], xyz[1], xyz[2], r);
    rgb[0] = quantize(srgb_gamma(r[0]));
    rgb[1] = quantize(srgb_gamma(r[1]));
    rgb[2] = quantize(srgb_gamma(r[2]));
}

// Runs fn(row) for every row, over host threads.
template <typename F>
void for_rows(uint32_t rows, uint64_t row_pixels, F fn) {
    const uint64_t work = (uint64_t)rows * row_pixels;
    unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    nt = (unsigned)std::min<uint64_t>(nt, std::max<uint64_t>(1, work / 65536));
    nt = std::min<unsigned>(nt, std::max(1u, rows));
    if (nt <= 1) {
        for (uint32_t y = 0; y < rows; y++) fn(y);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            for (uint32_t y = t; y < rows; y += nt) fn(y);
        });
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------- PNG

uint32_t crc_table[256];
bool crc_ready = [] {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
    return true;
}();

uint32_t crc32(const uint8_t* p, uint64_t n, uint32_t c = 0xffffffffu) {
    for (uint64_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void put_be32(std::string& s, uint32_t v) {
    s.push_back((char)(v >> 24));
    s.push_back((char)(v >> 16));
    s.push_back((char)(v >> 8));
    s.push_back((char)v);
}

void png_chunk(std::string& out, const char* type, const std::string& data) {
    put_be32(out, (uint32_t)data.size());
    std::string body(type, 4);
    body += data;
    out += body;
    put_be32(out, crc32(reinterpret_cast<const uint8_t*>(body.data()), body.size()) ^ 0xffffffffu);
}

// A zlib stream of stored (uncompressed) deflate blocks.
std::string zlib_stored(const std::string& raw) {
    std::string z;
    z.push_back((char)0x78);
    z.push_back((char)0x01);
    uint64_t pos = 0;
    do {
        const uint64_t n = std::min<uint64_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back((char)(last ? 1 : 0));
        z.push_back((char)(n & 0xff));
        z.push_back((char)(n >> 8));
        z.push_back((char)(~n & 0xff));
        z.push_back((char)((~n >> 8) & 0xff));
        z.append(raw, pos, n);
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    return z;
}

std::string encode_png(const float* xyz, uint32_t w, uint32_t h) {
    const uint64_t stride = 1 + 4ull * w;
    std::string raw(stride * h, '\0');
    for_rows(h, w, [&](uint32_t y) {
        // Rows from the film's last to its first (image.rs:146-147).
        const float* src = xyz + 3ull * w * (h - 1 - y);
        uint8_t* dst = reinterpret_cast<uint8_t*>(&raw[stride * y]);
        dst[0] = 0;  // filter: none
        for (uint32_t x = 0; x < w; x++) {
            srgb8(src + 3ull * x, dst + 1 + 4ull * x);
            dst[1 + 4ull * x + 3] = 255;
        }
    });
    std::string out("\x89PNG\r\n\x1a\n", 8);
    std::string ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr += std::string("\x08\x06\x00\x00\x00", 5);  // 8-bit RGBA
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", zlib_stored(raw));
    png_chunk(out, "IEND", std::string());
    return out;
}

// ---------------------------------------------------------------- PPM

std::string encode_ppm(const float* xyz, uint32_t w, uint32_t h, bool ascii) {
    std::vector<uint8_t> px(3ull * w * h);
    for_rows(h, w, [&](uint32_t y) {
        for (uint32_t x = 0; x < w; x++) srgb8(xyz + 3ull * ((uint64_t)y * w + x), &px[3ull * ((uint64_t)y * w + x)]);
    });
    std::string out = std::string(ascii ? "P3\n" : "P6\n") + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    if (!ascii) {
        out.append(reinterpret_cast<const char*>(px.data()), px.size());
        return out;
    }
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            const uint8_t* p = &px[3ull * ((uint64_t)y * w + x)];
            out += std::to_string(p[0]) + " " + std::to_string(p[1]) + " " + std::to_string(p[2]) + " ";
        }
        out += "\n";
    }
    return out;
}

// ---------------------------------------------------------------- EXR

// f32 -> f16, round to nearest even, with infinities, NaN and subnormals.
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u));
    const int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int shift = 14 - e;
        uint32_t h = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // may carry into the exponent, up to infinity
    return (uint16_t)(sign | h);
}

void put_le32(std::string& s, uint32_t v) {
    for (int i = 0; i < 4; i++) s.push_back((char)(v >> (8 * i)));
}

void exr_attr(std::string& out, const char* name, const char* type, const std::string& value) {
    out += name;
    out.push_back('\0');
    out += type;
    out.push_back('\0');
    put_le32(out, (uint32_t)value.size());
    out += value;
}

std::string encode_exr(const float* xyz, uint32_t w, uint32_t h) {
    // Pixels: xyz_to_rec709_e as half, stored per scanline channel by
    // channel in the file's alphabetical channel order B, G, R.
    const uint64_t line_bytes = 6ull * w;
    std::string lines(line_bytes * h, '\0');
    for_rows(h, w, [&](uint32_t y) {
        uint8_t* d = reinterpret_cast<uint8_t*>(&lines[line_bytes * y]);
        for (uint32_t x = 0; x < w; x++) {
            float r[3];
            const float* p = xyz + 3ull * ((uint64_t)y * w + x);
            apply(matrices().xyz_to_rec709_e, p[0], p[1], p[2], r);
            const uint16_t hv[3] = {to_half(r[2]), to_half(r[1]), to_half(r[0])};
            for (int c = 0; c < 3; c++) std::memcpy(d + 2ull * ((uint64_t)c * w + x), &hv[c], 2);
        }
    });
    std::string out;
    put_le32(out, 20000630u);  // magic
    put_le32(out, 2u);         // version 2, single-part scanline
    std::string ch;
    for (const char* name : {"B", "G", "R"}) {
        ch += name;
        ch.push_back('\0');
        put_le32(ch, 1u);                      // HALF
        ch += std::string("\0\0\0\0", 4);      // pLinear, reserved
        put_le32(ch, 1u);                      // xSampling
        put_le32(ch, 1u);                      // ySampling
    }
    ch.push_back('\0');
    exr_attr(out, "channels", "chlist", ch);
    exr_attr(out, "compression", "compression", std::string("\0", 1));  // NO_COMPRESSION
    std::string box;
    put_le32(box, 0);
    put_le32(box, 0);
    put_le32(box, w - 1);
    put_le32(box, h - 1);
    exr_attr(out, "dataWindow", "box2i", box);
    exr_attr(out, "displayWindow", "box2i", box);
    exr_attr(out, "lineOrder", "lineOrder", std::string("\0", 1));  // INCREASING_Y
    const float one = 1.0f, zero = 0.0f;
    std::string f1(reinterpret_cast<const char*>(&one), 4), f0(reinterpret_cast<const char*>(&zero), 4);
    exr_attr(out, "pixelAspectRatio", "float", f1);
    exr_attr(out, "screenWindowCenter", "v2f", f0 + f0);
    exr_attr(out, "screenWindowWidth", "float", f1);
    out.push_back('\0');
    // Offset table, then one chunk per scanline: y, byte count, data.
    uint64_t off = out.size() + 8ull * h;
    for (uint32_t y = 0; y < h; y++) {
        for (int i = 0; i < 8; i++) out.push_back((char)(off >> (8 * i)));
        off += 8 + line_bytes;
    }
    for (uint32_t y = 0; y < h; y++) {
        put_le32(out, y);
        put_le32(out, (uint32_t)line_bytes);
        out.append(lines, line_bytes * y, line_bytes);
    }
    return out;
}

// ---------------------------------------------------------------- base64

std::string base64(const uint8_t* p, uint64_t n) {
    static const char* A = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string s;
    s.reserve((n + 2) / 3 * 4);
    uint64_t i = 0;
    for (; i + 3 <= n; i += 3) {
        const uint32_t v = ((uint32_t)p[i] << 16) | ((uint32_t)p[i + 1] << 8) | p[i + 2];
        s += A[v >> 18];
        s += A[(v >> 12) & 63];
        s += A[(v >> 6) & 63];
        s += A[v & 63];
    }
    if (n - i == 1) {
        const uint32_t v = (uint32_t)p[i] << 16;
        s += A[v >> 18];
        s += A[(v >> 12) & 63];
        s += "==";
    } else if (n - i == 2) {
        const uint32_t v = ((uint32_t)p[i] << 16) | ((uint32_t)p[i + 1] << 8);
        s += A[v >> 18];
        s += A[(v >> 12) & 63];
        s += A[(v >> 6) & 63];
        s += '=';
    }
    return s;
}

int64_t deliver(const std::string& s, uint8_t* out, uint64_t cap) {
    if (cap) std::memcpy(out, s.data(), std::min<uint64_t>(cap, s.size()));
    return (int64_t)s.size();
}

}  // namespace

extern "C" {

int64_t psy_image_encode(const float* xyz, uint32_t width, uint32_t height, uint32_t format, uint8_t* out,
                         uint64_t cap) {
    using psy::api::fail;
    if (!xyz && (uint64_t)width * height != 0) return fail(PSY_E_INVALID, "xyz is NULL");
    if (!out && cap) return fail(PSY_E_INVALID, "out is NULL");
    if (width == 0 || height == 0) return fail(PSY_E_INVALID, "empty image");
    if ((uint64_t)width * height > (1ull << 32)) return fail(PSY_E_RANGE, "image too large");
    try {
        switch (format) {
            case PSY_IMAGE_PPM_ASCII: return deliver(encode_ppm(xyz, width, height, true), out, cap);
            case PSY_IMAGE_PPM: return deliver(encode_ppm(xyz, width, height, false), out, cap);
            case PSY_IMAGE_PNG: return deliver(encode_png(xyz, width, height), out, cap);
            case PSY_IMAGE_EXR: return deliver(encode_exr(xyz, width, height), out, cap);
            default: return fail(PSY_E_INVALID, "unknown image format");
        }
    } catch (const std::bad_alloc&) {
        return fail(PSY_E_NOMEM, "out of host memory while encoding the image");
    }
}

int64_t psy_image_bucket_base64(const float* xyz, uint32_t width, uint32_t height, uint32_t x0, uint32_t y0,
                                uint32_t x1, uint32_t y1, char* out, uint64_t cap) {
    using psy::api::fail;
    if (!xyz && (uint64_t)width * height != 0) return fail(PSY_E_INVALID, "xyz is NULL");
    if (!out && cap) return fail(PSY_E_INVALID, "out is NULL");
    // Image::get_bucket clips the bucket's far corner to the image
    // (image.rs:76-80).
    x1 = std::min(x1, width);
    y1 = std::min(y1, height);
    std::vector<float> data;
    if (x0 < x1 && y0 < y1) {
        data.reserve(4ull * (x1 - x0) * (y1 - y0));
        for (uint32_t y = y0; y < y1; y++)
            for (uint32_t x = x0; x < x1; x++) {
                float r[3];
                const float* p = xyz + 3ull * ((uint64_t)y * width + x);
                apply(matrices().xyz_to_rec709_e, p[0], p[1], p[2], r);
                data.insert(data.end(), {r[0], r[1], r[2], 1.0f});
            }
    }
    const std::string s = base64(reinterpret_cast<const uint8_t*>(data.data()), data.size() * 4);
    return deliver(s, reinterpret_cast<uint8_t*>(out), cap);
}

}  // extern "C"
