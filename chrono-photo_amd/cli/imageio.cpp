// imageio.cpp -- PNG and PPM/PGM codecs for the CLI, and the format dispatch (see imageio.h).
#include "imageio.h"

#include "jpeg.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>

namespace {

bool ends_with(std::string s, const char* suf) {
    std::transform(s.begin(), s.end(), s.begin(), ::tolower);
    const size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

bool read_file(const std::string& path, std::vector<uint8_t>* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(x >> 24); v.push_back(x >> 16); v.push_back(x >> 8); v.push_back(x);
}

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

bool read_png(const std::vector<uint8_t>& f, Image* img, std::string* err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || memcmp(f.data(), sig, 8) != 0) { *err = "not a PNG file"; return false; }
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat;
    while (pos + 8 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        const std::string type((const char*)&f[pos + 4], 4);
        if (pos + 12 + len > f.size()) { *err = "truncated PNG chunk"; return false; }
        const uint8_t* d = &f[pos + 8];
        if (type == "IHDR") {
            w = be32(d); h = be32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12];
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        pos += 12 + len;
    }
    int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (!ch || (depth != 8 && depth != 16)) { *err = "PNG must be 8/16-bit gray, gray+alpha, RGB or RGBA"; return false; }
    if (interlace) { *err = "interlaced PNG is not supported"; return false; }
    const size_t bpp = (size_t)ch * (depth / 8), stride = (size_t)w * bpp;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf rlen = raw.size();
    if (uncompress(raw.data(), &rlen, idat.data(), idat.size()) != Z_OK || rlen != raw.size()) {
        *err = "PNG image data does not inflate to the expected size"; return false;
    }
    img->width = w; img->height = h; img->channels = ch; img->bytes_per_sample = depth / 8;
    img->data.assign(stride * h, 0);
    std::vector<uint8_t> prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        uint8_t* row = &raw[y * (stride + 1) + 1];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = row[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: *err = "bad PNG filter type"; return false;
            }
            row[i] = (uint8_t)v;
        }
        memcpy(prev.data(), row, stride);
        if (depth == 8) {
            memcpy(&img->data[y * stride], row, stride);
        } else {  // big-endian samples to host order
            uint16_t* o = (uint16_t*)&img->data[y * stride];
            for (size_t i = 0; i < stride / 2; ++i) o[i] = (uint16_t)(row[2 * i] << 8 | row[2 * i + 1]);
        }
    }
    return true;
}

bool write_png(const std::string& path, const Image& img, std::string* err) {
    const int ctype = img.channels == 1 ? 0 : img.channels == 2 ? 4 : img.channels == 3 ? 2 : 6;
    const size_t stride = (size_t)img.width * img.channels * img.bytes_per_sample;
    std::vector<uint8_t> raw((stride + 1) * img.height);
    for (uint32_t y = 0; y < img.height; ++y) {
        raw[y * (stride + 1)] = 0;
        uint8_t* dst = &raw[y * (stride + 1) + 1];
        const uint8_t* src = &img.data[y * stride];
        if (img.bytes_per_sample == 1) memcpy(dst, src, stride);
        else
            for (size_t i = 0; i < stride / 2; ++i) {
                const uint16_t v = ((const uint16_t*)src)[i];
                dst[2 * i] = v >> 8;
                dst[2 * i + 1] = v & 0xFF;
            }
    }
    uLongf clen = compressBound(raw.size());
    std::vector<uint8_t> comp(clen);
    if (compress2(comp.data(), &clen, raw.data(), raw.size(), 6) != Z_OK) { *err = "deflate failed"; return false; }
    comp.resize(clen);
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    auto chunk = [&](const char* type, const std::vector<uint8_t>& body) {
        put32(out, (uint32_t)body.size());
        const size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), body.begin(), body.end());
        put32(out, (uint32_t)crc32(0, &out[start], (uInt)(out.size() - start)));
    };
    std::vector<uint8_t> ihdr;
    put32(ihdr, img.width);
    put32(ihdr, img.height);
    ihdr.push_back((uint8_t)(8 * img.bytes_per_sample));
    ihdr.push_back((uint8_t)ctype);
    ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    chunk("IHDR", ihdr);
    chunk("IDAT", comp);
    chunk("IEND", {});
    std::ofstream f(path, std::ios::binary);
    if (!f) { *err = "cannot create " + path; return false; }
    f.write((const char*)out.data(), (std::streamsize)out.size());
    return (bool)f;
}

bool read_pnm(const std::vector<uint8_t>& f, Image* img, std::string* err) {
    size_t pos = 0;
    auto token = [&]() -> std::string {
        while (pos < f.size()) {
            if (f[pos] == '#') { while (pos < f.size() && f[pos] != '\n') ++pos; }
            else if (isspace(f[pos])) ++pos;
            else break;
        }
        std::string t;
        while (pos < f.size() && !isspace(f[pos])) t += (char)f[pos++];
        return t;
    };
    const std::string magic = token();
    if (magic != "P5" && magic != "P6") { *err = "only binary PGM (P5) / PPM (P6) are supported"; return false; }
    const long w = atol(token().c_str()), h = atol(token().c_str()), maxv = atol(token().c_str());
    ++pos;  // single whitespace before the raster
    if (w <= 0 || h <= 0 || maxv <= 0 || maxv > 65535) { *err = "bad PNM header"; return false; }
    img->width = (uint32_t)w; img->height = (uint32_t)h;
    img->channels = magic == "P6" ? 3 : 1;
    img->bytes_per_sample = maxv > 255 ? 2 : 1;
    const size_t n = (size_t)w * h * img->channels, bytes = n * img->bytes_per_sample;
    if (pos + bytes > f.size()) { *err = "truncated PNM raster"; return false; }
    img->data.assign(bytes, 0);
    if (img->bytes_per_sample == 1) memcpy(img->data.data(), &f[pos], bytes);
    else
        for (size_t i = 0; i < n; ++i) ((uint16_t*)img->data.data())[i] = (uint16_t)(f[pos + 2 * i] << 8 | f[pos + 2 * i + 1]);
    return true;
}

bool write_pnm(const std::string& path, const Image& img, std::string* err) {
    if (img.channels != 1 && img.channels != 3) { *err = "PPM/PGM output needs 1 or 3 channels"; return false; }
    std::ofstream f(path, std::ios::binary);
    if (!f) { *err = "cannot create " + path; return false; }
    f << (img.channels == 3 ? "P6" : "P5") << "\n" << img.width << " " << img.height << "\n"
      << (img.bytes_per_sample == 2 ? 65535 : 255) << "\n";
    const size_t n = (size_t)img.width * img.height * img.channels;
    if (img.bytes_per_sample == 1) f.write((const char*)img.data.data(), (std::streamsize)n);
    else
        for (size_t i = 0; i < n; ++i) {
            const uint16_t v = ((const uint16_t*)img.data.data())[i];
            const char b[2] = {(char)(v >> 8), (char)(v & 0xFF)};
            f.write(b, 2);
        }
    return (bool)f;
}

}  // namespace

bool read_image(const std::string& path, Image* img, std::string* err) {
    std::vector<uint8_t> f;
    if (!read_file(path, &f)) { *err = "cannot read " + path; return false; }
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P') return read_png(f, img, err);
    if (f.size() >= 2 && f[0] == 'P') return read_pnm(f, img, err);
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8) {
        if (read_jpeg(f, img, err)) return true;
        *err = path + ": " + *err;
        return false;
    }
    *err = path + ": unsupported format (PNG, PPM, PGM and baseline JPEG are read)";
    return false;
}

bool write_image(const std::string& path, const Image& img, std::string* err, int quality) {
    if (ends_with(path, ".png")) return write_png(path, img, err);
    if (ends_with(path, ".ppm") || ends_with(path, ".pgm") || ends_with(path, ".pnm")) return write_pnm(path, img, err);
    if (ends_with(path, ".jpg") || ends_with(path, ".jpeg")) return write_jpeg(path, img, quality, false, 0, err);
    *err = path + ": output must be .png, .ppm, .pgm or .jpg";
    return false;
}
