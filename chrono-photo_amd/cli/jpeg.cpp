// jpeg.cpp -- baseline JPEG decoder and encoder after ITU-T T.81 (sequential DCT, Huffman
// coding), for the CLI's frames.
#include "jpeg.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace {

// kZigzag[k]: the row-major index of the k-th coefficient in zigzag order (anti-diagonals,
// alternating direction, T.81 figure A.6).
struct ZigzagTable {
    int z[64];
    constexpr ZigzagTable() : z{} {
        int k = 0;
        for (int s = 0; s < 15; ++s)
            for (int i = 0; i <= s; ++i) {
                const int r = s % 2 == 0 ? s - i : i, c = s - r;
                if (r < 8 && c < 8) z[k++] = r * 8 + c;
            }
    }
};
constexpr ZigzagTable kZigzag;

// T.81 annex K: quantization tables (row-major) and Huffman tables (code counts per length,
// symbols in code order).
const uint8_t kQuantLuma[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kQuantChroma[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                  24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                  99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                  99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
const uint8_t kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumaVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChromaVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

constexpr double kPi = 3.14159265358979323846;

// cos((2x + 1) u pi / 16) * C(u) / 2, C(0) = 1/sqrt(2): the 1-D DCT basis, both directions.
struct DctBasis {
    float m[8][8];  // [u][x]
    DctBasis() {
        for (int u = 0; u < 8; ++u)
            for (int x = 0; x < 8; ++x)
                m[u][x] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * kPi / 16.0));
    }
};
const DctBasis& basis() {
    static const DctBasis b;
    return b;
}

// coefficients F[v][u] (row-major) -> samples f[y][x]
void idct8x8(const float* F, float* f) {
    const DctBasis& B = basis();
    float t[64];
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < 8; ++x) {
            float s = 0.0f;
            for (int u = 0; u < 8; ++u) s += B.m[u][x] * F[v * 8 + u];
            t[v * 8 + x] = s;
        }
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            float s = 0.0f;
            for (int v = 0; v < 8; ++v) s += B.m[v][y] * t[v * 8 + x];
            f[y * 8 + x] = s;
        }
}

// samples f[y][x] -> coefficients F[v][u]
void fdct8x8(const float* f, float* F) {
    const DctBasis& B = basis();
    float t[64];
    for (int y = 0; y < 8; ++y)
        for (int u = 0; u < 8; ++u) {
            float s = 0.0f;
            for (int x = 0; x < 8; ++x) s += B.m[u][x] * f[y * 8 + x];
            t[y * 8 + u] = s;
        }
    for (int v = 0; v < 8; ++v)
        for (int u = 0; u < 8; ++u) {
            float s = 0.0f;
            for (int y = 0; y < 8; ++y) s += B.m[v][y] * t[y * 8 + u];
            F[v * 8 + u] = s;
        }
}

uint8_t clamp8(float v) {
    const float r = std::floor(v + 0.5f);
    return (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
}

// ------------------------------------------------------------------------------ decoder

struct HuffDecoder {
    bool present = false;
    uint8_t vals[256] = {};
    int32_t mincode[17] = {}, maxcode[18] = {}, valptr[17] = {};
    bool build(const uint8_t* bits, const uint8_t* v, int nv) {
        int code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            valptr[l] = k;
            mincode[l] = code;
            code += bits[l - 1];
            k += bits[l - 1];
            maxcode[l] = bits[l - 1] ? code - 1 : -1;
            if (code > (1 << l)) return false;  // more codes than the length holds
            code <<= 1;
        }
        if (k != nv || k > 256) return false;
        memcpy(vals, v, (size_t)k);
        present = true;
        return true;
    }
};

struct BitReader {
    const uint8_t* p = nullptr;
    size_t n = 0, pos = 0;
    uint32_t acc = 0;
    int cnt = 0;
    bool at_marker = false;  // a marker stops the data: zero bits are supplied past it
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!at_marker && pos < n) {
                b = p[pos];
                if (b == 0xFF) {
                    const uint8_t nx = pos + 1 < n ? p[pos + 1] : 0xD9;
                    if (nx == 0x00) pos += 2;
                    else {
                        at_marker = true;
                        b = 0;
                    }
                } else {
                    ++pos;
                }
            }
            acc |= b << (24 - cnt);
            cnt += 8;
        }
    }
    uint32_t bits(int k) {  // k <= 16
        if (k == 0) return 0;
        if (cnt < k) fill();
        const uint32_t v = acc >> (32 - k);
        acc <<= k;
        cnt -= k;
        return v;
    }
    int decode(const HuffDecoder& h, bool* ok) {
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | (int)bits(1);
            if (code <= h.maxcode[l]) return h.vals[h.valptr[l] + code - h.mincode[l]];
        }
        *ok = false;
        return 0;
    }
    // At a restart: drop the interval's padding bits and step over the RSTn marker.
    bool restart() {
        acc = 0;
        cnt = 0;
        while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) ++pos;
        if (pos + 1 < n) {
            pos += 2;
            at_marker = false;
            return true;
        }
        return false;
    }
};

int32_t extend(uint32_t v, int s) { return s == 0 ? 0 : (v < (1u << (s - 1)) ? (int32_t)v - (1 << s) + 1 : (int32_t)v); }

struct Component {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int pw = 0, ph = 0;  // plane size in samples (whole blocks)
    std::vector<uint8_t> plane;
    int32_t pred = 0;
};

uint16_t be16(const uint8_t* p) { return (uint16_t)(p[0] << 8 | p[1]); }

}  // namespace

bool read_jpeg(const std::vector<uint8_t>& f, Image* img, std::string* err) {
    auto bad = [&](const char* what) {
        *err = std::string("JPEG: ") + what;
        return false;
    };
    if (f.size() < 4 || f[0] != 0xFF || f[1] != 0xD8) return bad("not a JPEG file");
    uint16_t qt[4][64] = {};
    bool qt_present[4] = {};
    HuffDecoder dc[4], ac[4];
    std::vector<Component> comps;
    int X = 0, Y = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0, restart_interval = 0;
    bool frame = false, adobe_rgb = false, any_scan = false;
    size_t pos = 2;
    while (pos + 4 <= f.size()) {
        if (f[pos] != 0xFF) return bad("marker expected");
        const uint8_t m = f[pos + 1];
        if (m == 0xFF) {
            ++pos;
            continue;
        }
        if (m == 0xD9) break;  // EOI
        const size_t len = be16(&f[pos + 2]);
        if (len < 2 || pos + 2 + len > f.size()) return bad("truncated segment");
        const uint8_t* s = &f[pos + 4];
        const size_t sl = len - 2;
        if (m == 0xDB) {  // DQT
            size_t i = 0;
            while (i < sl) {
                const int pq = s[i] >> 4, tq = s[i] & 15;
                if (tq > 3 || i + 1 + 64 * (pq + 1) > sl) return bad("bad quantization table");
                for (int k = 0; k < 64; ++k)
                    qt[tq][kZigzag.z[k]] = pq ? be16(&s[i + 1 + 2 * k]) : s[i + 1 + k];
                qt_present[tq] = true;
                i += 1 + 64 * (pq + 1);
            }
        } else if (m == 0xC4) {  // DHT
            size_t i = 0;
            while (i + 17 <= sl) {
                const int tc = s[i] >> 4, th = s[i] & 15;
                int nv = 0;
                for (int l = 0; l < 16; ++l) nv += s[i + 1 + l];
                if (tc > 1 || th > 3 || i + 17 + nv > sl) return bad("bad Huffman table");
                HuffDecoder& h = tc ? ac[th] : dc[th];
                if (!h.build(&s[i + 1], &s[i + 17], nv)) return bad("bad Huffman table");
                i += 17 + nv;
            }
        } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1: sequential Huffman
            if (sl < 6 || s[0] != 8) return bad("only 8-bit samples are supported");
            Y = be16(&s[1]);
            X = be16(&s[3]);
            const int nf = s[5];
            if (X == 0 || Y == 0) return bad("image size 0 (DNL) is not supported");
            if ((nf != 1 && nf != 3) || sl < 6 + 3 * (size_t)nf) return bad("1 or 3 components are supported");
            comps.resize(nf);
            for (int c = 0; c < nf; ++c) {
                comps[c].id = s[6 + 3 * c];
                comps[c].h = s[7 + 3 * c] >> 4;
                comps[c].v = s[7 + 3 * c] & 15;
                comps[c].tq = s[8 + 3 * c];
                if (comps[c].h < 1 || comps[c].h > 2 || comps[c].v < 1 || comps[c].v > 2 || comps[c].tq > 3)
                    return bad("unsupported sampling factors");
                hmax = std::max(hmax, comps[c].h);
                vmax = std::max(vmax, comps[c].v);
            }
            mcux = (X + 8 * hmax - 1) / (8 * hmax);
            mcuy = (Y + 8 * vmax - 1) / (8 * vmax);
            for (auto& c : comps) {
                c.pw = mcux * c.h * 8;
                c.ph = mcuy * c.v * 8;
                c.plane.assign((size_t)c.pw * c.ph, 0);
            }
            frame = true;
        } else if ((m >= 0xC2 && m <= 0xC3) || (m >= 0xC5 && m <= 0xC7) || (m >= 0xC9 && m <= 0xCB) ||
                   (m >= 0xCD && m <= 0xCF)) {
            return bad("progressive, lossless and arithmetic-coded JPEG are not supported");
        } else if (m == 0xDD) {  // DRI
            if (sl < 2) return bad("bad restart interval");
            restart_interval = be16(s);
        } else if (m == 0xEE) {  // APP14 (Adobe): transform 0 means RGB components
            if (sl >= 12 && !memcmp(s, "Adobe", 5)) adobe_rgb = s[11] == 0;
        } else if (m == 0xDA) {  // SOS
            if (!frame) return bad("scan before frame header");
            const int ns = s[0];
            if (ns < 1 || ns > 4 || sl < 4 + 2 * (size_t)ns) return bad("bad scan header");
            std::vector<Component*> sc;
            for (int i = 0; i < ns; ++i) {
                Component* c = nullptr;
                for (auto& k : comps)
                    if (k.id == s[1 + 2 * i]) c = &k;
                if (!c) return bad("scan names an unknown component");
                c->td = s[2 + 2 * i] >> 4;
                c->ta = s[2 + 2 * i] & 15;
                if (c->td > 3 || c->ta > 3 || !dc[c->td].present || !ac[c->ta].present || !qt_present[c->tq])
                    return bad("scan uses a missing table");
                c->pred = 0;
                sc.push_back(c);
            }
            BitReader br;
            br.p = f.data();
            br.n = f.size();
            br.pos = pos + 2 + len;
            // blocks of the scan: interleaved MCUs, or one component's blocks in raster order
            int units_x, units_y;
            if (ns == 1) {
                const Component& c = *sc[0];
                units_x = ((X * c.h + hmax - 1) / hmax + 7) / 8;
                units_y = ((Y * c.v + vmax - 1) / vmax + 7) / 8;
            } else {
                units_x = mcux;
                units_y = mcuy;
            }
            bool ok = true;
            float F[64], out[64];
            auto block = [&](Component& c, int bx, int by) {
                int32_t coef[64] = {};
                const int t = br.decode(dc[c.td], &ok);
                if (!ok || t > 11) return (ok = false);
                // a corrupt stream could walk the predictor out of int32 range: valid data stays
                // within a few thousand, so the clamp changes nothing for it
                c.pred = std::max(-(1 << 20), std::min(1 << 20, c.pred + extend(br.bits(t), t)));
                coef[0] = c.pred;
                for (int k = 1; k < 64;) {
                    const int rs = br.decode(ac[c.ta], &ok);
                    if (!ok) return false;
                    const int r = rs >> 4, sz = rs & 15;
                    if (sz == 0) {
                        if (r != 15) break;  // EOB
                        k += 16;
                        continue;
                    }
                    k += r;
                    if (k > 63 || sz > 10) return (ok = false);
                    coef[kZigzag.z[k]] = extend(br.bits(sz), sz);
                    ++k;
                }
                // the product of two exact floats rounds once, like the float of the integer product,
                // without the integer overflow a corrupt coefficient could cause
                for (int i = 0; i < 64; ++i) F[i] = (float)coef[i] * (float)qt[c.tq][i];
                idct8x8(F, out);
                for (int y = 0; y < 8; ++y)
                    for (int x = 0; x < 8; ++x)
                        c.plane[(size_t)(by * 8 + y) * c.pw + bx * 8 + x] = clamp8(out[y * 8 + x] + 128.0f);
                return true;
            };
            int since_restart = 0;
            for (int uy = 0; uy < units_y && ok; ++uy)
                for (int ux = 0; ux < units_x && ok; ++ux) {
                    if (restart_interval && since_restart == restart_interval) {
                        if (!br.restart()) return bad("missing restart marker");
                        for (Component* c : sc) c->pred = 0;
                        since_restart = 0;
                    }
                    if (ns == 1) {
                        block(*sc[0], ux, uy);
                    } else {
                        for (Component* c : sc)
                            for (int v = 0; v < c->v && ok; ++v)
                                for (int h = 0; h < c->h && ok; ++h) block(*c, ux * c->h + h, uy * c->v + v);
                    }
                    ++since_restart;
                }
            if (!ok) return bad("corrupt entropy-coded data");
            any_scan = true;
            // continue after the scan's data: the next marker that is not RSTn or a stuffed byte
            size_t q = br.pos;
            while (q + 1 < f.size() && !(f[q] == 0xFF && f[q + 1] != 0x00 && (f[q + 1] < 0xD0 || f[q + 1] > 0xD7)))
                ++q;
            pos = q;
            continue;
        }
        pos += 2 + len;
    }
    if (!frame || !any_scan) return bad("no image data");
    img->width = (uint32_t)X;
    img->height = (uint32_t)Y;
    img->bytes_per_sample = 1;
    img->channels = comps.size() == 1 ? 1 : 3;
    img->data.assign((size_t)X * Y * img->channels, 0);
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            uint8_t* o = &img->data[((size_t)y * X + x) * img->channels];
            float s[3];
            for (size_t c = 0; c < comps.size(); ++c) {
                const Component& k = comps[c];
                const int cx = x * k.h / hmax, cy = y * k.v / vmax;  // nearest sample of the plane
                s[c] = k.plane[(size_t)cy * k.pw + cx];
            }
            if (comps.size() == 1) {
                o[0] = (uint8_t)s[0];
            } else if (adobe_rgb) {
                for (int c = 0; c < 3; ++c) o[c] = (uint8_t)s[c];
            } else {  // JFIF YCbCr -> RGB
                const float cb = s[1] - 128.0f, cr = s[2] - 128.0f;
                o[0] = clamp8(s[0] + 1.402f * cr);
                o[1] = clamp8(s[0] - 0.344136f * cb - 0.714136f * cr);
                o[2] = clamp8(s[0] + 1.772f * cb);
            }
        }
    return true;
}

// ------------------------------------------------------------------------------ encoder

namespace {

struct HuffEncoder {
    uint16_t code[256] = {};
    uint8_t size[256] = {};
    void build(const uint8_t* bits, const uint8_t* vals) {
        int c = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            for (int i = 0; i < bits[l - 1]; ++i, ++k) {
                code[vals[k]] = (uint16_t)c++;
                size[vals[k]] = (uint8_t)l;
            }
            c <<= 1;
        }
    }
};

struct BitWriter {
    std::vector<uint8_t>* out;
    uint32_t acc = 0;
    int cnt = 0;
    void put(uint32_t v, int k) {  // k <= 16
        acc = (acc << k) | (v & ((1u << k) - 1u));
        cnt += k;
        while (cnt >= 8) {
            const uint8_t b = (uint8_t)(acc >> (cnt - 8));
            out->push_back(b);
            if (b == 0xFF) out->push_back(0x00);
            cnt -= 8;
        }
    }
    void flush() {  // pad with one bits
        if (cnt > 0) put(0x7F, 8 - cnt);
        acc = 0;
    }
};

int bit_length(int32_t v) {
    uint32_t a = (uint32_t)(v < 0 ? -v : v);
    int n = 0;
    while (a) {
        ++n;
        a >>= 1;
    }
    return n;
}

void put16(std::vector<uint8_t>& o, int v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

}  // namespace

bool write_jpeg(const std::string& path, const Image& img, int quality, bool subsample, int restart,
                std::string* err) {
    if (img.bytes_per_sample != 1) {
        *err = path + ": JPEG holds 8-bit samples";
        return false;
    }
    if (img.width == 0 || img.height == 0 || img.width > 65535 || img.height > 65535) {
        *err = path + ": image size out of JPEG's range";
        return false;
    }
    const int W = (int)img.width, H = (int)img.height, C = (int)img.channels;
    const int nc = C >= 3 ? 3 : 1;
    const bool sub = subsample && nc == 3;
    quality = std::min(100, std::max(1, quality));
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;  // IJG scaling
    uint8_t qt[2][64];
    for (int i = 0; i < 64; ++i) {
        qt[0][i] = (uint8_t)std::min(255, std::max(1, (kQuantLuma[i] * scale + 50) / 100));
        qt[1][i] = (uint8_t)std::min(255, std::max(1, (kQuantChroma[i] * scale + 50) / 100));
    }
    // level-shifted planes: Y, Cb, Cr (JFIF) or gray
    std::vector<float> plane[3];
    for (int c = 0; c < nc; ++c) plane[c].resize((size_t)W * H);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const uint8_t* p = &img.data[i * C];
        if (nc == 1) {
            plane[0][i] = p[0] - 128.0f;
        } else {
            const float r = p[0], g = p[1], b = p[2];
            plane[0][i] = 0.299f * r + 0.587f * g + 0.114f * b - 128.0f;
            plane[1][i] = -0.168736f * r - 0.331264f * g + 0.5f * b;
            plane[2][i] = 0.5f * r - 0.418688f * g - 0.081312f * b;
        }
    }
    HuffEncoder hdc[2], hac[2];
    hdc[0].build(kDcLumaBits, kDcVals);
    hdc[1].build(kDcChromaBits, kDcVals);
    hac[0].build(kAcLumaBits, kAcLumaVals);
    hac[1].build(kAcChromaBits, kAcChromaVals);

    std::vector<uint8_t> o;
    o.insert(o.end(), {0xFF, 0xD8, 0xFF, 0xE0, 0x00, 0x10, 'J', 'F', 'I', 'F', 0x00, 0x01, 0x01, 0x00, 0x00, 0x01,
                       0x00, 0x01, 0x00, 0x00});
    for (int t = 0; t < (nc == 3 ? 2 : 1); ++t) {  // DQT
        o.insert(o.end(), {0xFF, 0xDB});
        put16(o, 67);
        o.push_back((uint8_t)t);
        for (int k = 0; k < 64; ++k) o.push_back(qt[t][kZigzag.z[k]]);
    }
    o.insert(o.end(), {0xFF, 0xC0});  // SOF0
    put16(o, 8 + 3 * nc);
    o.push_back(8);
    put16(o, H);
    put16(o, W);
    o.push_back((uint8_t)nc);
    for (int c = 0; c < nc; ++c) {
        o.push_back((uint8_t)(c + 1));
        o.push_back(c == 0 && sub ? 0x22 : 0x11);
        o.push_back(c == 0 ? 0 : 1);
    }
    auto dht = [&](int tc, int th, const uint8_t* bits, const uint8_t* vals) {
        int nv = 0;
        for (int l = 0; l < 16; ++l) nv += bits[l];
        o.insert(o.end(), {0xFF, 0xC4});
        put16(o, 19 + nv);
        o.push_back((uint8_t)(tc << 4 | th));
        o.insert(o.end(), bits, bits + 16);
        o.insert(o.end(), vals, vals + nv);
    };
    dht(0, 0, kDcLumaBits, kDcVals);
    dht(1, 0, kAcLumaBits, kAcLumaVals);
    if (nc == 3) {
        dht(0, 1, kDcChromaBits, kDcVals);
        dht(1, 1, kAcChromaBits, kAcChromaVals);
    }
    if (restart > 0) {
        o.insert(o.end(), {0xFF, 0xDD, 0x00, 0x04});
        put16(o, std::min(restart, 65535));
    }
    o.insert(o.end(), {0xFF, 0xDA});  // SOS
    put16(o, 6 + 2 * nc);
    o.push_back((uint8_t)nc);
    for (int c = 0; c < nc; ++c) {
        o.push_back((uint8_t)(c + 1));
        o.push_back(c == 0 ? 0x00 : 0x11);
    }
    o.insert(o.end(), {0x00, 0x3F, 0x00});

    BitWriter bw{&o};
    int32_t pred[3] = {0, 0, 0};
    const int mcu = sub ? 16 : 8;
    const int mx = (W + mcu - 1) / mcu, my = (H + mcu - 1) / mcu;
    float f[64], F[64];
    // one 8x8 block of component c whose top-left sample is (x0, y0) in the component's grid
    // (chroma of 4:2:0 at half resolution: the mean of each 2x2 group); edges repeat
    auto encode_block = [&](int c, int x0, int y0) {
        const int step = sub && c > 0 ? 2 : 1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                float s = 0.0f;
                for (int dy = 0; dy < step; ++dy)
                    for (int dx = 0; dx < step; ++dx) {
                        const int px = std::min(W - 1, (x0 + x) * step + dx), py = std::min(H - 1, (y0 + y) * step + dy);
                        s += plane[c][(size_t)py * W + px];
                    }
                f[y * 8 + x] = s / (float)(step * step);
            }
        fdct8x8(f, F);
        const int t = c == 0 ? 0 : 1;
        int32_t q[64];
        for (int k = 0; k < 64; ++k) {
            const float v = F[kZigzag.z[k]] / (float)qt[t][kZigzag.z[k]];
            q[k] = (int32_t)(v < 0.0f ? -std::floor(-v + 0.5f) : std::floor(v + 0.5f));
        }
        const int32_t diff = q[0] - pred[c];
        pred[c] = q[0];
        const int s = bit_length(diff);
        bw.put(hdc[t].code[s], hdc[t].size[s]);
        if (s) bw.put((uint32_t)(diff < 0 ? diff - 1 : diff), s);
        int run = 0;
        for (int k = 1; k < 64; ++k) {
            if (q[k] == 0) {
                ++run;
                continue;
            }
            while (run > 15) {
                bw.put(hac[t].code[0xF0], hac[t].size[0xF0]);
                run -= 16;
            }
            const int sz = bit_length(q[k]);
            const int rs = run << 4 | sz;
            bw.put(hac[t].code[rs], hac[t].size[rs]);
            bw.put((uint32_t)(q[k] < 0 ? q[k] - 1 : q[k]), sz);
            run = 0;
        }
        if (run) bw.put(hac[t].code[0x00], hac[t].size[0x00]);
    };
    int n = 0, rst = 0;
    for (int by = 0; by < my; ++by)
        for (int bx = 0; bx < mx; ++bx) {
            if (restart > 0 && n == restart) {
                bw.flush();
                o.push_back(0xFF);
                o.push_back((uint8_t)(0xD0 + (rst++ & 7)));
                pred[0] = pred[1] = pred[2] = 0;
                n = 0;
            }
            if (sub) {
                for (int v = 0; v < 2; ++v)
                    for (int h = 0; h < 2; ++h) encode_block(0, bx * 16 + h * 8, by * 16 + v * 8);
                encode_block(1, bx * 8, by * 8);
                encode_block(2, bx * 8, by * 8);
            } else {
                for (int c = 0; c < nc; ++c) encode_block(c, bx * 8, by * 8);
            }
            ++n;
        }
    bw.flush();
    o.push_back(0xFF);
    o.push_back(0xD9);
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp || fwrite(o.data(), 1, o.size(), fp) != o.size()) {
        if (fp) fclose(fp);
        *err = "cannot write " + path;
        return false;
    }
    fclose(fp);
    return true;
}
