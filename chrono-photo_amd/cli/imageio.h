// imageio.h -- frame I/O for the CLI: PNG (8/16-bit gray, gray+alpha, RGB, RGBA,
// non-interlaced) through zlib, binary PPM/PGM (P5/P6, maxval <= 65535), and baseline JPEG
// (jpeg.h; lossy, 8-bit).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

struct Image {
    uint32_t width = 0, height = 0, channels = 0, bytes_per_sample = 1;
    std::vector<uint8_t> data;  // [height][width][channels], 16-bit samples in host byte order
};

bool read_image(const std::string& path, Image* img, std::string* err);
// The format follows the extension; quality applies to .jpg/.jpeg (1-100).
bool write_image(const std::string& path, const Image& img, std::string* err, int quality = 95);
