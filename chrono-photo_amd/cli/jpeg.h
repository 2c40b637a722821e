// jpeg.h -- baseline JPEG for the CLI's frames: the reference reads its inputs through the
// `image` crate (JPEG among them) and writes `.jpg` outputs at `--quality` (src/main.rs:541-551).
// Decoding: sequential Huffman (SOF0/SOF1), 8-bit, 1 or 3 components, any 1-2 sampling factors,
// restart intervals.  Encoding: baseline, standard tables scaled by quality.  Lossy: pixel values
// after a decode depend on the decoder's IDCT and upsampling, so composites of JPEG frames are
// not bit-comparable with another decoder's.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "imageio.h"

bool read_jpeg(const std::vector<uint8_t>& file, Image* img, std::string* err);

// quality 1..100 (IJG scaling of the standard tables); subsample: 2x2 chroma (4:2:0) instead of
// full-resolution chroma; restart: MCUs per restart interval (0: none).  8-bit images with 1-4
// channels: gray+alpha keeps the gray, RGBA the RGB.
bool write_jpeg(const std::string& path, const Image& img, int quality, bool subsample, int restart,
                std::string* err);
