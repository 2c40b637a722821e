// Test driver for the CLI's image codec: reads argv[1], prints "w h channels bytes", and
// writes the decoded samples to argv[2] through the encoder chosen by its extension.
#include <cstdio>

#include "../../chrono-photo_amd/cli/imageio.h"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    Image img;
    std::string err;
    if (!read_image(argv[1], &img, &err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    printf("%u %u %u %u\n", img.width, img.height, img.channels, img.bytes_per_sample);
    if (!write_image(argv[2], img, &err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    return 0;
}
