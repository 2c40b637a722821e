// Test driver for the CLI's image codec: reads argv[1], prints "w h channels bytes", and
// writes the decoded samples to argv[2] through the encoder chosen by its extension.  For a
// JPEG output, optional argv[3..5] = quality, 4:2:0 chroma (0/1), MCUs per restart interval.
#include <cstdio>
#include <cstdlib>

#include "../../chrono-photo_amd/cli/imageio.h"
#include "../../chrono-photo_amd/cli/jpeg.h"

int main(int argc, char** argv) {
    if (argc != 3 && argc != 6) return 2;
    Image img;
    std::string err;
    if (!read_image(argv[1], &img, &err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    printf("%u %u %u %u\n", img.width, img.height, img.channels, img.bytes_per_sample);
    const bool ok = argc == 6 ? write_jpeg(argv[2], img, atoi(argv[3]), atoi(argv[4]) != 0, atoi(argv[5]), &err)
                              : write_image(argv[2], img, &err);
    if (!ok) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    return 0;
}
