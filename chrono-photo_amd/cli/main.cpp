// chrono-photo-hip -- command-line drop-in for chrono-photo's compositing modes on MI355X.
//
// Options follow the reference CLI (src/cli.rs:19-136) with its defaults (src/cli.rs:174-231):
// --mode outlier, --threshold abs/0.05/0.2, --background random, --outlier extreme,
// --weights 1 1 1 1, no fade.  A single non-flag argument is read as an options file, as in
// src/main.rs:33-43 with the quote handling of src/cli.rs:245-266.  The flow is
// src/main.rs:21-98: list the files (glob order, --frames), optional shake analysis and crop,
// then one composite or one per video window (src/main.rs:214-429).  Frames are decoded once
// into a device-resident stack; the time-slicing flags (--temp-dir, --slice, --compression)
// and the thread-count flags only tune the reference's disk pipeline and are accepted with a
// warning.  Images: PNG (8/16-bit), binary PPM/PGM and baseline JPEG (outputs at --quality).
// One option is this tool's own:
// --hbm-budget <size>[K|M|G] caps the device memory of the run; a stack larger than that
// stays in host memory and the composites stream it through the GPU in row bands.
#include <glob.h>
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/chrono_hip.h"
#include "imageio.h"

namespace {

[[noreturn]] void die(const std::string& msg) {
    fprintf(stderr, "error: %s\n", msg.c_str());
    exit(1);
}

void check(int rc, const char* what) {
    if (rc != CP_OK) die(std::string(what) + ": " + cp_last_error());
}

std::vector<std::string> split(const std::string& s, char sep) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (c == sep) { out.push_back(cur); cur.clear(); }
        else cur += c;
    }
    out.push_back(cur);
    return out;
}

struct Range {  // FrameRange, src/flist.rs:11-79
    bool has_start = false, has_end = false;
    int start = 0, end = 0;
    unsigned step = 1;
};

Range parse_range(const std::string& s, const char* opt) {
    auto p = split(s, '/');
    if (p.size() != 3) die(std::string("Option ") + opt + " expects 3 elements: start/end/step");
    Range r;
    if (p[0] != ".") { r.has_start = true; r.start = std::stoi(p[0]); }
    if (p[1] != ".") { r.has_end = true; r.end = std::stoi(p[1]); }
    if (p[2] != ".") r.step = (unsigned)std::stoi(p[2]);
    if (r.step == 0) die(std::string("Option ") + opt + ": step must be positive");
    return r;
}

struct Args {
    std::string pattern, output, output_blend, mode = "outlier", background = "random", outlier = "extreme";
    std::string threshold = "abs/0.05/0.2", fade, shake;
    std::vector<std::string> anchors;
    bool has_frames = false, has_vin = false, has_vout = false, debug = false;
    Range frames, vin, vout;
    float weights[4] = {1, 1, 1, 1};
    long sample = -1;
    uint64_t seed = 0;
    uint64_t hbm_budget = 0;  // bytes of device memory the run may use (0: no limit)
    int quality = 95;         // JPEG outputs (src/cli.rs:82-84)
    std::vector<std::string> unused;
};

// "<number>[K|M|G]" in bytes (powers of 1024).
uint64_t parse_bytes(const std::string& s) {
    if (s.empty()) die("--hbm-budget expects a size");
    uint64_t mul = 1;
    std::string num = s;
    const char u = (char)toupper((unsigned char)s.back());
    if (u == 'K' || u == 'M' || u == 'G') {
        mul = u == 'K' ? 1ull << 10 : u == 'M' ? 1ull << 20 : 1ull << 30;
        num = s.substr(0, s.size() - 1);
    }
    size_t used = 0;
    uint64_t v = 0;
    try {
        v = (uint64_t)std::stoull(num, &used);
    } catch (...) {
        used = 0;
    }
    if (used == 0 || used != num.size()) die("--hbm-budget expects a size such as 512M or 8G, got " + s);
    return v * mul;
}

// The reference's option-file tokenizer (src/cli.rs:245-266): quotes group, spaces split.
std::vector<std::string> tokenize_options(const std::string& text) {
    std::vector<std::string> out;
    auto parts = split(text, '"');
    for (size_t i = 0; i < parts.size(); ++i) {
        std::string part = parts[i];
        if (i % 2 == 1) { out.push_back(part); continue; }
        std::istringstream ss(part);
        std::string tok;
        while (ss >> tok) out.push_back(tok);
    }
    return out;
}

Args parse_args(std::vector<std::string> argv) {
    Args a;
    bool has_thr = false, has_bg = false, has_out = false;
    for (size_t i = 0; i < argv.size(); ++i) {
        const std::string& o = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argv.size()) die("missing value for " + o);
            return argv[++i];
        };
        if (o == "-p" || o == "--pattern") a.pattern = val();
        else if (o == "-o" || o == "--output") a.output = val();
        else if (o == "--output-blend") a.output_blend = val();
        else if (o == "-m" || o == "--mode") a.mode = val();
        else if (o == "-t" || o == "--threshold") { a.threshold = val(); has_thr = true; }
        else if (o == "-b" || o == "--background") { a.background = val(); has_bg = true; }
        else if (o == "-l" || o == "--outlier") { a.outlier = val(); has_out = true; }
        else if (o == "-f" || o == "--frames") { a.frames = parse_range(val(), "--frames"); a.has_frames = true; }
        else if (o == "--video-in") { a.vin = parse_range(val(), "--video-in"); a.has_vin = true; }
        else if (o == "--video-out") { a.vout = parse_range(val(), "--video-out"); a.has_vout = true; }
        else if (o == "--fade") a.fade = val();
        else if (o == "--sample") a.sample = std::stol(val());
        else if (o == "--seed") a.seed = std::stoull(val());
        else if (o == "--hbm-budget") a.hbm_budget = parse_bytes(val());
        else if (o == "--weights") {
            for (int k = 0; k < 4; ++k) a.weights[k] = std::stof(val());
        } else if (o == "--shake") a.shake = val();
        else if (o == "--shake-anchors") {
            while (i + 1 < argv.size() && argv[i + 1][0] != '-') a.anchors.push_back(argv[++i]);
        } else if (o == "-d" || o == "--debug") a.debug = true;
        else if (o == "-w" || o == "--wait") a.unused.push_back("--wait");
        else if (o == "-q" || o == "--quality") {
            const std::string v = val();
            if (std::stoi(v) < 1 || std::stoi(v) > 100) die("Expected 0 < qualtiy <= 100. Got value " + v);
            a.quality = std::stoi(v);
        } else if (o == "--temp-dir" || o == "-c" || o == "--compression" || o == "-s" || o == "--slice" ||
                   o == "--threads" || o == "--video-threads" || o == "--shake-threads") {
            (void)val();
            a.unused.push_back(o);
        } else die("unknown option " + o);
    }
    if (a.pattern.empty() || a.output.empty()) die("--pattern and --output are required");
    if (a.shake.empty() != a.anchors.empty())
        die("Provide both options or none: `--shake` and `--shake-anchors`");
    if (a.mode != "outlier" && (has_thr || has_bg || has_out || !a.output_blend.empty() || a.sample >= 0))
        printf("WARNING! The following options are not used, as they are required only for `--mode outlier`\n");
    return a;
}

// FileLister::files_vec (src/flist.rs:121-143): glob order, take(end).skip(start), every step-th.
std::vector<std::string> list_files(const Args& a) {
    glob_t g;
    memset(&g, 0, sizeof(g));
    std::vector<std::string> all;
    if (glob(a.pattern.c_str(), 0, nullptr, &g) == 0)
        for (size_t i = 0; i < g.gl_pathc; ++i) {
            struct stat sb;
            if (stat(g.gl_pathv[i], &sb) == 0 && S_ISREG(sb.st_mode)) all.push_back(g.gl_pathv[i]);
        }
    globfree(&g);
    std::sort(all.begin(), all.end());
    if (!a.has_frames) return all;
    // negative bounds wrap to huge counts in the reference's `as usize` casts
    const size_t end = a.frames.has_end && a.frames.end >= 0 ? std::min<size_t>(all.size(), a.frames.end) : all.size();
    const size_t start = !a.frames.has_start ? 0 : a.frames.start >= 0 ? (size_t)a.frames.start : all.size();
    std::vector<std::string> out;
    for (size_t i = start, k = 0; i < end; ++i, ++k)
        if (k % a.frames.step == 0) out.push_back(all[i]);
    return out;
}

std::pair<std::string, std::string> stem_ext(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    if (dot == std::string::npos) die("Unexpected format in " + path);
    return {dir + name.substr(0, dot), name.substr(dot + 1)};
}

}  // namespace

int main(int argc, char** argv) {
    const auto t_start = std::chrono::steady_clock::now();
    std::vector<std::string> args(argv + 1, argv + argc);
    if (args.size() == 1 && args[0][0] != '-') {  // options file, src/main.rs:33-43
        std::ifstream f(args[0]);
        if (!f) die("Something went wrong reading the options file " + args[0]);
        std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        std::replace(text.begin(), text.end(), '\n', ' ');
        std::replace(text.begin(), text.end(), '\r', ' ');
        args = tokenize_options(text);
        if (!args.empty() && args[0] == "chrono-photo") args.erase(args.begin());
    }
    Args a = parse_args(args);
    for (const auto& u : a.unused)
        printf("WARNING: %s has no effect here (frames stay resident on the GPU)\n", u.c_str());

    std::vector<std::string> files = list_files(a);
    if (files.empty()) die("No files found for pattern " + a.pattern);
    Image first;
    std::string err;
    if (!read_image(files[0], &first, &err)) die(err);

    cp_ctx* ctx = nullptr;
    check(cp_ctx_create(0, a.hbm_budget, &ctx), "device");
    const uint32_t N = (uint32_t)files.size(), W = first.width, H = first.height, C = first.channels,
                   B = first.bytes_per_sample;
    // Stacks larger than --hbm-budget stay in host memory and outlier composites stream row
    // bands through the device (the reference's time slices, src/slicer.rs).
    const uint64_t stack_bytes = (uint64_t)N * W * H * C * B;
    const bool out_of_core = a.hbm_budget && stack_bytes + (stack_bytes / 4) > a.hbm_budget;
    if (out_of_core) {
        if (!a.shake.empty()) die("--shake needs the whole stack on the GPU: raise --hbm-budget");
        printf("Stack of %llu bytes exceeds --hbm-budget: compositing row bands from host memory\n",
               (unsigned long long)stack_bytes);
    }
    cp_stack* st = nullptr;
    std::vector<Image> host;  // out of core: the decoded frames
    if (!out_of_core)
        check(cp_stack_create(ctx, W, H, C, B, N, &st), "stack");
    for (size_t i = 0; i < files.size(); ++i) {
        Image img;
        if (i == 0) img = std::move(first);
        else if (!read_image(files[i], &img, &err)) die(err);
        if (img.width != W || img.height != H || img.channels != C || img.bytes_per_sample != B)
            die("Image layout does not fit! (" + files[i] + ")");
        if (out_of_core) host.push_back(std::move(img));
        else check(cp_stack_push(st, img.data.data(), 1, nullptr), "upload");
    }
    std::vector<const void*> host_ptrs;
    for (const Image& im : host) host_ptrs.push_back(im.data.data());
    if (a.debug) printf("%u frames %ux%u, %u channels, %u-bit\n", N, W, H, C, 8 * B);

    // shake correction, src/main.rs:61-85
    std::vector<int32_t> crop;
    uint32_t out_w = W, out_h = H;
    if (!a.shake.empty()) {
        auto rs = split(a.shake, '/');
        if (rs.size() != 2) die("Unexpected format in shake parameters, expected <rad>/<search-rad>: " + a.shake);
        std::vector<int32_t> anchors;
        for (auto& an : a.anchors) {
            auto xy = split(an, '/');
            if (xy.size() != 2) die("Unexpected format in shake anchor, expected x/y: " + an);
            anchors.push_back(std::stoi(xy[0]));
            anchors.push_back(std::stoi(xy[1]));
        }
        std::vector<int32_t> offsets(2 * N);
        check(cp_shake(ctx, st, nullptr, N, anchors.data(), (uint32_t)anchors.size() / 2,
                       (uint32_t)std::stoul(rs[0]), (uint32_t)std::stoul(rs[1]), offsets.data()), "shake analysis");
        crop.resize(2 * N);
        int32_t has = 0;
        check(cp_crop_create(offsets.data(), N, W, H, crop.data(), &out_w, &out_h, &has), "crop");
        if (has) printf("Camera shake detected. Images will be corrected.\n");
        else {
            printf("No camera shake detected. Images will not be corrected.\n");
            crop.clear();
        }
    }

    // options
    cp_outlier_params p{};
    {
        auto t = split(a.threshold, '/');
        bool absolute;
        if (t[0] == "abs" || t[0] == "absolute") absolute = true;
        else if (t[0] == "rel" || t[0] == "relative") absolute = false;
        else die("Not a pixel outlier detection mode: " + a.threshold);
        if (t.size() < 2) die("Unexpected format in " + a.threshold);
        const float lo = std::stof(t[1]), hi = t.size() > 2 ? std::stof(t[2]) : lo;
        check(cp_threshold_new(absolute, lo, hi, B, &p.threshold), "threshold");
    }
    const std::map<std::string, int> bgs = {{"first", CP_BG_FIRST}, {"random", CP_BG_RANDOM},
                                            {"average", CP_BG_AVERAGE}, {"median", CP_BG_MEDIAN}};
    const std::map<std::string, int> sels = {{"first", CP_SEL_FIRST}, {"last", CP_SEL_LAST},
                                             {"extreme", CP_SEL_EXTREME}, {"average", CP_SEL_AVERAGE},
                                             {"forward", CP_SEL_FORWARD}, {"backward", CP_SEL_BACKWARD}};
    if (!bgs.count(a.background)) die("Not a background pixel selection mode: " + a.background);
    if (!sels.count(a.outlier)) die("Not an outlier selection mode: " + a.outlier);
    if (a.mode != "outlier" && a.mode != "darker" && a.mode != "lighter") die("Not a pixel selection mode: " + a.mode);
    p.background = bgs.at(a.background);
    p.outlier = sels.at(a.outlier);
    for (int i = 0; i < 4; ++i) p.weights[i] = a.weights[i];
    p.seed = a.seed;
    std::vector<float> fade_table;
    cp_fade fade{};
    fade.is_none = 1;
    if (!a.fade.empty()) {  // Fade::from_str, src/options.rs:141-186
        auto parts = split(a.fade, '/');
        if (parts.size() < 4 || (parts[0] != "clamp" && parts[0] != "repeat")) die("Not a frame fade spec: " + a.fade);
        std::vector<int32_t> fr;
        std::vector<float> va;
        for (size_t i = 2; i < parts.size(); ++i) {
            auto fv = split(parts[i], ',');
            if (fv.size() != 2) die("Expected (int,float) per frame for fade. Got: " + a.fade);
            fr.push_back(std::stoi(fv[0]));
            va.push_back(std::stof(fv[1]));
        }
        uint32_t len = 0;
        check(cp_fade_table(fr.data(), va.data(), (uint32_t)fr.size(), nullptr, 0, &len), "fade");
        fade_table.resize(len);
        check(cp_fade_table(fr.data(), va.data(), (uint32_t)fr.size(), fade_table.data(), len, &len), "fade");
        fade.is_none = 0;
        fade.repeat = parts[0] == "repeat";
        fade.absolute = parts[1] == "abs" || parts[1] == "absolute";
        fade.offset = fr[0];
        fade.len = len;
        fade.values = fade_table.data();
    }
    p.fade = fade;

    auto composite = [&](const std::vector<uint32_t>& idx, const std::string& out_path, const std::string& blend_path) {
        const uint32_t n = (uint32_t)idx.size();
        std::vector<int32_t> cr;
        if (!crop.empty())
            for (uint32_t t : idx) { cr.push_back(crop[2 * t]); cr.push_back(crop[2 * t + 1]); }
        cp_frames fr{idx.data(), n, (int32_t)idx[0], cr.empty() ? nullptr : cr.data(), out_w, out_h};
        Image img;
        img.width = out_w; img.height = out_h; img.channels = C; img.bytes_per_sample = B;
        img.data.assign((size_t)out_w * out_h * C * B, 0);
        if (a.mode == "outlier") {
            std::vector<uint32_t> sample;
            if (a.sample >= 0 && (uint32_t)a.sample < n) {  // seeded stand-in for the reference's ThreadRng draw
                std::mt19937_64 rng(a.seed);
                std::vector<uint32_t> all(n);
                for (uint32_t i = 0; i < n; ++i) all[i] = i;
                std::shuffle(all.begin(), all.end(), rng);
                sample.assign(all.begin(), all.begin() + a.sample);
                std::sort(sample.begin(), sample.end());
            }
            cp_outlier_params pp = p;
            pp.sample_pos = sample.empty() ? nullptr : sample.data();
            pp.n_sample = (uint32_t)sample.size();
            Image mask = img;
            mask.bytes_per_sample = 1;
            mask.data.assign((size_t)out_w * out_h * C, 0);
            uint64_t warnings = 0;
            cp_outputs o{img.data.data(), mask.data.data(), &warnings, nullptr};
            if (out_of_core)
                check(cp_outlier_host(ctx, host_ptrs.data(), N, W, H, C, B, &fr, &pp, &o, 0), "outlier composite");
            else
                check(cp_outlier(ctx, st, &fr, &pp, &o, 0), "outlier composite");
            if (warnings) printf("Warning: %llu pixels seem to consist of only outliers\n", (unsigned long long)warnings);
            if (!blend_path.empty() && !write_image(blend_path, mask, &err, a.quality)) die(err);
        } else {
            if (out_of_core)
                check(cp_simple_host(ctx, host_ptrs.data(), N, W, H, C, B, &fr, a.weights, &fade, a.mode == "darker",
                                     img.data.data(), 0),
                      "composite");
            else
                check(cp_simple(ctx, st, &fr, a.weights, &fade, a.mode == "darker", img.data.data(), 0), "composite");
        }
        if (!write_image(out_path, img, &err, a.quality)) die(err);
    };

    if (a.has_vin || a.has_vout) {  // create_video, src/main.rs:214-331
        // a missing range is filled with FrameRange::empty(), src/main.rs:163-170
        cp_frame_range vi{a.vin.has_start, a.vin.start, a.vin.has_end, a.vin.end, a.vin.step};
        cp_frame_range vo{a.vout.has_start, a.vout.start, a.vout.has_end, a.vout.end, a.vout.step};
        int32_t nwin = 0;
        check(cp_video_windows(&vi, &vo, N, nullptr, nullptr, nullptr, 0, &nwin), "video windows");
        std::vector<int32_t> ws(nwin), we(nwin), wf(nwin);
        check(cp_video_windows(&vi, &vo, N, ws.data(), we.data(), wf.data(), nwin, &nwin), "video windows");
        auto [stem, ext] = stem_ext(a.output);
        for (int32_t i = 0; i < nwin; ++i) {
            std::vector<uint32_t> idx;
            for (int32_t f = ws[i]; f < we[i]; f += (int32_t)vi.step) idx.push_back((uint32_t)f);
            char num[16];
            snprintf(num, sizeof(num), "-%05d.", wf[i]);
            if (idx.empty()) { printf("Skipping frame %d/%d\n", wf[i], nwin); continue; }
            printf("Processing frame %d/%d -> \n", wf[i], nwin);
            std::string blend;
            if (!a.output_blend.empty()) {  // placed beside --output, as src/main.rs:290-300 does
                auto [bs, be] = stem_ext(a.output_blend);
                const size_t bslash = bs.find_last_of('/');
                const size_t oslash = stem.find_last_of('/');
                const std::string odir = oslash == std::string::npos ? "" : stem.substr(0, oslash + 1);
                blend = odir + (bslash == std::string::npos ? bs : bs.substr(bslash + 1)) + num + be;
            }
            composite(idx, stem + num + ext, blend);
        }
    } else {
        std::vector<uint32_t> idx(N);
        for (uint32_t i = 0; i < N; ++i) idx[i] = i;
        composite(idx, a.output, a.output_blend);
    }
    if (st) cp_stack_destroy(st);
    cp_ctx_destroy(ctx);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    printf("Total time: %.3fs\n", s);
    return 0;
}
