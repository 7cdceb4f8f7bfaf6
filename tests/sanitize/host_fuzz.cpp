// Bounded mutation run over the library's host-side file code, built with
// AddressSanitizer and UndefinedBehaviorSanitizer on the CPU (test
// infrastructure; see the Makefile beside this file).
//
// Every seed file (TIFF, PNG, JPEG) is decoded once as it is, then `rounds`
// times after 1-4 seeded mutations (byte and bit flips, truncation, 32-bit
// header fields set to boundary values, chunk duplication), through both
// qlp_image_read and qlp_image_read16.  A sanitizer report or a crash fails
// the run; a refusal status is the expected outcome for a broken file.  The
// MATLAB writer is driven with valid and out-of-range array sizes.
//
// usage: host_fuzz <rounds> <work_dir> <seed files...>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qlpiv.h"

namespace qlpiv {
// The library's error reporting (qlpiv_api.hip), restated for this host-only build.
static thread_local std::string g_err;
int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
int fail_exception()
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(QLP_ERR_COMPUTE, "host memory allocation failed");
    } catch (const std::exception &e) {
        return fail(QLP_ERR_COMPUTE, std::string("internal error: ") + e.what());
    } catch (...) {
        return fail(QLP_ERR_COMPUTE, "internal error: unknown exception");
    }
}
}  // namespace qlpiv

static std::vector<uint8_t> slurp(const char *path)
{
    std::vector<uint8_t> b;
    FILE *f = fopen(path, "rb");
    if (!f) return b;
    uint8_t tmp[4096];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) b.insert(b.end(), tmp, tmp + n);
    fclose(f);
    return b;
}

static void spit(const std::string &path, const std::vector<uint8_t> &b)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        perror(path.c_str());
        exit(2);
    }
    fwrite(b.data(), 1, b.size(), f);
    fclose(f);
}

static int counts[8];

static void decode_both(const std::string &path)
{
    int H = 0, W = 0;
    int rc = qlp_image_read(path.c_str(), nullptr, 0, &H, &W);
    counts[rc < 8 ? rc : 7]++;
    if (rc == QLP_OK && (size_t)H * W <= (64u << 20)) {
        std::vector<uint8_t> out((size_t)H * W);
        rc = qlp_image_read(path.c_str(), out.data(), out.size(), &H, &W);
        if (rc != QLP_OK) {
            fprintf(stderr, "second decode of %s failed (%d)\n", path.c_str(), rc);
            exit(3);
        }
    }
    rc = qlp_image_read16(path.c_str(), nullptr, 0, &H, &W);
    counts[rc < 8 ? rc : 7]++;
    if (rc == QLP_OK && (size_t)H * W <= (32u << 20)) {
        std::vector<uint16_t> out((size_t)H * W);
        (void)qlp_image_read16(path.c_str(), out.data(), out.size(), &H, &W);
    }
}

// The frame-size fields of a file (TIFF ImageWidth / ImageLength values, PNG
// IHDR width / height, JPEG SOF height / width): offsets and widths in bytes.
static std::vector<std::pair<size_t, int>> size_fields(const std::vector<uint8_t> &b)
{
    std::vector<std::pair<size_t, int>> f;
    if (b.size() >= 8 && (b[0] == 'I' || b[0] == 'M')) {
        const bool le = b[0] == 'I';
        auto rd = [&](size_t o, int w) {
            uint32_t v = 0;
            for (int i = 0; i < w && o + i < b.size(); i++)
                v |= (uint32_t)b[o + i] << (le ? 8 * i : 8 * (w - 1 - i));
            return v;
        };
        const size_t ifd = rd(4, 4);
        const uint32_t n = rd(ifd, 2);
        for (uint32_t k = 0; k < n && ifd + 2 + 12 * (size_t)k + 12 <= b.size(); k++) {
            const size_t e = ifd + 2 + 12 * (size_t)k;
            const uint32_t tag = rd(e, 2), type = rd(e + 2, 2);
            if (tag == 256 || tag == 257) f.push_back({e + 8, type == 3 ? 2 : 4});
        }
    } else if (b.size() >= 24 && b[1] == 'P') {
        f.push_back({16, 4});
        f.push_back({20, 4});
    } else {
        for (size_t o = 0; o + 9 < b.size(); o++)
            if (b[o] == 0xff && (b[o + 1] == 0xc0 || b[o + 1] == 0xc1)) {
                f.push_back({o + 5, 2});
                f.push_back({o + 7, 2});
            }
    }
    return f;
}

static void mutate(std::vector<uint8_t> &b, std::mt19937 &rng)
{
    if (rng() % 4 == 0) {  // a frame-size field set to a boundary value
        static const uint32_t big[] = {0u, 1u, 0xffffu, 0x10000u, 0x7fffffffu, 0xffffffffu, 46341u};
        const auto f = size_fields(b);
        if (!f.empty()) {
            const auto [at, w] = f[rng() % f.size()];
            const uint32_t v = big[rng() % (sizeof big / sizeof big[0])];
            const bool le = b[0] == 'I';
            for (int i = 0; i < w && at + i < b.size(); i++)
                b[at + i] = (uint8_t)(le ? v >> (8 * i) : v >> (8 * (w - 1 - i)));
            return;
        }
    }
    static const uint32_t edge[] = {0u,          1u,          0x7fu,       0xffu,
                                    0x7fffu,     0xffffu,     0x10000u,    0x7fffffffu,
                                    0x80000000u, 0xfffffffeu, 0xffffffffu};
    const int m = 1 + (int)(rng() % 4);
    for (int k = 0; k < m && !b.empty(); k++) {
        const size_t at = rng() % b.size();
        switch (rng() % 6) {
        case 0: b[at] = (uint8_t)rng(); break;
        case 1: b[at] ^= (uint8_t)(1u << (rng() % 8)); break;
        case 2: b.resize(at + 1); break;
        case 3: {  // a 32-bit field (either byte order) set to a boundary value
            const uint32_t v = edge[rng() % (sizeof edge / sizeof edge[0])];
            const bool be = rng() & 1;
            for (int i = 0; i < 4 && at + i < b.size(); i++)
                b[at + i] = (uint8_t)(be ? v >> (24 - 8 * i) : v >> (8 * i));
            break;
        }
        case 4: {  // 16-bit field
            const uint32_t v = edge[rng() % 6];
            for (int i = 0; i < 2 && at + i < b.size(); i++) b[at + i] = (uint8_t)(v >> (8 * i));
            break;
        }
        default: {  // duplicate a chunk in place
            const size_t len = 1 + rng() % 64;
            const size_t from = rng() % b.size();
            std::vector<uint8_t> chunk(b.begin() + from, b.begin() + std::min(b.size(), from + len));
            b.insert(b.begin() + at, chunk.begin(), chunk.end());
            break;
        }
        }
    }
}

static void export_cases(const std::string &dir)
{
    const std::string path = dir + "/fuzz_export.mat";
    qlp_export_params p{};
    p.image_width = 256;
    p.image_height = 256;
    p.ia_start_i = p.ia_start_j = 32;
    p.ia_end_i = p.ia_end_j = 16;
    p.overlap_factor = 0.5f;
    p.total_maps = 3;
    qlp_exporter *e = nullptr;
    if (qlp_export_open(path.c_str(), &p, "Thu Jan 01 00:00:00 UTC 1970", &e) != QLP_OK) exit(4);
    const int margins[4] = {0, 0, 0, 0};
    std::vector<float> u(15 * 15, 1.5f), v(15 * 15, -0.5f);
    std::vector<uint8_t> bad(15 * 15, 0);
    if (qlp_export_frame(e, 1, 32, 32, 0, margins, 15, 15, u.data(), v.data(), bad.data()) != QLP_OK)
        exit(5);
    // out of range: the array would exceed a MAT-file element; must be refused
    if (qlp_export_frame(e, 0, 16, 16, 1, margins, 65536, 65536, u.data(), v.data(), nullptr) ==
        QLP_OK)
        exit(6);
    const int idx[1] = {2};
    if (qlp_export_results(e, 64, 64, margins, 0x7fffffff, 2, 4, 1, idx, u.data(), v.data(), 0) ==
        QLP_OK)
        exit(7);
    if (qlp_export_close(e) != QLP_OK) exit(8);
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: host_fuzz <rounds> <work_dir> <seed files...>\n");
        return 2;
    }
    const int rounds = atoi(argv[1]);
    const std::string dir = argv[2];
    std::mt19937 rng(12345);
    for (int s = 3; s < argc; s++) {
        const std::vector<uint8_t> seed = slurp(argv[s]);
        if (seed.empty()) {
            fprintf(stderr, "cannot read %s\n", argv[s]);
            return 2;
        }
        decode_both(argv[s]);
        for (int r = 0; r < rounds; r++) {
            std::vector<uint8_t> b = seed;
            mutate(b, rng);
            const std::string path = dir + "/fuzz_case.bin";
            spit(path, b);
            decode_both(path);
        }
    }
    export_cases(dir);
    printf("host_fuzz ok: status counts");
    for (int k = 0; k < 8; k++) printf(" %d", counts[k]);
    printf("\n");
    return 0;
}
