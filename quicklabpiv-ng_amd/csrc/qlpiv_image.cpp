// Frame ingest (host code): 8-bit grayscale TIFF and PNG files to uint8 frames.
//
// The reference reads frames with javax.imageio (jobs/ImageReaderJob.compute
// :80-110) and, for 8-bit images, copies a TYPE_BYTE_GRAY raster as it is
// (images/Image.readImageToBuffer :146-156).  This reader decodes exactly that
// case: one 8-bit sample per pixel, black is zero.  Other layouts (colour,
// palette, alpha, 1/2/4/16-bit, tiles, JPEG) go through Java2D conversions in
// the reference and are rejected with QLP_ERR_UNSUPPORTED.
//   TIFF: baseline strips, little or big endian, compression none, PackBits,
//         LZW or Deflate, with or without horizontal differencing (predictor 2).
//   PNG:  colour type 0, bit depth 8, not interlaced; the five row filters.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include <zlib.h>

#include "../../include/qlpiv.h"

namespace qlpiv {
int fail(int code, const std::string &msg);
}

using qlpiv::fail;

namespace {

bool read_file(const char *path, std::vector<uint8_t> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

struct Image {
    int H = 0, W = 0;
    std::vector<uint8_t> px;
};

// ---- zlib -----------------------------------------------------------------------

bool inflate_all(const uint8_t *src, size_t n, std::vector<uint8_t> &out, size_t expect)
{
    out.assign(expect, 0);
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit(&z) != Z_OK) return false;
    z.next_in = (Bytef *)src;
    z.avail_in = (uInt)n;
    z.next_out = out.data();
    z.avail_out = (uInt)expect;
    const int rc = inflate(&z, Z_FINISH);
    const size_t got = expect - z.avail_out;
    inflateEnd(&z);
    return (rc == Z_STREAM_END || rc == Z_OK || rc == Z_BUF_ERROR) && got == expect;
}

// ---- TIFF ---------------------------------------------------------------------

struct TiffReader {
    const std::vector<uint8_t> &b;
    bool le = true;
    bool ok = true;
    uint32_t u16(size_t o)
    {
        if (o + 2 > b.size()) return ok = false, 0;
        return le ? (uint32_t)b[o] | (uint32_t)b[o + 1] << 8 : (uint32_t)b[o] << 8 | b[o + 1];
    }
    uint32_t u32(size_t o)
    {
        if (o + 4 > b.size()) return ok = false, 0;
        return le ? (uint32_t)b[o] | (uint32_t)b[o + 1] << 8 | (uint32_t)b[o + 2] << 16 |
                        (uint32_t)b[o + 3] << 24
                  : (uint32_t)b[o] << 24 | (uint32_t)b[o + 1] << 16 | (uint32_t)b[o + 2] << 8 |
                        b[o + 3];
    }
    // values of an IFD entry (SHORT or LONG) at entry offset e
    std::vector<uint32_t> values(size_t e)
    {
        const uint32_t type = u16(e + 2), count = u32(e + 4);
        const size_t sz = type == 3 ? 2 : 4;
        std::vector<uint32_t> v;
        if ((type != 3 && type != 4) || count > (1u << 24)) return ok = false, v;
        const size_t at = sz * count <= 4 ? e + 8 : u32(e + 8);
        for (uint32_t k = 0; k < count && ok; k++) v.push_back(sz == 2 ? u16(at + 2 * k) : u32(at + 4 * k));
        return v;
    }
};

bool unpackbits(const uint8_t *s, size_t n, std::vector<uint8_t> &out, size_t expect)
{
    size_t i = 0;
    while (i < n && out.size() < expect) {
        const int8_t c = (int8_t)s[i++];
        if (c >= 0) {
            const size_t len = (size_t)c + 1;
            if (i + len > n) return false;
            out.insert(out.end(), s + i, s + i + len);
            i += len;
        } else if (c != -128) {
            if (i >= n) return false;
            out.insert(out.end(), (size_t)(1 - c), s[i++]);
        }
    }
    return out.size() >= expect;
}

// TIFF LZW: MSB-first codes of 9..12 bits, Clear 256, EOI 257, the code width
// growing one code early (TIFF 6.0 section 13).
bool unlzw(const uint8_t *s, size_t n, std::vector<uint8_t> &out, size_t expect)
{
    std::vector<std::vector<uint8_t>> table;
    auto reset = [&]() {
        table.assign(258, {});
        for (int k = 0; k < 256; k++) table[k] = {(uint8_t)k};
    };
    reset();
    size_t bitpos = 0;
    int width = 9;
    int prev = -1;
    auto next_code = [&]() -> int {
        if (bitpos + width > n * 8) return 257;
        int v = 0;
        for (int k = 0; k < width; k++, bitpos++) v = (v << 1) | ((s[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
        return v;
    };
    while (out.size() < expect) {
        const int code = next_code();
        if (code == 257) break;
        if (code == 256) {
            reset();
            width = 9;
            prev = -1;
            continue;
        }
        std::vector<uint8_t> entry;
        if (code < (int)table.size()) {
            entry = table[code];
            if (prev >= 0) {
                std::vector<uint8_t> add = table[prev];
                add.push_back(entry[0]);
                table.push_back(add);
            }
        } else if (code == (int)table.size() && prev >= 0) {
            entry = table[prev];
            entry.push_back(table[prev][0]);
            table.push_back(entry);
        } else {
            return false;
        }
        out.insert(out.end(), entry.begin(), entry.end());
        prev = code;
        if (table.size() + 1 >= (1u << width) && width < 12) width++;
    }
    return out.size() >= expect;
}

int read_tiff(const std::vector<uint8_t> &b, Image &img)
{
    TiffReader r{b};
    r.le = b[0] == 'I';
    if (r.u16(2) != 42) return fail(QLP_ERR_ARG, "image: not a TIFF file");
    const size_t ifd = r.u32(4);
    const uint32_t entries = r.u16(ifd);
    uint32_t W = 0, H = 0, bps = 1, comp = 1, photo = 1, spp = 1, rps = 0xffffffffu, planar = 1,
             pred = 1;
    bool tiled = false;
    std::vector<uint32_t> offs, counts;
    for (uint32_t k = 0; k < entries && r.ok; k++) {
        const size_t e = ifd + 2 + 12 * (size_t)k;
        const uint32_t tag = r.u16(e);
        auto first = [&]() {
            std::vector<uint32_t> v = r.values(e);
            return v.empty() ? 0u : v[0];
        };
        switch (tag) {
        case 256: W = first(); break;
        case 257: H = first(); break;
        case 258: bps = first(); break;
        case 259: comp = first(); break;
        case 262: photo = first(); break;
        case 273: offs = r.values(e); break;
        case 277: spp = first(); break;
        case 278: rps = first(); break;
        case 279: counts = r.values(e); break;
        case 284: planar = first(); break;
        case 317: pred = first(); break;
        case 322: tiled = true; break;
        default: break;
        }
    }
    if (!r.ok || W == 0 || H == 0 || offs.empty() || offs.size() != counts.size())
        return fail(QLP_ERR_ARG, "image: malformed TIFF directory");
    if (bps != 8 || spp != 1 || photo != 1 || planar != 1 || tiled || (pred != 1 && pred != 2) ||
        (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773))
        return fail(QLP_ERR_UNSUPPORTED,
                    "image: only 8-bit grayscale (black is zero) TIFF strips, uncompressed, "
                    "PackBits, LZW or Deflate, are read");
    if (rps > H) rps = H;
    img.W = (int)W;
    img.H = (int)H;
    img.px.clear();
    img.px.reserve((size_t)W * H);
    for (size_t s = 0; s < offs.size() && img.px.size() < (size_t)W * H; s++) {
        const size_t rows = std::min<size_t>(rps, H - img.px.size() / W);
        const size_t expect = rows * W;
        if ((size_t)offs[s] + counts[s] > b.size()) return fail(QLP_ERR_ARG, "image: TIFF strip out of file");
        const uint8_t *src = b.data() + offs[s];
        std::vector<uint8_t> strip;
        bool good = true;
        if (comp == 1) {
            good = counts[s] >= expect;
            if (good) strip.assign(src, src + expect);
        } else if (comp == 32773) {
            good = unpackbits(src, counts[s], strip, expect);
        } else if (comp == 5) {
            good = unlzw(src, counts[s], strip, expect);
        } else {
            good = inflate_all(src, counts[s], strip, expect);
        }
        if (!good) return fail(QLP_ERR_ARG, "image: corrupt TIFF strip");
        strip.resize(expect);
        if (pred == 2)
            for (size_t y = 0; y < rows; y++)
                for (size_t x = 1; x < W; x++)
                    strip[y * W + x] = (uint8_t)(strip[y * W + x] + strip[y * W + x - 1]);
        img.px.insert(img.px.end(), strip.begin(), strip.end());
    }
    if (img.px.size() != (size_t)W * H) return fail(QLP_ERR_ARG, "image: TIFF strips too short");
    return QLP_OK;
}

// ---- PNG ----------------------------------------------------------------------

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int read_png(const std::vector<uint8_t> &b, Image &img)
{
    size_t o = 8;
    uint32_t W = 0, H = 0;
    std::vector<uint8_t> z;
    bool header = false;
    while (o + 12 <= b.size()) {
        const uint32_t len = be32(&b[o]);
        if (o + 12 + (size_t)len > b.size()) return fail(QLP_ERR_ARG, "image: truncated PNG chunk");
        const char *type = (const char *)&b[o + 4];
        const uint8_t *d = &b[o + 8];
        if (!memcmp(type, "IHDR", 4)) {
            if (len < 13) return fail(QLP_ERR_ARG, "image: bad PNG header");
            W = be32(d);
            H = be32(d + 4);
            if (d[8] != 8 || d[9] != 0 || d[10] != 0 || d[11] != 0 || d[12] != 0)
                return fail(QLP_ERR_UNSUPPORTED,
                            "image: only 8-bit grayscale, non-interlaced PNG files are read");
            header = true;
        } else if (!memcmp(type, "IDAT", 4)) {
            z.insert(z.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        o += 12 + len;
    }
    if (!header || W == 0 || H == 0) return fail(QLP_ERR_ARG, "image: PNG without header");
    std::vector<uint8_t> raw;
    if (!inflate_all(z.data(), z.size(), raw, (size_t)H * (W + 1)))
        return fail(QLP_ERR_ARG, "image: corrupt PNG data");
    img.W = (int)W;
    img.H = (int)H;
    img.px.assign((size_t)W * H, 0);
    for (size_t y = 0; y < H; y++) {
        const uint8_t f = raw[y * (W + 1)];
        const uint8_t *in = &raw[y * (W + 1) + 1];
        uint8_t *row = &img.px[y * W];
        const uint8_t *up = y ? &img.px[(y - 1) * W] : nullptr;
        for (size_t x = 0; x < W; x++) {
            const int a = x ? row[x - 1] : 0, bb = up ? up[x] : 0, c = (x && up) ? up[x - 1] : 0;
            int p;
            switch (f) {
            case 0: p = 0; break;
            case 1: p = a; break;
            case 2: p = bb; break;
            case 3: p = (a + bb) >> 1; break;
            case 4: {
                const int q = a + bb - c, pa = abs(q - a), pb = abs(q - bb), pc = abs(q - c);
                p = (pa <= pb && pa <= pc) ? a : (pb <= pc ? bb : c);
                break;
            }
            default: return fail(QLP_ERR_ARG, "image: bad PNG row filter");
            }
            row[x] = (uint8_t)(in[x] + p);
        }
    }
    return QLP_OK;
}

int decode(const char *path, Image &img)
{
    if (!path) return fail(QLP_ERR_ARG, "image: null path");
    std::vector<uint8_t> b;
    if (!read_file(path, b)) return fail(QLP_ERR_ARG, std::string("Cannot find image file: ") + path);
    static const uint8_t png_sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (b.size() >= 8 && !memcmp(b.data(), png_sig, 8)) return read_png(b, img);
    if (b.size() >= 8 && ((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M')))
        return read_tiff(b, img);
    return fail(QLP_ERR_UNSUPPORTED, std::string("image: unsupported file format: ") + path);
}

}  // namespace

extern "C" {

int qlp_image_read(const char *path, uint8_t *out, size_t out_bytes, int *H, int *W)
{
    Image img;
    if (int rc = decode(path, img)) return rc;
    if (H) *H = img.H;
    if (W) *W = img.W;
    if (!out) return QLP_OK;
    if (out_bytes < img.px.size()) return fail(QLP_ERR_ARG, "image: output buffer too small");
    memcpy(out, img.px.data(), img.px.size());
    return QLP_OK;
}

}  // extern "C"
