// Frame ingest (host code): 8-bit grayscale TIFF, PNG and JPEG files to uint8
// frames, and 16-bit grayscale TIFF and PNG files to uint16 frames.
//
// The reference reads frames with javax.imageio (jobs/ImageReaderJob.compute
// :80-110) and, for 8-bit images, copies a TYPE_BYTE_GRAY raster as it is
// (images/Image.readImageToBuffer :146-156).  This reader decodes exactly that
// case: one 8-bit sample per pixel, black is zero; and for 10/12/16-bit
// frames (images/ImageInt16.readImageToBuffer :183-184) a TYPE_USHORT_GRAY
// raster copied as it is: one 16-bit sample per pixel.  Other layouts (colour,
// palette, alpha, 1/2/4/16-bit, tiles, progressive JPEG) go through Java2D
// conversions or other decoders in the reference and are rejected with
// QLP_ERR_UNSUPPORTED.
//   TIFF: baseline strips, little or big endian, compression none, PackBits,
//         LZW or Deflate, with or without horizontal differencing (predictor 2).
//   PNG:  colour type 0, bit depth 8, not interlaced; the five row filters.
//   JPEG: baseline or extended sequential Huffman, 8-bit, one component, with
//         restart intervals (ITU-T T.81); the inverse DCT is the integer
//         Loeffler-Ligtenberg-Moschytz factorisation with 13-bit constants and
//         2 extra bits between the passes that IJG-derived decoders (the one
//         javax.imageio wraps included) use by default, so the samples are the
//         ones those decoders produce.
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
int fail_exception();
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

// A frame is one Java array in the reference (images/Image.java: byte[] or
// short[] of height * width elements), so it holds at most 2^31 - 1 bytes;
// larger, empty or overflowing header dimensions are rejected before anything
// is allocated from them.
bool dims_ok(uint64_t W, uint64_t H, uint64_t bpp)
{
    return W >= 1 && H >= 1 && W <= 0x7fffffffu && H <= 0x7fffffffu &&
           W * H <= 0x7fffffffu / bpp;
}

struct Image {
    int H = 0, W = 0;
    bool wide = false;           // 16-bit samples requested (ImageInt16 frames)
    std::vector<uint8_t> px;     // 8-bit frames
    std::vector<uint16_t> px16;  // 16-bit frames
};

// ---- zlib -----------------------------------------------------------------------

bool inflate_all(const uint8_t *src, size_t n, std::vector<uint8_t> &out, size_t expect)
{
    // The output grows with what the stream actually inflates to (doubling from
    // 1 MB), so a small file that declares a huge frame allocates little.
    out.clear();
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit(&z) != Z_OK) return false;
    z.next_in = (Bytef *)src;
    z.avail_in = (uInt)std::min<size_t>(n, 0xffffffffu);
    bool ok = true;
    while (out.size() < expect) {
        const size_t old = out.size();
        const size_t grow = std::min(expect - old, std::max<size_t>(old, 1u << 20));
        out.resize(old + grow);
        z.next_out = out.data() + old;
        z.avail_out = (uInt)grow;
        const int rc = inflate(&z, Z_NO_FLUSH);
        const size_t got = grow - z.avail_out;
        out.resize(old + got);
        if (rc == Z_STREAM_END) break;
        if (rc != Z_OK && rc != Z_BUF_ERROR) {
            ok = false;
            break;
        }
        if (got == 0) break;  // the stream ends short of the frame
    }
    inflateEnd(&z);
    return ok && out.size() == expect;
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
            if (prev >= 0 && table.size() < 4096) {  // a full table waits for Clear
                std::vector<uint8_t> add = table[prev];
                add.push_back(entry[0]);
                table.push_back(add);
            }
        } else if (code == (int)table.size() && prev >= 0 && table.size() < 4096) {
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
    if (bps != (img.wide ? 16u : 8u) || spp != 1 || photo != 1 || planar != 1 || tiled ||
        (pred != 1 && pred != 2) ||
        (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773))
        return fail(QLP_ERR_UNSUPPORTED,
                    img.wide ? "image: only 16-bit grayscale (black is zero) TIFF strips, uncompressed, "
                               "PackBits, LZW or Deflate, are read as 16-bit frames"
                             : "image: only 8-bit grayscale (black is zero) TIFF strips, uncompressed, "
                               "PackBits, LZW or Deflate, are read");
    const size_t bpp = img.wide ? 2 : 1;
    if (!dims_ok(W, H, bpp)) return fail(QLP_ERR_ARG, "image: TIFF dimensions out of range");
    if (comp == 1) {  // uncompressed strips hold the samples as they are
        uint64_t have = 0;
        for (uint32_t c : counts) have += c;
        if (have < (uint64_t)W * H * bpp) return fail(QLP_ERR_ARG, "image: TIFF strips too short");
    }
    if (rps > H) rps = H;
    img.W = (int)W;
    img.H = (int)H;
    img.px.clear();
    for (size_t s = 0; s < offs.size() && img.px.size() < (size_t)W * H * bpp; s++) {
        const size_t rows = std::min<size_t>(rps, H - img.px.size() / (W * bpp));
        const size_t expect = rows * W * bpp;
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
        if (pred == 2 && !img.wide)
            for (size_t y = 0; y < rows; y++)
                for (size_t x = 1; x < W; x++)
                    strip[y * W + x] = (uint8_t)(strip[y * W + x] + strip[y * W + x - 1]);
        img.px.insert(img.px.end(), strip.begin(), strip.end());
    }
    if (img.px.size() != (size_t)W * H * bpp) return fail(QLP_ERR_ARG, "image: TIFF strips too short");
    if (img.wide) {
        // samples in the file's byte order; horizontal differencing on whole samples
        img.px16.resize((size_t)W * H);
        for (size_t k = 0; k < img.px16.size(); k++)
            img.px16[k] = r.le ? (uint16_t)(img.px[2 * k] | img.px[2 * k + 1] << 8)
                               : (uint16_t)(img.px[2 * k] << 8 | img.px[2 * k + 1]);
        if (pred == 2)
            for (size_t y = 0; y < H; y++)
                for (size_t x = 1; x < W; x++)
                    img.px16[y * W + x] = (uint16_t)(img.px16[y * W + x] + img.px16[y * W + x - 1]);
        img.px.clear();
    }
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
            if (d[8] != (img.wide ? 16 : 8) || d[9] != 0 || d[10] != 0 || d[11] != 0 || d[12] != 0)
                return fail(QLP_ERR_UNSUPPORTED,
                            img.wide ? "image: only 16-bit grayscale, non-interlaced PNG files are "
                                       "read as 16-bit frames"
                                     : "image: only 8-bit grayscale, non-interlaced PNG files are read");
            header = true;
        } else if (!memcmp(type, "IDAT", 4)) {
            z.insert(z.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        o += 12 + len;
    }
    if (!header) return fail(QLP_ERR_ARG, "image: PNG without header");
    // rows of RB bytes; the filters work on bytes, a neighbour is one sample (bpp bytes) back
    const size_t bpp = img.wide ? 2 : 1, RB = (size_t)W * bpp;
    if (!dims_ok(W, H, bpp) || (uint64_t)H * (RB + 1) > 0xffffffffu)
        return fail(QLP_ERR_ARG, "image: PNG dimensions out of range");
    std::vector<uint8_t> raw;
    if (!inflate_all(z.data(), z.size(), raw, (size_t)H * (RB + 1)))
        return fail(QLP_ERR_ARG, "image: corrupt PNG data");
    img.W = (int)W;
    img.H = (int)H;
    img.px.assign((size_t)H * RB, 0);
    for (size_t y = 0; y < H; y++) {
        const uint8_t f = raw[y * (RB + 1)];
        const uint8_t *in = &raw[y * (RB + 1) + 1];
        uint8_t *row = &img.px[y * RB];
        const uint8_t *up = y ? &img.px[(y - 1) * RB] : nullptr;
        for (size_t x = 0; x < RB; x++) {
            const int a = x >= bpp ? row[x - bpp] : 0, bb = up ? up[x] : 0,
                      c = (x >= bpp && up) ? up[x - bpp] : 0;
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
    if (img.wide) {  // PNG samples are big-endian
        img.px16.resize((size_t)W * H);
        for (size_t k = 0; k < img.px16.size(); k++)
            img.px16[k] = (uint16_t)(img.px[2 * k] << 8 | img.px[2 * k + 1]);
        img.px.clear();
    }
    return QLP_OK;
}

// ---- JPEG (ITU-T T.81, baseline / extended sequential Huffman, one component) ----

struct HuffTable {
    bool present = false;
    // canonical code of each length (F.2.2.3): codes of length l are
    // mincode[l] .. maxcode[l], their values start at valptr[l] in val
    int mincode[17], maxcode[18], valptr[17];
    uint8_t val[256];
};

bool build_huffman(const uint8_t *bits, const uint8_t *vals, int nvals, HuffTable &t)
{
    int code = 0, k = 0;
    for (int l = 1; l <= 16; l++) {
        t.valptr[l] = k;
        t.mincode[l] = code;
        code += bits[l - 1];
        k += bits[l - 1];
        t.maxcode[l] = bits[l - 1] ? code - 1 : -1;
        if (code > (1 << l)) return false;  // over-subscribed code lengths
        code <<= 1;
    }
    t.maxcode[17] = 0x7fffffff;
    if (k != nvals) return false;
    memcpy(t.val, vals, (size_t)nvals);
    t.present = true;
    return true;
}

struct BitReader {
    const uint8_t *p, *end;
    uint32_t acc = 0;
    int n = 0;
    bool marker = false;  // a marker was met: the segment's data is exhausted
    bool past = false;    // bits were read beyond the segment's data
    int fill_byte()
    {
        if (marker || p >= end) {
            past = true;
            return 0;
        }
        int b = *p;
        if (b == 0xff) {
            if (p + 1 < end && p[1] == 0x00) {
                p += 2;
                return 0xff;
            }
            marker = true;  // leave p at the marker
            return 0;
        }
        p++;
        return b;
    }
    int bit()
    {
        if (n == 0) {
            acc = (uint32_t)fill_byte();
            n = 8;
        }
        n--;
        return (int)((acc >> n) & 1u);
    }
    int bits(int s)
    {
        int v = 0;
        for (int i = 0; i < s; i++) v = (v << 1) | bit();
        return v;
    }
    void align() { n = 0; }
};

int huff_decode(BitReader &br, const HuffTable &t)
{
    int code = br.bit();
    int l = 1;
    while (l <= 16 && code > t.maxcode[l]) {
        code = (code << 1) | br.bit();
        l++;
    }
    if (l > 16) return -1;
    return t.val[t.valptr[l] + code - t.mincode[l]];
}

// EXTEND (F.2.2.1): the s-bit magnitude v as a signed coefficient difference.
inline int extend(int v, int s) { return (s && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Inverse DCT of one dequantised 8 x 8 block.  The step sequence (the rotation
// constants, CONST_BITS 13, PASS1_BITS 2, the DC-only column shortcut and the
// final descale) is the Independent JPEG Group's "islow" integer IDCT
// (jidctint.c, IJG libjpeg), which javax.imageio's decoder uses: matching its
// samples bit for bit needs the same fixed-point steps.  It is the
// Loeffler-Ligtenberg-Moschytz factorisation in fixed point with 13-bit constants (c = round(x *
// 2^13) of the rotation factors below), columns first keeping 2 extra bits,
// then rows, rounding each pass by adding half the discarded weight; the
// result is offset by 128 and clamped to 0..255.
void idct8x8(const int *in, uint8_t *out, int stride)
{
    constexpr int CB = 13, P1 = 2;
    auto fix = [](double x) { return (int64_t)(x * (1 << CB) + 0.5); };
    static const int64_t F0298 = fix(0.298631336), F0390 = fix(0.390180644),
                         F0541 = fix(0.541196100), F0765 = fix(0.765366865),
                         F0899 = fix(0.899976223), F1175 = fix(1.175875602),
                         F1501 = fix(1.501321110), F1847 = fix(1.847759065),
                         F1961 = fix(1.961570560), F2053 = fix(2.053119869),
                         F2562 = fix(2.562915447), F3072 = fix(3.072711026);
    int64_t ws[64];
    // one 1D transform of x[0..7] (stride sx) into 8 values before scaling
    auto rot = [&](const int64_t *x, int sx, int64_t *o) {
        // even part
        int64_t z2 = x[2 * sx], z3 = x[6 * sx];
        int64_t z1 = (z2 + z3) * F0541;
        const int64_t t2 = z1 - z3 * F1847;
        const int64_t t3 = z1 + z2 * F0765;
        z2 = x[0];
        z3 = x[4 * sx];
        const int64_t t0 = (int64_t)((uint64_t)(z2 + z3) << CB);
        const int64_t t1 = (int64_t)((uint64_t)(z2 - z3) << CB);
        const int64_t a10 = t0 + t3, a13 = t0 - t3, a11 = t1 + t2, a12 = t1 - t2;
        // odd part
        int64_t b0 = x[7 * sx], b1 = x[5 * sx], b2 = x[3 * sx], b3 = x[1 * sx];
        int64_t y1 = b0 + b3, y2 = b1 + b2, y3 = b0 + b2, y4 = b1 + b3;
        const int64_t y5 = (y3 + y4) * F1175;
        b0 *= F0298;
        b1 *= F2053;
        b2 *= F3072;
        b3 *= F1501;
        y1 *= -F0899;
        y2 *= -F2562;
        y3 *= -F1961;
        y4 *= -F0390;
        y3 += y5;
        y4 += y5;
        b0 += y1 + y3;
        b1 += y2 + y4;
        b2 += y2 + y3;
        b3 += y1 + y4;
        o[0] = a10 + b3;
        o[7] = a10 - b3;
        o[1] = a11 + b2;
        o[6] = a11 - b2;
        o[2] = a12 + b1;
        o[5] = a12 - b1;
        o[3] = a13 + b0;
        o[4] = a13 - b0;
    };
    auto descale = [](int64_t x, int n) { return (x + ((int64_t)1 << (n - 1))) >> n; };
    int64_t col[8], o[8];
    for (int c = 0; c < 8; c++) {
        for (int r = 0; r < 8; r++) col[r] = in[r * 8 + c];
        bool ac = false;
        for (int r = 1; r < 8; r++) ac |= col[r] != 0;
        if (!ac) {
            for (int r = 0; r < 8; r++) ws[r * 8 + c] = (int64_t)((uint64_t)col[0] << P1);
            continue;
        }
        rot(col, 1, o);
        for (int r = 0; r < 8; r++) ws[r * 8 + c] = descale(o[r], CB - P1);
    }
    for (int r = 0; r < 8; r++) {
        rot(ws + r * 8, 1, o);
        for (int c = 0; c < 8; c++) {
            const int64_t v = descale(o[c], CB + P1 + 3) + 128;
            out[r * stride + c] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }
}

int read_jpeg(const std::vector<uint8_t> &b, Image &img)
{
    const uint8_t *p = b.data(), *end = b.data() + b.size();
    uint16_t q[4][64];
    bool qok[4] = {false, false, false, false};
    HuffTable dc[4], ac[4];
    int H = 0, W = 0, qsel = 0, restart = 0;
    bool frame = false;
    p += 2;  // SOI
    auto be16 = [](const uint8_t *x) { return (int)x[0] << 8 | x[1]; };
    while (p + 4 <= end) {
        if (p[0] != 0xff) return fail(QLP_ERR_ARG, "jpeg: corrupt marker stream");
        const int m = p[1];
        if (m == 0xff) {  // fill byte
            p++;
            continue;
        }
        if (m == 0xd9) break;  // EOI before a scan
        const int len = be16(p + 2);
        const uint8_t *seg = p + 4, *segend = p + 2 + len;
        if (len < 2 || segend > end) return fail(QLP_ERR_ARG, "jpeg: truncated segment");
        if (m == 0xc0 || m == 0xc1) {  // baseline / extended sequential, Huffman
            if (len < 8 || seg[0] != 8)
                return fail(QLP_ERR_UNSUPPORTED, "jpeg: only 8-bit samples are read");
            H = be16(seg + 1);
            W = be16(seg + 3);
            const int nc = seg[5];
            if (nc != 1)
                return fail(QLP_ERR_UNSUPPORTED,
                            "jpeg: only one-component (grayscale) images are read; the reference "
                            "converts colour through Java2D");
            if (len < 8 + 3) return fail(QLP_ERR_ARG, "jpeg: truncated frame header");
            qsel = seg[8];
            if (qsel > 3) return fail(QLP_ERR_ARG, "jpeg: bad quantisation table index");
            if (H < 1 || W < 1) return fail(QLP_ERR_UNSUPPORTED, "jpeg: image height is defined later");
            frame = true;
        } else if (m >= 0xc2 && m <= 0xcf && m != 0xc4 && m != 0xc8 && m != 0xcc) {
            return fail(QLP_ERR_UNSUPPORTED, "jpeg: progressive, lossless and arithmetic-coded images are not read");
        } else if (m == 0xdb) {  // quantisation tables
            const uint8_t *x = seg;
            while (x < segend) {
                const int pq = x[0] >> 4, tq = x[0] & 15;
                if (tq > 3 || pq > 1 || x + 1 + 64 * (pq + 1) > segend)
                    return fail(QLP_ERR_ARG, "jpeg: bad quantisation table");
                for (int k = 0; k < 64; k++)
                    q[tq][kZigzag[k]] = pq ? (uint16_t)be16(x + 1 + 2 * k) : x[1 + k];
                qok[tq] = true;
                x += 1 + 64 * (pq + 1);
            }
        } else if (m == 0xc4) {  // Huffman tables
            const uint8_t *x = seg;
            while (x + 17 <= segend) {
                const int tc = x[0] >> 4, th = x[0] & 15;
                int nv = 0;
                for (int l = 0; l < 16; l++) nv += x[1 + l];
                if (tc > 1 || th > 3 || nv > 256 || x + 17 + nv > segend)
                    return fail(QLP_ERR_ARG, "jpeg: bad Huffman table");
                if (!build_huffman(x + 1, x + 17, nv, tc ? ac[th] : dc[th]))
                    return fail(QLP_ERR_ARG, "jpeg: bad Huffman code lengths");
                x += 17 + nv;
            }
        } else if (m == 0xdd) {  // restart interval
            if (len < 4) return fail(QLP_ERR_ARG, "jpeg: bad restart interval");
            restart = be16(seg);
        } else if (m == 0xda) {  // start of scan
            if (!frame) return fail(QLP_ERR_ARG, "jpeg: scan before the frame header");
            if (len < 8 || seg[0] != 1) return fail(QLP_ERR_ARG, "jpeg: bad scan header");
            const int td = seg[2] >> 4, ta = seg[2] & 15;
            if (td > 3 || ta > 3 || !dc[td].present || !ac[ta].present || !qok[qsel])
                return fail(QLP_ERR_ARG, "jpeg: scan refers to a missing table");
            const int bw = (W + 7) / 8, bh = (H + 7) / 8;
            if (!dims_ok((uint64_t)bw * 8, (uint64_t)bh * 8, 1))
                return fail(QLP_ERR_ARG, "jpeg: image dimensions out of range");
            // decoded block rows; grown one row at a time, so the buffer follows
            // the scan data actually present
            std::vector<uint8_t> full;
            BitReader br{segend, end};
            int pred = 0, coef[64];
            const int nblocks = bw * bh;
            for (int blk = 0; blk < nblocks; blk++) {
                if (restart && blk > 0 && blk % restart == 0) {
                    // RSTn: byte-align, skip the marker, reset the DC prediction
                    br.align();
                    while (br.p + 1 < end && !(br.p[0] == 0xff && br.p[1] >= 0xd0 && br.p[1] <= 0xd7))
                        br.p++;
                    if (br.p + 1 >= end) return fail(QLP_ERR_ARG, "jpeg: missing restart marker");
                    br.p += 2;
                    br.marker = false;
                    pred = 0;
                }
                for (int k = 0; k < 64; k++) coef[k] = 0;
                const int s = huff_decode(br, dc[td]);
                if (s < 0 || s > 11) return fail(QLP_ERR_ARG, "jpeg: corrupt DC code");
                pred += extend(br.bits(s), s);
                // a valid 8-bit image's DC coefficient (pred times its quantiser)
                // lies within +-1024, so |pred| <= 32767 accepts every valid file
                // and keeps pred * q inside an int
                if (pred < -32767 || pred > 32767) return fail(QLP_ERR_ARG, "jpeg: DC coefficient out of range");
                coef[0] = pred * q[qsel][0];
                for (int k = 1; k < 64;) {
                    const int rs = huff_decode(br, ac[ta]);
                    if (rs < 0) return fail(QLP_ERR_ARG, "jpeg: corrupt AC code");
                    const int r = rs >> 4, sz = rs & 15;
                    if (sz == 0) {
                        if (r != 15) break;  // end of block
                        k += 16;             // sixteen zeros
                        continue;
                    }
                    k += r;
                    if (k > 63) return fail(QLP_ERR_ARG, "jpeg: coefficient index past the block");
                    const int z = kZigzag[k];
                    coef[z] = extend(br.bits(sz), sz) * q[qsel][z];
                    k++;
                }
                // a valid scan never reads past its data (Huffman codes are read bit
                // by bit); a truncated one is refused rather than zero-filled
                if (br.past) return fail(QLP_ERR_ARG, "jpeg: scan data ends before the last block");
                const int by = blk / bw, bx = blk - by * bw;
                if (bx == 0) full.resize((size_t)(by + 1) * 8 * bw * 8);
                idct8x8(coef, full.data() + (size_t)by * 8 * bw * 8 + bx * 8, bw * 8);
            }
            img.H = H;
            img.W = W;
            img.px.resize((size_t)H * W);
            for (int y = 0; y < H; y++)
                memcpy(img.px.data() + (size_t)y * W, full.data() + (size_t)y * bw * 8, (size_t)W);
            return QLP_OK;
        }
        p = segend;
    }
    return fail(QLP_ERR_ARG, "jpeg: no image data");
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
    if (b.size() >= 4 && b[0] == 0xff && b[1] == 0xd8 && b[2] == 0xff) {
        if (img.wide) return fail(QLP_ERR_UNSUPPORTED, "image: JPEG files hold 8-bit samples");
        return read_jpeg(b, img);
    }
    return fail(QLP_ERR_UNSUPPORTED, std::string("image: unsupported file format: ") + path);
}

}  // namespace

extern "C" {

int qlp_image_read16(const char *path, uint16_t *out, size_t out_elems, int *H, int *W)
try {
    Image img;
    img.wide = true;
    if (int rc = decode(path, img)) return rc;
    if (H) *H = img.H;
    if (W) *W = img.W;
    if (!out) return QLP_OK;
    if (out_elems < img.px16.size()) return fail(QLP_ERR_ARG, "image: output buffer too small");
    memcpy(out, img.px16.data(), img.px16.size() * sizeof(uint16_t));
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_image_read(const char *path, uint8_t *out, size_t out_bytes, int *H, int *W)
try {
    Image img;
    if (int rc = decode(path, img)) return rc;
    if (H) *H = img.H;
    if (W) *W = img.W;
    if (!out) return QLP_OK;
    if (out_bytes < img.px.size()) return fail(QLP_ERR_ARG, "image: output buffer too small");
    memcpy(out, img.px.data(), img.px.size());
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
