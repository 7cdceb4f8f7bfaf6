// MATLAB Level 5 export of the velocity maps (host code).
//
// Writes the file of exporter/StructMultiFrameFloatVelocityExporter.java
// (openFile :54-113, exportDataToFile(frame, IterationStepTiles) :131-180,
// closeFile :115-129): the 128-byte header, the char variable "exportDate",
// the 1 x 1 struct "parameters", and per IA size a 1 x 1 struct
// "velocities<H>x<W>" with iaHeight, iaWidth, adaptiveStep, margins (struct:
// top, left, bottom, right) and u, v (single, nI x nJ x frames); or, from whole
// results (exportDataToFile(PIVResults) :185-238), per velocity map a struct
// with iaHeight, iaWidth, multiVolumeExport (and multiVolumeStartOffset),
// margins and u, v (nI x nJ x maps, the values as they are).  The element
// layout is the reference's (exporter/MatlabLevel5Matrix.java :417-481,
// MatlabLevel5Struct.java :50-66, MatlabLevel5Header.java :62-108):
// uncompressed, little endian, every sub-element padded to 8 bytes, scalars as
// int32 or single, text as UTF-8 chars, struct field names as wide as the
// longest name plus one.  Everything is held in memory and written on close, as
// the reference writes its element chain at closeFile.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/qlpiv.h"

namespace qlpiv {
int fail(int code, const std::string &msg);
int fail_exception();
}

namespace {

// MAT-file data types and array classes (MatlabMiTypesEnum, MatlabMxTypesEnum)
enum { miINT8 = 1, miINT32 = 5, miUINT32 = 6, miSINGLE = 7, miMATRIX = 14, miUTF8 = 16 };
enum { mxSTRUCT = 2, mxCHAR = 4, mxSINGLE = 7, mxINT32 = 12 };

size_t pad8(size_t n) { return (n + 7) & ~(size_t)7; }

void put32(std::vector<uint8_t> &o, uint32_t v)
{
    for (int k = 0; k < 4; k++) o.push_back((uint8_t)(v >> (8 * k)));
}

void put_bytes(std::vector<uint8_t> &o, const void *p, size_t n)
{
    const uint8_t *b = (const uint8_t *)p;
    o.insert(o.end(), b, b + n);
    o.resize(o.size() + (pad8(n) - n), 0);
}

// A miMATRIX element: numeric or char data, or a struct of named fields.
struct Element {
    int cls = mxINT32;
    std::vector<int> dims{1, 1};
    int data_type = miINT32;
    std::vector<uint8_t> data;  // unpadded
    std::vector<std::pair<std::string, std::unique_ptr<Element>>> fields;

    Element *field(const std::string &name)
    {
        for (auto &f : fields)
            if (f.first == name) return f.second.get();
        return nullptr;
    }
    Element &add(const std::string &name, std::unique_ptr<Element> e)
    {
        fields.emplace_back(name, std::move(e));
        return *fields.back().second;
    }

    // the element with its tag, named 'name' ("" for struct fields)
    void write(std::vector<uint8_t> &o, const std::string &name) const
    {
        std::vector<uint8_t> b;
        put32(b, miUINT32);  // array flags
        put32(b, 8);
        put32(b, (uint32_t)cls);
        put32(b, 0);  // nzmax
        put32(b, miINT32);  // dimensions, padded to an even count
        put32(b, (uint32_t)(4 * dims.size()));
        for (int d : dims) put32(b, (uint32_t)d);
        if (dims.size() % 2) put32(b, 0);
        put32(b, miINT8);  // array name
        put32(b, (uint32_t)name.size());
        put_bytes(b, name.data(), name.size());
        if (cls == mxSTRUCT) {
            size_t w = 0;
            for (auto &f : fields) w = f.first.size() > w ? f.first.size() : w;
            w += 1;
            put32(b, (4u << 16) | miINT32);  // field name length, small element
            put32(b, (uint32_t)w);
            put32(b, miINT8);
            put32(b, (uint32_t)(w * fields.size()));
            std::vector<uint8_t> names(w * fields.size(), 0);
            for (size_t k = 0; k < fields.size(); k++)
                memcpy(&names[k * w], fields[k].first.data(), fields[k].first.size());
            put_bytes(b, names.data(), names.size());
            for (auto &f : fields) f.second->write(b, "");
        } else {
            put32(b, (uint32_t)data_type);
            put32(b, (uint32_t)data.size());
            put_bytes(b, data.data(), data.size());
        }
        put32(o, miMATRIX);
        put32(o, (uint32_t)b.size());
        o.insert(o.end(), b.begin(), b.end());
    }
};

std::unique_ptr<Element> int32_scalar(int v)
{
    auto e = std::make_unique<Element>();
    e->cls = mxINT32;
    e->data_type = miINT32;
    put32(e->data, (uint32_t)v);
    return e;
}

std::unique_ptr<Element> single_scalar(float v)
{
    auto e = std::make_unique<Element>();
    e->cls = mxSINGLE;
    e->data_type = miSINGLE;
    uint32_t bits;
    memcpy(&bits, &v, 4);
    put32(e->data, bits);
    return e;
}

std::unique_ptr<Element> text(const std::string &s)
{
    auto e = std::make_unique<Element>();
    e->cls = mxCHAR;
    e->data_type = miUTF8;
    e->dims = {1, (int)s.size()};
    e->data.assign(s.begin(), s.end());
    return e;
}

std::unique_ptr<Element> structure()
{
    auto e = std::make_unique<Element>();
    e->cls = mxSTRUCT;
    e->data_type = 0;
    return e;
}

// java.util.Date.toString(): "EEE MMM dd HH:mm:ss zzz yyyy"
std::string java_date_now()
{
    const time_t t = time(nullptr);
    struct tm tmv;
    localtime_r(&t, &tmv);
    char buf[64];
    strftime(buf, sizeof buf, "%a %b %d %H:%M:%S %Z %Y", &tmv);
    return buf;
}

}  // namespace

struct qlp_exporter {
    std::string filename, date;
    int total_maps = 1;
    int mark_invalid_nan = 0;
    int multi_volume = 0;
    std::unique_ptr<Element> parameters;
    // velocity structs by IA height, in creation order
    std::vector<std::pair<std::string, std::unique_ptr<Element>>> maps;
    std::map<int, size_t> by_height;
};

using qlpiv::fail;

// A MAT-file element records its byte size in 32 bits (and the reference's
// float[] arrays hold at most 2^31 - 1 values): an nI x nJ x maps single array
// must stay below 2^31 bytes.
static bool array_fits(int nI, int nJ, int maps)
{
    return (uint64_t)nI * (uint64_t)nJ * (uint64_t)maps <= 0x7fffffffu / 4;
}

extern "C" {

int qlp_export_open(const char *filename, const qlp_export_params *p, const char *date_text,
                    qlp_exporter **out)
try {
    if (!filename || !p || !out) return fail(QLP_ERR_ARG, "export_open: null argument");
    *out = nullptr;
    if (p->total_maps < 1) return fail(QLP_ERR_ARG, "export_open: total_maps must be >= 1");
    FILE *f = fopen(filename, "wb");  // the reference opens (and truncates) the file here
    if (!f) return fail(QLP_ERR_ARG, std::string("Failed to create file for writing: ") + filename);
    fclose(f);
    auto e = std::make_unique<qlp_exporter>();
    e->filename = filename;
    e->date = date_text ? date_text : java_date_now();
    e->total_maps = p->total_maps;
    e->mark_invalid_nan = p->mark_invalid_nan;
    e->multi_volume = p->multi_volume ? 1 : 0;
    e->parameters = structure();
    Element &s = *e->parameters;
    s.add("imageWidth", int32_scalar(p->image_width));
    s.add("imageHeight", int32_scalar(p->image_height));
    s.add("iaStartI", int32_scalar(p->ia_start_i));
    s.add("iaEndI", int32_scalar(p->ia_end_i));
    s.add("iaStartJ", int32_scalar(p->ia_start_j));
    s.add("iaEndJ", int32_scalar(p->ia_end_j));
    s.add("overlapFactor", single_scalar(p->overlap_factor));
    s.add("hasSwappedUVVectors", int32_scalar(p->swap_uv ? 1 : 0));
    s.add("markInvalidAsNaN", int32_scalar(p->mark_invalid_nan ? 1 : 0));
    s.add("totalNumberOfVelocityMaps", int32_scalar(p->total_maps));
    s.add("multiVolumeExport", int32_scalar(p->multi_volume ? 1 : 0));
    if (p->multi_volume && p->next_filename) {
        s.add("multiVolumeLastVolume", int32_scalar(0));
        s.add("multiVolumeNextFilename", text(p->next_filename));
    } else {
        s.add("multiVolumeLastVolume", int32_scalar(1));
    }
    *out = e.release();
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_export_frame(qlp_exporter *e, int frame, int ia_height, int ia_width, int adaptive_step,
                     const int margins[4], int nI, int nJ, const float *u, const float *v,
                     const uint8_t *invalid)
try {
    if (!e || !margins || !u || !v) return fail(QLP_ERR_ARG, "export_frame: null argument");
    if (nI < 1 || nJ < 1 || ia_height < 1 || ia_width < 1)
        return fail(QLP_ERR_ARG, "export_frame: bad sizes");
    if (frame < 0 || frame >= e->total_maps)
        return fail(QLP_ERR_ARG, "Invalid frame number was specified");
    if (!array_fits(nI, nJ, e->total_maps))
        return fail(QLP_ERR_ARG, "export_frame: velocity array exceeds a MAT-file element");
    auto it = e->by_height.find(ia_height);
    Element *vel;
    if (it == e->by_height.end()) {
        auto s = structure();
        s->add("iaHeight", int32_scalar(ia_height));
        s->add("iaWidth", int32_scalar(ia_width));
        s->add("adaptiveStep", int32_scalar(adaptive_step));
        auto m = structure();
        m->add("top", int32_scalar(margins[0]));
        m->add("left", int32_scalar(margins[1]));
        m->add("bottom", int32_scalar(margins[2]));
        m->add("right", int32_scalar(margins[3]));
        s->add("margins", std::move(m));
        for (const char *c : {"u", "v"}) {
            auto a = std::make_unique<Element>();
            a->cls = mxSINGLE;
            a->data_type = miSINGLE;
            a->dims = {nI, nJ, e->total_maps};
            a->data.assign((size_t)nI * nJ * e->total_maps * 4, 0);
            s->add(c, std::move(a));
        }
        e->by_height[ia_height] = e->maps.size();
        e->maps.emplace_back("velocities" + std::to_string(ia_height) + "x" + std::to_string(ia_width),
                             std::move(s));
        vel = e->maps.back().second.get();
    } else {
        vel = e->maps[it->second].second.get();
    }
    Element *eu = vel->field("u"), *ev = vel->field("v");
    if (eu->dims[0] != nI || eu->dims[1] != nJ)
        return fail(QLP_ERR_ARG, "export_frame: grid differs from the map's first frame");
    // column-major per frame: i fastest (the reference's j-outer, i-inner loop)
    const float qnan = nanf("");
    size_t idx = (size_t)frame * nI * nJ * 4;
    for (int j = 0; j < nJ; j++)
        for (int i = 0; i < nI; i++, idx += 4) {
            const size_t t = (size_t)i * nJ + j;
            const bool bad = e->mark_invalid_nan && invalid && invalid[t];
            const float fu = bad ? qnan : u[t], fv = bad ? qnan : v[t];
            memcpy(&eu->data[idx], &fu, 4);
            memcpy(&ev->data[idx], &fv, 4);
        }
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_export_results(qlp_exporter *e, int ia_height, int ia_width, const int margins[4], int nI,
                       int nJ, int n_maps, int n_frames, const int *frame_index, const float *u,
                       const float *v, int start_offset)
try {
    if (!e || !margins || (n_frames > 0 && (!frame_index || !u || !v)))
        return fail(QLP_ERR_ARG, "export_results: null argument");
    if (nI < 1 || nJ < 1 || ia_height < 1 || ia_width < 1 || n_maps < 1 || n_frames < 0)
        return fail(QLP_ERR_ARG, "export_results: bad sizes");
    if (!array_fits(nI, nJ, n_maps))
        return fail(QLP_ERR_ARG, "export_results: velocity array exceeds a MAT-file element");
    auto s = structure();
    s->add("iaHeight", int32_scalar(ia_height));
    s->add("iaWidth", int32_scalar(ia_width));
    s->add("multiVolumeExport", int32_scalar(e->multi_volume));
    if (e->multi_volume) s->add("multiVolumeStartOffset", int32_scalar(start_offset));
    auto m = structure();
    m->add("top", int32_scalar(margins[0]));
    m->add("left", int32_scalar(margins[1]));
    m->add("bottom", int32_scalar(margins[2]));
    m->add("right", int32_scalar(margins[3]));
    s->add("margins", std::move(m));
    for (const char *c : {"u", "v"}) {
        auto a = std::make_unique<Element>();
        a->cls = mxSINGLE;
        a->data_type = miSINGLE;
        a->dims = {nI, nJ, n_maps};
        a->data.assign((size_t)nI * nJ * n_maps * 4, 0);
        s->add(c, std::move(a));
    }
    Element *eu = s->field("u"), *ev = s->field("v");
    for (int k = 0; k < n_frames; k++) {
        const int fr = frame_index[k];
        if (fr < 0 || fr >= n_maps) return fail(QLP_ERR_ARG, "Invalid frame number was specified");
        // column-major per frame: i fastest (the reference's j-outer, i-inner loop)
        size_t idx = (size_t)fr * nI * nJ * 4;
        const float *pu = u + (size_t)k * nI * nJ, *pv = v + (size_t)k * nI * nJ;
        for (int j = 0; j < nJ; j++)
            for (int i = 0; i < nI; i++, idx += 4) {
                memcpy(&eu->data[idx], &pu[(size_t)i * nJ + j], 4);
                memcpy(&ev->data[idx], &pv[(size_t)i * nJ + j], 4);
            }
    }
    e->maps.emplace_back("velocities" + std::to_string(ia_height) + "x" + std::to_string(ia_width),
                         std::move(s));
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_export_close(qlp_exporter *e)
try {
    if (!e) return fail(QLP_ERR_ARG, "export_close: null exporter");
    std::unique_ptr<qlp_exporter> own(e);
    std::vector<uint8_t> o;
    // header: 116 bytes of text, subsystem offset 0, version 0x0100, "IM"
    std::string title = "MATLAB 5.0 - QuickLabPIV-ng Velocities exported on: " + e->date;
    std::vector<uint8_t> head(116, 0);
    memcpy(head.data(), title.data(), title.size() < 116 ? title.size() : 116);
    o.insert(o.end(), head.begin(), head.end());
    put32(o, 0);
    put32(o, 0);
    o.push_back(0x00);
    o.push_back(0x01);
    o.push_back('I');
    o.push_back('M');
    text(e->date)->write(o, "exportDate");
    e->parameters->write(o, "parameters");
    for (auto &m : e->maps) m.second->write(o, m.first);
    FILE *f = fopen(e->filename.c_str(), "wb");
    if (!f) return fail(QLP_ERR_ARG, "Failed to create file for writing");
    const size_t w = fwrite(o.data(), 1, o.size(), f);
    const int c = fclose(f);
    if (w != o.size() || c != 0) return fail(QLP_ERR_COMPUTE, "Failed to export data to file");
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
