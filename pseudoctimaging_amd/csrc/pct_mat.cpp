// MAT-file (level 5) input and output without MATLAB's libmat: the
// reference program's MATLAB build reads its inputs from a .mat file
// (src/CPU/NewPatchedBrainPureC.cpp:584-704 ReadMatFile, 715-801
// FindDimension, 359-474 Choice / ConvertToRowBased / AnalyzeStructArray)
// and writes the pseudo-CT as a single-precision array (cpp:835-906
// ConvertMatrixMatlab, WriteOutputMatFile).
//
// Input layout, as the reference expects it: a struct `atlas` whose fields
// (MRI1 .. MRIn) are nx x ny x nz arrays, a struct `atlas_label` with the
// same number of fields, and a non-struct variable holding the target image.
// MATLAB arrays are column-major: element (x, y, z) of an nx x ny x nz array
// sits at x + nx (y + ny z); the library layout is (z nx + x) ny + y.
//
// Supported: little-endian files, numeric classes (converted to float),
// 1x1 structs, zlib-compressed elements (miCOMPRESSED), the small data
// element format.  Big-endian files and complex, sparse, cell or object
// arrays give PCT_ERR_FORMAT.
#include <zlib.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>
#include <string>
#include <vector>

#include "pct_internal.h"

using pct::checked_volume_elements;
using pct::kMaxFileElements;

namespace {

enum : uint32_t {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6,
    miSINGLE = 7, miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15,
};
enum : int { mxSTRUCT = 2, mxDOUBLE = 6, mxSINGLE = 7, mxINT8 = 8, mxUINT64 = 15 };
constexpr uint32_t kComplexFlag = 0x800;

using Store = std::vector<std::unique_ptr<std::vector<uint8_t>>>;

// One data element: its type and payload (inside a buffer that outlives it).
struct Element {
    uint32_t type = 0;
    const uint8_t* data = nullptr;
    uint32_t size = 0;
};

// Decompresses a zlib stream completely.
bool inflate_all(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    out.resize(std::max<size_t>(4096, n * 4));
    zs.next_in = const_cast<Bytef*>(in);
    zs.avail_in = (uInt)n;
    int rc = Z_OK;
    while (rc != Z_STREAM_END) {
        const size_t used = (size_t)(zs.next_out ? zs.next_out - out.data() : 0);
        if (used == out.size()) out.resize(out.size() * 2);
        zs.next_out = out.data() + used;
        zs.avail_out = (uInt)std::min<size_t>(out.size() - used, (size_t)1 << 30);
        rc = inflate(&zs, Z_NO_FLUSH);
        if (rc != Z_OK && rc != Z_STREAM_END) break;
        if (rc == Z_OK && zs.avail_in == 0 && zs.avail_out != 0) {   // input ended early
            rc = Z_DATA_ERROR;
            break;
        }
    }
    const size_t used = (size_t)(zs.next_out - out.data());
    inflateEnd(&zs);
    if (rc != Z_STREAM_END) return false;
    out.resize(used);
    return true;
}

// Sequential reader over a byte range.
struct Cursor {
    const uint8_t* p;
    size_t n;
    size_t pos;
    Store* store;

    // The next element; miCOMPRESSED elements are replaced by the element
    // their zlib stream holds.
    bool next(Element& e) {
        if (pos + 8 > n) return false;
        uint32_t w0, w1;
        std::memcpy(&w0, p + pos, 4);
        std::memcpy(&w1, p + pos + 4, 4);
        if (w0 >> 16) {   // small data element: type and size share a word
            e.type = w0 & 0xffff;
            e.size = w0 >> 16;
            if (e.size > 4) return false;
            e.data = p + pos + 4;
            pos += 8;
            return true;
        }
        e.type = w0;
        e.size = w1;
        if ((size_t)e.size > n - pos - 8) return false;
        e.data = p + pos + 8;
        pos += 8 + (size_t)e.size;
        if (e.type != miCOMPRESSED) {
            pos = (pos + 7) & ~(size_t)7;   // 64-bit alignment
            return true;
        }
        auto buf = std::make_unique<std::vector<uint8_t>>();
        if (!inflate_all(e.data, e.size, *buf)) return false;
        Cursor inner{buf->data(), buf->size(), 0, store};
        Element ie;
        if (!inner.next(ie) || ie.type == miCOMPRESSED) return false;
        store->push_back(std::move(buf));
        e = ie;
        return true;
    }
};

// A parsed array: numeric payload, or the fields of a 1x1 struct.
struct MatArray {
    int cls = 0;
    std::vector<long long> dims;
    std::string name;
    Element real;
    std::vector<std::string> fields;
    std::vector<MatArray> values;
};

bool parse_matrix(const Element& m, Store* store, MatArray& a);

// Element count of an array, -1 when the product of its dimensions
// overflows or exceeds kMaxFileElements.
long long count(const MatArray& a) {
    long long c = 1;
    for (long long d : a.dims)
        if (__builtin_mul_overflow(c, d, &c) || c > kMaxFileElements) return -1;
    return c;
}

// Array flags, dimensions, name, then the class's own subelements.
bool parse_matrix(const Element& m, Store* store, MatArray& a) {
    if (m.type != miMATRIX) return false;
    Cursor c{m.data, m.size, 0, store};
    if (m.size == 0) return false;
    Element flags, dims, name;
    if (!c.next(flags) || flags.type != miUINT32 || flags.size < 8) return false;
    uint32_t f0;
    std::memcpy(&f0, flags.data, 4);
    a.cls = (int)(f0 & 0xff);
    if (f0 & kComplexFlag) return false;
    if (!c.next(dims) || dims.type != miINT32 || dims.size < 8 || dims.size % 4) return false;
    for (uint32_t i = 0; i < dims.size / 4; ++i) {
        int32_t d;
        std::memcpy(&d, dims.data + 4 * i, 4);
        if (d < 0) return false;
        a.dims.push_back(d);
    }
    if (!c.next(name) || name.type != miINT8) return false;
    a.name.assign(reinterpret_cast<const char*>(name.data), name.size);
    if (a.cls == mxSTRUCT) {
        if (count(a) != 1) return false;
        Element flen, fnames;
        if (!c.next(flen) || flen.type != miINT32 || flen.size != 4) return false;
        int32_t len;
        std::memcpy(&len, flen.data, 4);
        if (!c.next(fnames) || fnames.type != miINT8 || len <= 0 || fnames.size % len) return false;
        for (uint32_t i = 0; i < fnames.size / (uint32_t)len; ++i) {
            const char* s = reinterpret_cast<const char*>(fnames.data) + (size_t)i * len;
            a.fields.emplace_back(s, strnlen(s, (size_t)len));
        }
        for (size_t i = 0; i < a.fields.size(); ++i) {
            Element fe;
            MatArray v;
            if (!c.next(fe) || !parse_matrix(fe, store, v)) return false;
            a.values.push_back(std::move(v));
        }
        return true;
    }
    if (a.cls < mxDOUBLE || a.cls > mxUINT64) return false;   // numeric classes only
    if (count(a) < 0) return false;
    if (count(a) == 0) {
        a.real = Element{miSINGLE, nullptr, 0};
        return true;
    }
    return c.next(a.real);
}

size_t mi_size(uint32_t t) {
    switch (t) {
        case miINT8: case miUINT8: return 1;
        case miINT16: case miUINT16: return 2;
        case miINT32: case miUINT32: case miSINGLE: return 4;
        case miDOUBLE: case miINT64: case miUINT64: return 8;
        default: return 0;
    }
}

// Element i of a numeric payload as float.
float value_at(const Element& e, size_t i) {
    const uint8_t* q = e.data + i * mi_size(e.type);
    switch (e.type) {
        case miINT8: return (float)*reinterpret_cast<const int8_t*>(q);
        case miUINT8: return (float)*q;
        case miINT16: { int16_t v; std::memcpy(&v, q, 2); return (float)v; }
        case miUINT16: { uint16_t v; std::memcpy(&v, q, 2); return (float)v; }
        case miINT32: { int32_t v; std::memcpy(&v, q, 4); return (float)v; }
        case miUINT32: { uint32_t v; std::memcpy(&v, q, 4); return (float)v; }
        case miSINGLE: { float v; std::memcpy(&v, q, 4); return v; }
        case miDOUBLE: { double v; std::memcpy(&v, q, 8); return (float)v; }
        case miINT64: { int64_t v; std::memcpy(&v, q, 8); return (float)v; }
        case miUINT64: { uint64_t v; std::memcpy(&v, q, 8); return (float)v; }
        default: return 0.f;
    }
}

// A 3D numeric array into `out` (nx*ny*nz floats) in the library layout
// (cpp:377-411 ConvertToRowBased).
bool to_row_major(const MatArray& a, int nx, int ny, int nz, float* out) {
    if (a.dims.size() != 3 || a.dims[0] != nx || a.dims[1] != ny || a.dims[2] != nz) return false;
    const size_t n = (size_t)nx * ny * nz;
    if (mi_size(a.real.type) == 0 || (size_t)a.real.size < n * mi_size(a.real.type)) return false;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                out[((size_t)z * nx + x) * ny + y] =
                    value_at(a.real, (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z));
    return true;
}

bool read_file(const char* path, std::vector<uint8_t>& bytes) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (len < 0) {
        std::fclose(f);
        return false;
    }
    bytes.resize((size_t)len);
    const bool ok = len == 0 || std::fread(bytes.data(), 1, (size_t)len, f) == (size_t)len;
    std::fclose(f);
    return ok;
}

void put_u32(std::vector<uint8_t>& b, uint32_t v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + 4);
}

void pad8(std::vector<uint8_t>& b) {
    while (b.size() % 8) b.push_back(0);
}

}  // namespace

extern "C" int pct_mat_read(const char* path, pct_volume_set* out) {
    if (!path || !out) return PCT_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) return PCT_ERR_IO;
    if (bytes.size() < 128) return PCT_ERR_FORMAT;
    uint16_t version;
    std::memcpy(&version, bytes.data() + 124, 2);
    if (bytes[126] != 'I' || bytes[127] != 'M' || version != 0x0100) return PCT_ERR_FORMAT;
    Store store;
    Cursor c{bytes.data(), bytes.size(), 128, &store};
    std::vector<MatArray> vars;
    while (c.pos < c.n) {
        Element e;
        if (!c.next(e)) return PCT_ERR_FORMAT;
        MatArray a;
        if (!parse_matrix(e, &store, a)) return PCT_ERR_FORMAT;
        vars.push_back(std::move(a));
    }
    // cpp:359-374 Choice: "atlas" and "atlas_label" are the structs; any
    // other variable is the image (cpp:687-689)
    const MatArray *atlas = nullptr, *label = nullptr, *img = nullptr;
    for (const MatArray& v : vars) {
        if (v.cls == mxSTRUCT && v.name == "atlas") atlas = &v;
        else if (v.cls == mxSTRUCT && v.name == "atlas_label") label = &v;
        else if (v.cls != mxSTRUCT) img = &v;
    }
    if (!atlas || !label || !img) return PCT_ERR_FORMAT;
    // cpp:715-801 FindDimension: one shape for every volume, one atlas count
    if (img->dims.size() != 3 || atlas->values.empty() ||
        atlas->values.size() != label->values.size())
        return PCT_ERR_DIMS;
    const long long num_ll = (long long)atlas->values.size();
    long long per_volume = 0, total = 0;
    // the atlases and labels hold 2 * num volumes; every product is checked
    // before the allocations, so no count can wrap into a short buffer
    if (!checked_volume_elements(2 * num_ll + 1, img->dims[0], img->dims[1], img->dims[2],
                                 &per_volume, &total))
        return img->dims[0] < 1 || img->dims[1] < 1 || img->dims[2] < 1 ? PCT_ERR_DIMS
                                                                       : PCT_ERR_TOO_LARGE;
    const int nx = (int)img->dims[0], ny = (int)img->dims[1], nz = (int)img->dims[2];
    const int num = (int)num_ll;
    for (int k = 0; k < num; ++k)
        if (atlas->values[k].dims != img->dims || label->values[k].dims != img->dims)
            return PCT_ERR_DIMS;
    const size_t n = (size_t)per_volume;
    float* im = static_cast<float*>(std::malloc(sizeof(float) * n));
    float* at = static_cast<float*>(std::malloc(sizeof(float) * n * num));
    float* lb = static_cast<float*>(std::malloc(sizeof(float) * n * num));
    bool ok = im && at && lb && to_row_major(*img, nx, ny, nz, im);
    // fields in file order (cpp:670-684 AnalyzeStructArray per field index)
    for (int k = 0; ok && k < num; ++k)
        ok = to_row_major(atlas->values[k], nx, ny, nz, at + n * k) &&
             to_row_major(label->values[k], nx, ny, nz, lb + n * k);
    if (!ok) {
        std::free(im);
        std::free(at);
        std::free(lb);
        return (im && at && lb) ? PCT_ERR_FORMAT : PCT_ERR_NO_MEMORY;
    }
    out->nx = nx;
    out->ny = ny;
    out->nz = nz;
    out->n_atlas = num;
    out->img = im;
    out->atlas = at;
    out->label = lb;
    return PCT_OK;
}

extern "C" int pct_mat_write(const char* path, const char* name, const float* data, int nx,
                             int ny, int nz) {
    if (!path || !name || !data || nx < 1 || ny < 1 || nz < 1) return PCT_ERR_INVALID;
    const size_t n = (size_t)nx * ny * nz;
    if (n * 4 > 0xffffffffull - 256) return PCT_ERR_TOO_LARGE;
    const size_t name_len = std::strlen(name);
    if (name_len == 0 || name_len > 63) return PCT_ERR_INVALID;
    std::vector<uint8_t> b;
    // 116-byte descriptive text, 8-byte subsystem offset, version, endian
    char text[117];
    std::snprintf(text, sizeof(text), "MATLAB 5.0 MAT-file, written by libpct");
    b.insert(b.end(), text, text + std::strlen(text));
    b.resize(116, ' ');
    b.resize(124, 0);
    b.push_back(0x00);
    b.push_back(0x01);
    b.push_back('I');
    b.push_back('M');
    // one miMATRIX element: flags, dimensions, name, single-precision data
    std::vector<uint8_t> m;
    put_u32(m, miUINT32);
    put_u32(m, 8);
    put_u32(m, (uint32_t)mxSINGLE);
    put_u32(m, 0);
    put_u32(m, miINT32);
    put_u32(m, 12);
    put_u32(m, (uint32_t)nx);
    put_u32(m, (uint32_t)ny);
    put_u32(m, (uint32_t)nz);
    pad8(m);
    put_u32(m, miINT8);
    put_u32(m, (uint32_t)name_len);
    m.insert(m.end(), name, name + name_len);
    pad8(m);
    put_u32(m, miSINGLE);
    put_u32(m, (uint32_t)(n * 4));
    const size_t data_at = m.size();
    m.resize(data_at + n * 4);
    // cpp:835-846 ConvertMatrixMatlab: column-major order
    float* col = reinterpret_cast<float*>(m.data() + data_at);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                std::memcpy(col + (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z),
                            data + ((size_t)z * nx + x) * ny + y, 4);
    pad8(m);
    put_u32(b, miMATRIX);
    put_u32(b, (uint32_t)m.size());
    b.insert(b.end(), m.begin(), m.end());
    FILE* f = std::fopen(path, "wb");
    if (!f) return PCT_ERR_IO;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return (std::fclose(f) == 0 && ok) ? PCT_OK : PCT_ERR_IO;
}
