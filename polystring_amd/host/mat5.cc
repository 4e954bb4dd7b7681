// mat5.cc -- MAT v5 I/O (see mat5.h).  Format: 128-byte header (text,
// subsystem offset, version 0x0100, endian "IM"), then tagged data elements
// (type, byte count, payload padded to 8 bytes; small elements pack type and
// count in one 8-byte tag).  A matrix element holds array flags, dimensions,
// name and real part.
#include "mat5.h"

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

namespace mat5 {

namespace {

enum {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6, miSINGLE = 7,
    miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15, miUTF8 = 16
};
enum { mxCHAR_CLASS = 4, mxDOUBLE_CLASS = 6 };

struct Cursor {
    const uint8_t *p, *end;
    bool tag(uint32_t &type, uint32_t &n, const uint8_t *&payload) {
        if (end - p < 8) return false;
        uint32_t w0, w1;
        std::memcpy(&w0, p, 4);
        std::memcpy(&w1, p + 4, 4);
        if (w0 >> 16) {  // small data element: count in the upper half of word 0
            type = w0 & 0xffff;
            n = w0 >> 16;
            payload = p + 4;
            p += 8;
            return n <= 4;
        }
        type = w0;
        n = w1;
        payload = p + 8;
        const size_t padded = type == miCOMPRESSED ? n : (n + 7u) & ~7u;
        if ((size_t)(end - payload) < n) return false;
        p = payload + (padded <= (size_t)(end - payload) ? padded : (size_t)(end - payload));
        return true;
    }
};

size_t elem_size(uint32_t t) {
    switch (t) {
    case miINT8: case miUINT8: case miUTF8: return 1;
    case miINT16: case miUINT16: return 2;
    case miINT32: case miUINT32: case miSINGLE: return 4;
    case miDOUBLE: case miINT64: case miUINT64: return 8;
    default: return 0;
    }
}

double value_at(uint32_t t, const uint8_t *b, size_t i) {
    switch (t) {
    case miINT8: return (double)((const int8_t *)b)[i];
    case miUINT8: case miUTF8: return (double)b[i];
    case miINT16: { int16_t v; std::memcpy(&v, b + 2 * i, 2); return v; }
    case miUINT16: { uint16_t v; std::memcpy(&v, b + 2 * i, 2); return v; }
    case miINT32: { int32_t v; std::memcpy(&v, b + 4 * i, 4); return v; }
    case miUINT32: { uint32_t v; std::memcpy(&v, b + 4 * i, 4); return v; }
    case miSINGLE: { float v; std::memcpy(&v, b + 4 * i, 4); return v; }
    case miDOUBLE: { double v; std::memcpy(&v, b + 8 * i, 8); return v; }
    case miINT64: { int64_t v; std::memcpy(&v, b + 8 * i, 8); return (double)v; }
    case miUINT64: { uint64_t v; std::memcpy(&v, b + 8 * i, 8); return (double)v; }
    default: return 0.0;
    }
}

bool parse_matrix(const uint8_t *p, size_t n, std::map<std::string, Var> &out) {
    Cursor c{p, p + n};
    uint32_t t, cnt;
    const uint8_t *pl;
    if (!c.tag(t, cnt, pl) || t != miUINT32 || cnt < 8) return false;  // array flags
    uint32_t flags;
    std::memcpy(&flags, pl, 4);
    const int cls = flags & 0xff;
    const bool complex_ = (flags >> 11) & 1;
    if (!c.tag(t, cnt, pl) || t != miINT32) return false;  // dimensions
    Var v;
    for (uint32_t i = 0; i < cnt / 4; ++i) v.dims.push_back((int)value_at(miINT32, pl, i));
    if (!c.tag(t, cnt, pl) || t != miINT8) return false;  // name
    std::string name((const char *)pl, cnt);
    if (!c.tag(t, cnt, pl)) return true;  // empty array
    const size_t es = elem_size(t);
    if (es == 0 || complex_) return true;  // cells, structs, complex: skipped
    const size_t count = cnt / es;
    v.data.resize(count);
    for (size_t i = 0; i < count; ++i) v.data[i] = value_at(t, pl, i);
    if (cls == mxCHAR_CLASS) {
        v.is_char = true;
        for (double d : v.data) v.text.push_back((char)d);
    }
    out[name] = std::move(v);
    return true;
}

bool parse_elements(const uint8_t *p, size_t n, std::map<std::string, Var> &out, std::string &err) {
    Cursor c{p, p + n};
    uint32_t t, cnt;
    const uint8_t *pl;
    while (c.p < c.end) {
        if (!c.tag(t, cnt, pl)) {
            err = "truncated data element";
            return false;
        }
        if (t == miMATRIX) {
            if (!parse_matrix(pl, cnt, out)) {
                err = "malformed matrix element";
                return false;
            }
        } else if (t == miCOMPRESSED) {
            std::vector<uint8_t> buf(cnt * 4 + 1024);
            z_stream zs{};
            if (inflateInit(&zs) != Z_OK) { err = "zlib init failed"; return false; }
            zs.next_in = const_cast<uint8_t *>(pl);
            zs.avail_in = cnt;
            int rc;
            do {
                if (zs.total_out == buf.size()) buf.resize(buf.size() * 2);
                zs.next_out = buf.data() + zs.total_out;
                zs.avail_out = (uInt)(buf.size() - zs.total_out);
                rc = inflate(&zs, Z_NO_FLUSH);
            } while (rc == Z_OK);
            const size_t got = zs.total_out;
            inflateEnd(&zs);
            if (rc != Z_STREAM_END) { err = "zlib inflate failed"; return false; }
            if (!parse_elements(buf.data(), got, out, err)) return false;
        }
    }
    return true;
}

void write_tag(FILE *f, uint32_t type, uint32_t n) {
    std::fwrite(&type, 4, 1, f);
    std::fwrite(&n, 4, 1, f);
}

void pad(FILE *f, size_t n) {
    static const uint8_t z[8] = {0};
    if (n % 8) std::fwrite(z, 1, 8 - n % 8, f);
}

}  // namespace

bool read(const std::string &path, std::map<std::string, Var> &out, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open MAT-file: " + path;
        return false;
    }
    std::vector<uint8_t> all;
    uint8_t chunk[65536];
    size_t k;
    while ((k = std::fread(chunk, 1, sizeof(chunk), f)) > 0) all.insert(all.end(), chunk, chunk + k);
    std::fclose(f);
    if (all.size() < 128 || all[126] != 'I' || all[127] != 'M') {
        err = "not a little-endian MAT v5 file: " + path;
        return false;
    }
    return parse_elements(all.data() + 128, all.size() - 128, out, err);
}

Writer::Writer(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "wb");
    f_ = f;
    if (!f) return;
    char hdr[128];
    std::memset(hdr, ' ', sizeof(hdr));
    const char *text = "MATLAB 5.0 MAT-file, written by polystring_amd";
    std::memcpy(hdr, text, std::strlen(text));
    std::memset(hdr + 116, 0, 8);
    hdr[124] = 0x00;
    hdr[125] = 0x01;
    hdr[126] = 'I';
    hdr[127] = 'M';
    std::fwrite(hdr, 1, 128, f);
}

Writer::~Writer() {
    if (f_) std::fclose((FILE *)f_);
}

void Writer::put(const std::string &name, const std::vector<int> &dims, const double *colmajor) {
    FILE *f = (FILE *)f_;
    if (!f) return;
    size_t count = 1;
    for (int d : dims) count *= (size_t)d;
    const uint32_t nlen = (uint32_t)name.size();
    const uint32_t dbytes = 4 * (uint32_t)dims.size();
    const uint32_t body = 16 + 8 + ((dbytes + 7) & ~7u) + 8 + ((nlen + 7) & ~7u) + 8 + (uint32_t)(8 * count);
    write_tag(f, miMATRIX, body);
    write_tag(f, miUINT32, 8);
    const uint32_t flags[2] = {mxDOUBLE_CLASS, 0};
    std::fwrite(flags, 4, 2, f);
    write_tag(f, miINT32, dbytes);
    for (int d : dims) { int32_t v = d; std::fwrite(&v, 4, 1, f); }
    pad(f, dbytes);
    write_tag(f, miINT8, nlen);
    std::fwrite(name.data(), 1, nlen, f);
    pad(f, nlen);
    write_tag(f, miDOUBLE, (uint32_t)(8 * count));
    std::fwrite(colmajor, 8, count, f);
}

void Writer::put_string(const std::string &name, const std::string &s) {
    FILE *f = (FILE *)f_;
    if (!f) return;
    const uint32_t nlen = (uint32_t)name.size(), cbytes = 2 * (uint32_t)s.size();
    const uint32_t body = 16 + 16 + 8 + ((nlen + 7) & ~7u) + 8 + ((cbytes + 7) & ~7u);
    write_tag(f, miMATRIX, body);
    write_tag(f, miUINT32, 8);
    const uint32_t flags[2] = {mxCHAR_CLASS, 0};
    std::fwrite(flags, 4, 2, f);
    write_tag(f, miINT32, 8);
    const int32_t dims[2] = {1, (int32_t)s.size()};
    std::fwrite(dims, 4, 2, f);
    write_tag(f, miINT8, nlen);
    std::fwrite(name.data(), 1, nlen, f);
    pad(f, nlen);
    write_tag(f, miUINT16, cbytes);
    for (char ch : s) { uint16_t v = (uint8_t)ch; std::fwrite(&v, 2, 1, f); }
    pad(f, cbytes);
}

std::vector<double> c_to_fortran(const double *c, int Lx, int Ly, int Lz) {
    std::vector<double> f((size_t)Lx * Ly * Lz);
    for (int i = 0; i < Lx; ++i)
        for (int j = 0; j < Ly; ++j)
            for (int k = 0; k < Lz; ++k) f[i + (size_t)Lx * (j + (size_t)Ly * k)] = c[((size_t)i * Ly + j) * Lz + k];
    return f;
}

std::vector<double> fortran_to_c(const double *f, int Lx, int Ly, int Lz) {
    std::vector<double> c((size_t)Lx * Ly * Lz);
    for (int i = 0; i < Lx; ++i)
        for (int j = 0; j < Ly; ++j)
            for (int k = 0; k < Lz; ++k) c[((size_t)i * Ly + j) * Lz + k] = f[i + (size_t)Lx * (j + (size_t)Ly * k)];
    return c;
}

}  // namespace mat5
