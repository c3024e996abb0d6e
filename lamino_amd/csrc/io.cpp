// Request ingest and result output on the host: projection stacks and vector
// fields as TIFF, vector fields as VTK ImageData (.vti), angle files.
//
//   TIFF read   include/tiff.h:31-85   (32-bit float pages, the odd-ification of
//                                       the detector with its width/height checks
//                                       crossed, :54-59)
//   TIFF write  include/tiff.h:87-126  (float pages, one strip, no compression)
//   vectors     include/tiff.h:128-174 (3 float samples per pixel, RGB, LZW)
//   write3      include/tiff.h:176-188 (one file per component: base + i + ".tiff")
//   VTI         include/write_vti.h:36-111 (appended raw Float64, 32-bit size header)
//   angles      include/config.h:61-83 (whitespace separated; degrees when the
//                                       largest value exceeds 2 pi in magnitude)
//
// The TIFF codec is this file's own: baseline little- or big-endian files with
// strips (uncompressed or LZW, TIFF 6.0 section 13), which covers what the
// reference's libtiff calls write.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "capi_util.h"

namespace lam {
namespace {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------------ LZW (TIFF variant)
// Codes start 9 bits wide; 256 clears the table, 257 ends the strip, new codes
// from 258.  The width grows one code early (at 511, 1023, 2047), and the
// table is cleared before it would need a 13th bit.  Bits are packed MSB first.
struct BitWriter {
    std::vector<uint8_t> &out;
    uint32_t acc = 0;
    int nbits = 0;
    void put(unsigned code, int width) {
        acc = (acc << width) | code;
        nbits += width;
        while (nbits >= 8) {
            out.push_back((uint8_t)(acc >> (nbits - 8)));
            nbits -= 8;
        }
        acc &= (1u << nbits) - 1u;
    }
    void flush() {
        if (nbits > 0) out.push_back((uint8_t)(acc << (8 - nbits)));
        nbits = 0;
        acc = 0;
    }
};

std::vector<uint8_t> lzw_encode(const uint8_t *data, size_t n) {
    constexpr unsigned kClear = 256, kEoi = 257, kFirst = 258, kFull = 4094;
    std::vector<uint8_t> out;
    out.reserve(n / 2 + 16);
    BitWriter bw{out};
    // child table: (prefix code, byte) -> code, as a flat hash of 2^14 slots
    constexpr unsigned kSlots = 1u << 14;
    std::vector<int32_t> key(kSlots), val(kSlots);
    auto reset = [&]() { std::fill(key.begin(), key.end(), -1); };
    reset();
    unsigned next = kFirst;
    int width = 9;
    bw.put(kClear, width);
    if (n == 0) {
        bw.put(kEoi, width);
        bw.flush();
        return out;
    }
    unsigned cur = data[0];
    for (size_t i = 1; i < n; ++i) {
        const uint8_t b = data[i];
        const int32_t k = (int32_t)((cur << 8) | b);
        unsigned h = ((unsigned)k * 2654435761u) >> 18;
        int found = -1;
        while (key[h] != -1) {
            if (key[h] == k) { found = val[h]; break; }
            h = (h + 1) & (kSlots - 1);
        }
        if (found >= 0) {
            cur = (unsigned)found;
            continue;
        }
        bw.put(cur, width);
        key[h] = k;
        val[h] = (int32_t)next;
        ++next;
        // The decoder builds each entry one code later than the encoder and widens
        // when its next free code reaches 2^width - 1 ("early change"), so the
        // encoder widens when its own reaches 2^width; a full table (4094) clears.
        if (next == kFull) {
            bw.put(kClear, width);
            reset();
            next = kFirst;
            width = 9;
        } else if (next >= (1u << width)) {
            ++width;
        }
        cur = b;
    }
    bw.put(cur, width);
    ++next;
    if (next == kFull) {
        bw.put(kClear, width);
        width = 9;
    } else if (next >= (1u << width)) {
        ++width;
    }
    bw.put(kEoi, width);
    bw.flush();
    return out;
}

std::vector<uint8_t> lzw_decode(const uint8_t *src, size_t n, size_t expect) {
    constexpr unsigned kClear = 256, kEoi = 257;
    std::vector<uint8_t> out;
    out.reserve(expect);
    std::vector<std::vector<uint8_t>> table;
    auto reset = [&]() {
        table.assign(258, {});
        for (unsigned c = 0; c < 256; ++c) table[c] = {(uint8_t)c};
    };
    reset();
    int width = 9;
    size_t bitpos = 0;
    int prev = -1;
    auto get = [&](unsigned &code) {
        if (bitpos + (size_t)width > 8 * n) return false;
        unsigned v = 0;
        for (int i = 0; i < width; ++i, ++bitpos) v = (v << 1) | ((src[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
        code = v;
        return true;
    };
    unsigned code;
    while (out.size() < expect && get(code)) {
        if (code == kEoi) break;
        if (code == kClear) {
            reset();
            width = 9;
            prev = -1;
            continue;
        }
        std::vector<uint8_t> entry;
        if (code < table.size()) {
            entry = table[code];
            if (prev >= 0) {
                std::vector<uint8_t> add = table[(size_t)prev];
                add.push_back(entry[0]);
                table.push_back(std::move(add));
            }
        } else {
            if (prev < 0 || code != table.size())
                throw Error(LAM_ERR_ARG, "TIFF: corrupt LZW strip");
            entry = table[(size_t)prev];
            entry.push_back(entry[0]);
            table.push_back(entry);
        }
        out.insert(out.end(), entry.begin(), entry.end());
        prev = (int)code;
        if (table.size() + 1 >= (1u << width) && width < 12) ++width;
    }
    return out;
}

// ------------------------------------------------------------------ TIFF reading
struct Page {
    uint32_t w = 0, h = 0, rows_per_strip = 0;
    uint16_t bps = 0, spp = 1, fmt = 1, comp = 1, planar = 1, predictor = 1;
    std::vector<uint64_t> offs, counts;
    bool tiled = false;
};

struct TiffFile {
    std::vector<uint8_t> buf;
    bool swap = false;
    std::vector<Page> pages;

    uint16_t u16(size_t o) const {
        if (o + 2 > buf.size()) throw Error(LAM_ERR_ARG, "TIFF: truncated file");
        uint16_t v;
        std::memcpy(&v, &buf[o], 2);
        return swap ? (uint16_t)((v >> 8) | (v << 8)) : v;
    }
    uint32_t u32(size_t o) const {
        if (o + 4 > buf.size()) throw Error(LAM_ERR_ARG, "TIFF: truncated file");
        uint32_t v;
        std::memcpy(&v, &buf[o], 4);
        return swap ? __builtin_bswap32(v) : v;
    }
    // value i of an IFD entry of type SHORT (3) or LONG (4)
    uint32_t entry_value(size_t e, uint32_t i) const {
        const uint16_t type = u16(e + 2);
        const uint32_t count = u32(e + 4);
        const size_t sz = type == 3 ? 2 : 4;
        const size_t base = (size_t)count * sz <= 4 ? e + 8 : (size_t)u32(e + 8);
        if (i >= count) throw Error(LAM_ERR_ARG, "TIFF: tag value index out of range");
        return type == 3 ? u16(base + i * 2) : u32(base + i * 4);
    }

    explicit TiffFile(const std::string &path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw Error(LAM_ERR_ARG, "cannot open TIFF file: " + path);
        buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        if (buf.size() < 8) throw Error(LAM_ERR_ARG, "not a TIFF file: " + path);
        if (buf[0] == 'I' && buf[1] == 'I') swap = false;
        else if (buf[0] == 'M' && buf[1] == 'M') swap = true;
        else throw Error(LAM_ERR_ARG, "not a TIFF file: " + path);
        if (u16(2) != 42) throw Error(LAM_ERR_ARG, "not a classic TIFF file (BigTIFF is not read): " + path);
        size_t ifd = u32(4);
        size_t guard_pages = 0;
        while (ifd != 0) {
            if (++guard_pages > (1u << 24)) throw Error(LAM_ERR_ARG, "TIFF: directory loop");
            const uint16_t ne = u16(ifd);
            Page pg;
            for (uint16_t q = 0; q < ne; ++q) {
                const size_t e = ifd + 2 + 12 * (size_t)q;
                const uint16_t tag = u16(e);
                const uint32_t count = u32(e + 4);
                switch (tag) {
                case 256: pg.w = entry_value(e, 0); break;
                case 257: pg.h = entry_value(e, 0); break;
                case 258: pg.bps = (uint16_t)entry_value(e, 0); break;
                case 259: pg.comp = (uint16_t)entry_value(e, 0); break;
                case 273:
                    pg.offs.resize(count);
                    for (uint32_t i = 0; i < count; ++i) pg.offs[i] = entry_value(e, i);
                    break;
                case 277: pg.spp = (uint16_t)entry_value(e, 0); break;
                case 278: pg.rows_per_strip = entry_value(e, 0); break;
                case 279:
                    pg.counts.resize(count);
                    for (uint32_t i = 0; i < count; ++i) pg.counts[i] = entry_value(e, i);
                    break;
                case 284: pg.planar = (uint16_t)entry_value(e, 0); break;
                case 317: pg.predictor = (uint16_t)entry_value(e, 0); break;
                case 322: case 323: case 324: case 325: pg.tiled = true; break;
                case 339: pg.fmt = (uint16_t)entry_value(e, 0); break;
                default: break;
                }
            }
            if (pg.rows_per_strip == 0) pg.rows_per_strip = pg.h;
            pages.push_back(std::move(pg));
            ifd = u32(ifd + 2 + 12 * (size_t)ne);
        }
        if (pages.empty()) throw Error(LAM_ERR_ARG, "TIFF file has no pages: " + path);
    }

    // rows [0, nrows) of page p, first ncols samples of each (one sample per pixel)
    void read_page(size_t p, size_t nrows, size_t ncols, float *out) const {
        const Page &pg = pages[p];
        if (pg.tiled) throw Error(LAM_ERR_ARG, "TIFF: tiled images are not read");
        // like the reference's scanline reads, any 32-bit sample is taken as a float
        if (pg.bps != 32 || pg.spp != 1 || pg.planar != 1)
            throw Error(LAM_ERR_ARG, "TIFF: projections must be one 32-bit sample per pixel");
        if (pg.comp != 1 && pg.comp != 5) throw Error(LAM_ERR_ARG, "TIFF: only uncompressed and LZW strips are read");
        if (pg.predictor != 1) throw Error(LAM_ERR_ARG, "TIFF: predictors are not read");
        if (pg.offs.size() != pg.counts.size() || pg.offs.empty()) throw Error(LAM_ERR_ARG, "TIFF: bad strip tables");
        const size_t line = (size_t)pg.w * 4;
        size_t row = 0;
        for (size_t s = 0; s < pg.offs.size() && row < nrows; ++s) {
            const size_t rows = std::min<size_t>(pg.rows_per_strip, pg.h - row);
            const size_t want = rows * line;
            if (pg.offs[s] + pg.counts[s] > buf.size()) throw Error(LAM_ERR_ARG, "TIFF: strip past end of file");
            std::vector<uint8_t> dec;
            const uint8_t *strip = &buf[pg.offs[s]];
            if (pg.comp == 5) {
                dec = lzw_decode(strip, pg.counts[s], want);
                strip = dec.data();
                if (dec.size() < want) throw Error(LAM_ERR_ARG, "TIFF: short LZW strip");
            } else if (pg.counts[s] < want) {
                throw Error(LAM_ERR_ARG, "TIFF: short strip");
            }
            for (size_t r = 0; r < rows && row < nrows; ++r, ++row) {
                const uint8_t *src = strip + r * line;
                for (size_t k = 0; k < ncols; ++k) {
                    uint32_t v;
                    std::memcpy(&v, src + 4 * k, 4);
                    if (swap) v = __builtin_bswap32(v);
                    std::memcpy(&out[row * ncols + k], &v, 4);
                }
            }
        }
        if (row < nrows) throw Error(LAM_ERR_ARG, "TIFF: fewer rows than the image length");
    }
};

// The stack dims the reference reads (include/tiff.h:51-60): an even width
// drops the last row and an even height the last column.
void stack_dims(const TiffFile &t, int64_t dims[3]) {
    const Page &p0 = t.pages[0];
    for (const Page &p : t.pages)
        if (p.w != p0.w || p.h != p0.h) throw Error(LAM_ERR_ARG, "TIFF: pages differ in size");
    if ((size_t)p0.w * 4 * p0.spp != (size_t)p0.w * sizeof(float) || p0.bps != 32)
        throw Error(LAM_ERR_ARG, "Error: line_size, width mismatch");
    dims[0] = (int64_t)t.pages.size();
    dims[1] = (int64_t)p0.h - (p0.w % 2 == 0 ? 1 : 0);
    dims[2] = (int64_t)p0.w - (p0.h % 2 == 0 ? 1 : 0);
}

// ------------------------------------------------------------------ TIFF writing
struct TiffWriter {
    std::vector<uint8_t> out;
    size_t last_next = 4;   // where the previous "next IFD" offset goes

    TiffWriter() {
        const uint8_t hdr[8] = {'I', 'I', 42, 0, 0, 0, 0, 0};
        out.assign(hdr, hdr + 8);
    }
    void put16(size_t at, uint16_t v) { std::memcpy(&out[at], &v, 2); }
    void put32(size_t at, uint32_t v) { std::memcpy(&out[at], &v, 4); }
    size_t append(const void *p, size_t n) {
        if (out.size() % 2) out.push_back(0);   // word alignment
        const size_t at = out.size();
        const uint8_t *b = static_cast<const uint8_t *>(p);
        out.insert(out.end(), b, b + n);
        if (out.size() > 0xFFFFFFF0ull) throw Error(LAM_ERR_ARG, "TIFF: file exceeds 4 GiB (classic TIFF)");
        return at;
    }
    struct Ent {
        uint16_t tag, type;
        uint32_t count, value;
    };
    // one page: spp 32-bit float samples per pixel, one strip
    void page(uint32_t w, uint32_t h, uint16_t spp, const std::vector<uint8_t> &strip, bool lzw, const char *desc) {
        const size_t strip_at = append(strip.data(), strip.size());
        uint32_t bps_at = 0, fmt_at = 0, desc_at = 0;
        if (spp > 1) {
            std::vector<uint16_t> b(spp, 32), f(spp, 3);
            bps_at = (uint32_t)append(b.data(), 2 * spp);
            fmt_at = (uint32_t)append(f.data(), 2 * spp);
        }
        uint32_t desc_n = 0;
        if (desc) {
            desc_n = (uint32_t)std::strlen(desc) + 1;
            desc_at = (uint32_t)append(desc, desc_n);
        }
        std::vector<Ent> e;
        e.push_back({256, 4, 1, w});
        e.push_back({257, 4, 1, h});
        e.push_back({258, 3, spp, spp > 1 ? bps_at : 32u});
        e.push_back({259, 3, 1, lzw ? 5u : 1u});
        e.push_back({262, 3, 1, spp == 3 ? 2u : 1u});
        if (desc) e.push_back({270, 2, desc_n, desc_at});
        e.push_back({273, 4, 1, (uint32_t)strip_at});
        e.push_back({274, 3, 1, 1u});
        e.push_back({277, 3, 1, spp});
        e.push_back({278, 4, 1, h});
        e.push_back({279, 4, 1, (uint32_t)strip.size()});
        e.push_back({284, 3, 1, 1u});
        e.push_back({339, 3, spp, spp > 1 ? fmt_at : 3u});
        if (out.size() % 2) out.push_back(0);
        const size_t ifd = out.size();
        out.resize(ifd + 2 + 12 * e.size() + 4, 0);
        put16(ifd, (uint16_t)e.size());
        for (size_t q = 0; q < e.size(); ++q) {
            const size_t at = ifd + 2 + 12 * q;
            put16(at, e[q].tag);
            put16(at + 2, e[q].type);
            put32(at + 4, e[q].count);
            if (e[q].type == 3 && e[q].count == 1) put16(at + 8, (uint16_t)e[q].value);
            else put32(at + 8, e[q].value);
        }
        put32(last_next, (uint32_t)ifd);
        last_next = ifd + 2 + 12 * e.size();
    }
    void save(const std::string &path) {
        std::ofstream f(path, std::ios::binary);
        if (!f) throw Error(LAM_ERR_ARG, "cannot open for writing: " + path);
        f.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)out.size());
        if (!f) throw Error(LAM_ERR_ARG, "write failed: " + path);
    }
};

void check_dims(const int64_t *dims) {
    need(dims, "dims");
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 1 || dims[a] > 0x7FFFFFFF) throw Error(LAM_ERR_ARG, "dimensions must be positive");
}

void write_scalar_tiff(const std::string &path, const float *data, const int64_t dims[3]) {
    TiffWriter tw;
    const size_t plane = (size_t)dims[1] * dims[2];
    for (int64_t i = 0; i < dims[0]; ++i) {
        std::vector<uint8_t> strip(plane * 4);
        std::memcpy(strip.data(), data + (size_t)i * plane, plane * 4);
        tw.page((uint32_t)dims[2], (uint32_t)dims[1], 1, strip, false, nullptr);
    }
    tw.save(path);
}

// shortest round-trip text of a double (what std::format("{}") prints)
std::string num(double v) {
    char b[64];
    auto r = std::to_chars(b, b + sizeof(b), v);
    return std::string(b, r.ptr);
}

}  // namespace
}  // namespace lam

using namespace lam;

extern "C" {

lam_status lam_tiff_info(const char *path, int64_t dims[3]) {
    return guard([&] {
        need(path, "path"); need(dims, "dims");
        TiffFile t(path);
        stack_dims(t, dims);
    });
}

lam_status lam_tiff_read(const char *path, float *out_host, const int64_t dims[3]) {
    return guard([&] {
        need(path, "path"); need(out_host, "out_host"); need(dims, "dims");
        TiffFile t(path);
        int64_t d[3];
        stack_dims(t, d);
        LAM_REQUIRE(d[0] == dims[0] && d[1] == dims[1] && d[2] == dims[2], LAM_ERR_SHAPE,
                    "TIFF stack dims differ from the caller's (lam_tiff_info gives them)");
        const size_t plane = (size_t)d[1] * d[2];
        for (int64_t i = 0; i < d[0]; ++i) t.read_page((size_t)i, (size_t)d[1], (size_t)d[2], out_host + i * plane);
    });
}

lam_status lam_tiff_write(const char *path, const float *data_host, const int64_t dims[3]) {
    return guard([&] {
        need(path, "path"); need(data_host, "data_host");
        check_dims(dims);
        write_scalar_tiff(path, data_host, dims);
    });
}

lam_status lam_tiff_write_vectors(const char *path, const float *c0, const float *c1, const float *c2,
                                  const int64_t dims[3]) {
    return guard([&] {
        need(path, "path"); need(c0, "c0"); need(c1, "c1"); need(c2, "c2");
        check_dims(dims);
        TiffWriter tw;
        const size_t plane = (size_t)dims[1] * dims[2];
        std::vector<float> inter(3 * plane);
        for (int64_t i = 0; i < dims[0]; ++i) {
            const size_t o = (size_t)i * plane;
            for (size_t q = 0; q < plane; ++q) {
                inter[3 * q] = c0[o + q];
                inter[3 * q + 1] = c1[o + q];
                inter[3 * q + 2] = c2[o + q];
            }
            const auto strip = lzw_encode(reinterpret_cast<const uint8_t *>(inter.data()), inter.size() * 4);
            tw.page((uint32_t)dims[2], (uint32_t)dims[1], 3, strip, true,
                    "3D vector field: x,y,z components; layout: [slice, row, col]");
        }
        tw.save(path);
    });
}

lam_status lam_tiff_write3(const char *name, const float *c0, const float *c1, const float *c2,
                           const int64_t dims[3]) {
    return guard([&] {
        need(name, "name"); need(c0, "c0"); need(c1, "c1"); need(c2, "c2");
        check_dims(dims);
        std::string base = name;
        const size_t at = base.rfind(".tif");
        if (at != std::string::npos) base = base.substr(0, at);
        const float *c[3] = {c0, c1, c2};
        for (int i = 0; i < 3; ++i) write_scalar_tiff(base + std::to_string(i) + ".tiff", c[i], dims);
    });
}

lam_status lam_vti_write_vectors(const char *path, const float *c0, const float *c1, const float *c2,
                                 const int64_t dims[3], const char *field_name) {
    return guard([&] {
        need(path, "path"); need(c0, "c0"); need(c1, "c1"); need(c2, "c2");
        check_dims(dims);
        const std::string field = field_name ? field_name : "magnetization";
        std::string fp = path;
        {   // std::filesystem::path::replace_extension(".vti")
            const size_t slash = fp.find_last_of('/');
            const size_t dot = fp.find_last_of('.');
            const bool has_ext = dot != std::string::npos && (slash == std::string::npos || dot > slash + 1) &&
                                 fp.substr(dot) != "." && fp.substr(dot) != "..";
            if (has_ext) fp = fp.substr(0, dot);
            fp += ".vti";
        }
        const size_t nz = (size_t)dims[0], ny = (size_t)dims[1], nx = (size_t)dims[2];
        const size_t npoints = nx * ny * nz;
        std::string h;
        h += "<?xml version=\"1.0\"?>\n";
        h += "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        h += "  <ImageData WholeExtent=\"0 " + std::to_string(nx - 1) + " 0 " + std::to_string(ny - 1) + " 0 " +
             std::to_string(nz - 1) + "\" Origin=\"" + num(-0.5 * (double)nx) + " " + num(-0.5 * (double)ny) + " " +
             num(-0.5 * (double)nz) + "\" Spacing=\"" + num(1.0) + " " + num(1.0) + " " + num(1.0) + "\">\n";
        h += "    <Piece Extent=\"0 " + std::to_string(nx - 1) + " 0 " + std::to_string(ny - 1) + " 0 " +
             std::to_string(nz - 1) + "\">\n";
        h += "      <PointData Vectors=\"" + field + "\">\n";
        h += "        <DataArray type=\"Float64\" Name=\"" + field +
             "\" NumberOfComponents=\"3\" format=\"appended\" offset=\"0\"/>\n";
        h += "      </PointData>\n";
        h += "      <CellData></CellData>\n";
        h += "    </Piece>\n";
        h += "  </ImageData>\n";
        h += "  <AppendedData encoding=\"raw\">\n";
        h += "   _";
        std::ofstream f(fp, std::ios::binary);
        if (!f) throw Error(LAM_ERR_ARG, "Could not open file for writing");
        f.write(h.data(), (std::streamsize)h.size());
        // 32-bit byte count, as the reference writes it (it wraps beyond 4 GiB)
        const uint32_t byte_count = (uint32_t)(npoints * 3 * sizeof(double));
        f.write(reinterpret_cast<const char *>(&byte_count), 4);
        std::vector<double> chunk;
        constexpr size_t kChunkPoints = 1 << 16;
        for (size_t i0 = 0; i0 < npoints; i0 += kChunkPoints) {
            const size_t m = std::min(kChunkPoints, npoints - i0);
            chunk.resize(3 * m);
            for (size_t q = 0; q < m; ++q) {
                chunk[3 * q] = (double)c0[i0 + q];
                chunk[3 * q + 1] = (double)c1[i0 + q];
                chunk[3 * q + 2] = (double)c2[i0 + q];
            }
            f.write(reinterpret_cast<const char *>(chunk.data()), (std::streamsize)(chunk.size() * sizeof(double)));
        }
        const std::string tail = "\n  </AppendedData>\n</VTKFile>\n";
        f.write(tail.data(), (std::streamsize)tail.size());
        if (!f) throw Error(LAM_ERR_ARG, "write failed: " + fp);
    });
}

// Angle file (include/config.h:61-83), parsed as float like the reference's
// recon (read_angles_file<float>); *n receives the count.  out may be NULL to
// query the count; otherwise it holds cap values.
lam_status lam_read_angles(const char *path, double *out, int64_t cap, int64_t *n) {
    return guard([&] {
        need(path, "path"); need(n, "n");
        std::ifstream f(path);
        if (!f.is_open()) throw Error(LAM_ERR_ARG, std::string("Could not open angles file: ") + path);
        std::vector<float> a;
        float v;
        while (f >> v) a.push_back(v);
        if (a.empty()) throw Error(LAM_ERR_ARG, std::string("No angles found in file: ") + path);
        const float mx = *std::max_element(a.begin(), a.end());
        if (std::abs(mx) > 2 * kPi)
            for (float &x : a) x = (float)((double)x * kPi / (double)(float)180.0);
        *n = (int64_t)a.size();
        if (out) {
            LAM_REQUIRE(cap >= (int64_t)a.size(), LAM_ERR_ARG, "angles buffer too small");
            for (size_t i = 0; i < a.size(); ++i) out[i] = (double)a[i];
        }
    });
}

}  // extern "C"
