/* I/O contract of the reference CLI, natively (no libtiff): single-band
 * float32 GeoTIFF read/write (geotiff_io.cpp:50-178), the TMY series reader
 * (main.cpp:30-40) and the voxel dump (main.cpp:44-67).
 *
 * Reader: classic TIFF, either byte order; strips or tiles; compression none
 * (1) or deflate (8 / 32946) via zlib; predictor none (1), horizontal (2) or
 * floating point (3, byte planes + byte differencing per row).  GeoTIFF tags
 * 33550 / 33922 / 34735 / 34736 / 34737 and GDAL_NODATA 42113 are kept.
 * Writer: little-endian, one strip, compression copied from the input
 * (deflate or none), no predictor, the geo tags carried over -- the
 * reference's write_tiff behaviour (geotiff_io.cpp:145-178). */
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "sgpu_io.h"

namespace sgpu {
namespace {

struct Reader {
    const std::vector<uint8_t>& b;
    bool le;
    uint16_t u16(size_t o) const {
        if (o + 2 > b.size()) fail(SGPU_ERR_FORMAT, "TIFF: read past end");
        return le ? (uint16_t)(b[o] | (b[o + 1] << 8)) : (uint16_t)((b[o] << 8) | b[o + 1]);
    }
    uint32_t u32(size_t o) const {
        if (o + 4 > b.size()) fail(SGPU_ERR_FORMAT, "TIFF: read past end");
        return le ? (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8) | ((uint32_t)b[o + 2] << 16) | ((uint32_t)b[o + 3] << 24)
                  : ((uint32_t)b[o] << 24) | ((uint32_t)b[o + 1] << 16) | ((uint32_t)b[o + 2] << 8) | b[o + 3];
    }
    double f64(size_t o) const {
        uint64_t v = 0;
        if (o + 8 > b.size()) fail(SGPU_ERR_FORMAT, "TIFF: read past end");
        for (int i = 0; i < 8; ++i) v |= (uint64_t)b[o + (le ? i : 7 - i)] << (8 * i);
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
};

size_t type_size(uint16_t t) {
    switch (t) {
        case 1: case 2: case 6: case 7: return 1;   // BYTE ASCII SBYTE UNDEFINED
        case 3: case 8: return 2;                   // SHORT SSHORT
        case 4: case 9: case 11: return 4;          // LONG SLONG FLOAT
        case 5: case 10: case 12: return 8;         // RATIONAL SRATIONAL DOUBLE
        default: return 0;
    }
}

struct Entry {
    uint16_t tag, type;
    uint32_t count;
    size_t data;   // byte offset of the value(s) in the file
};

std::vector<uint64_t> ints(const Reader& r, const Entry& e) {
    std::vector<uint64_t> v(e.count);
    for (uint32_t i = 0; i < e.count; ++i) {
        if (e.type == 3) v[i] = r.u16(e.data + 2 * i);
        else if (e.type == 4) v[i] = r.u32(e.data + 4 * i);
        else if (e.type == 1) v[i] = r.b.at(e.data + i);
        else fail(SGPU_ERR_FORMAT, "TIFF: tag %u has non-integer type %u", e.tag, e.type);
    }
    return v;
}

std::vector<uint8_t> inflate_all(const uint8_t* src, size_t n, size_t expect) {
    std::vector<uint8_t> out(expect);
    z_stream z{};
    if (inflateInit(&z) != Z_OK) fail(SGPU_ERR_FORMAT, "TIFF: zlib init failed");
    z.next_in = const_cast<uint8_t*>(src);
    z.avail_in = (uInt)n;
    z.next_out = out.data();
    z.avail_out = (uInt)out.size();
    const int rc = inflate(&z, Z_FINISH);
    const size_t got = out.size() - z.avail_out;
    inflateEnd(&z);
    if (rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) fail(SGPU_ERR_FORMAT, "TIFF: corrupt deflate data");
    if (got != expect) fail(SGPU_ERR_FORMAT, "TIFF: deflate block holds %zu bytes, expected %zu", got, expect);
    return out;
}

/* Undo a predictor on one decoded block of `rows` rows of `w` float32 samples. */
void unpredict(std::vector<uint8_t>& blk, int w, int rows, int predictor, bool le) {
    const size_t row_bytes = (size_t)w * 4;
    if (predictor == 3) {
        std::vector<uint8_t> tmp(row_bytes);
        for (int r = 0; r < rows; ++r) {
            uint8_t* p = blk.data() + r * row_bytes;
            for (size_t i = 1; i < row_bytes; ++i) p[i] = (uint8_t)(p[i] + p[i - 1]);
            /* byte planes, most significant first -> native little-endian floats */
            for (int i = 0; i < w; ++i)
                for (int k = 0; k < 4; ++k) tmp[(size_t)i * 4 + (3 - k)] = p[(size_t)k * w + i];
            std::memcpy(p, tmp.data(), row_bytes);
        }
    } else if (predictor == 2) {
        for (int r = 0; r < rows; ++r) {
            uint8_t* p = blk.data() + r * row_bytes;
            uint32_t prev = 0;
            for (int i = 0; i < w; ++i) {
                uint32_t v;
                std::memcpy(&v, p + 4 * i, 4);
                if (!le) v = __builtin_bswap32(v);
                v += prev;
                prev = v;
                std::memcpy(p + 4 * i, &v, 4);
            }
        }
    } else if (!le) {
        for (size_t i = 0; i + 4 <= blk.size(); i += 4) {
            std::swap(blk[i], blk[i + 3]);
            std::swap(blk[i + 1], blk[i + 2]);
        }
    }
}

}  // namespace

GeoTiff read_geotiff(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(SGPU_ERR_IO, "cannot open %s", path.c_str());
    std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (b.size() < 8) fail(SGPU_ERR_FORMAT, "%s: not a TIFF file", path.c_str());
    Reader r{b, b[0] == 'I'};
    if (!((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M')) || r.u16(2) != 42)
        fail(SGPU_ERR_FORMAT, "%s: not a classic TIFF file", path.c_str());
    const size_t ifd = r.u32(4);
    const uint16_t n = r.u16(ifd);
    std::vector<Entry> es;
    for (uint16_t i = 0; i < n; ++i) {
        const size_t o = ifd + 2 + 12 * (size_t)i;
        Entry e{r.u16(o), r.u16(o + 2), r.u32(o + 4), 0};
        const size_t sz = type_size(e.type) * e.count;
        e.data = (sz <= 4) ? o + 8 : r.u32(o + 8);
        /* a tag of a type this reader does not know is kept but never read (typed() refuses it) */
        if (type_size(e.type) && e.data + sz > b.size()) fail(SGPU_ERR_FORMAT, "%s: tag %u out of file", path.c_str(), e.tag);
        es.push_back(e);
    }
    /* A tag whose values are read as `type`: its type must be that one (so its byte range was
       checked above against the file size). */
    auto typed = [&](const Entry* e, uint16_t type) -> const Entry* {
        if (e && e->type != type)
            fail(SGPU_ERR_FORMAT, "%s: tag %u has type %u, expected %u", path.c_str(), e->tag, e->type, type);
        return e;
    };
    auto find = [&](uint16_t tag) -> const Entry* {
        for (const Entry& e : es)
            if (e.tag == tag) return &e;
        return nullptr;
    };
    auto one = [&](uint16_t tag, uint64_t dflt) -> uint64_t {
        const Entry* e = find(tag);
        return e ? ints(r, *e).at(0) : dflt;
    };
    GeoTiff g;
    g.width = (int)one(256, 0);
    g.height = (int)one(257, 0);
    const int bps = (int)one(258, 1), fmt = (int)one(339, 1), spp = (int)one(277, 1);
    g.compression = (int)one(259, 1);
    const int predictor = (int)one(317, 1);
    if (g.width <= 0 || g.height <= 0) fail(SGPU_ERR_FORMAT, "%s: missing image size", path.c_str());
    if (bps != 32 || fmt != 3 || spp != 1)
        fail(SGPU_ERR_FORMAT, "Expected 32-bit float single-band TIFF, got bps=%d fmt=%d spp=%d (%s)", bps, fmt, spp,
             path.c_str());
    if (g.compression != 1 && g.compression != 8 && g.compression != 32946)
        fail(SGPU_ERR_FORMAT, "%s: unsupported compression %d", path.c_str(), g.compression);
    if (const Entry* e = typed(find(33550), 12))
        for (uint32_t i = 0; i < e->count; ++i) g.pixel_scale.push_back(r.f64(e->data + 8 * i));
    if (const Entry* e = typed(find(33922), 12))
        for (uint32_t i = 0; i < e->count; ++i) g.tiepoints.push_back(r.f64(e->data + 8 * i));
    if (const Entry* e = typed(find(34735), 3))
        for (uint32_t i = 0; i < e->count; ++i) g.geo_keys.push_back(r.u16(e->data + 2 * i));
    if (const Entry* e = typed(find(34736), 12))
        for (uint32_t i = 0; i < e->count; ++i) g.geo_doubles.push_back(r.f64(e->data + 8 * i));
    if (const Entry* e = typed(find(34737), 2)) g.geo_ascii.assign((const char*)b.data() + e->data, e->count);
    if (const Entry* e = typed(find(42113), 2)) g.nodata.assign((const char*)b.data() + e->data, e->count);
    g.data.assign((size_t)g.width * g.height, 0.0f);
    const bool tiled = find(322) != nullptr;
    const int bw = tiled ? (int)one(322, 0) : g.width;
    const int bh = tiled ? (int)one(323, 0) : (int)std::min<uint64_t>(one(278, (uint64_t)g.height), (uint64_t)g.height);
    g.rows_per_strip = tiled ? g.height : bh;
    if (bw <= 0 || bh <= 0) fail(SGPU_ERR_FORMAT, "%s: bad block size", path.c_str());
    const Entry* eo = find(tiled ? 324 : 273);
    const Entry* ec = find(tiled ? 325 : 279);
    if (!eo || !ec) fail(SGPU_ERR_FORMAT, "%s: missing block offsets", path.c_str());
    const std::vector<uint64_t> offs = ints(r, *eo), cnts = ints(r, *ec);
    const int across = tiled ? (g.width + bw - 1) / bw : 1;
    const int down = (g.height + bh - 1) / bh;
    if ((int)offs.size() < across * down || cnts.size() < offs.size())
        fail(SGPU_ERR_FORMAT, "%s: too few blocks", path.c_str());
    for (int by = 0; by < down; ++by) {
        for (int bx = 0; bx < across; ++bx) {
            const int k = by * across + bx;
            const int rows = tiled ? bh : std::min(bh, g.height - by * bh);
            const size_t expect = (size_t)bw * rows * 4;
            if (offs[k] + cnts[k] > b.size()) fail(SGPU_ERR_FORMAT, "%s: block %d out of file", path.c_str(), k);
            if (g.compression == 1 && cnts[k] < expect)   // before any byte of the block is read
                fail(SGPU_ERR_FORMAT, "%s: block %d holds %llu bytes, expected %zu", path.c_str(), k,
                     (unsigned long long)cnts[k], expect);
            std::vector<uint8_t> blk = (g.compression == 1)
                                           ? std::vector<uint8_t>(b.begin() + offs[k], b.begin() + offs[k] + expect)
                                           : inflate_all(b.data() + offs[k], cnts[k], expect);
            unpredict(blk, bw, rows, predictor, r.le);
            for (int y = 0; y < rows && by * bh + y < g.height; ++y) {
                const int cw = std::min(bw, g.width - bx * bw);
                std::memcpy(&g.data[(size_t)(by * bh + y) * g.width + bx * bw], &blk[(size_t)y * bw * 4], (size_t)cw * 4);
            }
        }
    }
    return g;
}

namespace {

struct Writer {
    std::vector<uint8_t> b;
    void u16(uint16_t v) { b.push_back(v & 0xFF); b.push_back(v >> 8); }
    void u32(uint32_t v) { for (int i = 0; i < 4; ++i) b.push_back((v >> (8 * i)) & 0xFF); }
    void put_at32(size_t o, uint32_t v) { for (int i = 0; i < 4; ++i) b[o + i] = (v >> (8 * i)) & 0xFF; }
};

}  // namespace

void write_geotiff(const std::string& path, const float* data, int w, int h, const GeoTiff& meta) {
    const size_t raw_bytes = (size_t)w * h * 4;
    const bool deflate = meta.compression == 8 || meta.compression == 32946;
    std::vector<uint8_t> strip;
    if (deflate) {
        uLongf cap = compressBound((uLong)raw_bytes);
        strip.resize(cap);
        if (compress2(strip.data(), &cap, (const Bytef*)data, (uLong)raw_bytes, 6) != Z_OK)
            fail(SGPU_ERR_IO, "deflate failed for %s", path.c_str());
        strip.resize(cap);
    } else {
        strip.assign((const uint8_t*)data, (const uint8_t*)data + raw_bytes);
    }
    struct Tag { uint16_t tag, type; uint32_t count; std::vector<uint8_t> bytes; };
    std::vector<Tag> tags;
    auto short_tag = [&](uint16_t t, uint16_t v) { tags.push_back({t, 3, 1, {(uint8_t)(v & 0xFF), (uint8_t)(v >> 8), 0, 0}}); };
    auto long_tag = [&](uint16_t t, uint32_t v) {
        tags.push_back({t, 4, 1, {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)}});
    };
    auto doubles = [&](uint16_t t, const std::vector<double>& v) {
        if (v.empty()) return;
        std::vector<uint8_t> bytes(v.size() * 8);
        std::memcpy(bytes.data(), v.data(), bytes.size());
        tags.push_back({t, 12, (uint32_t)v.size(), bytes});
    };
    long_tag(256, (uint32_t)w);
    long_tag(257, (uint32_t)h);
    short_tag(258, 32);
    short_tag(259, deflate ? 8 : 1);
    short_tag(262, 1);                // min-is-black
    long_tag(273, 0);                 // strip offset, patched below
    short_tag(277, 1);
    long_tag(278, (uint32_t)h);       // one strip
    long_tag(279, (uint32_t)strip.size());
    short_tag(284, 1);
    short_tag(339, 3);                // IEEE float
    doubles(33550, meta.pixel_scale);
    doubles(33922, meta.tiepoints);
    if (!meta.geo_keys.empty()) {
        std::vector<uint8_t> bytes(meta.geo_keys.size() * 2);
        std::memcpy(bytes.data(), meta.geo_keys.data(), bytes.size());
        tags.push_back({34735, 3, (uint32_t)meta.geo_keys.size(), bytes});
    }
    doubles(34736, meta.geo_doubles);
    if (!meta.geo_ascii.empty()) {
        std::vector<uint8_t> bytes(meta.geo_ascii.begin(), meta.geo_ascii.end());
        tags.push_back({34737, 2, (uint32_t)bytes.size(), bytes});
    }
    Writer o;
    o.b = {'I', 'I'};
    o.u16(42);
    o.u32(8);
    const size_t ifd_size = 2 + 12 * tags.size() + 4;
    size_t extra = 8 + ifd_size;
    std::vector<size_t> where(tags.size());
    for (size_t i = 0; i < tags.size(); ++i) {
        if (tags[i].bytes.size() > 4) {
            where[i] = extra;
            extra += (tags[i].bytes.size() + 1) & ~(size_t)1;
        }
    }
    const size_t strip_at = extra;
    o.u16((uint16_t)tags.size());
    for (size_t i = 0; i < tags.size(); ++i) {
        const Tag& t = tags[i];
        o.u16(t.tag);
        o.u16(t.type);
        o.u32(t.count);
        if (t.bytes.size() > 4) o.u32((uint32_t)where[i]);
        else {
            std::vector<uint8_t> v = t.bytes;
            v.resize(4, 0);
            if (t.tag == 273) { v = {(uint8_t)strip_at, (uint8_t)(strip_at >> 8), (uint8_t)(strip_at >> 16), (uint8_t)(strip_at >> 24)}; }
            o.b.insert(o.b.end(), v.begin(), v.end());
        }
    }
    o.u32(0);   // no next IFD
    for (size_t i = 0; i < tags.size(); ++i)
        if (tags[i].bytes.size() > 4) {
            o.b.insert(o.b.end(), tags[i].bytes.begin(), tags[i].bytes.end());
            if (tags[i].bytes.size() & 1) o.b.push_back(0);
        }
    o.b.insert(o.b.end(), strip.begin(), strip.end());
    std::ofstream f(path, std::ios::binary);
    if (!f) fail(SGPU_ERR_IO, "cannot create %s", path.c_str());
    f.write((const char*)o.b.data(), (std::streamsize)o.b.size());
    if (!f) fail(SGPU_ERR_IO, "write failed for %s", path.c_str());
}

std::vector<float> read_tmy(const std::string& path, size_t n) {
    std::ifstream f(path);
    if (!f) fail(SGPU_ERR_IO, "Cannot open TMY irradiance file: %s", path.c_str());
    std::vector<float> v;
    v.reserve(n);
    float x;
    while (v.size() < n && (f >> x)) v.push_back(x);
    if (v.size() != n) fail(SGPU_ERR_FORMAT, "%s: %zu values, expected %zu", path.c_str(), v.size(), n);
    return v;
}

void write_voxel_dump(const std::string& path, int n_total, int n_top, int w, const int32_t* pp, const float* grid,
                      const float* wall_z, const float* irr) {
    std::ofstream f(path, std::ios::binary);
    if (!f) fail(SGPU_ERR_IO, "cannot write %s", path.c_str());
    f.write((const char*)&n_total, 4);
    f.write((const char*)&n_top, 4);
    std::vector<float> rec(4 * (size_t)n_total);
    for (int i = 0; i < n_total; ++i) {
        const int top = pp[i];
        const int gx = top % w, gy = top / w;
        rec[4 * (size_t)i] = (float)gx;
        rec[4 * (size_t)i + 1] = (float)gy;
        rec[4 * (size_t)i + 2] = (i < n_top) ? grid[(size_t)gy * w + gx] : wall_z[i - n_top];
        rec[4 * (size_t)i + 3] = irr[i];
    }
    f.write((const char*)rec.data(), (std::streamsize)(rec.size() * 4));
    if (!f) fail(SGPU_ERR_IO, "write failed for %s", path.c_str());
}

}  // namespace sgpu
