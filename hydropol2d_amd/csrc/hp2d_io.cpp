// hp2d_io.cpp - the on-disk formats either side of the step loop (host code).
//
// GeoTIFF ingest: the rasters HydroPol2D_preprocessing.m:192-194 reads with
// GRIDobj (DEM, LULC, SOIL and the optional maps), one band, strips or tiles,
// uncompressed, PackBits, LZW or Deflate, horizontal or floating-point
// predictor, 8-64-bit integers or IEEE floats, either byte order; the
// georeference from ModelPixelScale / ModelTiepoint, the EPSG code from the
// GeoKey directory, GDAL_NODATA cells as NaN.  Z comes back as MATLAB holds
// it: ny x nx column-major, row 1 the northern image row.
//
// MAT-file writer: MATLAB Level 5 MAT-files (the format `load` reads), for the
// temporary outputs of save_output_maps.m:575-580 and 620-830
// (save_map_hydro_N.mat with Maps.Hydro / Maps.WQ_States, and
// current_time_series_outputs.mat with TempTS): variables named by dotted
// paths are written as nested 1 x 1 structs.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "hp2d.h"

namespace hp2d {
void set_global_error(const std::string &m);
}

namespace {

// ---------------------------------------------------------------------------
// TIFF
// ---------------------------------------------------------------------------
struct Tiff {
    std::vector<uint8_t> buf;
    bool be = false;  // big-endian ("MM")
    uint32_t width = 0, height = 0, bits = 0, spp = 1, fmt = 1, comp = 1, pred = 1, planar = 1;
    uint32_t rows_per_strip = 0, tile_w = 0, tile_h = 0;
    std::vector<uint64_t> offsets, counts;
    double scale[3] = {0, 0, 0}, tie[6] = {0, 0, 0, 0, 0, 0};
    double nodata = NAN;
    int epsg = 0;
    std::string err;

    uint16_t u16(size_t o) const {
        return be ? (uint16_t)(buf[o] << 8 | buf[o + 1]) : (uint16_t)(buf[o] | buf[o + 1] << 8);
    }
    uint32_t u32(size_t o) const {
        return be ? ((uint32_t)buf[o] << 24 | (uint32_t)buf[o + 1] << 16 | (uint32_t)buf[o + 2] << 8 | buf[o + 3])
                  : ((uint32_t)buf[o] | (uint32_t)buf[o + 1] << 8 | (uint32_t)buf[o + 2] << 16 |
                     (uint32_t)buf[o + 3] << 24);
    }
    double f64(size_t o) const {
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v |= (uint64_t)buf[o + (be ? 7 - i : i)] << (8 * i);
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
};

// values of one IFD entry as doubles (integers and doubles; ASCII separately)
bool entry_values(const Tiff &t, size_t e, std::vector<double> &out, std::string *ascii) {
    const uint16_t type = t.u16(e + 2);
    const uint32_t count = t.u32(e + 4);
    static const int size_of[] = {0, 1, 1, 2, 4, 8, 1, 1, 2, 4, 8, 4, 8};
    if (type < 1 || type > 12) return false;
    const size_t bytes = (size_t)size_of[type] * count;
    const size_t at = bytes <= 4 ? e + 8 : t.u32(e + 8);
    if (at + bytes > t.buf.size()) return false;
    out.clear();
    if (type == 2) {  // ASCII
        if (ascii) ascii->assign((const char *)&t.buf[at], strnlen((const char *)&t.buf[at], bytes));
        return true;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const size_t o = at + (size_t)i * size_of[type];
        switch (type) {
            case 1: case 7: out.push_back(t.buf[o]); break;
            case 3: out.push_back(t.u16(o)); break;
            case 4: out.push_back(t.u32(o)); break;
            case 6: out.push_back((int8_t)t.buf[o]); break;
            case 8: out.push_back((int16_t)t.u16(o)); break;
            case 9: out.push_back((int32_t)t.u32(o)); break;
            case 5: out.push_back((double)t.u32(o) / (double)t.u32(o + 4)); break;
            case 10: out.push_back((double)(int32_t)t.u32(o) / (double)(int32_t)t.u32(o + 4)); break;
            case 11: { float f; uint32_t v = t.u32(o); std::memcpy(&f, &v, 4); out.push_back(f); } break;
            case 12: out.push_back(t.f64(o)); break;
            default: return false;
        }
    }
    return true;
}

bool tiff_parse(const char *path, Tiff &t) {
    FILE *f = std::fopen(path, "rb");
    if (!f) { t.err = std::string("cannot open ") + path; return false; }
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    t.buf.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(t.buf.data(), 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if (got != t.buf.size() || t.buf.size() < 8) { t.err = "short file"; return false; }
    if (t.buf[0] == 'I' && t.buf[1] == 'I') t.be = false;
    else if (t.buf[0] == 'M' && t.buf[1] == 'M') t.be = true;
    else { t.err = "not a TIFF file"; return false; }
    if (t.u16(2) != 42) { t.err = "BigTIFF and other variants are not read"; return false; }
    const size_t ifd = t.u32(4);
    if (ifd + 2 > t.buf.size()) { t.err = "bad IFD offset"; return false; }
    const uint16_t n = t.u16(ifd);
    std::vector<double> v;
    std::string ascii;
    std::vector<double> geokeys;
    for (uint16_t k = 0; k < n; ++k) {
        const size_t e = ifd + 2 + 12 * (size_t)k;
        if (e + 12 > t.buf.size()) { t.err = "truncated IFD"; return false; }
        const uint16_t tag = t.u16(e);
        ascii.clear();
        if (!entry_values(t, e, v, &ascii)) continue;
        switch (tag) {
            case 256: t.width = (uint32_t)v[0]; break;
            case 257: t.height = (uint32_t)v[0]; break;
            case 258: t.bits = (uint32_t)v[0]; break;
            case 259: t.comp = (uint32_t)v[0]; break;
            case 273: case 324: t.offsets.assign(v.begin(), v.end()); break;
            case 277: t.spp = (uint32_t)v[0]; break;
            case 278: t.rows_per_strip = (uint32_t)v[0]; break;
            case 279: case 325: t.counts.assign(v.begin(), v.end()); break;
            case 284: t.planar = (uint32_t)v[0]; break;
            case 317: t.pred = (uint32_t)v[0]; break;
            case 322: t.tile_w = (uint32_t)v[0]; break;
            case 323: t.tile_h = (uint32_t)v[0]; break;
            case 339: t.fmt = (uint32_t)v[0]; break;
            case 33550: for (size_t i = 0; i < 3 && i < v.size(); ++i) t.scale[i] = v[i]; break;
            case 33922: for (size_t i = 0; i < 6 && i < v.size(); ++i) t.tie[i] = v[i]; break;
            case 34735: geokeys = v; break;
            case 42113: t.nodata = ascii.empty() ? NAN : std::strtod(ascii.c_str(), nullptr); break;
            default: break;
        }
    }
    // GeoKey directory: header of 4 shorts, then (key, location, count, value)
    for (size_t i = 4; i + 3 < geokeys.size(); i += 4) {
        const int key = (int)geokeys[i], loc = (int)geokeys[i + 1];
        if (loc == 0 && (key == 3072 || (key == 2048 && t.epsg == 0))) t.epsg = (int)geokeys[i + 3];
    }
    if (!t.width || !t.height || !t.bits) { t.err = "missing image dimensions"; return false; }
    if (t.spp != 1) { t.err = "only single-band rasters are read"; return false; }
    if (t.offsets.empty() || t.offsets.size() != t.counts.size()) { t.err = "missing strip/tile offsets"; return false; }
    if (t.comp != 1 && t.comp != 5 && t.comp != 8 && t.comp != 32946 && t.comp != 32773) {
        t.err = "compression " + std::to_string(t.comp) + " is not read (none, LZW, Deflate, PackBits)";
        return false;
    }
    if (t.pred != 1 && t.pred != 2 && t.pred != 3) { t.err = "unknown predictor"; return false; }
    if (!(t.bits == 8 || t.bits == 16 || t.bits == 32 || t.bits == 64)) { t.err = "bit depth not read"; return false; }
    if (t.fmt == 3 && t.bits < 32) { t.err = "half floats are not read"; return false; }
    if (!t.rows_per_strip) t.rows_per_strip = t.height;
    return true;
}

bool unpack_packbits(const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t want) {
    out.clear();
    size_t i = 0;
    while (i < n && out.size() < want) {
        const int8_t h = (int8_t)in[i++];
        if (h >= 0) {
            const size_t len = (size_t)h + 1;
            if (i + len > n) return false;
            out.insert(out.end(), in + i, in + i + len);
            i += len;
        } else if (h != -128) {
            if (i >= n) return false;
            out.insert(out.end(), (size_t)(1 - h), in[i++]);
        }
    }
    return out.size() >= want;
}

// TIFF LZW (MSB-first codes, 9-12 bits, ClearCode 256, EOI 257, early change)
bool unpack_lzw(const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t want) {
    out.clear();
    std::vector<std::vector<uint8_t>> dict;
    auto reset = [&]() {
        dict.assign(258, {});
        for (int c = 0; c < 256; ++c) dict[c] = {(uint8_t)c};
    };
    reset();
    uint64_t bitbuf = 0;
    int nbits = 0, width = 9;
    size_t pos = 0;
    int prev = -1;
    while (out.size() < want) {
        while (nbits < width && pos < n) {
            bitbuf = (bitbuf << 8) | in[pos++];
            nbits += 8;
        }
        if (nbits < width) break;
        const int code = (int)((bitbuf >> (nbits - width)) & ((1u << width) - 1));
        nbits -= width;
        if (code == 256) {
            reset();
            width = 9;
            prev = -1;
            continue;
        }
        if (code == 257) break;
        std::vector<uint8_t> entry;
        if (code < (int)dict.size()) {
            entry = dict[code];
            if (prev >= 0) {
                std::vector<uint8_t> add = dict[prev];
                add.push_back(entry[0]);
                dict.push_back(std::move(add));
            }
        } else if (prev >= 0 && code == (int)dict.size()) {
            entry = dict[prev];
            entry.push_back(dict[prev][0]);
            dict.push_back(entry);
        } else {
            return false;
        }
        out.insert(out.end(), entry.begin(), entry.end());
        prev = code;
        const size_t next = dict.size() + 1;  // early change
        if (next >= 4096) width = 12;
        else if (next >= 2048) width = 12;
        else if (next >= 1024) width = 11;
        else if (next >= 512) width = 10;
    }
    return out.size() >= want;
}

bool inflate_all(const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t want) {
    out.assign(want, 0);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef *>(in);
    zs.avail_in = (uInt)n;
    zs.next_out = out.data();
    zs.avail_out = (uInt)want;
    const int r = inflate(&zs, Z_FINISH);
    const bool ok = (r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR) && zs.total_out == want;
    inflateEnd(&zs);
    return ok;
}

// sample i of a decoded block, by format and width, in the file's byte order
double sample(const Tiff &t, const uint8_t *p) {
    const int nb = (int)t.bits / 8;
    uint64_t v = 0;
    for (int i = 0; i < nb; ++i) v |= (uint64_t)p[t.be ? nb - 1 - i : i] << (8 * i);
    if (t.fmt == 3) {
        if (nb == 4) { float f; uint32_t u = (uint32_t)v; std::memcpy(&f, &u, 4); return f; }
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
    if (t.fmt == 2) {  // two's complement
        const int sh = 64 - 8 * nb;
        return (double)((int64_t)(v << sh) >> sh);
    }
    return (double)v;
}

// undo the predictor on one decoded block of w columns x h rows
void unpredict(const Tiff &t, std::vector<uint8_t> &blk, uint32_t w, uint32_t h) {
    const int nb = (int)t.bits / 8;
    if (t.pred == 2) {  // horizontal differencing of samples
        for (uint32_t r = 0; r < h; ++r) {
            uint8_t *row = blk.data() + (size_t)r * w * nb;
            for (uint32_t c = 1; c < w; ++c) {
                uint64_t a = 0, b = 0;
                for (int i = 0; i < nb; ++i) {
                    a |= (uint64_t)row[(size_t)(c - 1) * nb + (t.be ? nb - 1 - i : i)] << (8 * i);
                    b |= (uint64_t)row[(size_t)c * nb + (t.be ? nb - 1 - i : i)] << (8 * i);
                }
                const uint64_t s = a + b;
                for (int i = 0; i < nb; ++i) row[(size_t)c * nb + (t.be ? nb - 1 - i : i)] = (uint8_t)(s >> (8 * i));
            }
        }
    } else if (t.pred == 3) {  // floating point: byte differencing, then the bytes regrouped
        std::vector<uint8_t> tmp((size_t)w * nb);
        for (uint32_t r = 0; r < h; ++r) {
            uint8_t *row = blk.data() + (size_t)r * w * nb;
            for (size_t i = 1; i < (size_t)w * nb; ++i) row[i] = (uint8_t)(row[i] + row[i - 1]);
            for (uint32_t c = 0; c < w; ++c)
                for (int b = 0; b < nb; ++b)  // byte b of the sample, most significant first
                    tmp[(size_t)c * nb + (t.be ? b : nb - 1 - b)] = row[(size_t)b * w + c];
            std::memcpy(row, tmp.data(), tmp.size());
        }
    }
}

bool decode_block(const Tiff &t, size_t k, uint32_t w, uint32_t h, std::vector<uint8_t> &blk) {
    const size_t want = (size_t)w * h * (t.bits / 8);
    const uint64_t off = t.offsets[k], cnt = t.counts[k];
    if (off + cnt > t.buf.size()) return false;
    const uint8_t *in = t.buf.data() + off;
    bool ok;
    switch (t.comp) {
        case 1:
            if (cnt < want) return false;
            blk.assign(in, in + want);
            ok = true;
            break;
        case 32773: ok = unpack_packbits(in, cnt, blk, want); break;
        case 5: ok = unpack_lzw(in, cnt, blk, want); break;
        default: ok = inflate_all(in, cnt, blk, want); break;
    }
    if (!ok) return false;
    blk.resize(want);
    unpredict(t, blk, w, h);
    return true;
}

bool tiff_read(const Tiff &t, double *Z) {
    const uint32_t W = t.width, H = t.height;
    const int nb = (int)t.bits / 8;
    std::vector<uint8_t> blk;
    auto put = [&](uint32_t row, uint32_t col, const uint8_t *p) {
        double v = sample(t, p);
        if (!std::isnan(t.nodata) && v == t.nodata) v = NAN;
        Z[(size_t)row + (size_t)col * H] = v;  // MATLAB column-major, row 1 = first image row
    };
    if (t.tile_w && t.tile_h) {
        const uint32_t tx = (W + t.tile_w - 1) / t.tile_w, ty = (H + t.tile_h - 1) / t.tile_h;
        if (t.offsets.size() < (size_t)tx * ty) return false;
        for (uint32_t by = 0; by < ty; ++by)
            for (uint32_t bx = 0; bx < tx; ++bx) {
                if (!decode_block(t, (size_t)by * tx + bx, t.tile_w, t.tile_h, blk)) return false;
                for (uint32_t r = 0; r < t.tile_h && by * t.tile_h + r < H; ++r)
                    for (uint32_t c = 0; c < t.tile_w && bx * t.tile_w + c < W; ++c)
                        put(by * t.tile_h + r, bx * t.tile_w + c,
                            blk.data() + ((size_t)r * t.tile_w + c) * nb);
            }
    } else {
        const uint32_t rps = t.rows_per_strip;
        const size_t ns = (H + rps - 1) / rps;
        if (t.offsets.size() < ns) return false;
        for (size_t s = 0; s < ns; ++s) {
            const uint32_t rows = std::min<uint32_t>(rps, H - (uint32_t)s * rps);
            if (!decode_block(t, s, W, rows, blk)) return false;
            for (uint32_t r = 0; r < rows; ++r)
                for (uint32_t c = 0; c < W; ++c)
                    put((uint32_t)s * rps + r, c, blk.data() + ((size_t)r * W + c) * nb);
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// MAT-file Level 5
// ---------------------------------------------------------------------------
enum : uint32_t { miINT8 = 1, miUINT8 = 2, miINT32 = 5, miUINT32 = 6, miSINGLE = 7, miDOUBLE = 9,
                  miUINT16 = 4, miMATRIX = 14 };
enum : uint32_t { mxSTRUCT = 2, mxCHAR = 4, mxDOUBLE = 6, mxSINGLE = 7, mxINT32 = 12, mxUINT8 = 9 };

struct Node {
    const hp2d_mat_var *leaf = nullptr;
    std::vector<std::pair<std::string, std::unique_ptr<Node>>> kids;  // insertion order
    Node *child(const std::string &name) {
        for (auto &k : kids)
            if (k.first == name) return k.second.get();
        kids.emplace_back(name, std::make_unique<Node>());
        return kids.back().second.get();
    }
};

void put_u32(std::vector<uint8_t> &o, uint32_t v) {
    for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void pad8(std::vector<uint8_t> &o) {
    while (o.size() % 8) o.push_back(0);
}
void element(std::vector<uint8_t> &o, uint32_t type, const void *data, size_t bytes) {
    put_u32(o, type);
    put_u32(o, (uint32_t)bytes);
    const uint8_t *p = (const uint8_t *)data;
    o.insert(o.end(), p, p + bytes);
    pad8(o);
}

bool matrix(std::vector<uint8_t> &o, const std::string &name, const Node &n, std::string &err) {
    std::vector<uint8_t> body;
    uint32_t mclass, flags_dims[2];
    std::vector<int32_t> dims;
    if (n.leaf) {
        const hp2d_mat_var &v = *n.leaf;
        switch (v.mclass) {
            case HP2D_MAT_DOUBLE: mclass = mxDOUBLE; break;
            case HP2D_MAT_SINGLE: mclass = mxSINGLE; break;
            case HP2D_MAT_INT32: mclass = mxINT32; break;
            case HP2D_MAT_UINT8: mclass = mxUINT8; break;
            case HP2D_MAT_CHAR: mclass = mxCHAR; break;
            default: err = "unknown class of " + std::string(v.name); return false;
        }
        for (int d = 0; d < v.ndims; ++d) dims.push_back((int32_t)v.dims[d]);
        if (dims.size() == 1) dims.push_back(1);
        if (dims.empty()) dims = {1, 1};
    } else {
        mclass = mxSTRUCT;
        dims = {1, 1};
    }
    flags_dims[0] = mclass;
    flags_dims[1] = 0;
    element(body, miUINT32, flags_dims, 8);
    element(body, miINT32, dims.data(), dims.size() * 4);
    element(body, miINT8, name.data(), name.size());
    if (n.leaf) {
        const hp2d_mat_var &v = *n.leaf;
        size_t count = 1;
        for (int32_t d : dims) count *= (size_t)d;
        switch (v.mclass) {
            case HP2D_MAT_DOUBLE: element(body, miDOUBLE, v.data, count * 8); break;
            case HP2D_MAT_SINGLE: element(body, miSINGLE, v.data, count * 4); break;
            case HP2D_MAT_INT32: element(body, miINT32, v.data, count * 4); break;
            case HP2D_MAT_UINT8: element(body, miUINT8, v.data, count); break;
            case HP2D_MAT_CHAR: {  // ASCII bytes as UTF-16 code units
                std::vector<uint16_t> u(count);
                for (size_t i = 0; i < count; ++i) u[i] = ((const uint8_t *)v.data)[i];
                element(body, miUINT16, u.data(), count * 2);
            } break;
        }
    } else {
        // field names: a fixed length, then the names NUL-padded to it
        int32_t len = 1;
        for (auto &k : n.kids) len = std::max<int32_t>(len, (int32_t)k.first.size() + 1);
        if (len > 64) { err = "a field name passes 63 characters"; return false; }
        put_u32(body, (4u << 16) | miINT32);  // small element: 4 bytes of data
        put_u32(body, (uint32_t)len);
        std::vector<char> names((size_t)len * n.kids.size(), 0);
        for (size_t i = 0; i < n.kids.size(); ++i)
            std::memcpy(&names[i * len], n.kids[i].first.data(), n.kids[i].first.size());
        element(body, miINT8, names.data(), names.size());
        for (auto &k : n.kids)
            if (!matrix(body, "", *k.second, err)) return false;
    }
    if (body.size() > 0x7fffffffu) { err = "a variable passes the 2 GB of a Level 5 MAT-file element"; return false; }
    put_u32(o, miMATRIX);
    put_u32(o, (uint32_t)body.size());
    o.insert(o.end(), body.begin(), body.end());
    return true;
}

}  // namespace

extern "C" {

hp2d_status hp2d_geotiff_info(const char *path, int32_t *ny, int32_t *nx, double *geo6) {
    if (!path || !ny || !nx) return HP2D_E_ARG;
    Tiff t;
    if (!tiff_parse(path, t)) {
        hp2d::set_global_error("hp2d_geotiff_info: " + t.err);
        return HP2D_E_ARG;
    }
    *ny = (int32_t)t.height;
    *nx = (int32_t)t.width;
    if (geo6) {
        // upper-left corner of the upper-left pixel (tie point at raster (I, J))
        geo6[0] = t.tie[3] - t.tie[0] * t.scale[0];
        geo6[1] = t.tie[4] + t.tie[1] * t.scale[1];
        geo6[2] = t.scale[0];
        geo6[3] = t.scale[1];
        geo6[4] = t.nodata;
        geo6[5] = t.epsg;
    }
    return HP2D_OK;
}

hp2d_status hp2d_geotiff_read(const char *path, double *Z) {
    if (!path || !Z) return HP2D_E_ARG;
    Tiff t;
    if (!tiff_parse(path, t)) {
        hp2d::set_global_error("hp2d_geotiff_read: " + t.err);
        return HP2D_E_ARG;
    }
    if (!tiff_read(t, Z)) {
        hp2d::set_global_error(std::string("hp2d_geotiff_read: corrupt image data in ") + path);
        return HP2D_E_ARG;
    }
    return HP2D_OK;
}

hp2d_status hp2d_mat_write(const char *path, int32_t n_vars, const hp2d_mat_var *vars) {
    if (!path || n_vars < 1 || !vars) return HP2D_E_ARG;
    Node root;
    for (int32_t i = 0; i < n_vars; ++i) {
        if (!vars[i].name || !vars[i].data || vars[i].ndims < 0 || vars[i].ndims > 8 ||
            (vars[i].ndims > 0 && !vars[i].dims)) {
            hp2d::set_global_error("hp2d_mat_write: bad variable " + std::to_string(i));
            return HP2D_E_ARG;
        }
        Node *n = &root;
        std::string p = vars[i].name;
        size_t a = 0;
        while (true) {
            const size_t b = p.find('.', a);
            const std::string part = p.substr(a, b == std::string::npos ? std::string::npos : b - a);
            if (part.empty()) {
                hp2d::set_global_error("hp2d_mat_write: bad name " + p);
                return HP2D_E_ARG;
            }
            n = n->child(part);
            if (b == std::string::npos) break;
            a = b + 1;
        }
        if (n->leaf || !n->kids.empty()) {
            hp2d::set_global_error("hp2d_mat_write: " + p + " is given twice or is also a struct");
            return HP2D_E_ARG;
        }
        n->leaf = &vars[i];
    }
    std::vector<uint8_t> out(128, ' ');
    const char *text = "MATLAB 5.0 MAT-file, written by hp2d-gfx950";
    std::memcpy(out.data(), text, std::strlen(text));
    std::memset(out.data() + 116, 0, 8);  // no subsystem data
    out[124] = 0x00;
    out[125] = 0x01;  // version 0x0100
    out[126] = 'I';
    out[127] = 'M';   // little-endian
    std::string err;
    for (auto &k : root.kids)
        if (!matrix(out, k.first, *k.second, err)) {
            hp2d::set_global_error("hp2d_mat_write: " + err);
            return HP2D_E_ARG;
        }
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        hp2d::set_global_error(std::string("hp2d_mat_write: cannot write ") + tmp);
        return HP2D_E_ARG;
    }
    const size_t w = std::fwrite(out.data(), 1, out.size(), f);
    const int c = std::fclose(f);
    if (w != out.size() || c != 0 || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        hp2d::set_global_error(std::string("hp2d_mat_write: write failed for ") + path);
        return HP2D_E_ARG;
    }
    return HP2D_OK;
}

}  // extern "C"
