// mkde_tiff.cpp -- GeoTIFF reader/writer and CSV point I/O for the drop-in
// command line (the reference reads and writes rasters through GDAL,
// GeoTIFF.cpp:20-577, and points through ReadSamplePoints / WriteSamplePoints,
// tiled:1623-1711, 1857-1870; GDAL is not available here).
//
// Reader: classic TIFF and BigTIFF, little/big endian, strips or tiles,
// compression none / LZW / Deflate, predictor 1 / 2 / 3, sample formats
// uint / int / float of 8-64 bits, one band; values are converted to float
// as GeoTIFFReader::GetArray2D does.  Georeferencing from ModelPixelScale +
// ModelTiepoint or ModelTransformation; nodata from GDAL_NODATA.
// Writer: BigTIFF, 256x256 tiles, float32, LZW (as the reference's
// GeoTIFFWriter: BIGTIFF=YES, COMPRESS=LZW), georeferencing, GDAL_NODATA,
// GeoKey tags copied from a template raster, band statistics.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "mkde.h"

extern "C" int mkde_internal_fail(int code, const char *msg);

static int tfail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return mkde_internal_fail(code, buf);
}

// ------------------------------------------------------------------ LZW --
// TIFF LZW (TIFF 6.0 section 13): MSB-first codes of 9-12 bits, Clear 256,
// EOI 257, first free code 258, code width grows one code early.

namespace {

const int LZW_CLEAR = 256, LZW_EOI = 257, LZW_FIRST = 258, LZW_MAXCODE = 4095;

struct BitWriter {
    std::vector<uint8_t> &out;
    uint64_t acc = 0;
    int nbits = 0;
    explicit BitWriter(std::vector<uint8_t> &o) : out(o) {}
    void put(unsigned code, int width)
    {
        acc = (acc << width) | code;
        nbits += width;
        while (nbits >= 8) {
            out.push_back((uint8_t)(acc >> (nbits - 8)));
            nbits -= 8;
        }
    }
    void flush()
    {
        if (nbits > 0) out.push_back((uint8_t)(acc << (8 - nbits)));
        nbits = 0;
        acc = 0;
    }
};

void lzw_encode(const uint8_t *in, size_t n, std::vector<uint8_t> &out)
{
    out.clear();
    BitWriter bw(out);
    std::unordered_map<uint32_t, int> table;
    table.reserve(8192);
    int width = 9, next = LZW_FIRST;
    bw.put(LZW_CLEAR, width);
    if (n == 0) {
        bw.put(LZW_EOI, width);
        bw.flush();
        return;
    }
    int w = in[0];
    for (size_t i = 1; i < n; ++i) {
        const uint8_t k = in[i];
        const uint32_t key = ((uint32_t)w << 8) | k;
        auto it = table.find(key);
        if (it != table.end()) {
            w = it->second;
            continue;
        }
        bw.put((unsigned)w, width);
        table.emplace(key, next);
        ++next;
        if (next == LZW_MAXCODE - 1) {
            bw.put(LZW_CLEAR, width);
            table.clear();
            next = LZW_FIRST;
            width = 9;
        } else if (next > (1 << width) - 1) {
            ++width;
        }
        w = k;
    }
    bw.put((unsigned)w, width);
    ++next;
    if (next == LZW_MAXCODE - 1) {
        bw.put(LZW_CLEAR, width);
        width = 9;
    } else if (next > (1 << width) - 1) {
        ++width;
    }
    bw.put(LZW_EOI, width);
    bw.flush();
}

bool lzw_decode(const uint8_t *in, size_t n, uint8_t *out, size_t cap, size_t *produced)
{
    std::vector<uint16_t> prefix(4096);
    std::vector<uint8_t> suffix(4096), first(4096);
    std::vector<uint16_t> length(4096);
    for (int c = 0; c < 256; ++c) {
        prefix[c] = 0xffff;
        suffix[c] = (uint8_t)c;
        first[c] = (uint8_t)c;
        length[c] = 1;
    }
    size_t bitpos = 0, o = 0;
    const size_t nbits_total = n * 8;
    int width = 9, next = LZW_FIRST, old = -1;
    std::vector<uint8_t> tmp(4096);
    auto get = [&](int wdt) -> int {
        if (bitpos + wdt > nbits_total) return LZW_EOI;
        unsigned v = 0;
        for (int b = 0; b < wdt; ++b) {
            const size_t p = bitpos + b;
            v = (v << 1) | ((in[p >> 3] >> (7 - (p & 7))) & 1u);
        }
        bitpos += wdt;
        return (int)v;
    };
    auto emit = [&](int code) -> bool {
        int len = length[code];
        if (o + len > cap) return false;
        int c = code;
        for (int k = len - 1; k >= 0; --k) {
            out[o + k] = suffix[c];
            c = prefix[c];
        }
        o += len;
        return true;
    };
    for (;;) {
        int code = get(width);
        if (code == LZW_EOI) break;
        if (code == LZW_CLEAR) {
            width = 9;
            next = LZW_FIRST;
            code = get(width);
            if (code == LZW_EOI) break;
            if (code > 255) return false;
            if (!emit(code)) return false;
            old = code;
            continue;
        }
        if (old < 0) return false;
        if (code < next) {
            if (!emit(code)) return false;
            if (next < 4096) {
                prefix[next] = (uint16_t)old;
                suffix[next] = first[code];
                first[next] = first[old];
                length[next] = (uint16_t)(length[old] + 1);
                ++next;
            }
        } else if (code == next) {
            if (next >= 4096) return false;
            prefix[next] = (uint16_t)old;
            suffix[next] = first[old];
            first[next] = first[old];
            length[next] = (uint16_t)(length[old] + 1);
            ++next;
            if (!emit(code)) return false;
        } else {
            return false;
        }
        old = code;
        if (next + 1 > (1 << width) - 1 && width < 12) ++width;
    }
    *produced = o;
    return true;
}

// --------------------------------------------------------------- reader --

struct Tag {
    uint16_t type = 0;
    uint64_t count = 0;
    std::vector<uint8_t> data;   // raw bytes, file byte order
};

struct Tiff {
    FILE *f = nullptr;
    bool be = false, big = false;
    std::map<uint16_t, Tag> tags;
    ~Tiff()
    {
        if (f) fclose(f);
    }
    bool read_at(uint64_t off, void *dst, size_t n)
    {
        if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
        return fread(dst, 1, n, f) == n;
    }
    uint64_t u(const uint8_t *p, int bytes) const
    {
        uint64_t v = 0;
        if (be)
            for (int i = 0; i < bytes; ++i) v = (v << 8) | p[i];
        else
            for (int i = bytes - 1; i >= 0; --i) v = (v << 8) | p[i];
        return v;
    }
    static int type_size(int t)
    {
        switch (t) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: case 13: return 4;
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;
        default: return 0;
        }
    }
    bool has(uint16_t t) const { return tags.count(t) != 0; }
    // numeric value k of a tag as double / uint64
    double num(uint16_t t, uint64_t k = 0) const
    {
        auto it = tags.find(t);
        if (it == tags.end() || k >= it->second.count) return 0.0;
        const Tag &g = it->second;
        const int s = type_size(g.type);
        const uint8_t *p = g.data.data() + k * s;
        switch (g.type) {
        case 1: case 7: return p[0];
        case 6: return (int8_t)p[0];
        case 3: return (double)u(p, 2);
        case 8: return (double)(int16_t)u(p, 2);
        case 4: case 13: return (double)u(p, 4);
        case 9: return (double)(int32_t)u(p, 4);
        case 16: case 18: return (double)u(p, 8);
        case 17: return (double)(int64_t)u(p, 8);
        case 11: {
            uint32_t b = (uint32_t)u(p, 4);
            float fv;
            memcpy(&fv, &b, 4);
            return fv;
        }
        case 12: {
            uint64_t b = u(p, 8);
            double dv;
            memcpy(&dv, &b, 8);
            return dv;
        }
        case 5: return (double)u(p, 4) / (double)std::max<uint64_t>(1, u(p + 4, 4));
        case 10: return (double)(int32_t)u(p, 4) / (double)std::max<int32_t>(1, (int32_t)u(p + 4, 4));
        default: return 0.0;
        }
    }
    uint64_t unum(uint16_t t, uint64_t k = 0) const { return (uint64_t)num(t, k); }
    std::string str(uint16_t t) const
    {
        auto it = tags.find(t);
        if (it == tags.end()) return std::string();
        std::string s(it->second.data.begin(), it->second.data.end());
        while (!s.empty() && s.back() == '\0') s.pop_back();
        return s;
    }
};

int open_tiff(const char *path, Tiff &T)
{
    T.f = fopen(path, "rb");
    if (!T.f) return tfail(MKDE_E_IO, "cannot open '%s'", path);
    uint8_t h[16];
    if (fread(h, 1, 8, T.f) != 8) return tfail(MKDE_E_IO, "'%s': short file", path);
    if (h[0] == 'I' && h[1] == 'I') T.be = false;
    else if (h[0] == 'M' && h[1] == 'M') T.be = true;
    else return tfail(MKDE_E_IO, "'%s' is not a TIFF file", path);
    const uint64_t magic = T.u(h + 2, 2);
    uint64_t ifd;
    if (magic == 42) {
        T.big = false;
        ifd = T.u(h + 4, 4);
    } else if (magic == 43) {
        T.big = true;
        if (fread(h + 8, 1, 8, T.f) != 8) return tfail(MKDE_E_IO, "'%s': short BigTIFF header", path);
        ifd = T.u(h + 8, 8);
    } else {
        return tfail(MKDE_E_IO, "'%s': bad TIFF magic %llu", path, (unsigned long long)magic);
    }
    uint8_t cb[8];
    const int csz = T.big ? 8 : 2;
    if (!T.read_at(ifd, cb, csz)) return tfail(MKDE_E_IO, "'%s': bad IFD offset", path);
    const uint64_t nent = T.u(cb, csz);
    const int esz = T.big ? 20 : 12;
    std::vector<uint8_t> ent(nent * esz);
    if (nent > 4096 || !T.read_at(ifd + csz, ent.data(), ent.size())) return tfail(MKDE_E_IO, "'%s': bad IFD", path);
    for (uint64_t k = 0; k < nent; ++k) {
        const uint8_t *e = ent.data() + k * esz;
        Tag g;
        const uint16_t tag = (uint16_t)T.u(e, 2);
        g.type = (uint16_t)T.u(e + 2, 2);
        g.count = T.big ? T.u(e + 4, 8) : T.u(e + 4, 4);
        const int ts = Tiff::type_size(g.type);
        if (ts == 0) continue;
        const uint64_t bytes = g.count * (uint64_t)ts;
        if (bytes > ((uint64_t)1 << 31)) return tfail(MKDE_E_IO, "'%s': tag %u too large", path, tag);
        g.data.resize(bytes);
        const int inl = T.big ? 8 : 4;
        const uint8_t *vp = e + (T.big ? 12 : 8);
        if (bytes <= (uint64_t)inl) {
            memcpy(g.data.data(), vp, bytes);
        } else {
            const uint64_t off = T.u(vp, inl);
            if (!T.read_at(off, g.data.data(), bytes)) return tfail(MKDE_E_IO, "'%s': bad tag %u data", path, tag);
        }
        T.tags[tag] = std::move(g);
    }
    return MKDE_OK;
}

struct Layout {
    int64_t ncols = 0, nrows = 0;
    int bits = 0, fmt = 1, spp = 1, comp = 1, pred = 1, planar = 1;
    bool tiled = false;
    int64_t bw = 0, bh = 0;   // block (tile or strip) width/height
    int64_t nbx = 0, nby = 0;
};

int layout_of(const Tiff &T, Layout &L)
{
    L.ncols = (int64_t)T.unum(256);
    L.nrows = (int64_t)T.unum(257);
    L.bits = T.has(258) ? (int)T.unum(258) : 1;
    L.fmt = T.has(339) ? (int)T.unum(339) : 1;
    L.spp = T.has(277) ? (int)T.unum(277) : 1;
    L.comp = T.has(259) ? (int)T.unum(259) : 1;
    L.pred = T.has(317) ? (int)T.unum(317) : 1;
    L.planar = T.has(284) ? (int)T.unum(284) : 1;
    if (L.ncols <= 0 || L.nrows <= 0) return tfail(MKDE_E_IO, "TIFF without dimensions");
    if (L.bits != 8 && L.bits != 16 && L.bits != 32 && L.bits != 64)
        return tfail(MKDE_E_IO, "unsupported BitsPerSample %d", L.bits);
    if (L.fmt == 3 && L.bits < 32) return tfail(MKDE_E_IO, "unsupported float sample of %d bits", L.bits);
    if (L.comp != 1 && L.comp != 5 && L.comp != 8 && L.comp != 32946)
        return tfail(MKDE_E_IO, "unsupported TIFF compression %d", L.comp);
    if (L.tiled = T.has(322), L.tiled) {
        L.bw = (int64_t)T.unum(322);
        L.bh = (int64_t)T.unum(323);
    } else {
        L.bw = L.ncols;
        L.bh = T.has(278) ? std::min<int64_t>((int64_t)T.unum(278), L.nrows) : L.nrows;
    }
    if (L.bw <= 0 || L.bh <= 0) return tfail(MKDE_E_IO, "bad TIFF block size");
    L.nbx = (L.ncols + L.bw - 1) / L.bw;
    L.nby = (L.nrows + L.bh - 1) / L.bh;
    return MKDE_OK;
}

// band-1 samples of one block, decoded to float (row-major bw x bh)
int read_block(Tiff &T, const Layout &L, int64_t bx, int64_t by, std::vector<float> &vals)
{
    const uint16_t offtag = L.tiled ? 324 : 273, cnttag = L.tiled ? 325 : 279;
    const int64_t idx = by * L.nbx + bx;
    const uint64_t off = T.unum(offtag, idx), cnt = T.unum(cnttag, idx);
    const int bps = L.bits / 8;
    const int spp_chunk = L.planar == 2 ? 1 : L.spp;
    const int64_t rows = L.tiled ? L.bh : std::min(L.bh, L.nrows - by * L.bh);
    const size_t rawsize = (size_t)L.bw * (size_t)rows * (size_t)spp_chunk * (size_t)bps;
    std::vector<uint8_t> comp(cnt), raw(rawsize, 0);
    if (cnt > 0 && !T.read_at(off, comp.data(), cnt)) return tfail(MKDE_E_IO, "TIFF block read failed");
    if (L.comp == 1) {
        memcpy(raw.data(), comp.data(), std::min<size_t>(cnt, rawsize));
    } else if (L.comp == 5) {
        size_t got = 0;
        if (!lzw_decode(comp.data(), comp.size(), raw.data(), raw.size(), &got))
            return tfail(MKDE_E_IO, "corrupt LZW block");
    } else {
        uLongf dl = (uLongf)rawsize;
        int zr = uncompress(raw.data(), &dl, comp.data(), (uLong)cnt);
        if (zr != Z_OK && zr != Z_BUF_ERROR) return tfail(MKDE_E_IO, "corrupt Deflate block (%d)", zr);
    }
    const int64_t rowbytes = L.bw * spp_chunk * bps;
    // predictors
    if (L.pred == 2) {
        for (int64_t r = 0; r < rows; ++r) {
            uint8_t *p = raw.data() + r * rowbytes;
            for (int64_t c = spp_chunk; c < L.bw * spp_chunk; ++c) {
                if (bps == 1) p[c] = (uint8_t)(p[c] + p[c - spp_chunk]);
                else {
                    uint64_t a = T.u(p + c * bps, bps), b = T.u(p + (c - spp_chunk) * bps, bps);
                    uint64_t s = a + b;
                    for (int k = 0; k < bps; ++k) {
                        int sh = T.be ? 8 * (bps - 1 - k) : 8 * k;
                        p[c * bps + k] = (uint8_t)(s >> sh);
                    }
                }
            }
        }
    } else if (L.pred == 3) {
        // floating-point predictor: bytes of each row were split into planes
        // (most significant first) and byte-differenced
        std::vector<uint8_t> tmp(rowbytes);
        const int64_t nsamp = L.bw * spp_chunk;
        for (int64_t r = 0; r < rows; ++r) {
            uint8_t *p = raw.data() + r * rowbytes;
            for (int64_t c = spp_chunk; c < rowbytes; ++c) p[c] = (uint8_t)(p[c] + p[c - spp_chunk]);
            for (int64_t s = 0; s < nsamp; ++s)
                for (int k = 0; k < bps; ++k) {
                    // plane k holds byte k of the big-endian representation
                    const uint8_t b = p[k * nsamp + s];
                    const int dst = T.be ? k : (bps - 1 - k);
                    tmp[s * bps + dst] = b;
                }
            memcpy(p, tmp.data(), rowbytes);
        }
    }
    vals.assign((size_t)L.bw * rows, 0.0f);
    for (int64_t r = 0; r < rows; ++r) {
        const uint8_t *p = raw.data() + r * rowbytes;
        for (int64_t c = 0; c < L.bw; ++c) {
            const uint8_t *s = p + c * spp_chunk * bps;
            // host byte order is little endian; predictor 3 already produced
            // the file's byte order
            uint64_t v = T.u(s, bps);
            float out;
            if (L.fmt == 3) {
                if (bps == 4) {
                    uint32_t b = (uint32_t)v;
                    memcpy(&out, &b, 4);
                } else {
                    double d;
                    memcpy(&d, &v, 8);
                    out = (float)d;
                }
            } else if (L.fmt == 2) {
                int64_t sv = bps == 1 ? (int8_t)v : bps == 2 ? (int16_t)v : bps == 4 ? (int32_t)v : (int64_t)v;
                out = (float)sv;
            } else {
                out = (float)v;
            }
            vals[(size_t)r * L.bw + c] = out;
        }
    }
    return MKDE_OK;
}

void geotransform_of(const Tiff &T, double gt[6])
{
    gt[0] = 0.0;
    gt[1] = 1.0;
    gt[2] = 0.0;
    gt[3] = 0.0;
    gt[4] = 0.0;
    gt[5] = 1.0;
    if (T.has(34264) && T.tags.at(34264).count >= 16) {
        gt[0] = T.num(34264, 3);
        gt[1] = T.num(34264, 0);
        gt[2] = T.num(34264, 1);
        gt[3] = T.num(34264, 7);
        gt[4] = T.num(34264, 4);
        gt[5] = T.num(34264, 5);
    } else if (T.has(33550) && T.has(33922)) {
        const double sx = T.num(33550, 0), sy = T.num(33550, 1);
        const double ti = T.num(33922, 0), tj = T.num(33922, 1), tx = T.num(33922, 3), ty = T.num(33922, 4);
        gt[1] = sx;
        gt[5] = -sy;
        gt[0] = tx - ti * sx;
        gt[3] = ty + tj * sy;
    }
    // GTRasterTypeGeoKey (1025) == RasterPixelIsPoint: shift to the corner
    if (T.has(34735)) {
        const uint64_t nk = T.unum(34735, 3);
        for (uint64_t k = 0; k < nk; ++k) {
            if (T.unum(34735, 4 + 4 * k) == 1025 && T.unum(34735, 5 + 4 * k) == 0 &&
                T.unum(34735, 7 + 4 * k) == 2) {
                gt[0] -= 0.5 * gt[1];
                gt[3] -= 0.5 * gt[5];
            }
        }
    }
}

}  // namespace

// ------------------------------------------------------------- C ABI ----

extern "C" int mkde_tiff_info(const char *path, int64_t *ncols, int64_t *nrows, double *geotransform, double *nodata,
                              int *has_nodata, int *block_x, int *block_y)
{
    if (!path) return tfail(MKDE_E_ARG, "path is NULL");
    try {
        Tiff T;
        int s = open_tiff(path, T);
        if (s) return s;
        Layout L;
        if ((s = layout_of(T, L))) return s;
        if (ncols) *ncols = L.ncols;
        if (nrows) *nrows = L.nrows;
        if (geotransform) geotransform_of(T, geotransform);
        const std::string nd = T.str(42113);
        if (has_nodata) *has_nodata = nd.empty() ? 0 : 1;
        if (nodata) *nodata = nd.empty() ? 0.0 : atof(nd.c_str());
        if (block_x) *block_x = (int)L.bw;
        if (block_y) *block_y = (int)L.bh;
    } catch (const std::bad_alloc &) {
        return tfail(MKDE_E_NOMEM, "out of host memory");
    }
    return MKDE_OK;
}

extern "C" int mkde_tiff_read(const char *path, int64_t xoff, int64_t yoff, int64_t ncols, int64_t nrows, float *out)
{
    if (!path || !out) return tfail(MKDE_E_ARG, "null argument");
    try {
        Tiff T;
        int s = open_tiff(path, T);
        if (s) return s;
        Layout L;
        if ((s = layout_of(T, L))) return s;
        if (xoff < 0 || yoff < 0 || ncols <= 0 || nrows <= 0 || xoff + ncols > L.ncols || yoff + nrows > L.nrows)
            return tfail(MKDE_E_ARG, "window outside the raster");
        std::vector<float> blk;
        for (int64_t by = yoff / L.bh; by <= (yoff + nrows - 1) / L.bh; ++by)
            for (int64_t bx = xoff / L.bw; bx <= (xoff + ncols - 1) / L.bw; ++bx) {
                if ((s = read_block(T, L, bx, by, blk))) return s;
                const int64_t rows_in_blk = (int64_t)(blk.size() / L.bw);
                for (int64_t r = 0; r < rows_in_blk; ++r) {
                    const int64_t gr = by * L.bh + r;
                    if (gr < yoff || gr >= yoff + nrows) continue;
                    for (int64_t c = 0; c < L.bw; ++c) {
                        const int64_t gc = bx * L.bw + c;
                        if (gc < xoff || gc >= xoff + ncols) continue;
                        out[(gr - yoff) * ncols + (gc - xoff)] = blk[(size_t)r * L.bw + c];
                    }
                }
            }
    } catch (const std::bad_alloc &) {
        return tfail(MKDE_E_NOMEM, "out of host memory");
    }
    return MKDE_OK;
}

// ------------------------------------------------------------- writer ----

struct mkde_raster_writer {
    std::string path;
    int64_t ncols = 0, nrows = 0;
    double gt[6];
    float nodata = 0.0f;
    std::vector<float> data;
    std::vector<uint8_t> geokeys, geodoubles, geoascii;   // raw tag payloads (little endian)
    uint16_t geokeys_type = 3, geodoubles_type = 12;
    uint64_t geokeys_count = 0, geodoubles_count = 0;
};

namespace {

void put16(std::vector<uint8_t> &b, uint16_t v)
{
    b.push_back((uint8_t)v);
    b.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t> &b, uint32_t v)
{
    for (int k = 0; k < 4; ++k) b.push_back((uint8_t)(v >> (8 * k)));
}
void put64(std::vector<uint8_t> &b, uint64_t v)
{
    for (int k = 0; k < 8; ++k) b.push_back((uint8_t)(v >> (8 * k)));
}

}  // namespace

extern "C" int mkde_writer_open(const char *path, int64_t ncols, int64_t nrows, const double *geotransform,
                                float nodata, const char *georef_from, mkde_raster_writer **out)
{
    if (!path || !out || ncols <= 0 || nrows <= 0) return tfail(MKDE_E_ARG, "mkde_writer_open: bad arguments");
    *out = nullptr;
    try {
        mkde_raster_writer *w = new mkde_raster_writer();
        w->path = path;
        w->ncols = ncols;
        w->nrows = nrows;
        const double dflt[6] = {0.0, 1.0, 0.0, 0.0, 0.0, -1.0};
        memcpy(w->gt, geotransform ? geotransform : dflt, sizeof w->gt);
        w->nodata = nodata;
        w->data.assign((size_t)ncols * nrows, nodata);
        if (georef_from) {
            Tiff T;
            int s = open_tiff(georef_from, T);
            if (s) {
                delete w;
                return s;
            }
            // GeoKey tags re-encoded little endian (the writer's byte order)
            auto copy_tag = [&](uint16_t tag, std::vector<uint8_t> &dst, uint16_t &type, uint64_t &count) {
                if (!T.has(tag)) return;
                const Tag &g = T.tags.at(tag);
                type = g.type;
                count = g.count;
                const int ts = Tiff::type_size(g.type);
                dst.clear();
                for (uint64_t k = 0; k < g.count; ++k) {
                    uint64_t v = T.u(g.data.data() + k * ts, ts);
                    for (int b = 0; b < ts; ++b) dst.push_back((uint8_t)(v >> (8 * b)));
                }
            };
            copy_tag(34735, w->geokeys, w->geokeys_type, w->geokeys_count);
            copy_tag(34736, w->geodoubles, w->geodoubles_type, w->geodoubles_count);
            if (T.has(34737)) {
                const Tag &g = T.tags.at(34737);
                w->geoascii.assign(g.data.begin(), g.data.end());
            }
        }
        *out = w;
    } catch (const std::bad_alloc &) {
        return tfail(MKDE_E_NOMEM, "out of host memory for a %lld x %lld raster", (long long)ncols, (long long)nrows);
    }
    return MKDE_OK;
}

// WriteGeoTIFF_NextTile (GeoTIFF.cpp:552-575): one window of the raster.
extern "C" int mkde_writer_put(mkde_raster_writer *w, int64_t xoff, int64_t yoff, int64_t ncols, int64_t nrows,
                               const float *data)
{
    if (!w || !data) return tfail(MKDE_E_ARG, "null argument");
    if (xoff < 0 || yoff < 0 || xoff + ncols > w->ncols || yoff + nrows > w->nrows)
        return tfail(MKDE_E_ARG, "window outside the raster");
    for (int64_t r = 0; r < nrows; ++r)
        memcpy(&w->data[(size_t)(yoff + r) * w->ncols + xoff], data + r * ncols, (size_t)ncols * sizeof(float));
    return MKDE_OK;
}

// Writes the BigTIFF (256x256 tiles, LZW) with statistics of the valid cells
// (the reference's ~GeoTIFFWriter computes and stores them, GeoTIFF.cpp:505-515)
// and frees the writer.
extern "C" int mkde_writer_close(mkde_raster_writer *w, int compress)
{
    if (!w) return MKDE_OK;
    int status = MKDE_OK;
    try {
        const int64_t T = 256;
        const int64_t nbx = (w->ncols + T - 1) / T, nby = (w->nrows + T - 1) / T;
        // statistics
        double mn = INFINITY, mx = -INFINITY, sum = 0.0, sum2 = 0.0;
        int64_t cnt = 0;
        for (float v : w->data) {
            if (v == w->nodata || std::isnan(v)) continue;
            mn = std::min(mn, (double)v);
            mx = std::max(mx, (double)v);
            sum += v;
            sum2 += (double)v * v;
            ++cnt;
        }
        const double mean = cnt ? sum / cnt : 0.0;
        const double sd = cnt ? std::sqrt(std::max(0.0, sum2 / cnt - mean * mean)) : 0.0;
        char meta[1024];
        snprintf(meta, sizeof meta,
                 "<GDALMetadata>\n  <Item name=\"STATISTICS_MAXIMUM\" sample=\"0\">%.17g</Item>\n"
                 "  <Item name=\"STATISTICS_MEAN\" sample=\"0\">%.17g</Item>\n"
                 "  <Item name=\"STATISTICS_MINIMUM\" sample=\"0\">%.17g</Item>\n"
                 "  <Item name=\"STATISTICS_STDDEV\" sample=\"0\">%.17g</Item>\n</GDALMetadata>\n",
                 cnt ? mx : 0.0, mean, cnt ? mn : 0.0, sd);
        char nd[64];
        snprintf(nd, sizeof nd, "%.9g", (double)w->nodata);

        FILE *f = fopen(w->path.c_str(), "wb");
        if (!f) {
            status = tfail(MKDE_E_IO, "cannot create '%s'", w->path.c_str());
        } else {
            // header (16 bytes), then tile data, then the tag payloads and the IFD
            std::vector<uint8_t> hdr;
            hdr.push_back('I');
            hdr.push_back('I');
            put16(hdr, 43);
            put16(hdr, 8);
            put16(hdr, 0);
            put64(hdr, 0);   // IFD offset, patched at the end
            fwrite(hdr.data(), 1, hdr.size(), f);
            uint64_t pos = hdr.size();
            std::vector<uint64_t> offs, cnts;
            std::vector<uint8_t> raw((size_t)T * T * 4), enc;
            std::vector<float> tile((size_t)T * T);
            for (int64_t by = 0; by < nby; ++by)
                for (int64_t bx = 0; bx < nbx; ++bx) {
                    for (int64_t r = 0; r < T; ++r)
                        for (int64_t c = 0; c < T; ++c) {
                            const int64_t gr = by * T + r, gc = bx * T + c;
                            tile[r * T + c] =
                                (gr < w->nrows && gc < w->ncols) ? w->data[(size_t)gr * w->ncols + gc] : w->nodata;
                        }
                    memcpy(raw.data(), tile.data(), raw.size());
                    const uint8_t *src = raw.data();
                    size_t len = raw.size();
                    if (compress) {
                        lzw_encode(raw.data(), raw.size(), enc);
                        src = enc.data();
                        len = enc.size();
                    }
                    offs.push_back(pos);
                    cnts.push_back(len);
                    fwrite(src, 1, len, f);
                    pos += len;
                }
            // tag payloads
            struct E {
                uint16_t tag, type;
                uint64_t count;
                std::vector<uint8_t> val;
            };
            std::vector<E> es;
            auto add = [&](uint16_t tag, uint16_t type, uint64_t count, std::vector<uint8_t> v) {
                es.push_back(E{tag, type, count, std::move(v)});
            };
            auto u16 = [](uint16_t v) {
                std::vector<uint8_t> b;
                put16(b, v);
                return b;
            };
            auto u32 = [](uint32_t v) {
                std::vector<uint8_t> b;
                put32(b, v);
                return b;
            };
            auto dbl = [](std::initializer_list<double> vs) {
                std::vector<uint8_t> b;
                for (double d : vs) {
                    uint64_t u;
                    memcpy(&u, &d, 8);
                    put64(b, u);
                }
                return b;
            };
            auto ascii = [](const char *s) {
                std::vector<uint8_t> b(s, s + strlen(s) + 1);
                return b;
            };
            add(256, 4, 1, u32((uint32_t)w->ncols));
            add(257, 4, 1, u32((uint32_t)w->nrows));
            add(258, 3, 1, u16(32));
            add(259, 3, 1, u16(compress ? 5 : 1));
            add(262, 3, 1, u16(1));
            add(277, 3, 1, u16(1));
            add(284, 3, 1, u16(1));
            add(322, 3, 1, u16((uint16_t)T));
            add(323, 3, 1, u16((uint16_t)T));
            {
                std::vector<uint8_t> b;
                for (uint64_t o : offs) put64(b, o);
                add(324, 16, offs.size(), b);
            }
            {
                std::vector<uint8_t> b;
                for (uint64_t o : cnts) put64(b, o);
                add(325, 16, cnts.size(), b);
            }
            add(339, 3, 1, u16(3));
            if (w->gt[2] == 0.0 && w->gt[4] == 0.0) {
                add(33550, 12, 3, dbl({w->gt[1], -w->gt[5], 0.0}));
                add(33922, 12, 6, dbl({0.0, 0.0, 0.0, w->gt[0], w->gt[3], 0.0}));
            } else {
                add(34264, 12, 16,
                    dbl({w->gt[1], w->gt[2], 0, w->gt[0], w->gt[4], w->gt[5], 0, w->gt[3], 0, 0, 0, 0, 0, 0, 0, 1}));
            }
            if (!w->geokeys.empty()) add(34735, w->geokeys_type, w->geokeys_count, w->geokeys);
            if (!w->geodoubles.empty()) add(34736, w->geodoubles_type, w->geodoubles_count, w->geodoubles);
            if (!w->geoascii.empty()) add(34737, 2, w->geoascii.size(), w->geoascii);
            add(42112, 2, strlen(meta) + 1, ascii(meta));
            add(42113, 2, strlen(nd) + 1, ascii(nd));
            std::sort(es.begin(), es.end(), [](const E &a, const E &b) { return a.tag < b.tag; });
            // out-of-line payloads
            std::vector<uint64_t> valoff(es.size(), 0);
            for (size_t k = 0; k < es.size(); ++k) {
                if (es[k].val.size() > 8) {
                    if (pos & 1) {
                        fputc(0, f);
                        ++pos;
                    }
                    valoff[k] = pos;
                    fwrite(es[k].val.data(), 1, es[k].val.size(), f);
                    pos += es[k].val.size();
                }
            }
            if (pos & 1) {
                fputc(0, f);
                ++pos;
            }
            const uint64_t ifd = pos;
            std::vector<uint8_t> d;
            put64(d, es.size());
            for (size_t k = 0; k < es.size(); ++k) {
                put16(d, es[k].tag);
                put16(d, es[k].type);
                put64(d, es[k].count);
                if (es[k].val.size() <= 8) {
                    std::vector<uint8_t> v = es[k].val;
                    v.resize(8, 0);
                    d.insert(d.end(), v.begin(), v.end());
                } else {
                    put64(d, valoff[k]);
                }
            }
            put64(d, 0);
            fwrite(d.data(), 1, d.size(), f);
            fseeko(f, 8, SEEK_SET);
            std::vector<uint8_t> io;
            put64(io, ifd);
            fwrite(io.data(), 1, 8, f);
            if (fclose(f) != 0) status = tfail(MKDE_E_IO, "write error on '%s'", w->path.c_str());
        }
    } catch (const std::bad_alloc &) {
        status = tfail(MKDE_E_NOMEM, "out of host memory");
    }
    delete w;
    return status;
}

// ------------------------------------------------------------------ CSV --
// ReadSamplePoints (tiled:1623-1711): a header line, then "x,y,[h,]w" per
// line (tokens split on ',' and newline, parsed with atof); weights are used
// only with sample weights on, else 1.  with_h != 0 requires the h column.
// Returns the number of points in *n; pass x == NULL to count only.
extern "C" int mkde_csv_read_points(const char *path, int with_h, int sample_weights, int64_t *n, float *x, float *y,
                                    float *h, float *w)
{
    if (!path || !n) return tfail(MKDE_E_ARG, "null argument");
    FILE *f = fopen(path, "rt");
    if (!f) return tfail(MKDE_E_IO, "Error opening file '%s'", path);
    char line[4096];
    if (!fgets(line, sizeof line, f)) {
        fclose(f);
        return tfail(MKDE_E_IO, "No point in file '%s'", path);
    }
    int64_t k = 0;
    while (fgets(line, sizeof line, f)) {
        char *save = nullptr;
        char *tok = strtok_r(line, ",\n\r", &save);
        if (!tok) continue;
        const float xv = (float)atof(tok);
        tok = strtok_r(nullptr, ",\n\r", &save);
        const float yv = tok ? (float)atof(tok) : 0.0f;
        float hv = 0.0f, wv = 1.0f;
        if (with_h) {
            tok = strtok_r(nullptr, ",\n\r", &save);
            hv = tok ? (float)atof(tok) : 0.0f;
            tok = strtok_r(nullptr, ",\n\r", &save);
            if (!tok) {
                fclose(f);
                return tfail(MKDE_E_IO, "***Error - No bandwidth provided in the file (line %lld)", (long long)k + 2);
            }
            wv = (float)atof(tok);
        } else {
            tok = strtok_r(nullptr, ",\n\r", &save);
            wv = tok ? (float)atof(tok) : 1.0f;
        }
        if (x) {
            if (k >= *n) {
                fclose(f);
                return tfail(MKDE_E_RANGE, "more points than the buffer holds");
            }
            x[k] = xv;
            y[k] = yv;
            if (h) h[k] = hv;
            if (w) w[k] = sample_weights ? wv : 1.0f;
        }
        ++k;
    }
    fclose(f);
    if (k == 0) return tfail(MKDE_E_IO, "No point in file '%s'", path);
    *n = k;
    return MKDE_OK;
}

// WriteSamplePoints (tiled:1857-1870): "x, y, h, w" with %f fields.
extern "C" int mkde_csv_write_points(const char *path, int64_t n, const float *x, const float *y, const float *h,
                                     const float *w)
{
    FILE *f = fopen(path, "w");
    if (!f) return tfail(MKDE_E_IO, "Error opening file '%s'", path);
    fprintf(f, "x, y, h, w\n");
    for (int64_t p = 0; p < n; ++p) fprintf(f, "%f, %f, %f, %f\n", x[p], y[p], h[p], w[p]);
    if (fclose(f) != 0) return tfail(MKDE_E_IO, "write error on '%s'", path);
    return MKDE_OK;
}

// TIFF LZW codec entry points (tests).
extern "C" int64_t mkde_lzw_encode(const uint8_t *in, int64_t n, uint8_t *out, int64_t cap)
{
    std::vector<uint8_t> enc;
    lzw_encode(in, (size_t)n, enc);
    if ((int64_t)enc.size() > cap) return -1;
    memcpy(out, enc.data(), enc.size());
    return (int64_t)enc.size();
}

extern "C" int64_t mkde_lzw_decode(const uint8_t *in, int64_t n, uint8_t *out, int64_t cap)
{
    size_t got = 0;
    if (!lzw_decode(in, (size_t)n, out, (size_t)cap, &got)) return -1;
    return (int64_t)got;
}
