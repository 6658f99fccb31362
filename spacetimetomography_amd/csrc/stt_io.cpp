// stt_io.cpp — the reference's on-disk formats, host only.
//
// The reference moves all of its data through CImg built with libtiff:
//   - the projection stack: CImg(filename) -> load_tiff, split on z into
//     frames (ST_Tomography.cpp:355, CImg.h:31391-31420);
//   - initial volumes "Input/rose.-%03d-frame-9iter.tiff" (VolumeData.cpp:119);
//   - reconstructions and warped frames: save_tiff (ST_Tomography.cpp:464, 805);
//   - flows: save_analyze (ST_Tomography.cpp:467);
//   - start angles: "lds.txt" read with operator>> (ST_Tomography.cpp:258-267).
// libtiff is not linked here.  The TIFF codec below reads baseline TIFF 6.0
// and BigTIFF files in either byte order (strips or tiles, chunky or planar,
// no compression, PackBits or LZW, predictor 1 or 2) and converts samples
// the way CImg::_load_tiff does; it writes float pages with the tag set
// CImg::_save_tiff hands to libtiff.  Analyze 7.5 / NIfTI-1 follow
// CImg::save_analyze and CImg::_load_analyze byte for byte.
#include <fcntl.h>
#include <strings.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/stt.h"
#include "stt_host.hpp"

namespace {

using stt::set_error;

// ---------------------------------------------------------------- files --

struct Fd {
    int fd = -1;
    Fd() = default;
    Fd(const Fd&) = delete;
    Fd& operator=(const Fd&) = delete;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

bool read_at(int fd, uint64_t off, void* dst, size_t n) {
    char* p = static_cast<char*>(dst);
    while (n) {
        ssize_t r = ::pread(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

bool write_all(int fd, const void* src, size_t n) {
    const char* p = static_cast<const char*>(src);
    while (n) {
        ssize_t r = ::write(fd, p, n);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
    }
    return true;
}

int open_read(const char* path, Fd& f, uint64_t& size) {
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return set_error(STT_ERR_INVALID, std::string("cannot open '") + path + "': " + std::strerror(errno));
    struct stat st;
    if (::fstat(f.fd, &st) != 0) return set_error(STT_ERR_INVALID, std::string("cannot stat '") + path + "'");
    size = (uint64_t)st.st_size;
    return STT_OK;
}

int open_write(const char* path, Fd& f) {
    f.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (f.fd < 0)
        return set_error(STT_ERR_INVALID, std::string("cannot create '") + path + "': " + std::strerror(errno));
    return STT_OK;
}

// Extension as cimg::split_filename takes it: the text after the last '.'.
std::string extension(const char* path) {
    const char* dot = std::strrchr(path, '.');
    return dot ? std::string(dot + 1) : std::string();
}

bool ieq(const std::string& a, const char* b) { return strcasecmp(a.c_str(), b) == 0; }

template <typename T>
T bswap(T v) {
    uint8_t b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    std::reverse(b, b + sizeof(T));
    std::memcpy(&v, b, sizeof(T));
    return v;
}

void swap_words(uint8_t* p, size_t n_words, int width) {
    for (size_t i = 0; i < n_words; ++i, p += width) std::reverse(p, p + width);
}

// ------------------------------------------------------------ TIFF read --

enum : uint16_t {
    TAG_WIDTH = 256, TAG_LENGTH = 257, TAG_BPS = 258, TAG_COMPRESSION = 259, TAG_PHOTOMETRIC = 262,
    TAG_DESCRIPTION = 270, TAG_STRIP_OFFSETS = 273, TAG_SPP = 277, TAG_ROWS_PER_STRIP = 278,
    TAG_STRIP_COUNTS = 279, TAG_XRES = 282, TAG_YRES = 283, TAG_PLANAR = 284, TAG_PREDICTOR = 317,
    TAG_TILE_WIDTH = 322, TAG_TILE_LENGTH = 323, TAG_TILE_OFFSETS = 324, TAG_TILE_COUNTS = 325,
    TAG_SAMPLE_FORMAT = 339,
};

int type_size(uint16_t t) {
    switch (t) {
        case 1: case 2: case 6: case 7: return 1;     // BYTE ASCII SBYTE UNDEFINED
        case 3: case 8: return 2;                     // SHORT SSHORT
        case 4: case 9: case 11: case 13: return 4;   // LONG SLONG FLOAT IFD
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;  // RATIONALs DOUBLE LONG8 SLONG8 IFD8
        default: return 0;
    }
}

struct Tiff {
    int fd = -1;
    uint64_t size = 0;
    bool swap = false;  // file byte order differs from the host's (little endian)
    bool big = false;   // BigTIFF
    std::string path;

    template <typename T>
    T get(const uint8_t* p) const {
        T v;
        std::memcpy(&v, p, sizeof(T));
        return swap ? bswap(v) : v;
    }
    int err(const std::string& m) const { return set_error(STT_ERR_INVALID, "TIFF '" + path + "': " + m); }
};

struct Entry {
    uint16_t tag = 0, type = 0;
    uint64_t count = 0;
    uint8_t field[8] = {};
};

// Raw bytes of an entry's values (inline or at the offset it holds).
int entry_bytes(const Tiff& t, const Entry& e, std::vector<uint8_t>& out, uint64_t max_count) {
    const int ts = type_size(e.type);
    if (!ts) return t.err("tag " + std::to_string(e.tag) + " has unknown type " + std::to_string(e.type));
    if (e.count > max_count) return t.err("tag " + std::to_string(e.tag) + " has too many values");
    const uint64_t n = e.count * (uint64_t)ts;
    out.resize((size_t)n);
    if (n <= (t.big ? 8u : 4u)) {
        std::memcpy(out.data(), e.field, (size_t)n);
        return STT_OK;
    }
    const uint64_t off = t.big ? t.get<uint64_t>(e.field) : t.get<uint32_t>(e.field);
    if (off > t.size || n > t.size - off) return t.err("tag " + std::to_string(e.tag) + " points past the end");
    if (!read_at(t.fd, off, out.data(), (size_t)n)) return t.err("read error");
    return STT_OK;
}

// Integer values of an entry (BYTE, SHORT, LONG, LONG8 and the IFD types).
int entry_uints(const Tiff& t, const Entry& e, std::vector<uint64_t>& v, uint64_t max_count) {
    std::vector<uint8_t> b;
    int rc = entry_bytes(t, e, b, max_count);
    if (rc) return rc;
    v.resize((size_t)e.count);
    for (size_t i = 0; i < v.size(); ++i) {
        switch (e.type) {
            case 1: case 7: v[i] = b[i]; break;
            case 3: v[i] = t.get<uint16_t>(&b[2 * i]); break;
            case 4: case 13: v[i] = t.get<uint32_t>(&b[4 * i]); break;
            case 16: case 18: v[i] = t.get<uint64_t>(&b[8 * i]); break;
            default: return t.err("tag " + std::to_string(e.tag) + " is not an unsigned integer");
        }
    }
    return STT_OK;
}

int entry_uint(const Tiff& t, const Entry& e, uint64_t& v) {
    std::vector<uint64_t> vals;
    int rc = entry_uints(t, e, vals, 1u << 16);
    if (rc) return rc;
    if (vals.empty()) return t.err("tag " + std::to_string(e.tag) + " has no value");
    v = vals[0];
    return STT_OK;
}

double entry_real(const Tiff& t, const Entry& e) {
    std::vector<uint8_t> b;
    if (e.count < 1 || entry_bytes(t, e, b, 1u << 16)) return 0.0;
    switch (e.type) {
        case 5: {
            const uint32_t num = t.get<uint32_t>(&b[0]), den = t.get<uint32_t>(&b[4]);
            return den ? (double)num / den : 0.0;
        }
        case 11: return t.get<float>(&b[0]);
        case 12: return t.get<double>(&b[0]);
        case 3: return t.get<uint16_t>(&b[0]);
        case 4: return t.get<uint32_t>(&b[0]);
        default: return 0.0;
    }
}

struct Page {
    uint64_t w = 0, h = 0;
    uint64_t spp = 1, bps = 1, sf = 1, comp = 1, photo = 0, planar = 1, pred = 1;
    bool has_spp = false, tiled = false;
    uint64_t rps = 0xFFFFFFFFull, tw = 0, th = 0;
    std::vector<uint64_t> offs, cnts;
    double xres = 0, yres = 0;
    float vz = 0;
    int spectrum = 1;
};

// Reads the IFD at off into pg; next receives the following IFD's offset.
int read_ifd(const Tiff& t, uint64_t off, Page& pg, uint64_t& next) {
    const size_t cnt_bytes = t.big ? 8 : 2, ent_bytes = t.big ? 20 : 12, next_bytes = t.big ? 8 : 4;
    uint8_t hdr[8];
    if (off + cnt_bytes > t.size || !read_at(t.fd, off, hdr, cnt_bytes)) return t.err("IFD past the end");
    const uint64_t n = t.big ? t.get<uint64_t>(hdr) : t.get<uint16_t>(hdr);
    if (n == 0 || n > 4096) return t.err("implausible IFD entry count " + std::to_string(n));
    const uint64_t len = n * ent_bytes + next_bytes;
    if (off + cnt_bytes + len > t.size) return t.err("IFD past the end");
    std::vector<uint8_t> raw((size_t)len);
    if (!read_at(t.fd, off + cnt_bytes, raw.data(), raw.size())) return t.err("read error");
    next = t.big ? t.get<uint64_t>(&raw[n * ent_bytes]) : t.get<uint32_t>(&raw[n * ent_bytes]);

    bool have_bps = false;
    int rc = STT_OK;
    for (uint64_t i = 0; i < n && !rc; ++i) {
        const uint8_t* p = &raw[i * ent_bytes];
        Entry e;
        e.tag = t.get<uint16_t>(p);
        e.type = t.get<uint16_t>(p + 2);
        e.count = t.big ? t.get<uint64_t>(p + 4) : t.get<uint32_t>(p + 4);
        std::memcpy(e.field, p + (t.big ? 12 : 8), t.big ? 8 : 4);
        if (!type_size(e.type)) continue;  // unknown types are skipped, as readers must
        switch (e.tag) {
            case TAG_WIDTH: rc = entry_uint(t, e, pg.w); break;
            case TAG_LENGTH: rc = entry_uint(t, e, pg.h); break;
            case TAG_BPS: {
                std::vector<uint64_t> v;
                rc = entry_uints(t, e, v, 1u << 16);
                if (!rc && !v.empty()) {
                    pg.bps = v[0];
                    have_bps = true;
                    for (uint64_t b : v)
                        if (b != v[0]) rc = t.err("samples of different bit depths");
                }
                break;
            }
            case TAG_COMPRESSION: rc = entry_uint(t, e, pg.comp); break;
            case TAG_PHOTOMETRIC: rc = entry_uint(t, e, pg.photo); break;
            case TAG_SPP: rc = entry_uint(t, e, pg.spp); pg.has_spp = true; break;
            case TAG_ROWS_PER_STRIP: rc = entry_uint(t, e, pg.rps); break;
            case TAG_PLANAR: rc = entry_uint(t, e, pg.planar); break;
            case TAG_PREDICTOR: rc = entry_uint(t, e, pg.pred); break;
            case TAG_SAMPLE_FORMAT: rc = entry_uint(t, e, pg.sf); break;
            case TAG_TILE_WIDTH: rc = entry_uint(t, e, pg.tw); pg.tiled = true; break;
            case TAG_TILE_LENGTH: rc = entry_uint(t, e, pg.th); pg.tiled = true; break;
            case TAG_STRIP_OFFSETS: case TAG_TILE_OFFSETS: rc = entry_uints(t, e, pg.offs, 1u << 26); break;
            case TAG_STRIP_COUNTS: case TAG_TILE_COUNTS: rc = entry_uints(t, e, pg.cnts, 1u << 26); break;
            case TAG_XRES: pg.xres = entry_real(t, e); break;
            case TAG_YRES: pg.yres = entry_real(t, e); break;
            case TAG_DESCRIPTION: {
                // CImg writes "VX=%g VY=%g VZ=%g spacing=%g" (CImg.h:51610); the z
                // size is only recorded there.
                std::vector<uint8_t> b;
                if (entry_bytes(t, e, b, 1u << 20) == STT_OK) {
                    std::string s(b.begin(), b.end());
                    float vx, vy, vz;
                    const size_t k = s.find("VX=");
                    if (k != std::string::npos && std::sscanf(s.c_str() + k, "VX=%f VY=%f VZ=%f", &vx, &vy, &vz) == 3)
                        pg.vz = vz;
                }
                break;
            }
            default: break;
        }
    }
    if (rc) return rc;
    if (!have_bps) pg.bps = 1;  // TIFFGetFieldDefaulted
    if (pg.w == 0 || pg.h == 0 || pg.w > (1u << 24) || pg.h > (1u << 24)) return t.err("bad image size");
    if (pg.spp == 0 || pg.spp > 64) return t.err("bad samples per pixel");
    // CImg::_load_tiff's channel count (CImg.h:47649).
    pg.spectrum = (!pg.has_spp || pg.photo >= 3) ? (pg.photo > 1 ? 3 : 1) : (int)pg.spp;
    return STT_OK;
}

int open_tiff(const char* path, Tiff& t, Fd& f, std::vector<Page>& pages) {
    t.path = path;
    int rc = open_read(path, f, t.size);
    if (rc) return rc;
    t.fd = f.fd;
    uint8_t h[16];
    if (t.size < 8 || !read_at(t.fd, 0, h, 8)) return t.err("too short for a TIFF header");
    if (h[0] == 'I' && h[1] == 'I') t.swap = false;
    else if (h[0] == 'M' && h[1] == 'M') t.swap = true;
    else return t.err("not a TIFF file");
    const uint16_t magic = t.get<uint16_t>(h + 2);
    uint64_t first;
    if (magic == 42) {
        first = t.get<uint32_t>(h + 4);
    } else if (magic == 43) {
        t.big = true;
        if (t.size < 16 || !read_at(t.fd, 0, h, 16)) return t.err("too short for a BigTIFF header");
        if (t.get<uint16_t>(h + 4) != 8) return t.err("unsupported BigTIFF offset size");
        first = t.get<uint64_t>(h + 8);
    } else {
        return t.err("bad TIFF magic number");
    }
    std::unordered_set<uint64_t> seen;
    for (uint64_t off = first; off != 0;) {
        if (!seen.insert(off).second) return t.err("IFD chain loops");
        if (pages.size() >= (1u << 20)) return t.err("too many pages");
        Page pg;
        uint64_t next = 0;
        if ((rc = read_ifd(t, off, pg, next))) return rc;
        pages.push_back(std::move(pg));
        off = next;
    }
    if (pages.empty()) return t.err("no image");
    const Page& p0 = pages[0];
    for (const Page& p : pages)
        if (p.w != p0.w || p.h != p0.h || p.spectrum != p0.spectrum)
            return t.err("pages of different sizes are not supported");
    // W, H < 2^24, pages < 2^20, channels <= 64: bound the product before it is used
    const double samples = (double)p0.w * (double)p0.h * (double)pages.size() * (double)p0.spectrum;
    if (samples > 1099511627776.0) return t.err("image larger than 2^40 samples");
    return STT_OK;
}

// PackBits (TIFF 6.0 section 9): fills exactly n bytes of out.
bool unpack_bits(const uint8_t* in, size_t nin, uint8_t* out, size_t n) {
    size_t i = 0, o = 0;
    while (o < n && i < nin) {
        const int c = (int8_t)in[i++];
        if (c >= 0) {
            const size_t len = (size_t)c + 1;
            if (i + len > nin || o + len > n) return false;
            std::memcpy(out + o, in + i, len);
            i += len;
            o += len;
        } else if (c != -128) {
            const size_t len = (size_t)(1 - c);
            if (i >= nin || o + len > n) return false;
            std::memset(out + o, in[i++], len);
            o += len;
        }
    }
    return o == n;
}

// LZW as TIFF 6.0 section 13 defines it: MSB-first codes of 9 to 12 bits,
// Clear 256, EndOfInformation 257, code width growing one code early.
// Fills exactly n bytes of out.
bool unlzw(const uint8_t* in, size_t nin, uint8_t* out, size_t n) {
    constexpr int kClear = 256, kEoi = 257, kMax = 4096;
    std::vector<uint32_t> start(kMax), len(kMax);
    size_t o = 0;
    uint64_t bitpos = 0;
    const uint64_t nbits = (uint64_t)nin * 8;
    int width = 9, next = 258, prev = -1;
    auto read_code = [&](int& code) {
        if (bitpos + width > nbits) return false;
        uint32_t v = 0;
        for (int k = 0; k < width; ++k, ++bitpos) v = (v << 1) | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
        code = (int)v;
        return true;
    };
    // Table entries point at their first occurrence in the output.
    for (;;) {
        int code;
        if (!read_code(code)) break;
        if (code == kEoi) break;
        if (code == kClear) {
            width = 9;
            next = 258;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255 || o >= n) return o == n;
            out[o] = (uint8_t)code;
            start[code] = (uint32_t)o;
            len[code] = 1;
            ++o;
            prev = code;
        } else {
            const size_t at = o;
            uint32_t l;
            if (code < next) {
                l = code < 256 ? 1 : len[code];
                if (o + l > n) return false;
                if (code < 256) out[o] = (uint8_t)code;
                else std::memmove(out + o, out + start[code], l);
                o += l;
            } else if (code == next) {
                const uint32_t pl = prev < 256 ? 1 : len[prev];
                l = pl + 1;
                if (o + l > n) return false;
                if (prev < 256) out[o] = (uint8_t)prev;
                else std::memmove(out + o, out + start[prev], pl);
                out[o + pl] = out[o];
                o += l;
            } else {
                return false;
            }
            if (next < kMax) {
                // new entry = string(prev) + first byte of this code's string
                const uint32_t pl = prev < 256 ? 1 : len[prev];
                start[next] = (uint32_t)(at - pl);
                len[next] = pl + 1;
                ++next;
            }
            prev = code;
        }
        if (next + 1 >= (1 << width) && width < 12) ++width;
        if (o == n) break;
    }
    return o == n;
}

enum class SType { U8, I8, U16, I16, U32, I32, F32, F64 };

// CImg::_load_tiff's sample type (CImg.h:47732-47760); 64-bit float added.
bool sample_type(const Page& p, SType& st) {
    switch (p.bps) {
        case 8: st = p.sf == 1 ? SType::U8 : SType::I8; return true;
        case 16: st = p.sf == 1 ? SType::U16 : SType::I16; return true;
        case 32: st = p.sf == 1 ? SType::U32 : p.sf == 2 ? SType::I32 : SType::F32; return true;
        case 64: st = SType::F64; return p.sf == 3;
        default: return false;
    }
}

template <typename T>
void undo_predictor(uint8_t* buf, uint64_t rows, uint64_t row_samples, uint64_t stride) {
    for (uint64_t r = 0; r < rows; ++r) {
        T* v = reinterpret_cast<T*>(buf) + r * row_samples;
        for (uint64_t i = stride; i < row_samples; ++i) v[i] = (T)(v[i] + v[i - stride]);
    }
}

float sample_at(const uint8_t* p, SType st) {
    switch (st) {
        case SType::U8: return (float)*p;
        case SType::I8: return (float)(int8_t)*p;
        case SType::U16: { uint16_t v; std::memcpy(&v, p, 2); return (float)v; }
        case SType::I16: { int16_t v; std::memcpy(&v, p, 2); return (float)v; }
        case SType::U32: { uint32_t v; std::memcpy(&v, p, 4); return (float)v; }
        case SType::I32: { int32_t v; std::memcpy(&v, p, 4); return (float)v; }
        case SType::F32: { float v; std::memcpy(&v, p, 4); return v; }
        case SType::F64: { double v; std::memcpy(&v, p, 8); return (float)v; }
    }
    return 0.f;
}

// Decodes page z of the stack into out (CImg layout, W*H*D*C floats).
int decode_page(const Tiff& t, const Page& p, uint64_t z, uint64_t D, float* out, std::vector<uint8_t>& raw,
                std::vector<uint8_t>& buf) {
    SType st;
    if (!sample_type(p, st))
        return t.err("unsupported sample layout: " + std::to_string(p.bps) + " bits, format " + std::to_string(p.sf));
    if (p.photo >= 3) return t.err("palette, CMYK and YCbCr images are not supported");
    if (p.comp != 1 && p.comp != 5 && p.comp != 32773)
        return t.err("compression " + std::to_string(p.comp) + " is not supported (none, LZW, PackBits)");
    if (p.pred != 1 && p.pred != 2) return t.err("predictor " + std::to_string(p.pred) + " is not supported");
    if (p.pred == 2 && (st == SType::F32 || st == SType::F64)) return t.err("horizontal predictor on float samples");
    if (p.planar != 1 && p.planar != 2) return t.err("bad planar configuration");
    const uint64_t W = p.w, H = p.h, bytes = p.bps / 8;
    const uint64_t spp = p.spp;
    const uint64_t C = (uint64_t)p.spectrum;
    const uint64_t cw = p.tiled ? p.tw : W;                      // chunk width
    const uint64_t ch = p.tiled ? p.th : std::min(p.rps, H);     // chunk height
    if (cw == 0 || ch == 0 || cw > (1u << 24) || ch > (1u << 24)) return t.err("bad strip or tile size");
    const uint64_t across = (W + cw - 1) / cw, down = (H + ch - 1) / ch;
    const uint64_t planes = p.planar == 2 ? spp : 1;
    const uint64_t per_px = p.planar == 2 ? 1 : spp;              // samples per pixel within a chunk
    const uint64_t nchunks = across * down * planes;
    if (p.offs.size() < nchunks || p.cnts.size() < nchunks) return t.err("missing strip or tile offsets");
    const uint64_t row_samples = cw * per_px, row_bytes = row_samples * bytes;

    for (uint64_t s = 0; s < planes; ++s)
        for (uint64_t cy = 0; cy < down; ++cy)
            for (uint64_t cx = 0; cx < across; ++cx) {
                const uint64_t idx = (s * down + cy) * across + cx;
                const uint64_t y0 = cy * ch, x0 = cx * cw;
                // strips hold only the rows that exist; tiles are always full
                const uint64_t rows = p.tiled ? ch : std::min(ch, H - y0);
                const uint64_t need = rows * row_bytes;
                const uint64_t off = p.offs[idx], cnt = p.cnts[idx];
                if (off > t.size || cnt > t.size - off) return t.err("strip or tile past the end of the file");
                // native-order fp32 strips of one sample per pixel are rows of
                // the output already: read them in place
                if (!p.tiled && per_px == 1 && st == SType::F32 && p.comp == 1 && !t.swap && p.pred == 1) {
                    const uint64_t c = p.planar == 2 ? s : 0;
                    if (c >= C) continue;
                    if (cnt < need) return t.err("short strip or tile");
                    if (!read_at(t.fd, off, out + W * (y0 + H * (z + D * c)), (size_t)need)) return t.err("read error");
                    continue;
                }
                buf.resize((size_t)need);
                if (p.comp == 1) {
                    if (cnt < need) return t.err("short strip or tile");
                    if (!read_at(t.fd, off, buf.data(), (size_t)need)) return t.err("read error");
                } else {
                    raw.resize((size_t)cnt);
                    if (cnt && !read_at(t.fd, off, raw.data(), (size_t)cnt)) return t.err("read error");
                    const bool ok = p.comp == 5 ? unlzw(raw.data(), raw.size(), buf.data(), (size_t)need)
                                                : unpack_bits(raw.data(), raw.size(), buf.data(), (size_t)need);
                    if (!ok) return t.err("corrupt compressed strip or tile");
                }
                if (t.swap && bytes > 1) swap_words(buf.data(), (size_t)(need / bytes), (int)bytes);
                if (p.pred == 2) {
                    switch (bytes) {
                        case 1: undo_predictor<uint8_t>(buf.data(), rows, row_samples, per_px); break;
                        case 2: undo_predictor<uint16_t>(buf.data(), rows, row_samples, per_px); break;
                        case 4: undo_predictor<uint32_t>(buf.data(), rows, row_samples, per_px); break;
                        default: return t.err("horizontal predictor on 64-bit samples");
                    }
                }
                const uint64_t xn = std::min(cw, W - x0), yn = std::min(rows, H - y0);
                for (uint64_t r = 0; r < yn; ++r) {
                    const uint8_t* src = buf.data() + r * row_bytes;
                    const uint64_t y = y0 + r;
                    for (uint64_t x = 0; x < xn; ++x)
                        for (uint64_t v = 0; v < per_px; ++v) {
                            const uint64_t c = p.planar == 2 ? s : v;
                            if (c >= C) continue;
                            out[(x0 + x) + W * (y + H * (z + D * c))] =
                                sample_at(src + (x * per_px + v) * bytes, st);
                        }
                }
            }
    return STT_OK;
}

int tiff_info(const char* path, stt_image_info* info) {
    Tiff t;
    Fd f;
    std::vector<Page> pages;
    int rc = open_tiff(path, t, f, pages);
    if (rc) return rc;
    const Page& p = pages[0];
    info->width = (int)p.w;
    info->height = (int)p.h;
    info->depth = (int)pages.size();
    info->spectrum = p.spectrum;
    info->bits = (int)p.bps;
    info->sample_format = (int)p.sf;
    info->voxel[0] = p.xres > 0 ? (float)(1.0 / p.xres) : 1.f;
    info->voxel[1] = p.yres > 0 ? (float)(1.0 / p.yres) : 1.f;
    info->voxel[2] = p.vz > 0 ? p.vz : 1.f;
    return STT_OK;
}

int tiff_load(const char* path, float* out, size_t n) {
    Tiff t;
    Fd f;
    std::vector<Page> pages;
    int rc = open_tiff(path, t, f, pages);
    if (rc) return rc;
    const Page& p0 = pages[0];
    const uint64_t D = pages.size();
    const uint64_t total = p0.w * p0.h * D * (uint64_t)p0.spectrum;
    if (n < total) return set_error(STT_ERR_INVALID, "output buffer too small for '" + std::string(path) + "'");
    std::memset(out, 0, total * sizeof(float));
    std::atomic<int> first_rc{STT_OK};
    std::string first_msg;
#pragma omp parallel
    {
        std::vector<uint8_t> raw, buf;
#pragma omp for schedule(dynamic, 1)
        for (long long z = 0; z < (long long)D; ++z) {
            if (first_rc.load() != STT_OK) continue;
            const int r = decode_page(t, pages[(size_t)z], (uint64_t)z, D, out, raw, buf);
            if (r != STT_OK) {
                int expect = STT_OK;
                if (first_rc.compare_exchange_strong(expect, r)) first_msg = stt_last_error();
            }
        }
    }
    if (first_rc.load() != STT_OK) return set_error(first_rc.load(), first_msg);
    return STT_OK;
}

// ----------------------------------------------------------- TIFF write --

struct IfdWriter {
    bool big;
    std::vector<uint8_t> ifd, extra;  // entries; out-of-line values
    uint64_t extra_base = 0;          // file offset of `extra`
    int n = 0;

    void put(std::vector<uint8_t>& v, uint64_t x, int w) {
        for (int i = 0; i < w; ++i) v.push_back((uint8_t)(x >> (8 * i)));  // little endian
    }
    // One entry with `count` values of `type`, given as integers (or raw bytes for ASCII).
    void entry(uint16_t tag, uint16_t type, const std::vector<uint64_t>& vals, const char* ascii = nullptr) {
        const int ts = type_size(type);
        const uint64_t count = ascii ? std::strlen(ascii) + 1 : vals.size();
        std::vector<uint8_t> data;
        if (ascii) data.assign(ascii, ascii + count);
        else
            for (uint64_t v : vals) put(data, v, ts);
        put(ifd, tag, 2);
        put(ifd, type, 2);
        put(ifd, count, big ? 8 : 4);
        const size_t cap = big ? 8 : 4;
        if (data.size() <= cap) {
            data.resize(cap, 0);
            ifd.insert(ifd.end(), data.begin(), data.end());
        } else {
            put(ifd, extra_base + extra.size(), (int)cap);
            extra.insert(extra.end(), data.begin(), data.end());
            if (extra.size() & 1) extra.push_back(0);  // values start on a word boundary
        }
        ++n;
    }
};

int tiff_save(const char* path, const float* img, int w, int h, int d, int c) {
    if (!path || !img || w < 1 || h < 1 || d < 1 || c < 1 || c > 64)
        return set_error(STT_ERR_INVALID, "save_tiff: need a path, an image and positive sizes (c <= 64)");
    const uint64_t W = (uint64_t)w, H = (uint64_t)h, D = (uint64_t)d, C = (uint64_t)c;
    const uint64_t page_bytes = W * H * C * 4;
    // CImg::save_tiff switches to BigTIFF from 2 GiB of pixel data (CImg.h:51564).
    const bool big = W * H * D * C * 4 >= (1ull << 31);
    // libtiff's default strip size for a requested (uint32)-1: 8 KiB worth of rows.
    const uint64_t scanline = W * C * 4;
    const uint64_t rps = std::max<uint64_t>(8192 / scanline, 1);
    const uint64_t nstrips = (H + rps - 1) / rps;
    const uint16_t photometric = (c == 3 || c == 4) ? 2 : 1;

    // Layout: header, then per page its strips followed by its IFD and the IFD's
    // out-of-line values.  All pages have the same IFD size.
    const uint64_t hdr = big ? 16 : 8;
    auto build_ifd = [&](uint64_t data_off, uint64_t ifd_off, uint64_t next_ifd) {
        IfdWriter iw{big, {}, {}, 0, 0};
        const int nent = 14;
        const uint64_t ifd_len = (big ? 8 : 2) + (uint64_t)nent * (big ? 20 : 12) + (big ? 8 : 4);
        iw.extra_base = ifd_off + ifd_len;
        std::vector<uint64_t> offs(nstrips), cnts(nstrips);
        for (uint64_t s = 0; s < nstrips; ++s) {
            const uint64_t rows = std::min(rps, H - s * rps);
            offs[s] = data_off + s * rps * scanline;
            cnts[s] = rows * scanline;
        }
        const uint16_t LONGT = 4, SHORTT = 3, OFFT = big ? 16 : 4;
        iw.entry(TAG_WIDTH, LONGT, {W});
        iw.entry(TAG_LENGTH, LONGT, {H});
        iw.entry(TAG_BPS, SHORTT, std::vector<uint64_t>(C, 32));
        iw.entry(TAG_COMPRESSION, SHORTT, {1});
        iw.entry(TAG_PHOTOMETRIC, SHORTT, {photometric});
        iw.entry(266, SHORTT, {1});  // FillOrder MSB2LSB
        iw.entry(TAG_STRIP_OFFSETS, OFFT, offs);
        iw.entry(274, SHORTT, {1});  // Orientation TOPLEFT
        iw.entry(TAG_SPP, SHORTT, {C});
        iw.entry(TAG_ROWS_PER_STRIP, LONGT, {rps});
        iw.entry(TAG_STRIP_COUNTS, OFFT, cnts);
        iw.entry(TAG_PLANAR, SHORTT, {1});
        iw.entry(305, 2, {}, "CImg");  // Software
        iw.entry(TAG_SAMPLE_FORMAT, SHORTT, std::vector<uint64_t>(C, 3));
        std::vector<uint8_t> out;
        iw.put(out, (uint64_t)iw.n, big ? 8 : 2);
        out.insert(out.end(), iw.ifd.begin(), iw.ifd.end());
        iw.put(out, next_ifd, big ? 8 : 4);
        out.insert(out.end(), iw.extra.begin(), iw.extra.end());
        while (out.size() % 8) out.push_back(0);
        return out;
    };
    const uint64_t ifd_size = build_ifd(0, 0, 0).size();
    const uint64_t block = page_bytes + ifd_size;  // page_bytes is a multiple of 4; pad it to 8 below
    const uint64_t pad = (8 - page_bytes % 8) % 8;
    const uint64_t stride = block + pad;
    if (!big && hdr + D * stride > 0xFFFFFFFFull)
        return set_error(STT_ERR_INVALID, "save_tiff: image too large for classic TIFF");

    Fd f;
    int rc = open_write(path, f);
    if (rc) return rc;
    std::vector<uint8_t> head;
    auto put = [&](uint64_t x, int wdt) {
        for (int i = 0; i < wdt; ++i) head.push_back((uint8_t)(x >> (8 * i)));
    };
    head.push_back('I');
    head.push_back('I');
    const uint64_t first_ifd = hdr + page_bytes + pad;
    if (big) {
        put(43, 2);
        put(8, 2);
        put(0, 2);
        put(first_ifd, 8);
    } else {
        put(42, 2);
        put(first_ifd, 4);
    }
    if (!write_all(f.fd, head.data(), head.size())) return set_error(STT_ERR_INVALID, "save_tiff: write error");
    std::vector<float> inter;
    const uint8_t zeros[8] = {};
    for (uint64_t z = 0; z < D; ++z) {
        const uint64_t data_off = hdr + z * stride;
        const uint64_t ifd_off = data_off + page_bytes + pad;
        const uint64_t next = z + 1 < D ? ifd_off + stride : 0;
        const float* src = img + z * W * H;
        bool ok;
        if (C == 1) {
            ok = write_all(f.fd, src, (size_t)page_bytes);
        } else {
            // chunky: the samples of a pixel are adjacent (CImg.h:51628-51631)
            inter.resize((size_t)(W * H * C));
            for (uint64_t i = 0; i < W * H; ++i)
                for (uint64_t v = 0; v < C; ++v) inter[i * C + v] = img[i + W * H * (z + D * v)];
            ok = write_all(f.fd, inter.data(), (size_t)page_bytes);
        }
        if (ok && pad) ok = write_all(f.fd, zeros, (size_t)pad);
        const std::vector<uint8_t> ifd = build_ifd(data_off, ifd_off, next);
        if (ok) ok = write_all(f.fd, ifd.data(), ifd.size());
        if (!ok) return set_error(STT_ERR_INVALID, std::string("save_tiff: write error on '") + path + "'");
    }
    return STT_OK;
}

// ------------------------------------------------------------- Analyze --

struct AnalyzeNames {
    std::string hdr, img;  // img empty for a single-file .nii
};

// File names as CImg::_load_analyze pairs them (CImg.h:47838-47848).
AnalyzeNames analyze_read_names(const char* path) {
    const std::string p(path), ext = extension(path);
    const std::string body = ext.empty() && p.find('.') == std::string::npos ? p : p.substr(0, p.size() - ext.size() - 1);
    if (ieq(ext, "hdr")) return {p, body + ".img"};
    if (ieq(ext, "img")) return {body + ".hdr", p};
    return {p, ""};
}

struct AnalyzeHeader {
    bool swap = false;
    uint64_t dims[4] = {1, 1, 1, 1};
    uint16_t datatype = 0;
    float scale = 1.f;
    float voxel[3] = {1.f, 1.f, 1.f};
    uint64_t vox_offset = 0;
};

int analyze_header(const char* path, const AnalyzeNames& nm, AnalyzeHeader& ah) {
    Fd f;
    uint64_t size = 0;
    int rc = open_read(nm.hdr.c_str(), f, size);
    if (rc) return rc;
    uint32_t hsize = 0;
    if (size < 4 || !read_at(f.fd, 0, &hsize, 4)) return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': short header");
    if (hsize == 0) return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': zero-size header");
    // CImg's byte-order test: a header size >= 4096 is the swapped 348.
    if (hsize >= 4096) {
        ah.swap = true;
        hsize = bswap(hsize);
    }
    if (hsize < 116 || hsize > size) return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': bad header size");
    std::vector<uint8_t> h(hsize);
    if (!read_at(f.fd, 0, h.data(), hsize)) return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': read error");
    auto s16 = [&](size_t o) { uint16_t v; std::memcpy(&v, &h[o], 2); return ah.swap ? bswap(v) : v; };
    auto f32 = [&](size_t o) { float v; std::memcpy(&v, &h[o], 4); return ah.swap ? bswap(v) : v; };
    const uint16_t ndim = s16(40);
    for (int k = 0; k < 4; ++k) ah.dims[k] = ndim > k ? s16(42 + 2 * k) : 1;
    if (ndim == 0 || ah.dims[0] == 0 || ah.dims[1] == 0 || ah.dims[2] == 0 || ah.dims[3] == 0)
        return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': image with zero dimensions");
    if ((double)ah.dims[0] * ah.dims[1] * ah.dims[2] * ah.dims[3] > 1099511627776.0)
        return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': image larger than 2^40 samples");
    ah.datatype = s16(70);
    ah.scale = f32(112);
    if (ah.scale == 0.f) ah.scale = 1.f;
    for (int k = 0; k < 3; ++k) ah.voxel[k] = f32(80 + 4 * k);
    if (nm.img.empty()) {
        const float vo = f32(108);
        if (!(vo >= 0.f) || vo > 4294967295.f) return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': bad vox_offset");
        ah.vox_offset = (uint64_t)(uint32_t)vo;
    }
    return STT_OK;
}

int analyze_info(const char* path, stt_image_info* info) {
    const AnalyzeNames nm = analyze_read_names(path);
    AnalyzeHeader ah;
    int rc = analyze_header(path, nm, ah);
    if (rc) return rc;
    info->width = (int)ah.dims[0];
    info->height = (int)ah.dims[1];
    info->depth = (int)ah.dims[2];
    info->spectrum = (int)ah.dims[3];
    switch (ah.datatype) {
        case 2: info->bits = 8; info->sample_format = 1; break;
        case 4: info->bits = 16; info->sample_format = 2; break;
        case 8: info->bits = 32; info->sample_format = 2; break;
        case 16: info->bits = 32; info->sample_format = 3; break;
        case 64: info->bits = 64; info->sample_format = 3; break;
        default: info->bits = 0; info->sample_format = 0; break;
    }
    for (int k = 0; k < 3; ++k) info->voxel[k] = ah.voxel[k];
    return STT_OK;
}

template <typename T, typename M>
void analyze_convert(const uint8_t* raw, size_t n, bool swap, float scale, float* out) {
    for (size_t i = 0; i < n; ++i) {
        T v;
        std::memcpy(&v, raw + i * sizeof(T), sizeof(T));
        if (swap) v = bswap(v);
        // CImg: _data[off] = (T)(buffer[off]*scalefactor), in the promoted type M
        out[i] = (float)((M)v * (M)scale);
    }
}

int analyze_load(const char* path, float* out, size_t n) {
    const AnalyzeNames nm = analyze_read_names(path);
    AnalyzeHeader ah;
    int rc = analyze_header(path, nm, ah);
    if (rc) return rc;
    const uint64_t total = ah.dims[0] * ah.dims[1] * ah.dims[2] * ah.dims[3];
    if (n < total) return set_error(STT_ERR_INVALID, std::string("output buffer too small for '") + path + "'");
    int width;
    switch (ah.datatype) {
        case 2: width = 1; break;
        case 4: width = 2; break;
        case 8: case 16: width = 4; break;
        case 64: width = 8; break;
        default:
            return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': unsupported datatype " +
                                                  std::to_string(ah.datatype));
    }
    Fd f;
    uint64_t size = 0;
    rc = open_read(nm.img.empty() ? nm.hdr.c_str() : nm.img.c_str(), f, size);
    if (rc) return rc;
    const uint64_t bytes = total * (uint64_t)width;
    if (ah.vox_offset > size || bytes > size - ah.vox_offset)
        return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': data shorter than the header says");
    // x * 1.0f is x (a signalling NaN aside, which CImg would return quieted): read in place
    if (ah.datatype == 16 && !ah.swap && ah.scale == 1.f) {
        if (!read_at(f.fd, ah.vox_offset, out, (size_t)bytes))
            return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': read error");
        return STT_OK;
    }
    std::vector<uint8_t> raw;
    const uint64_t chunk = 1ull << 24;  // elements per read
    for (uint64_t i0 = 0; i0 < total; i0 += chunk) {
        const uint64_t m = std::min(chunk, total - i0);
        raw.resize((size_t)(m * width));
        if (!read_at(f.fd, ah.vox_offset + i0 * width, raw.data(), raw.size()))
            return set_error(STT_ERR_INVALID, std::string("Analyze '") + path + "': read error");
        float* o = out + i0;
        switch (ah.datatype) {
            case 2: analyze_convert<uint8_t, float>(raw.data(), m, false, ah.scale, o); break;
            case 4: analyze_convert<int16_t, float>(raw.data(), m, ah.swap, ah.scale, o); break;
            case 8: analyze_convert<int32_t, float>(raw.data(), m, ah.swap, ah.scale, o); break;
            case 16: analyze_convert<float, float>(raw.data(), m, ah.swap, ah.scale, o); break;
            case 64: analyze_convert<double, double>(raw.data(), m, ah.swap, ah.scale, o); break;
        }
    }
    return STT_OK;
}

int analyze_save(const char* path, const float* img, int w, int h, int d, int c, const float* voxel) {
    if (!path || !img || w < 1 || h < 1 || d < 1 || c < 1)
        return set_error(STT_ERR_INVALID, "save_analyze: need a path, an image and positive sizes");
    if (w > 32767 || h > 32767 || d > 32767 || c > 32767)
        return set_error(STT_ERR_INVALID, "save_analyze: dimensions must fit a 16-bit header field");
    // Names as CImg::save_analyze derives them (CImg.h:51722-51742).
    const std::string p(path), ext = extension(path);
    std::string hname, iname;
    if (ext.empty()) {
        hname = p + ".hdr";
        iname = p + ".img";
    }
    if (ieq(ext.substr(0, 3), "hdr")) {
        hname = p;
        iname = p.substr(0, p.size() - 3) + "img";
    }
    if (ieq(ext.substr(0, 3), "img")) {
        iname = p;
        hname = p.substr(0, p.size() - 3) + "hdr";
    }
    if (ieq(ext.substr(0, 3), "nii")) {
        hname = p;
        iname.clear();
    }
    if (hname.empty()) return set_error(STT_ERR_INVALID, "save_analyze: extension must be .hdr, .img, .nii or none");
    std::vector<uint8_t> hd(iname.empty() ? 352 : 348, 0);
    auto i32 = [&](size_t o, int32_t v) { std::memcpy(&hd[o], &v, 4); };
    auto i16 = [&](size_t o, int16_t v) { std::memcpy(&hd[o], &v, 2); };
    auto f32 = [&](size_t o, float v) { std::memcpy(&hd[o], &v, 4); };
    i32(0, 348);
    std::memcpy(&hd[4], "CImg", 5);
    std::memcpy(&hd[14], " ", 2);
    i16(36, 4096);
    hd[38] = 114;
    i16(40, 4);
    i16(42, (int16_t)w);
    i16(44, (int16_t)h);
    i16(46, (int16_t)d);
    i16(48, (int16_t)c);
    i16(70, 16);  // float
    i16(72, 4);
    f32(108, iname.empty() ? 352.f : 0.f);
    f32(112, 1.f);
    f32(76, 0.f);
    for (int k = 0; k < 3; ++k) f32(80 + 4 * k, voxel ? voxel[k] : 1.f);
    const size_t bytes = (size_t)w * h * d * c * sizeof(float);
    {
        Fd f;
        int rc = open_write(hname.c_str(), f);
        if (rc) return rc;
        bool ok = write_all(f.fd, hd.data(), hd.size());
        if (ok && iname.empty()) ok = write_all(f.fd, img, bytes);
        if (!ok) return set_error(STT_ERR_INVALID, "save_analyze: write error on '" + hname + "'");
    }
    if (!iname.empty()) {
        Fd f;
        int rc = open_write(iname.c_str(), f);
        if (rc) return rc;
        if (!write_all(f.fd, img, bytes)) return set_error(STT_ERR_INVALID, "save_analyze: write error on '" + iname + "'");
    }
    return STT_OK;
}

enum class Kind { Tiff, Analyze, Unknown };

Kind kind_of(const char* path) {
    const std::string ext = extension(path);
    if (ieq(ext, "tif") || ieq(ext, "tiff")) return Kind::Tiff;
    if (ieq(ext, "hdr") || ieq(ext, "nii") || ieq(ext, "img")) return Kind::Analyze;
    return Kind::Unknown;
}

}  // namespace

extern "C" {

int stt_image_query(const char* path, stt_image_info* info) {
    if (!path || !info) return set_error(STT_ERR_INVALID, "image_query: null argument");
    switch (kind_of(path)) {
        case Kind::Tiff: return tiff_info(path, info);
        case Kind::Analyze: return analyze_info(path, info);
        default: return set_error(STT_ERR_INVALID, std::string("unsupported image extension: '") + path + "'");
    }
}

int stt_load_image(const char* path, float* out, size_t n) {
    if (!path || !out) return set_error(STT_ERR_INVALID, "load_image: null argument");
    switch (kind_of(path)) {
        case Kind::Tiff: return tiff_load(path, out, n);
        case Kind::Analyze: return analyze_load(path, out, n);
        default: return set_error(STT_ERR_INVALID, std::string("unsupported image extension: '") + path + "'");
    }
}

int stt_save_tiff(const char* path, const float* img, int w, int h, int d, int c) {
    return tiff_save(path, img, w, h, d, c);
}

int stt_save_analyze(const char* path, const float* img, int w, int h, int d, int c, const float* voxel) {
    return analyze_save(path, img, w, h, d, c, voxel);
}

int stt_read_start_angles(const char* path, int nframes, int nrounds, float* deg) {
    if (!path || !deg || nframes < 1 || nrounds < 1)
        return set_error(STT_ERR_INVALID, "read_start_angles: need a path, an output and nframes, nrounds >= 1");
    std::FILE* fp = std::fopen(path, "rb");
    if (!fp) return set_error(STT_ERR_INVALID, std::string("cannot open '") + path + "': " + std::strerror(errno));
    std::string text;
    char chunk[4096];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof chunk, fp)) > 0) text.append(chunk, r);
    std::fclose(fp);
    const size_t need = (size_t)nframes * (size_t)nrounds;
    size_t got = 0, i = 0;
    auto is_space = [](char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f'; };
    while (got < need) {
        while (i < text.size() && is_space(text[i])) ++i;
        if (i >= text.size()) break;
        size_t j = i;
        while (j < text.size() && !is_space(text[j])) ++j;
        // operator>> takes an optional sign, digits, a point and an exponent;
        // from_chars rejects a leading '+', so skip it.
        const char* b = text.data() + i;
        const char* e = text.data() + j;
        if (b < e && *b == '+') ++b;
        float v = 0.f;
        const auto res = std::from_chars(b, e, v, std::chars_format::general);
        if (b == e || res.ec != std::errc() || res.ptr != e || !std::isfinite(v))
            return set_error(STT_ERR_INVALID, std::string("'") + path + "': malformed number '" + text.substr(i, j - i) + "'");
        deg[got++] = v;
        i = j;
    }
    if (got < need)
        return set_error(STT_ERR_INVALID, std::string("'") + path + "' holds " + std::to_string(got) + " numbers, " +
                                              std::to_string(need) + " needed (frames x rounds)");
    return STT_OK;
}

int stt_split_sizes(int size, int nparts, int* sizes) {
    if (!sizes || size < 1 || nparts < 1) return set_error(STT_ERR_INVALID, "split_sizes: need size, nparts >= 1");
    if (nparts > size) return set_error(STT_ERR_INVALID, "split_sizes: cannot split into more blocks than elements");
    if (nparts == 1) {
        sizes[0] = size;
        return STT_OK;
    }
    // CImg::get_split(axis, nb > 0) (CImg.h:31391-31420): a Bresenham walk
    // that closes a block whenever the running error drops to zero or below.
    int err = size, p0 = 0, k = 0;
    for (int p = 0; p < size; ++p)
        if ((err -= nparts) <= 0) {
            if (k >= nparts) return set_error(STT_ERR_INVALID, "split_sizes: internal block count mismatch");
            sizes[k++] = p - p0 + 1;
            err += size;
            p0 = p + 1;
        }
    if (k != nparts) return set_error(STT_ERR_INVALID, "split_sizes: internal block count mismatch");
    return STT_OK;
}

}  // extern "C"
