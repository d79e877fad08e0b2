// io.cpp -- host I/O around the walk: the `.bt` DEM format
// (input.rs:86-185), the greyscale heatmap (output/png.rs:5-37) and the
// per-sector band-data cache (cache.rs:6-76).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "engine.hpp"

namespace {

thread_local std::string g_io_error;

tvs_status io_fail(tvs_ctx *c, tvs_status st, const std::string &msg) {
    (c ? c->err : g_io_error) = msg;
    return st;
}

// Little-endian field readers over the 256-byte header.
template <typename T>
T le(const unsigned char *p) {
    T v;
    std::memcpy(&v, p, sizeof v);
    return v;
}

template <typename T>
void put_le(unsigned char *p, T v) {
    std::memcpy(p, &v, sizeof v);
}

constexpr size_t kBtHeaderSize = 256;

tvs_status read_bt(const char *path, tvs_bt_header *hdr, std::FILE **out_file) {
    if (!path || !hdr) return io_fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return io_fail(nullptr, TVS_ERR_IO, std::string("cannot open ") + path + ": " +
                                                    std::strerror(errno));
    unsigned char h[kBtHeaderSize];
    // input.rs:90-113: magic then the fixed fields (the reference reads
    // exactly 66 bytes of header before seeking to 256).
    if (std::fread(h, 1, 66, f) != 66) {
        std::fclose(f);
        return io_fail(nullptr, TVS_ERR_IO, "truncated .bt header");
    }
    if (std::memcmp(h, "binterr1.3", 10) != 0) {
        std::fclose(f);
        return io_fail(nullptr, TVS_ERR_IO, "Not a Binary Terrain v1.3 file");
    }
    hdr->width = le<uint32_t>(h + 10);
    hdr->height = le<uint32_t>(h + 14);
    hdr->data_size = le<uint16_t>(h + 18);
    hdr->is_float = le<uint16_t>(h + 20) != 0 ? 1 : 0;
    hdr->horizontal_units = le<uint16_t>(h + 22);
    hdr->utm_zone = le<uint16_t>(h + 24);
    hdr->datum = le<uint16_t>(h + 26);
    hdr->left = le<double>(h + 28);
    hdr->right = le<double>(h + 36);
    hdr->bottom = le<double>(h + 44);
    hdr->top = le<double>(h + 52);
    hdr->projection_source = le<uint16_t>(h + 60);
    hdr->vertical_scale = le<float>(h + 62);
    if (!hdr->is_float && hdr->data_size != 2) {  // input.rs:158-160
        std::fclose(f);
        return io_fail(nullptr, TVS_ERR_IO, "Unsupported `.bt` field value for data size: " +
                                                std::to_string(hdr->data_size));
    }
    if (out_file) *out_file = f; else std::fclose(f);
    return TVS_OK;
}

int mkdirs(const std::string &path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
    return 0;
}

// cache.rs:23-31: <state>/total-viewsheds/<width>/{band_deltas,distances}/<angle>.bin
std::string cache_file(const char *state_dir, uint32_t width, const char *kind,
                       uint32_t sector) {
    return std::string(state_dir) + "/total-viewsheds/" + std::to_string(width) + "/" +
           kind + "/" + std::to_string(sector) + ".bin";
}

// ---- PNG (stored deflate blocks; greyscale 8-bit) ----
struct CrcTable {
    uint32_t v[256];
    CrcTable() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t x = i;
            for (int k = 0; k < 8; ++k) x = (x & 1u) ? 0xedb88320u ^ (x >> 1) : x >> 1;
            v[i] = x;
        }
    }
};

uint32_t crc32(const unsigned char *p, size_t n, uint32_t c = 0xffffffffu) {
    // A function-local static is initialised once, thread-safely, on first
    // use: PNGs may be written from several threads at once.
    static const CrcTable table;
    for (size_t i = 0; i < n; ++i) c = table.v[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<unsigned char> &v, uint32_t x) {
    v.push_back(static_cast<unsigned char>(x >> 24));
    v.push_back(static_cast<unsigned char>(x >> 16));
    v.push_back(static_cast<unsigned char>(x >> 8));
    v.push_back(static_cast<unsigned char>(x));
}

void chunk(std::vector<unsigned char> &png, const char *type,
           const std::vector<unsigned char> &data) {
    put_be32(png, static_cast<uint32_t>(data.size()));
    const size_t start = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    put_be32(png, crc32(png.data() + start, png.size() - start) ^ 0xffffffffu);
}

}  // namespace

extern "C" {

const char *tvs_io_last_error(void) { return g_io_error.c_str(); }

tvs_status tvs_bt_read_header(const char *path, tvs_bt_header *out) {
    return read_bt(path, out, nullptr);
}

tvs_status tvs_bt_read_f32(const char *path, float *out, size_t count) {
    tvs_bt_header h;
    std::FILE *f = nullptr;
    tvs_status st = read_bt(path, &h, &f);
    if (st != TVS_OK) return st;
    const size_t n = static_cast<size_t>(h.width) * h.height;
    if (!out || count != n) {
        std::fclose(f);
        return io_fail(nullptr, TVS_ERR_SIZE, "buffer length != width * height");
    }
    std::fseek(f, static_cast<long>(kBtHeaderSize), SEEK_SET);   // input.rs:133
    const size_t elem = h.is_float ? 4 : 2;
    std::vector<unsigned char> buf(n * elem);
    const size_t got = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    if (got != buf.size()) return io_fail(nullptr, TVS_ERR_IO, "truncated .bt data");
    for (size_t i = 0; i < n; ++i)                                // main.rs:51-56
        out[i] = h.is_float ? le<float>(&buf[4 * i])
                            : static_cast<float>(le<int16_t>(&buf[2 * i]));
    return TVS_OK;
}

tvs_status tvs_bt_read_i16(const char *path, int16_t *out, size_t count) {
    tvs_bt_header h;
    std::FILE *f = nullptr;
    tvs_status st = read_bt(path, &h, &f);
    if (st != TVS_OK) return st;
    if (h.is_float) {
        std::fclose(f);
        return io_fail(nullptr, TVS_ERR_STATE, "file holds float samples; use tvs_bt_read_f32");
    }
    const size_t n = static_cast<size_t>(h.width) * h.height;
    if (!out || count != n) {
        std::fclose(f);
        return io_fail(nullptr, TVS_ERR_SIZE, "buffer length != width * height");
    }
    std::fseek(f, static_cast<long>(kBtHeaderSize), SEEK_SET);
    std::vector<unsigned char> buf(n * 2);
    const size_t got = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    if (got != buf.size()) return io_fail(nullptr, TVS_ERR_IO, "truncated .bt data");
    for (size_t i = 0; i < n; ++i) out[i] = le<int16_t>(&buf[2 * i]);
    return TVS_OK;
}

tvs_status tvs_bt_write(const char *path, const tvs_bt_header *h, const void *data) {
    if (!path || !h || !data) return io_fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    unsigned char hd[kBtHeaderSize];
    std::memset(hd, 0, sizeof hd);
    std::memcpy(hd, "binterr1.3", 10);
    put_le<uint32_t>(hd + 10, h->width);
    put_le<uint32_t>(hd + 14, h->height);
    put_le<uint16_t>(hd + 18, h->is_float ? 4 : 2);
    put_le<uint16_t>(hd + 20, h->is_float ? 1 : 0);
    put_le<uint16_t>(hd + 22, h->horizontal_units);
    put_le<uint16_t>(hd + 24, h->utm_zone);
    put_le<uint16_t>(hd + 26, h->datum);
    put_le<double>(hd + 28, h->left);
    put_le<double>(hd + 36, h->right);
    put_le<double>(hd + 44, h->bottom);
    put_le<double>(hd + 52, h->top);
    put_le<uint16_t>(hd + 60, h->projection_source);
    put_le<float>(hd + 62, h->vertical_scale);
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(nullptr, TVS_ERR_IO, std::string("cannot create ") + path);
    const size_t n = static_cast<size_t>(h->width) * h->height * (h->is_float ? 4 : 2);
    const bool ok = std::fwrite(hd, 1, sizeof hd, f) == sizeof hd &&
                    std::fwrite(data, 1, n, f) == n;
    std::fclose(f);
    return ok ? TVS_OK : io_fail(nullptr, TVS_ERR_IO, "short write");
}

// input.rs:178-185: Haversine distance between (top, right) and
// (top, left) read as (lon, lat) points, divided by the width.  Same
// longitude, so it is R * |right - left| in radians through the haversine
// formula, R the mean Earth radius 6371008.8 m.
double tvs_bt_scale(const tvs_bt_header *h) {
    if (!h || h->width == 0) return 0.0;
    const double R = 6371008.8;
    const double lat1 = h->right, lat2 = h->left, lon1 = h->top, lon2 = h->top;
    const double t1 = lat1 * (M_PI / 180.0), t2 = lat2 * (M_PI / 180.0);
    const double dt = (lat2 - lat1) * (M_PI / 180.0);
    const double dl = (lon2 - lon1) * (M_PI / 180.0);
    const double s1 = std::sin(dt / 2.0), s2 = std::sin(dl / 2.0);
    const double a = s1 * s1 + std::cos(t1) * std::cos(t2) * s2 * s2;
    const double c = 2.0 * std::asin(std::sqrt(a));
    return R * c / static_cast<double>(h->width);
}

// output/png.rs:5-37: min/max over the data, 8-bit
// `((v - min) / range * 255).clamp(0, 255) as u8`, range 1 when flat.
tvs_status tvs_write_heatmap_png(const char *path, const float *data, uint32_t width,
                                 uint32_t height) {
    tvs::TraceRange trace_range("tvs_write_heatmap_png");
    if (!path || (!data && width * height)) return io_fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n = static_cast<size_t>(width) * height;
    float mn = INFINITY, mx = -INFINITY;
    for (size_t i = 0; i < n; ++i) {
        mn = std::fmin(mn, data[i]);
        mx = std::fmax(mx, data[i]);
    }
    const float range = (mx - mn == 0.0f) ? 1.0f : mx - mn;
    std::vector<unsigned char> raw;
    raw.reserve(n + height);
    for (uint32_t r = 0; r < height; ++r) {
        raw.push_back(0);  // filter: none
        for (uint32_t x = 0; x < width; ++x) {
            // png.rs:16-25: `((value - min) / range * 255.0).clamp(0.0, 255.0)
            // as u8`.  A NaN stays NaN through clamp, and Rust's float-to-int
            // `as` maps NaN to 0: test it before clamping (fmin/fmax would
            // turn it into 255).
            float v = (data[static_cast<size_t>(r) * width + x] - mn) / range * 255.0f;
            if (v != v) v = 0.0f;
            v = std::fmax(0.0f, std::fmin(255.0f, v));
            raw.push_back(static_cast<unsigned char>(v));
        }
    }
    // zlib stream of stored blocks.
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(static_cast<unsigned char>(len));
        z.push_back(static_cast<unsigned char>(len >> 8));
        z.push_back(static_cast<unsigned char>(~len));
        z.push_back(static_cast<unsigned char>((~len) >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char ch : raw) {
        a = (a + ch) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 0, 0, 0, 0});  // 8-bit greyscale
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(nullptr, TVS_ERR_IO, std::string("cannot create ") + path);
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    std::fclose(f);
    return ok ? TVS_OK : io_fail(nullptr, TVS_ERR_IO, "short write");
}

// cache.rs:63-76: raw little-endian i32 deltas and f32 distances of one
// sector.  The reference keys the files on (width, angle) only; a file whose
// length is not band_size - 1 values belongs to another line of sight and
// is rejected on load rather than used (cache.rs:23-31 keeps no L).
tvs_status tvs_band_cache_save(tvs_ctx *c, const char *state_dir, uint32_t sector) {
    tvs::TraceRange trace_range("tvs_band_cache_save");
    if (!c || !state_dir) return io_fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n = c->geo.band_deltas_size;
    std::vector<int32_t> d(n ? n : 1);
    std::vector<float> x(n ? n : 1);
    tvs_status st = tvs_get_band(c, sector, d.data(), x.data(), n);
    if (st != TVS_OK) return st;
    const std::string base = std::string(state_dir) + "/total-viewsheds/" +
                             std::to_string(c->geo.dem_width);
    if (mkdirs(base + "/band_deltas") || mkdirs(base + "/distances"))
        return io_fail(c, TVS_ERR_IO, "cannot create " + base);
    const std::string pd = cache_file(state_dir, c->geo.dem_width, "band_deltas", sector);
    const std::string px = cache_file(state_dir, c->geo.dem_width, "distances", sector);
    for (int k = 0; k < 2; ++k) {
        const std::string &p = k ? px : pd;
        std::FILE *f = std::fopen(p.c_str(), "wb");
        if (!f) return io_fail(c, TVS_ERR_IO, "cannot create " + p);
        const size_t wrote = k ? std::fwrite(x.data(), 4, n, f) : std::fwrite(d.data(), 4, n, f);
        std::fclose(f);
        if (wrote != n) return io_fail(c, TVS_ERR_IO, "short write " + p);
    }
    return TVS_OK;
}

tvs_status tvs_band_cache_load(tvs_ctx *c, const char *state_dir, uint32_t sector) {
    tvs::TraceRange trace_range("tvs_band_cache_load");
    if (!c || !state_dir) return io_fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n = c->geo.band_deltas_size;
    std::vector<int32_t> d(n ? n : 1);
    std::vector<float> x(n ? n : 1);
    const std::string pd = cache_file(state_dir, c->geo.dem_width, "band_deltas", sector);
    const std::string px = cache_file(state_dir, c->geo.dem_width, "distances", sector);
    for (int k = 0; k < 2; ++k) {
        const std::string &p = k ? px : pd;
        std::FILE *f = std::fopen(p.c_str(), "rb");
        if (!f) return io_fail(c, TVS_ERR_IO, "no cache file " + p);
        std::fseek(f, 0, SEEK_END);
        const long len = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (len != static_cast<long>(4 * n)) {
            std::fclose(f);
            return io_fail(c, TVS_ERR_SIZE, p + " holds " + std::to_string(len) +
                                                " bytes, expected " + std::to_string(4 * n));
        }
        const size_t got = k ? std::fread(x.data(), 4, n, f) : std::fread(d.data(), 4, n, f);
        std::fclose(f);
        if (got != n) return io_fail(c, TVS_ERR_IO, "short read " + p);
    }
    return tvs_set_band(c, sector, d.data(), x.data(), n);
}

}  // extern "C"
