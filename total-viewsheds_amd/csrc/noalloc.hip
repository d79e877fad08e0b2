// noalloc.hip -- tvs_run_sectors_noalloc: the batched walk of
// tvs_run_sectors without any device allocation, for callers that cannot
// allocate (serving loops, code that must not call hipMalloc).
//
// The batched walk writes every band's surface into a per-(sector,
// direction, point) buffer that the reduction then adds per point in sector
// order.  Here that buffer is `scratch`, a file-scope device array of
// kScratchBytes: sectors are walked in chunks whose band surfaces fit in it
// (one walk launch and one reduction launch per chunk, in sector order, so
// the f32 sums are those of tvs_run_sectors bit for bit).  Each device that
// runs the library holds its own copy of the array until the process ends.
// Its device address comes from hipGetSymbolAddress, once per device, on the
// first call there; the array's name itself is only a host-side stand-in.
//
// Calls on one device share that device's array.  A call takes the
// device's lock and returns only once its launches have finished, so calls
// from several host threads (on any contexts of the device) run one after
// another instead of racing on the array.  Every slot the reduction reads is
// written by the walk of the same chunk first, so nothing needs clearing
// between calls.
#include <algorithm>
#include <mutex>
#include <string>

#include "engine.hpp"

static constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB

static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_device_lock[kMaxDevices];
float *g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device

tvs_status nfail(tvs_ctx *c, tvs_status st, const std::string &msg) {
    c->err = msg;
    return st;
}

tvs_status nhip(tvs_ctx *c, hipError_t e, const char *what) {
    return nfail(c, e == hipErrorOutOfMemory ? TVS_ERR_OUT_OF_MEMORY : TVS_ERR_HIP,
                 std::string(what) + ": " + hipGetErrorString(e));
}

#define TVS_NHIP(c, expr, what)                              \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return nhip((c), _e, (what));  \
    } while (0)

// With the device's lock held and the device current.
tvs_status scratch_address(tvs_ctx *c, float **out) {
    float *&p = g_scratch[c->device];
    if (!p) {
        void *addr = nullptr;
        TVS_NHIP(c, hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)), "scratch address");
        if (!addr) return nfail(c, TVS_ERR_HIP, "scratch address: null");
        p = static_cast<float *>(addr);
    }
    *out = p;
    return TVS_OK;
}

// Runs with the lock held and the context's device current.
tvs_status run_chunks(tvs_ctx *c, uint32_t first, uint32_t count) {
    const uint32_t N = c->geo.computable_points_count;
    float *buf = nullptr;
    tvs_status st = scratch_address(c, &buf);
    if (st != TVS_OK) return st;
    // The shape of tvs_run_sectors, in a block order that needs no device
    // table (the directional order is built in device memory on demand; the
    // XCD-contiguous order runs the same kernel and gives the same results).
    tvs::WalkConfig cfg = tvs::effective_walk_config(c);
    cfg.directional = false;
    const uint32_t per_chunk =
        static_cast<uint32_t>(kScratchBytes / (2ull * N * sizeof(float)));
    if ((st = tvs::ensure_short_div(c, c->tab, first, count)) != TVS_OK) return st;
    // The run's rings and overflow count: this run records none.
    TVS_NHIP(c, tvs::launch_fill_u32(c->stream, reinterpret_cast<uint32_t *>(c->overflow), 2, 0u),
             "overflow reset");
    c->ring_valid = 0;
    c->timing = tvs_timing{0.0f, 0.0f, c->timing.bands_ms, 0u,
                           static_cast<uint64_t>(count) * 2ull * N * c->geo.band_deltas_size};
    // One descriptor per chunk, kept until the stream has drained (also when
    // a launch fails part way: the copies already queued read them).
    tvs::TileDesc descs[tvs::kSectorSteps];
    uint32_t begins[tvs::kSectorSteps][4];
    uint32_t launches = 0;
    c->desc_cached = false;   // the chunks' descriptors go through d_tile / d_begin
    auto enqueue = [&]() -> tvs_status {
        TVS_NHIP(c, hipEventRecord(c->ev[0], c->stream), "event");
        for (uint32_t s = first, j = 0; s < first + count; s += per_chunk, ++j) {
            const uint32_t n = std::min(per_chunk, first + count - s);
            tvs::TileDesc &d = descs[j];
            const tvs_status ps = tvs::prepare_tile(c, s, n, 0, cfg, &d, false);
            if (ps != TVS_OK) return ps;
            d.partial = buf;
            const uint32_t n_items = d.blocks_per_pass * 2u * n;
            begins[j][0] = 0u;
            begins[j][1] = n_items;
            begins[j][2] = 0u;
            begins[j][3] = N;
            TVS_NHIP(c, hipMemcpyAsync(c->d_tile, &d, sizeof d, hipMemcpyHostToDevice, c->stream),
                     "tile descriptor");
            TVS_NHIP(c, hipMemcpyAsync(c->d_begin, begins[j], sizeof begins[j],
                                       hipMemcpyHostToDevice, c->stream),
                     "tile descriptor");
            TVS_NHIP(c, tvs::launch_walk_tiles(c->stream, c->d_tile, c->d_begin, 1, n_items,
                                               c->has_i16, false, d.fast == 0u, cfg,
                                               d.blocks_per_pass),
                     "batched walk");
            TVS_NHIP(c, tvs::launch_reduce_tiles(c->stream, c->d_tile, c->d_begin + 2, 1, N),
                     "sector reduction");
            ++launches;
        }
        TVS_NHIP(c, hipEventRecord(c->ev[1], c->stream), "event");
        TVS_NHIP(c, hipEventRecord(c->ev[2], c->stream), "event");
        return TVS_OK;
    };
    st = enqueue();
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (st != TVS_OK) return st;
    if (e != hipSuccess) return nhip(c, e, "noalloc run");
    c->timing.walk_launches = launches;
    return TVS_OK;
}

}  // namespace

extern "C" {

tvs_status tvs_run_sectors_noalloc(tvs_ctx *c, uint32_t first, uint32_t count) {
    tvs::TraceRange trace_range("tvs_run_sectors_noalloc");
    if (!c) return TVS_ERR_INVALID_ARGUMENT;
    if (first >= tvs::kSectorSteps || count > tvs::kSectorSteps - first)
        return nfail(c, TVS_ERR_INVALID_ARGUMENT, "sector range outside 0..180");
    const uint32_t N = c->geo.computable_points_count;
    const uint64_t per_sector = 2ull * N * sizeof(float);
    if (per_sector > kScratchBytes)
        return nfail(c, TVS_ERR_SIZE,
                     "one sector's band surfaces take " + std::to_string(per_sector) +
                         " bytes (2 x " + std::to_string(N) +
                         " points x 4), more than the " + std::to_string(kScratchBytes) +
                         "-byte static scratch of tvs_run_sectors_noalloc: use "
                         "tvs_run_sectors");
    if (c->params.ring_capture_sectors)
        return nfail(c, TVS_ERR_INVALID_ARGUMENT,
                     "ring capture needs a ring buffer: use tvs_run_sectors");
    if (!c->has_f32) return nfail(c, TVS_ERR_STATE, "no elevations uploaded");
    for (uint32_t s = first; s < first + count; ++s)
        if (!c->band_ready[s])
            return nfail(c, TVS_ERR_STATE, "sector " + std::to_string(s) + " has no band data");
    if (c->device < 0 || c->device >= kMaxDevices)
        return nfail(c, TVS_ERR_NO_DEVICE, "device index beyond the library's table");
    if (count == 0 || N == 0) return TVS_OK;
    std::lock_guard<std::mutex> lock(g_device_lock[c->device]);
    int prev = -1;
    TVS_NHIP(c, hipGetDevice(&prev), "current device");
    if (prev != c->device) TVS_NHIP(c, hipSetDevice(c->device), "set device");
    const tvs_status st = run_chunks(c, first, count);
    if (prev != c->device) (void)hipSetDevice(prev);
    return st;
}

}  // extern "C"
