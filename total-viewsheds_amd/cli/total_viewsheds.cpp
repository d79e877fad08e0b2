// total-viewsheds -- command line front end over libtvs_hip.so.
//
// Flags and defaults follow the reference CLI (config.rs:8-64, main.rs:21-71):
// --input, --max-line-of-sight, --rings-per-km, --observer-height,
// --compute, --output-dir, --scale.  `--compute hip` (the default here) is
// the batched device walk, `hip-plain` the one-sector-per-launch walk.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/tvs.h"

namespace {

struct Config {
    std::string input;
    long long max_line_of_sight = -1;   // metres; -1 = (w div 3) * scale
    float rings_per_km = 5.0f;
    float observer_height = 1.65f;
    std::string compute = "hip";
    std::string output_dir = "./";
    double scale = -1.0;
    std::string state_dir;              // band-data cache root
    bool use_cache = true;
    bool heatmap_each_sector = false;
    bool tune = false;                  // time the walk shapes before the run
    bool verbose = false;
};

void usage(std::FILE *f) {
    std::fprintf(f,
        "Generate _all_ the viewsheds for a given Digital Elevation Model, therefore the total "
        "viewsheds.\n\n"
        "Usage: total-viewsheds [OPTIONS] --input <Path to the DEM file>\n\n"
        "Options:\n"
        "      --input <path>                The input DEM file (.bt, Binary Terrain 1.3)\n"
        "      --max-line-of-sight <metres>  Maximum distance to search for visible points.\n"
        "                                    Defaults to one third of the DEM width\n"
        "      --rings-per-km <n>            Visible rings expected per km of band of sight; sizes\n"
        "                                    the ring records [default: 5]\n"
        "      --observer-height <metres>    Height of the observer [default: 1.65]\n"
        "      --compute <method>            hip (all sectors in one device launch) or hip-plain\n"
        "                                    (one sector per launch) [default: hip]\n"
        "      --output-dir <dir>            Where heatmap.png is written [default: ./]\n"
        "      --scale <metres>              Override the DEM point spacing from the file\n"
        "      --state-dir <dir>             Band-data cache root [default: $XDG_STATE_HOME or\n"
        "                                    ~/.local/state]\n"
        "      --no-cache                    Always compute band data (and, with --tune, the walk\n"
        "                                    shape) on the device\n"
        "      --heatmap-each-sector         Rewrite heatmap.png after every sector with the running\n"
        "                                    total, as the reference does\n"
        "      --tune                        Time the walk's launch shapes on this DEM first and\n"
        "                                    keep the fastest; the choice is cached per width, line\n"
        "                                    of sight and height type in the state directory\n"
        "  -v, --verbose                     Per-sector log lines\n"
        "  -h, --help                        Print help\n");
}

bool parse(int argc, char **argv, Config &c) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        std::string v;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
        }
        auto need = [&](const char *name) -> bool {
            if (!v.empty()) return true;
            if (i + 1 >= argc) {
                std::fprintf(stderr, "error: %s needs a value\n", name);
                return false;
            }
            v = argv[++i];
            return true;
        };
        if (a == "-h" || a == "--help") { usage(stdout); std::exit(0); }
        else if (a == "-v" || a == "--verbose") c.verbose = true;
        else if (a == "--no-cache") c.use_cache = false;
        else if (a == "--heatmap-each-sector") c.heatmap_each_sector = true;
        else if (a == "--tune") c.tune = true;
        else if (a == "--input") { if (!need("--input")) return false; c.input = v; }
        else if (a == "--max-line-of-sight") { if (!need(a.c_str())) return false; c.max_line_of_sight = std::atoll(v.c_str()); }
        else if (a == "--rings-per-km") { if (!need(a.c_str())) return false; c.rings_per_km = std::strtof(v.c_str(), nullptr); }
        else if (a == "--observer-height") { if (!need(a.c_str())) return false; c.observer_height = std::strtof(v.c_str(), nullptr); }
        else if (a == "--compute") { if (!need(a.c_str())) return false; c.compute = v; }
        else if (a == "--output-dir") { if (!need(a.c_str())) return false; c.output_dir = v; }
        else if (a == "--scale") { if (!need(a.c_str())) return false; c.scale = std::strtod(v.c_str(), nullptr); }
        else if (a == "--state-dir") { if (!need(a.c_str())) return false; c.state_dir = v; }
        else {
            std::fprintf(stderr, "error: unexpected argument '%s'\n", argv[i]);
            return false;
        }
    }
    if (c.input.empty()) {
        std::fprintf(stderr, "error: the following required arguments were not provided:\n  --input <Path to the DEM file>\n");
        return false;
    }
    if (c.compute != "hip" && c.compute != "hip-plain") {
        std::fprintf(stderr, "error: --compute %s is not provided by this engine (use hip or hip-plain)\n",
                     c.compute.c_str());
        return false;
    }
    return true;
}

int die(tvs_ctx *ctx, tvs_status st, const char *what) {
    std::fprintf(stderr, "Error: %s: %s (%s)\n", what,
                 ctx ? tvs_last_error(ctx) : tvs_last_error(nullptr), tvs_status_string(st));
    if (ctx) tvs_destroy(ctx);
    return 1;
}

// mkdir -p.
bool make_dirs(const std::string &path) {
    for (size_t i = 1; i <= path.size(); ++i) {
        if (i < path.size() && path[i] != '/') continue;
        const std::string part = path.substr(0, i);
        if (::mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    return true;
}

// The walk shape chosen by --tune for one geometry and height type, under
// <state>/total-viewsheds/<width>/walk_shape/<L>pts_<i16|f32>.txt (L the
// line of sight in points, which with the width fixes the walk's work) as
// the seven integers of tvs_walk_shape (block, unroll, xcd_contiguous,
// pipeline, persistent, both_directions, two_points).
std::string shape_path(const std::string &state, uint32_t width, uint32_t los_points,
                       bool is_float) {
    return state + "/total-viewsheds/" + std::to_string(width) + "/walk_shape/" +
           std::to_string(los_points) + (is_float ? "pts_f32" : "pts_i16") + ".txt";
}

bool load_shape(const std::string &path, tvs_walk_shape *out) {
    std::FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    tvs_walk_shape w{};
    const int n = std::fscanf(f, "%u %u %u %u %u %u %u", &w.block, &w.unroll, &w.xcd_contiguous,
                              &w.pipeline, &w.persistent, &w.both_directions, &w.two_points);
    std::fclose(f);
    if (n != 7) return false;
    *out = w;
    return true;
}

bool save_shape(const std::string &path, const tvs_walk_shape &w) {
    if (!make_dirs(path.substr(0, path.rfind('/')))) return false;
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "%u %u %u %u %u %u %u\n", w.block, w.unroll, w.xcd_contiguous, w.pipeline,
                 w.persistent, w.both_directions, w.two_points);
    return std::fclose(f) == 0;
}

std::string default_state_dir() {
    if (const char *x = std::getenv("XDG_STATE_HOME"); x && *x) return x;
    if (const char *h = std::getenv("HOME"); h && *h) return std::string(h) + "/.local/state";
    return "";
}

}  // namespace

int main(int argc, char **argv) {
    Config cfg;
    if (!parse(argc, argv, cfg)) {
        usage(stderr);
        return 2;
    }
    tvs_bt_header hdr;
    tvs_status st = tvs_bt_read_header(cfg.input.c_str(), &hdr);
    if (st != TVS_OK) {
        std::fprintf(stderr, "Error: %s\n", tvs_io_last_error());
        return 1;
    }
    const double scale = cfg.scale > 0 ? cfg.scale : tvs_bt_scale(&hdr);   // main.rs:28
    const uint32_t los = cfg.max_line_of_sight >= 0
                             ? static_cast<uint32_t>(cfg.max_line_of_sight)
                             : tvs_default_max_line_of_sight(hdr.width, scale);  // main.rs:36-38
    std::fprintf(stderr, "Initialising: input %s, width %u, scale %g m, max line of sight %u m, "
                         "rings/km %g, observer %g m, compute %s\n",
                 cfg.input.c_str(), hdr.width, scale, los, cfg.rings_per_km,
                 cfg.observer_height, cfg.compute.c_str());
    const size_t n = static_cast<size_t>(hdr.width) * hdr.height;
    if (hdr.height < hdr.width) {
        std::fprintf(stderr, "Error: the DEM must hold width*width samples (height %u < width %u)\n",
                     hdr.height, hdr.width);
        return 1;
    }
    tvs_params p{};
    p.dem_width = hdr.width;
    p.scale = static_cast<float>(scale);
    p.max_line_of_sight = los;
    p.rings_per_km = cfg.rings_per_km;
    p.observer_height = cfg.observer_height;
    p.ring_capture_sectors = 0;
    tvs_ctx *ctx = nullptr;
    if ((st = tvs_create(&p, &ctx)) != TVS_OK) return die(nullptr, st, "DEM");
    const size_t ww = static_cast<size_t>(hdr.width) * hdr.width;
    if (hdr.is_float) {
        std::vector<float> e(n);
        if ((st = tvs_bt_read_f32(cfg.input.c_str(), e.data(), n)) != TVS_OK) {
            std::fprintf(stderr, "Error: %s\n", tvs_io_last_error());
            tvs_destroy(ctx);
            return 1;
        }
        st = tvs_upload_dem_f32(ctx, e.data(), ww);
    } else {
        std::vector<int16_t> e(n);
        if ((st = tvs_bt_read_i16(cfg.input.c_str(), e.data(), n)) != TVS_OK) {
            std::fprintf(stderr, "Error: %s\n", tvs_io_last_error());
            tvs_destroy(ctx);
            return 1;
        }
        st = tvs_upload_dem_i16(ctx, e.data(), ww);
    }
    if (st != TVS_OK) return die(ctx, st, "elevation upload");
    tvs_geometry g;
    tvs_get_geometry(ctx, &g);

    // compute.rs:118-161: cached band data where valid, device compute else.
    const std::string state = cfg.state_dir.empty() ? default_state_dir() : cfg.state_dir;
    const bool cache = cfg.use_cache && !state.empty();
    std::vector<uint32_t> missing;
    for (uint32_t s = 0; s < 180; ++s) {
        if (cache && tvs_band_cache_load(ctx, state.c_str(), s) == TVS_OK) continue;
        missing.push_back(s);
    }
    if (!missing.empty()) {
        std::fprintf(stderr, "Computing band data for %zu sectors on the device\n", missing.size());
        for (size_t i = 0; i < missing.size();) {
            size_t j = i;
            while (j + 1 < missing.size() && missing[j + 1] == missing[j] + 1) ++j;
            if ((st = tvs_compute_bands(ctx, missing[i], static_cast<uint32_t>(j - i + 1))) != TVS_OK)
                return die(ctx, st, "band data");
            i = j + 1;
        }
        if (cache)
            for (uint32_t s : missing)
                if (tvs_band_cache_save(ctx, state.c_str(), s) != TVS_OK)
                    std::fprintf(stderr, "warning: band cache not saved: %s\n", tvs_last_error(ctx));
    }

    const tvs_method method = cfg.compute == "hip" ? TVS_METHOD_BATCHED : TVS_METHOD_PLAIN;
    if (cfg.tune && method == TVS_METHOD_BATCHED) {
        // A cached choice for this geometry and height type, else the tuner
        // (it keeps the fastest shape as the context's; surfaces untouched).
        const std::string sp =
            state.empty() ? "" : shape_path(state, hdr.width, g.max_los_as_points, hdr.is_float);
        tvs_walk_shape best{};
        if (cfg.use_cache && !sp.empty() && load_shape(sp, &best) &&
            tvs_set_walk_shape(ctx, &best) == TVS_OK) {
            std::fprintf(stderr, "Walk shape from %s: %u x %u\n", sp.c_str(), best.block, best.unroll);
        } else {
            std::fprintf(stderr, "Tuning the walk's launch shape on this DEM\n");
            if ((st = tvs_autotune(ctx, 0, 180, 2, nullptr, 0, nullptr, &best)) != TVS_OK)
                return die(ctx, st, "tuning");
            std::fprintf(stderr, "Walk shape: %u x %u, %.2f ms per pass\n", best.block, best.unroll,
                         best.ms);
            if (cfg.use_cache && !sp.empty() && !save_shape(sp, best))
                std::fprintf(stderr, "warning: walk shape not saved to %s\n", sp.c_str());
        }
    }
    std::vector<float> surfaces(g.computable_points_count);
    const std::string png = cfg.output_dir + "/heatmap.png";
    ::mkdir(cfg.output_dir.c_str(), 0755);
    if (cfg.heatmap_each_sector) {
        // compute.rs:106-112 + 206-214: after every sector, add the
        // cumulative surfaces into a running total and re-render it.
        std::vector<float> running(g.computable_points_count, 0.0f);
        for (uint32_t s = 0; s < 180; ++s) {
            if (cfg.verbose) std::fprintf(stderr, "Running kernel for %u°\n", s);
            if ((st = tvs_run_sectors(ctx, s, 1, method)) != TVS_OK) return die(ctx, st, "walk");
            if ((st = tvs_read_surfaces(ctx, surfaces.data(), surfaces.size())) != TVS_OK)
                return die(ctx, st, "readback");
            for (size_t i = 0; i < running.size(); ++i) running[i] += surfaces[i];
            if (tvs_write_heatmap_png(png.c_str(), running.data(), g.tvs_width, g.tvs_width) != TVS_OK)
                std::fprintf(stderr, "warning: %s\n", tvs_io_last_error());
        }
    } else {
        if ((st = tvs_run_sectors(ctx, 0, 180, method)) != TVS_OK) return die(ctx, st, "walk");
        if ((st = tvs_read_surfaces(ctx, surfaces.data(), surfaces.size())) != TVS_OK)
            return die(ctx, st, "readback");
        if (tvs_write_heatmap_png(png.c_str(), surfaces.data(), g.tvs_width, g.tvs_width) != TVS_OK) {
            std::fprintf(stderr, "Error: %s\n", tvs_io_last_error());
            tvs_destroy(ctx);
            return 1;
        }
    }
    tvs_timing t;
    tvs_get_timing(ctx, &t);
    std::fprintf(stderr, "Done: %u x %u total viewsheds, %.3e point-steps, walk %.2f ms; heatmap %s\n",
                 g.tvs_width, g.tvs_width, static_cast<double>(180.0 * 2 * g.computable_points_count *
                                                              g.band_deltas_size),
                 cfg.heatmap_each_sector ? -1.0f : t.walk_ms, png.c_str());
    tvs_destroy(ctx);
    return 0;
}
