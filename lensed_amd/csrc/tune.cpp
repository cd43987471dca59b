// Autotuner of the launch shapes (lensed_tuning): what the reference decides
// once from device queries -- the render's work-group size
// (src/lensed.c:917-948) and the convolution's, halved until the device
// accepts it (src/lensed.c:998-1027) -- is chosen here by timing the actual
// launch chain of a layout on the launch's own parameters.  No knob changes
// a result (the model planes and every logL are the same bits), so the
// search only compares times.
//
// Search, coordinate by coordinate from the defaults: the render stage
// (which includes the pixel cull) over the list/walk switch and the render
// and cull grid sizes; then the convolution stage over the tile height and
// the item order together, the column sweep, then the grid size (the
// sweep's grid is its item count, so the grid knob leaves it unchanged).  A stage keeps a setting
// only when it beats the stage's best time so far by more than the timing
// noise; at the end the chosen shapes must beat the defaults over the whole
// chain, timed alternately, or the defaults stay.  Choices are cached on
// disk per device, program and layout, next to the program cache.
#include "tune.h"

#include "../kernels/desc.h"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace lsd {

static uint64_t tune_hash(const std::string& s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

static std::string tune_path(const std::string& dir, const std::string& key) {
    char name[48];
    snprintf(name, sizeof name, "/tune_%016llx.txt", (unsigned long long)tune_hash(key));
    return dir + name;
}

bool tune_cache_read(const std::string& dir, const std::string& key, lensed_tuning* t) {
    if (dir.empty() || getenv("LENSED_NO_CACHE")) return false;
    std::ifstream f(tune_path(dir, key));
    std::string k;
    if (!f || !std::getline(f, k) || k != key) return false;
    lensed_tuning r{};
    if (!(f >> r.list_fraction >> r.render_grid >> r.cull_grid >> r.conv_grid >> r.conv_rows >> r.conv_order >>
          r.default_ms >> r.tuned_ms))
        return false;
    *t = r;
    return true;
}

void tune_cache_write(const std::string& dir, const std::string& key, const lensed_tuning& t) {
    if (dir.empty()) return;
    for (size_t p = 1; p <= dir.size(); ++p)
        if (p == dir.size() || dir[p] == '/') mkdir(dir.substr(0, p).c_str(), 0755);
    const std::string path = tune_path(dir, key);
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    std::ofstream f(tmp);
    if (!f) return;
    char line[256];
    snprintf(line, sizeof line, "%.17g %.17g %.17g %.17g %d %d %.17g %.17g\n", t.list_fraction, t.render_grid,
             t.cull_grid, t.conv_grid, t.conv_rows, t.conv_order, t.default_ms, t.tuned_ms);
    f << key << "\n" << line;
    f.close();
    if (f) rename(tmp.c_str(), path.c_str());
}

int autotune(const TuneTimer& timer, bool culls, lensed_tuning* best) {
    // the timer's own repetitions make each figure a minimum; a setting must
    // win by 2 % to displace the current one
    const double margin = 0.98;
    lensed_tuning cur{};
    // the GPU's clock comes up over the first launches after a pause (C3's
    // render 2.3 -> 2.0 ms over about ten launches): the defaults run until
    // a timing no longer beats the one before by the margin (at most eight
    // rounds), so that the search compares settings at the same clock
    TuneTiming t0;
    double prev = 0;
    for (int i = 0; i < 8; ++i) {
        if (timer(cur, &t0)) return -1;
        if (prev > 0 && t0.total_ms > margin * prev) break;
        prev = t0.total_ms;
    }
    double render_best = t0.render_ms, conv_best = t0.conv_ms;
    auto try_render = [&](const lensed_tuning& cand) -> int {
        TuneTiming t;
        if (timer(cand, &t)) return -1;
        if (t.render_ms < margin * render_best) {
            render_best = t.render_ms;
            cur.list_fraction = cand.list_fraction;
            cur.render_grid = cand.render_grid;
            cur.cull_grid = cand.cull_grid;
        }
        return 0;
    };
    auto try_conv = [&](const lensed_tuning& cand) -> int {
        TuneTiming t;
        if (timer(cand, &t)) return -1;
        if (t.conv_ms < margin * conv_best) {
            conv_best = t.conv_ms;
            cur.conv_rows = cand.conv_rows;
            cur.conv_order = cand.conv_order;
            cur.conv_grid = cand.conv_grid;
        }
        return 0;
    };
    // render stage
    if (culls)
        for (double lf : {-1.0, 2.0}) {
            lensed_tuning c = cur;
            c.list_fraction = lf;
            if (try_render(c)) return -1;
        }
    for (double g : {0.5, 2.0, 4.0}) {
        lensed_tuning c = cur;
        c.render_grid = g;
        if (try_render(c)) return -1;
    }
    if (culls)
        for (double g : {1.0, 2.0, 8.0}) {  // (the default is 4)
            lensed_tuning c = cur;
            c.cull_grid = g;
            if (try_render(c)) return -1;
        }
    // convolution stage: tile height with item order, the column sweep
    // (layouts it does not serve build the default tile, which then times
    // as the default), then the grid
    for (int rows : {8, 16, 2})
        for (int order : {1, 2}) {
            lensed_tuning c = cur;
            c.conv_rows = rows;
            c.conv_order = order;
            if (try_conv(c)) return -1;
        }
    {
        lensed_tuning c = cur;
        c.conv_rows = LENSED_CONV_SWEEP;
        if (try_conv(c)) return -1;
    }
    for (double g : {0.5, 2.0, 4.0}) {
        lensed_tuning c = cur;
        c.conv_grid = g;
        if (try_conv(c)) return -1;
    }
    // the whole chain, chosen against the defaults, alternately
    const lensed_tuning def{};
    double td = 1e300, tc = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
        TuneTiming a, b;
        if (timer(def, &a) || timer(cur, &b)) return -1;
        td = std::min(td, a.total_ms);
        tc = std::min(tc, b.total_ms);
    }
    if (!(tc < td)) {
        cur = def;
        tc = td;
    }
    cur.default_ms = td;
    cur.tuned_ms = tc;
    *best = cur;
    return 0;
}

}  // namespace lsd
