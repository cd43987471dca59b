// Autotuner of the launch shapes (lensed_tuning), see tune.cpp.
#pragma once

#include <functional>
#include <string>

#include "../../include/lensed.h"

namespace lsd {

// stage and chain times of one launch under given shapes (each the least of
// a few repetitions)
struct TuneTiming {
    double render_ms = 0;  // pixel cull + render
    double conv_ms = 0;    // convolution + chi^2
    double total_ms = 0;   // the whole chain, parameters in to logL out
};

// times the layout's launch chain under the given shapes; 0 on success
using TuneTimer = std::function<int(const lensed_tuning&, TuneTiming*)>;

// search the shapes for one layout (culls: the layout has a pixel-cull pass);
// *best receives the choice with default_ms / tuned_ms filled in
int autotune(const TuneTimer& timer, bool culls, lensed_tuning* best);

// the disk cache of choices, one small text file per key in `dir`
bool tune_cache_read(const std::string& dir, const std::string& key, lensed_tuning* t);
void tune_cache_write(const std::string& dir, const std::string& key, const lensed_tuning& t);

}  // namespace lsd
