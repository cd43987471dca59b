// The likelihood context (struct lensed) and its device runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/lensed.h"
#include "../kernels/desc.h"
#include "prior_impl.h"
#include "program.h"

namespace lsd {

void set_error(const std::string& msg);

// one parameter of a system, with its own copy of the prior
struct Param {
    std::string id;
    std::string label;
    int type = 0;
    double lower = 0, upper = 0;
    bool bounded = false;
    prior pri_copy{prior::DELTA, 0, 0};  // copy of the caller's prior
    bool has_prior = false;
    bool ipp = false, derived = false, wrap = false;
};

struct System {
    size_t width = 0, height = 0, npix = 0;
    size_t psfw = 0, psfh = 0;
    double pcs[4] = {1, 1, 1, 1};
    std::string rule;
    std::vector<float> quad;    // float4 nodes (dx, dy, w, e), scaled
    int nq = 0;
    float qrad = 0;             // largest node distance from the pixel centre
    std::vector<ModelObject> objs;
    std::vector<Param> pars;    // object order
    std::vector<size_t> pmap;   // sampler order -> object order
    // the batched path's prior transform when every prior is uniform or a
    // delta (xfast): value = xa + u * xs (xs = 0 for a delta), bounds xlo..xhi
    // (infinite when unbounded), all in sampler order
    bool xfast = false;
    std::vector<double> xa, xs, xlo, xhi;
    std::vector<char> xdelta;
    size_t ndims = 0;
    int model = 0;
    size_t dwords = 0;
    // offsets of this system's constant data on the device
    long long image_off = 0, psf_off = 0;
    long long dark_off = 0;  // doubles into the context's dark partials (lensed_dark_partials)
    int quad_off = 0;
    int ntile_r = 0;
    int sep_rank = 0;             // separable PSF rank, 0 = direct convolution
    long long sep_off = 0;        // offset of the PSF factors in the psf buffer
    long long psft_off = 0;       // offset of the transposed PSF taps
    // EPL angular table: index of the EPL object and of its q, t parameters
    int epl_obj = -1, epl_q = -1, epl_t = -1;
    bool table = false;
    bool cull = false;          // pixels culled before the render (model_culls)
    std::vector<float> h_image, h_weight;  // host copies for the output layers
    // dumper results (object order)
    std::vector<double> mean, sigma, ml, map;
    double logev = 0, logev_err = 0, logev_ins = 0, max_loglike = 0;
    // the candidates of the last lensed_loglike_batch() call whose EPL table
    // the emulator built (1) rather than the exact series (0); empty when
    // no table came from the network
    std::vector<char> emu_used;
};

// launch layout for a given number of candidates per system
struct Layout {
    std::vector<lensed_desc> desc;  // systems with ncand > 0
    std::vector<int> sys_of_desc;
    int n_set = 0, n_render = 0, n_cull = 0, n_conv = 0, n_final = 0, n_table = 0;
    int n_nodes = 0, n_tcands = 0;  // emulator: table nodes and candidates with a table
    size_t lds_conv = 0, lds_render = 0;
    bool conv_small = false;        // short tiles: launch too small to fill the GPU
    bool any_cull = false;          // some system culls pixels before the render (lensed_cull)
    int grid_conv = 0, grid_render = 0;  // resident blocks for this launch's LDS use
    int grid_cull = 0;
    int list_max = 0;               // the render walks the cull's list below this many lit quads
    int conv_rows = 8;              // rows per thread of the convolution tile (2, 8, 16)
    int conv_order = 0;             // convolution item order (kernel argument, 0 or 1)
    lensed_tuning tune{};           // the resolved knobs (what lensed_get_tuning reports)
    lensed_desc* d_desc = nullptr;
    hipGraphExec_t graph = nullptr;
    int id = 0;                     // tells the launches of this layout apart on the device (owner)
    bool emulate = true;            // EPL tables from the emulator where it is certified (else exact)
    bool tuned = false;             // the autotuner has set this layout's shapes
    unsigned long long last_use = 0;  // context's launch clock at its last use (eviction order)
};

// device buffers of one launch: the context's own, or carved from the static
// scratch by the entry point that allocates nothing
struct Buffers {
    lensed_desc* desc = nullptr;
    float* params = nullptr;
    unsigned* blocks = nullptr;
    float* model = nullptr;
    double* partials = nullptr;
    double* logl = nullptr;
    void* tables = nullptr;
    void* npre = nullptr;
    int* trace = nullptr;
    int* live = nullptr;
    int* owner = nullptr;
    int* written = nullptr;
    bool emulate = true;  // emulated EPL tables where the context has the emulator
    bool record = true;   // the context's statistics describe this launch (lensed_render_stats)
};

// cached launch layouts per context; beyond this many the least recently
// used one is freed (its descriptors and graph)
constexpr size_t kMaxLayouts = 32;

}  // namespace lsd

struct lensed {
    int device = 0;
    int flags = 0;
    size_t max_batch = 1;
    std::vector<lsd::System> sys;
    hipModule_t module = nullptr;
    hipFunction_t f_set = nullptr, f_render = nullptr, f_conv = nullptr, f_conv_small = nullptr, f_final = nullptr,
                  f_tables = nullptr, f_cull = nullptr, f_conv_tall = nullptr, f_conv_sweep = nullptr,
                  f_conv_sweep_out = nullptr, f_set_tables = nullptr, f_dark = nullptr;
    lensed_tuning tune_req{};    // requested launch shapes (zero fields: defaults)
    int cu_count = 1;
    hipStream_t stream = nullptr;
    int grid_render = 0, grid_conv = 0, grid_conv_small = 0;
    // device buffers
    float* d_image = nullptr;
    float* d_weight = nullptr;
    float* d_psf = nullptr;
    float* d_quad = nullptr;
    float* d_params = nullptr;
    unsigned* d_blocks = nullptr;
    float* d_model = nullptr;
    float* d_err = nullptr;     // error plane, one candidate per system (dumper)
    float* d_conv = nullptr;    // convolved image, one candidate (dumper)
    float* d_chi2 = nullptr;    // per-pixel chi^2, one candidate (dumper)
    double* d_partials = nullptr;
    double* d_logl = nullptr;
    void* d_tables = nullptr;   // EPL angular tables, float4 [candidates][LENSED_EPL_TABLE_F4]
    // EPL emulator: weights (fp32 and packed bf16), table nodes, per-candidate
    // choice (1 = exact series), the certified (q, t) cells, and the choice
    // of the last launch on the host
    void* emu_wf = nullptr;
    void* emu_wp = nullptr;
    void* d_nodes = nullptr;
    int* d_flags = nullptr;
    int* d_cells = nullptr;
    int* h_flags = nullptr;
    float emu_tol = 4e-6f;
    int emu_ncert = 0;            // certified cells (0: every table comes from the exact series)
    double emu_cert_err = 0;      // largest held-out error of the certification
    long emu_rechecked = 0, emu_changed = 0;  // exact re-evaluations of emulated candidates
    double emu_recheck_max = 0;   // largest |logL(emulated) - logL(exact)| among them
    const lsd::Layout* last_table_layout = nullptr;
    const lsd::Layout* last_layout = nullptr;  // of the last launch (lensed_render_stats)
    // pinned host staging
    float* h_params = nullptr;
    double* h_logl = nullptr;
    size_t cap_params = 0, cap_blocks = 0, cap_model = 0, cap_partials = 0, cap_out = 0, cap_tables = 0;
    std::map<std::vector<size_t>, lsd::Layout> layouts;
    std::string source;          // generated program (for dumps)
    std::string output;          // FITS file the dumper writes (empty: none)
    void* d_npre = nullptr;      // per candidate and node: EPL lens-frame pair of the node offset
    size_t cap_npre = 0;
    int* d_trace = nullptr;      // per render item and wave: trace its rays (lensed_cull)
    int* d_live = nullptr;       // count, then the render items lensed_cull kept
    int* d_owner = nullptr;      // [0] layout that last wrote the model images, [1] whether the
                                 // current launch's layout is that one (set by lensed_set_params)
    int* d_written = nullptr;    // per render square: its model values may be nonzero (lensed_cull)
    double* d_dark = nullptr;    // per system: the data's chi^2 block partials for a model of 0
    int next_layout_id = 0;
    unsigned long long layout_clock = 0;
    // the entry point that allocates nothing: this device's static scratch,
    // a launch layout per system built at creation (patched per call) and a
    // pinned staging slot for its descriptor
    void* scratch = nullptr;
    std::vector<lsd::Layout> noalloc;
    lensed_desc* h_desc = nullptr;
    size_t cap_trace = 0;
    size_t cap_cull = 0;   // cull work items (LENSED_CULL_GROUP render items each)
    // profiling: +params, -params, set_params, render, convolve, loglike, +loglike_mem, -loglike_mem
    hipEvent_t ev[10] = {};  // the eight stages' bounds, then the end of the cull pass
    double prof_ms[8] = {};
    double prof_cull_ms = 0;  // the cull pass's share of the render stage
    long prof_calls = 0;
};

namespace lsd {

// evaluate ncand[s] parameter vectors (object order, float) per system;
// phys[s] == nullptr: system s's parameters are already staged as floats in
// h_params at its descriptor's offset
int evaluate(::lensed* l, const std::vector<size_t>& ncand, const std::vector<const double*>& phys,
             const std::vector<double*>& lnew, bool emulate = true);

// exact logL for the emulated candidates of system s's last batch that lie
// above lmin (phys in sampler order, ll replaced in place); 0 on success
int recheck_emulated(::lensed* l, size_t s, const double* phys, int n, double lmin, double* ll);

}  // namespace lsd
