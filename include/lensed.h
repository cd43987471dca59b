/*
 * lensed-amd: C ABI of the MI355X-native Lensed likelihood.
 *
 * Every entry point below replaces one of the reference's host interfaces;
 * the comment on each cites the reference declaration (file:line) whose
 * signature and meaning it keeps.  Plain C types only: no HIP or torch
 * types cross this boundary.
 */
#ifndef LENSED_AMD_H
#define LENSED_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* object types, src/input.h:4-9 */
enum { OBJ_LENS = 'L', OBJ_SOURCE = 'S', OBJ_FOREGROUND = 'F' };

/* parameter types, src/input.h:12-21 */
enum {
    PAR_PARAMETER = 0,
    PAR_POSITION_X,
    PAR_POSITION_Y,
    PAR_RADIUS,
    PAR_MAGNITUDE,
    PAR_AXIS_RATIO,
    PAR_POS_ANGLE
};

/* opaque prior, src/input.h:73 */
typedef struct prior prior;

/* parameter of an object, src/input.h:76-109 */
typedef struct {
    const char* name;
    const char* id;
    int type;
    double lower;
    double upper;
    int bounded;
    prior* pri;
    int wrap;
    int ipp;
    int derived;
    int defval;
    const char* label;
} param;

/* object on the line of sight, src/input.h:112-129 (size in 4-byte words) */
typedef struct {
    int type;
    size_t size;
    const char* id;
    const char* name;
    size_t npars;
    param* pars;
} object;

/* ------------------------------------------------------ program generator */

/* src/kernel.h:5 -- sources of the per-object metadata program */
void object_program(const char* name, size_t* nkernels, const char*** kernels);

/* src/kernel.h:8 -- sources of the main program: prelude, objects, the
 * generated compute()/set_params() and the device kernels (HIP C++ here) */
void main_program(size_t nobjs, object objs[], size_t* nkernels, const char*** kernels);

/* src/kernel.h:11-14 -- build options; the -D names are the reference's */
const char* kernel_options(size_t width, size_t height, int psf, size_t psfw,
                           size_t psfh, size_t nq, const char* flags[]);

/* src/kernel.h:17 -- prefix + name, malloc'd */
char* kernel_name(const char* prefix, const char* name);

/* The model part of main_program() alone -- prelude, translated objects
 * and the generated compute_0()/set_params_0() -- which also compiles with a
 * host C++ compiler.  No counterpart in the reference; used to check the
 * plugin translation and the generated ray-shoot on the CPU.  malloc'd,
 * NULL on failure. */
char* lensed_model_source(size_t nobjs, const object* objs);
/* The same with the first EPL lens tabulated, as the table and emulator
 * modes generate it (epl_pre_0() and the table form of compute_0()).  No
 * counterpart in the reference; for the same CPU checks. */
char* lensed_model_source_tabulated(size_t nobjs, const object* objs);

/* Build check with no counterpart in the reference (it compiles in
 * clBuildProgram, src/lensed.c:755): compile a program from main_program()
 * with hipRTC for `arch` (NULL = gfx950) without needing a device.  Returns
 * the code-object size or -1; *log (malloc'd) receives the build log. */
long lensed_program_compile(const char* source, const char* arch, char** log);

/* ----------------------------------------------------------------- priors */

/* src/prior.h:4-25 */
prior* prior_read(const char* str);
prior* prior_default(double value);
void prior_free(prior* pri);
void prior_print(const prior* pri, char* buf, size_t n);
double prior_apply(const prior* pri, double u);
double prior_lower(const prior* pri);
double prior_upper(const prior* pri);
int prior_pseudo(const prior* pri);

/* -------------------------------------------------------- object metadata */

/* Replaces the device round trip of add_object (src/input/objects.c:15-266):
 * loads objects/<name>.cl, compiles its metadata program with hipRTC and
 * reads type, data size and parameter table from the code object.  Fills
 * obj (id, name, type, size, npars, pars with default pseudo-priors as
 * src/input/objects.c:225-226).  Returns 0 on success. */
int lensed_object_load(object* obj, const char* id, const char* name);
void lensed_object_free(object* obj);

/* ------------------------------------------------------------- quadrature */

/* QUAD_RULES / quad_rule, src/quadrature.c:21-43.  qq, ww: 2*n floats each
 * (node offset scaled by (sx, sy); value weight, error weight).  Returns the
 * number of nodes, or -1 for an unknown rule. */
int lensed_quad_size(const char* rule);
int lensed_quad_rule(const char* rule, float* qq, float* ww, double sx, double sy);
const char* lensed_quad_name(int index); /* NULL past the last rule */
/* the unscaled rule on the unit pixel in double: pts 2*n, wht n, err n */
int lensed_quad_table(const char* rule, double* pts, double* wht, double* err);

/* -------------------------------------------------------------- context */

/* the likelihood context, counterpart of struct lensed (src/lensed.h:3-72) */
typedef struct lensed lensed;

/* flags for lensed_config.flags */
enum {
    LENSED_PROFILE = 1, /* per-stage hipEvent timing, the --profile stages */
    LENSED_EMULATE = 2, /* EPL angular factor tables from the trained emulator where it is certified, else
                           from the exact series (LENSED_TABLE) */
    LENSED_TABLE = 4,   /* EPL angular factor from the exact-series table */
    LENSED_GRAPH = 8,   /* replay batched launches from a captured hipGraph */
    LENSED_NO_CULL = 16, /* trace every ray (no skipping of pixels whose rays provably all miss the source) */
    LENSED_AUTOTUNE = 32 /* time the launch shapes (lensed_tuning) at the first launch of each layout and keep
                            the fastest; choices cached on disk per device, program and layout */
};

/* One observed system: data as prepared by src/lensed.c:406-511. */
typedef struct {
    size_t width, height;
    const float* image;   /* [height][width], row-major, x fastest */
    const float* weight;  /* [height][width]; 0 = masked */
    const float* psf;     /* [psfh][psfw], normalised to unit sum; NULL = none */
    size_t psfw, psfh;
    double pcs[4];        /* pixel coordinates: origin (rx, ry), scale (sx, sy) */
    const char* rule;     /* quadrature rule, NULL = "g3k7" */
    size_t nobjs;
    object* objs;         /* objects with priors set (lensed_object_load) */
} lensed_system;

typedef struct {
    int device;           /* HIP device ordinal */
    int flags;            /* LENSED_* */
    size_t max_batch;     /* largest number of candidates per call */
    const char* emulator; /* weights file for LENSED_EMULATE, or NULL */
    double emulator_tol;  /* certification bound on the emulated angular factor's held-out error, 0 = 4e-6 */
} lensed_options;

/* Build the context for one or more systems (several systems are evaluated
 * together in one launch).  NULL on failure, see lensed_last_error(). */
lensed* lensed_create(const lensed_system* systems, size_t nsystems, const lensed_options* opts);
void lensed_destroy(lensed* l);
const char* lensed_last_error(void);

/* parameter bookkeeping of system s, as src/lensed.c:106-271 builds it */
size_t lensed_npars(const lensed* l, size_t s);
size_t lensed_ndims(const lensed* l, size_t s);
/* pmap[i] = object-order index of the i-th sampler parameter */
int lensed_pmap(const lensed* l, size_t s, size_t* pmap);

/* ---------------------------------------------------- sampler callbacks */

/* src/nested.h:3 -- the MultiNest likelihood callback; the context is
 * system 0 of a lensed*.  cube is in/out as in src/nested.c:43-61. */
void loglike(double cube[], int* ndim, int* npar, double* lnew, void* lensed_);

/* src/nested.h:4-6 -- progress callback; records the constraints and, when
 * a dump prefix is set with lensed_set_output(), re-renders the ML point. */
void dumper(int* nsamples, int* nlive, int* npar, double** physlive,
            double** posterior, double** constraints, double* maxloglike,
            double* logz, double* inslogz, double* logzerr, void* lensed_);

/* ------------------------------------------------------ batch extension */

/* ncand unit cubes (ndim each, sampler order) of system s -> physical
 * parameters (npar each, sampler order, may be NULL) and log-likelihoods,
 * evaluated in one batched launch.  Candidates whose transformed value
 * falls outside hard bounds get logZero (-DBL_MAX) instead of hanging the
 * way src/nested.c:55-57 does.  Returns 0 on success. */
int lensed_loglike_batch(lensed* l, size_t s, const double* cubes, size_t ncand,
                         double* phys_out, double* lnew_out);

/* No counterpart in the reference: lensed_loglike_batch() for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * other allocating routine: every per-launch buffer (descriptor, parameters,
 * object blocks, EPL tables, model images, the pixel cull's list and flags,
 * chi^2 partials, logL) is carved from one static device array of 128 MiB
 * per GPU, which the library holds until the process ends (its address is
 * found once per device when a context is created).  Limits: ncand <= the
 * context's max_batch, and everything must fit the 128 MiB (about 470
 * candidates at 256 x 256; larger counts are rejected with an error that
 * names the limit).  EPL tables are always built from the exact series here,
 * also for a context created with LENSED_EMULATE.  Every call on a GPU shares
 * its array, which keeps the last call's contents: calls that overlap on one
 * GPU (from several threads or contexts) race and must be serialised by the
 * caller.  Results equal lensed_loglike_batch()'s bit for bit. */
int lensed_loglike_batch_noalloc(lensed* l, size_t s, const double* cubes, size_t ncand, double* phys_out,
                                 double* lnew_out);

/* Physical parameters in object order ([ncand][npar]) -> log-likelihoods,
 * for every system at once: phys[s], ncand[s], lnew[s]. */
int lensed_loglike_systems(lensed* l, const double* const* phys, const size_t* ncand,
                           double* const* lnew);

/* One physical parameter vector (object order) of system s -> the planes
 * the dumper writes (src/nested.c:176-237): raw value, quadrature error,
 * PSF-convolved image and per-pixel chi^2.  Any output may be NULL. */
int lensed_render(lensed* l, size_t s, const double* phys, float* value, float* error,
                  float* image, float* chi2);

/* stage timings in ms accumulated under LENSED_PROFILE, in the order
 * +params, -params, set_params, render, convolve, loglike, +loglike_mem,
 * -loglike_mem (src/lensed.c:1125-1132); returns the number of calls */
long lensed_profile(const lensed* l, double* ms8);

/* No counterpart in the reference: what the render of the last launch
 * traced.  items = render items (16 x 16 pixel blocks of every candidate),
 * listed = items the pixel cull kept (all of them without culling), quads =
 * 4 x 4 pixel quads whose rays were traced (16 per item without culling);
 * cull_ms = time of the cull pass accumulated under LENSED_PROFILE (part of
 * the render stage of lensed_profile).  Any output may be NULL; listed and
 * quads wait for the launch and read its render list back. */
int lensed_render_stats(const lensed* l, long* items, long* listed, long* quads, double* cull_ms);

/* ----------------------------------------------------------- launch shapes */

/* Launch shapes of the likelihood kernels: the knobs the autotuner searches.
 * No counterpart in the reference, which sizes its work groups from device
 * queries (src/lensed.c:917-948 for render, the halving loop of
 * src/lensed.c:998-1027 for convolve).  A zero field keeps the default.  No
 * setting changes a result: the model planes and every logL are the same
 * bits under all of them. */
typedef struct {
    double list_fraction; /* the render walks the cull's list of lit 4 x 4 quads while under this fraction
                             of all quads is lit, else every item, skipping dark squares by their flag
                             (0.75; negative: never the list, above 1: always) */
    double render_grid;   /* render blocks as a multiple of the resident count (1) */
    double cull_grid;     /* pixel-cull blocks, the same (4) */
    double conv_grid;     /* convolution blocks, the same (1) */
    int conv_rows;        /* output rows per thread of the convolution tile (4 rows of threads): 2, 8 or 16
                             (8, or 2 when the launch is too small to fill the device); 16 needs 64 + PSF
                             height - 1 rows of LDS and falls back to 8 where that passes 64 KiB; or 1, the
                             column sweep: one wave per 64-column strip of an image runs down the strip,
                             for launches whose systems all have a rank-1 separable PSF of at most 21 x 21 and
                             at most 512 rows (otherwise the default tile) */
    int conv_order;       /* convolution items: 1 candidate-fastest over the grid, 2 each block a run of one
                             candidate's tiles down 64-column strips within its XCD's share (1) */
    /* read back by lensed_get_tuning() after an autotuned launch, else 0: */
    double default_ms;    /* the launch chain with the defaults, as the autotuner timed it */
    double tuned_ms;      /* with the knobs it chose */
} lensed_tuning;

/* Candidates per launch, tuned (the batch shape): lensed_loglike_batch() on
 * system s timed over K = 256, 512, ... up to max_batch (seeded unit-cube
 * draws, 3 launches each after one warm-up; launch shapes autotuned first
 * under LENSED_AUTOTUNE); *best_k = the smallest K whose logL evals/s are
 * within 5 % of the best.  ks and rates (may be NULL, room for 32) receive
 * each K tried and its evals/s, *nk their number.  No counterpart in the
 * reference. */
int lensed_autotune_batch(lensed* l, size_t s, size_t* best_k, size_t* ks, double* rates, int* nk);

/* knobs for launches from now on (layouts already built are rebuilt);
 * NULL restores the defaults.  Returns 0, or -1 for an invalid value. */
int lensed_set_tuning(lensed* l, const lensed_tuning* t);
/* the knobs of the last launch, resolved (defaults filled in; the
 * autotuner's choice under LENSED_AUTOTUNE), or the requested ones before
 * any launch */
int lensed_get_tuning(const lensed* l, lensed_tuning* t);

/* --------------------------------------------------------- EPL emulator */

/* No counterpart in the reference: the learned stand-in for the EPL
 * angular factor (objects/epl.cl:50-61 series).  MLP 4 -> 3 x 128 SiLU -> 2
 * trained on the GPU on nsamples (a power of two) seeded samples of the
 * series; writes the weights file; stats[0] = final training MSE,
 * stats[1] = seconds. */
int lensed_emulator_train(const char* path, long long nsamples, int epochs, int batch, double lr,
                          unsigned long long seed, double* stats);
/* read a weights file (kParams floats); 0 on success */
int lensed_emulator_read(const char* path, float* weights);
/* FP32 loss (mean squared error) and gradient of the training kernels */
double lensed_emulator_loss_grad(const float* weights, const float* X, const float* Y, int n, float* grad);
/* the network at X[n][4] with tangents dX[n][4]; bf16 = 1 runs the
 * BF16-MFMA inference kernel (outputs Y[n][2] and dY[n][2]) */
int lensed_emulator_eval(const float* weights, const float* X, const float* dX, int n, int bf16, float* Y,
                         float* dY);
/* Of the last launch with EPL tables under LENSED_EMULATE: the candidates
 * whose table the exact series built (outside the certified cells), the
 * candidates with a table, and the largest held-out error of the
 * certification.  No counterpart in the reference. */
int lensed_emulator_stats(const lensed* l, long* fallbacks, long* checked, double* max_err);
/* The certification, done once when the context is created: the network's
 * tables at the corners and centre of every (q, t) cell of a 32 x 32 grid
 * over q in [0.3, 1], t in [0.5, 1.5], checked against the FP64 series at
 * every interval midpoint; a cell is certified when all five are within
 * emulator_tol, and only candidates in certified cells take the network's
 * table.  Cells, certified cells, largest held-out error.  No counterpart in
 * the reference. */
int lensed_emulator_cells(const lensed* l, long* ncells, long* ncertified, double* max_err);
/* The exact recheck of lensed_nested_run(): candidates of an emulated launch
 * that may be accepted (logL above the current threshold) are evaluated
 * again with exact-series tables and take that logL.  Counts since the
 * context was created: rechecked, changed by it, largest |delta logL|.  No
 * counterpart in the reference. */
int lensed_emulator_rechecks(const lensed* l, long* rechecked, long* changed, double* max_abs_dlogl);

/* ------------------------------------------------------ nested sampling */

/* src/lensed.c:1270-1275 -- MultiNest's C entry point, with the arguments
 * Lensed passes.  Provided here by a single-ellipsoid nested sampler that
 * calls LogLike once per candidate and Dumper every updInt iterations; IS,
 * mmodal, nClsPar, maxModes, Ztol, initMPI and logZero are accepted and
 * ignored (no importance weighting, mode separation or MPI); fb = 1 prints
 * progress on stderr.  efr is the
 * target acceptance when ceff is set, else the shrinking factor.  With a
 * root and resume = 1 the run continues from <root>resume.dat; outfile = 1
 * writes <root>.txt (weight, -2 logL, parameters per posterior sample). */
void run(int IS, int mmodal, int ceff, int nlive, double tol, double efr, int ndims, int nPar, int nClsPar,
         int maxModes, int updInt, double Ztol, char root[], int seed, int* pWrap, int fb, int resume,
         int outfile, int initMPI, double logZero, int maxiter,
         void (*LogLike)(double*, int*, int*, double*, void*),
         void (*Dumper)(int*, int*, int*, double**, double**, double**, double*, double*, double*, double*,
                        void*),
         void* context);

/* The same sampler proposing `batch` candidates at a time; no counterpart in
 * the reference (MultiNest proposes one point per likelihood call). */
typedef struct {
    int nlive;       /* live points (300) */
    double tol;      /* stop when the live points can raise ln Z by < tol (0.1) */
    double efr;      /* target acceptance if ceff, else shrinking factor (0.8) */
    int ceff;        /* constant efficiency mode */
    int batch;       /* candidates proposed per likelihood call; in lensed_nested_run() -1 takes
                        lensed_autotune_batch()'s choice */
    int seed;        /* random seed */
    int updint;      /* iterations between dumper calls (0 = only at the end) */
    int maxiter;     /* iteration limit (0 = none) */
    int resume;      /* continue from <root>resume.dat */
    int outfile;     /* write <root>.txt */
    const int* wrap; /* npar periodic flags (sampler order), or NULL: none, or in
                        lensed_nested_run() the parameters' own `wrap` flags */
    const char* root;/* output prefix, or NULL */
    int feedback;    /* progress lines on stderr every updint (or 1000) iterations */
} lensed_ns_options;

typedef struct {
    double logz, logz_err, info, max_loglike;
    long nevals, niter;
} lensed_ns_result;

/* cubes[n][ndim] -> phys[n][npar] (sampler order) and lnew[n] */
typedef void (*lensed_batch_fn)(const double* cubes, int n, double* phys, double* lnew, void* ctx);

/* a generic likelihood; mean, sigma and ml (npar each, sampler order) may be NULL */
int lensed_nested_sample(int ndim, int npar, lensed_batch_fn fn, void* ctx, const lensed_ns_options* opts,
                         double* mean, double* sigma, double* ml, lensed_ns_result* result);
/* system s of a context through lensed_loglike_batch(); for system 0 the
 * results are also recorded as dumper() records them */
int lensed_nested_run(lensed* l, size_t s, const lensed_ns_options* opts, double* mean, double* sigma,
                      double* ml, lensed_ns_result* result);

/* ----------------------------------------------- input data and output */

/* src/data.c:54-118 -- the 2-D image of a FITS file specification in
 * CFITSIO's syntax (`file.fits`, `file.fits[ext]`, `file.fits[IMG]`,
 * `file.fits[x1:x2,y1:y2]`, steps `x1:x2:dx`, reversed ranges, `*`), as
 * floats (BSCALE/BZERO applied), row-major; *data is freed with
 * lensed_fits_free().  pcs (may be NULL) receives the pixel coordinate
 * system of src/data.c:236-278: origin (rx, ry) and step (sx, sy). */
int lensed_fits_read(const char* spec, size_t* width, size_t* height, float** data, double* pcs);
void lensed_fits_free(void* p);
/* src/data.c:236-278 -- only the pixel coordinate system of a specification */
int lensed_read_pcs(const char* spec, double* pcs);
/* src/data.c:120-165,372-391 -- n float images as IMAGE extensions named
 * names[k] after an empty primary HDU; a leading '!' is accepted */
int lensed_fits_write(const char* path, size_t width, size_t height, size_t n, const float* const* images,
                      const char* const* names);
/* src/data.c:421-501 -- mode and FWHM of the pixel values (mask: skip zeros) */
void lensed_find_mode(size_t n, const float* values, const float* mask, double* mode, double* fwhm);

/* the data options of src/input/options.c:63-146; weight, gain and xweight
 * take a FITS specification or a number, as src/input/options.c:506-532 */
typedef struct {
    const char* image;   /* required */
    const char* weight;  /* weight map or value; else weights = gain / (image + offset) */
    const char* gain;
    double offset;
    double bscale;       /* 0 = none */
    const char* xweight; /* extra weight multiplier, or NULL */
    const char* mask;    /* non-zero pixels are masked (weight 0), or NULL */
    const char* psf;     /* normalised to unit sum, or NULL */
} lensed_data_options;

#define LENSED_WARN_SKY_NOT_ZERO 1    /* background mode far from zero */
#define LENSED_WARN_NO_OFFSET 2       /* mode near zero but no offset given */
#define LENSED_WARN_NEGATIVE_WEIGHT 4 /* some weights not >= 0 */

typedef struct {
    size_t width, height;
    float* image;
    float* weight;
    float* psf;          /* NULL without a PSF */
    size_t psfw, psfh;
    double pcs[4];
    size_t masked;       /* pixels with weight 0 */
    double mode, fwhm;   /* background check of src/lensed.c:501-540 */
    int warnings;        /* LENSED_WARN_* */
} lensed_data;

/* src/lensed.c:406-549 -- read and prepare one observation */
int lensed_data_read(const lensed_data_options* opts, lensed_data* data);
void lensed_data_free(lensed_data* data);

/* src/lensed.c:572-594, src/nested.c:169-226 -- every dumper() call writes
 * the layers IMG (model), RES (data - model), RAW (unconvolved model), ERR
 * (quadrature error / model), WHT (weights), PVL (erfc(sqrt(chi^2 / 2)))
 * of system 0 at the maximum-likelihood point to this FITS file (NULL: off) */
int lensed_set_output(lensed* l, const char* path);
/* the same layers for one physical parameter vector (object order) */
int lensed_write_layers(lensed* l, size_t s, const double* phys, const char* path);

/* -------------------------------------------------------- configuration */

/* src/input.c, src/input/ini.c, src/input/options.c -- Lensed's INI files
 * ([options], [objects], [priors] with `wrap`/`image`, [labels]) and
 * `--name=value` options, with the option table's types, defaults and
 * conditional requirements; objects are loaded as they are named. */
typedef struct lensed_config lensed_config;
lensed_config* lensed_config_create(void);
void lensed_config_free(lensed_config* c);
/* read one INI file (relative paths are taken from its directory) */
int lensed_config_read_ini(lensed_config* c, const char* path);
/* set an option by (abbreviated) name, as an INI line or --name=value does */
int lensed_config_set(lensed_config* c, const char* name, const char* value);
/* one command-line argument "--name=value" (src/input.c:83-100) */
int lensed_config_arg(lensed_config* c, const char* arg);
/* src/input.c:214-239: required options, some objects, every prior given */
int lensed_config_check(const lensed_config* c);
/* an option's value as text (booleans "true"/"false"), NULL when unset */
const char* lensed_config_get(const lensed_config* c, const char* name);
/* the i-th entry of the option table; -1 past its end */
int lensed_config_option(int i, const char** name, const char** type, const char** help, const char** defval);
/* the objects in [objects] order, owned by the configuration */
size_t lensed_config_nobjects(const lensed_config* c);
object* lensed_config_objects(lensed_config* c);
/* lensed_data_read() with the configuration's data options */
int lensed_config_data(const lensed_config* c, lensed_data* data);
/* src/lensed.c:1175-1227 -- <root>.paramnames (id, label) and <root>.ranges
 * (id, prior lower and upper or N) in sampler order for system s */
int lensed_write_paramfiles(const lensed* l, size_t s, const char* root);

/* results recorded by dumper(): mean, sigma, ML, MAP per parameter in object
 * order (each may be NULL) and the log-evidence summary */
int lensed_results(const lensed* l, double* mean, double* sigma, double* ml, double* map,
                   double* logev, double* logev_err, double* max_loglike);

#ifdef __cplusplus
}
#endif

#endif /* LENSED_AMD_H */
