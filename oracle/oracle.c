/*
 * CPU oracle for the Lensed forward-model likelihood.
 *
 * TEST INFRASTRUCTURE: this library is the checker the parity tests compare
 * the HIP path against, and the "port" CPU baseline bench.py times.  Nothing
 * in the product (lensed_amd/) loads or calls it.
 *
 * It restates the reference algorithm in plain C, one function per reference
 * piece, each citing the file it follows:
 *   objects        objects/<name>.cl (the formulas exactly as written there,
 *                  e.g. atan2 + sin/cos for the EPL series)
 *   compute        src/kernel.c:65-111 (generated ray-shoot)
 *   set_params     src/kernel.c:113-150, 401-656 (incl. image-plane priors)
 *   render         kernel/lensed.cl:9-38
 *   convolve       kernel/lensed.cl:56-103
 *   chi^2 / logL   kernel/lensed.cl:41-53 + src/nested.c:101-115
 *   norm prior     src/prior/norm.c:9-15 (Abramowitz & Stegun 26.2.23)
 * The body is type-generic (oracle_impl.h) and built twice: _d in double
 * (the FP64 oracle) and _f in float (the FP32 restatement).
 *
 * Parity of every function is pinned by tests/test_oracle.py against the
 * reference's own golden images (tests/golden/), closed forms and brute
 * force; see the header of that file for which check pins which function.
 */
#include <math.h>
#include <string.h>
#include <tgmath.h>

enum {
    ORC_SIS, ORC_SIS_SHEAR, ORC_NSIS, ORC_POINT_MASS, ORC_SIE, ORC_SIE_SHEAR, ORC_NSIE,
    ORC_EPL, ORC_EPL_SHEAR, ORC_SERSIC, ORC_SERSIC_OLD, ORC_GAUSS, ORC_EXPONENTIAL, ORC_DEVAUC,
    ORC_SKY, ORC_NKINDS
};

static const char* ORC_NAMES[ORC_NKINDS] = {
    "sis", "sis_plus_shear", "nsis", "point_mass", "sie", "sie_plus_shear", "nsie",
    "epl", "epl_plus_shear", "sersic", "sersic-old", "gauss", "exponential", "devauc", "sky"};
static const int ORC_NPARS[ORC_NKINDS] = {3, 5, 4, 3, 5, 7, 6, 6, 8, 7, 7, 6, 6, 6, 3};
static const char ORC_TYPES[ORC_NKINDS] = "LLLLLLLLLSSSSSF";

int orc_kind(const char* name) {
    for (int k = 0; k < ORC_NKINDS; ++k)
        if (strcmp(name, ORC_NAMES[k]) == 0) return k;
    return -1;
}
int orc_npars(int kind) { return kind >= 0 && kind < ORC_NKINDS ? ORC_NPARS[kind] : -1; }
int orc_type(int kind) { return kind >= 0 && kind < ORC_NKINDS ? ORC_TYPES[kind] : 0; }

#define REAL double
#define SFX(n) n##_d
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(n) n##_f
#include "oracle_impl.h"
#undef REAL
#undef SFX

/* threads of the OpenMP loops (the fast build; 1 in the strict one) */
#ifdef _OPENMP
#include <omp.h>
int orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); return omp_get_max_threads(); }
#else
int orc_set_threads(int n) { (void)n; return 1; }
#endif

/* Sersic b(n) of objects/sersic.cl:32-33 in FP64, for the check against its
 * defining equation Gamma(2n, b) / Gamma(2n) = 1/2 (docs/sources.md:109-111) */
double orc_sersic_b(double n) { return sersic_b_d(n); }

/* chi^2 = sum w (m - d)^2 with every term in FP32 and the sum in FP64,
 * kernel/lensed.cl:50-51 + src/nested.c:107-109 */
double orc_chi2_f32terms(int n, const float* image, const float* weight, const float* model) {
    double chi2 = 0.0;
    for (int i = 0; i < n; ++i) {
        float d = model[i] - image[i];
        chi2 += (double)(weight[i] * d * d);
    }
    return chi2;
}

/* chi^2 entirely in FP64 */
double orc_chi2_d(int n, const double* image, const double* weight, const double* model) {
    double chi2 = 0.0;
    for (int i = 0; i < n; ++i) {
        double d = model[i] - image[i];
        chi2 += weight[i] * d * d;
    }
    return chi2;
}

/* src/prior/norm.c:9-15 */
double orc_gauss_quantile(double u) {
    double t = u < 0.5 ? sqrt(-2.0 * log(u)) : sqrt(-2.0 * log(1 - u));
    t = t - ((0.010328 * t + 0.802853) * t + 2.515517) / (((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.);
    return u < 0.5 ? -t : t;
}
