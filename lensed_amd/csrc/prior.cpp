// Priors: unit-cube -> physical transforms applied in the likelihood
// prologue (src/nested.c:43-61).  The registry mirrors src/prior.c: a single
// real is the `delta` pseudo-prior (fixed value, "derived" in the sampler's
// ordering), `unif a b` and `norm m s`.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lensed.h"
#include "prior_impl.h"

namespace {

bool parse_real(const std::string& s, double* v) {
    if (s.empty()) return false;
    char* end = nullptr;
    *v = strtod(s.c_str(), &end);
    return end && *end == '\0';
}

// Inverse of the standard normal CDF by the rational approximation of
// Abramowitz & Stegun 26.2.23 (|error| < 4.5e-4), the transform that
// src/prior/norm.c:9-15 applies.
double normal_quantile_as26223(double u) {
    const double c0 = 2.515517, c1 = 0.802853, c2 = 0.010328;
    const double d1 = 1.432788, d2 = 0.189269, d3 = 0.001308;
    double p = u < 0.5 ? u : 1 - u;
    double t = std::sqrt(-2.0 * std::log(p));
    double x = t - (c0 + t * (c1 + t * c2)) / (1.0 + t * (d1 + t * (d2 + t * d3)));
    return u < 0.5 ? -x : x;
}

}  // namespace

extern "C" {

prior* prior_read(const char* str) {
    if (!str) return nullptr;
    std::vector<std::string> args;
    const char* ws = " \t\n\v\f\r";
    size_t len = strlen(str), pos = strspn(str, ws);
    while (pos < len) {
        size_t spn = strcspn(str + pos, ws);
        args.emplace_back(str + pos, spn);
        pos += spn;
        pos += strspn(str + pos, ws);
    }
    if (args.empty()) return nullptr;
    double a = 0, b = 0;
    if (args.size() == 1 && parse_real(args[0], &a)) return new prior{prior::DELTA, a, 0};
    if (args.size() == 3 && parse_real(args[1], &a) && parse_real(args[2], &b)) {
        if (args[0] == "unif") return new prior{prior::UNIF, a, b};
        if (args[0] == "norm") return new prior{prior::NORM, a, b};
    }
    return nullptr;
}

prior* prior_default(double value) { return new prior{prior::DELTA, value, 0}; }

void prior_free(prior* pri) { delete pri; }

void prior_print(const prior* pri, char* buf, size_t n) {
    if (!pri || !buf || n == 0) return;
    switch (pri->kind) {
        case prior::DELTA: snprintf(buf, n, "%g", pri->a); break;
        case prior::UNIF: snprintf(buf, n, "U(%g, %g)", pri->a, pri->b); break;
        case prior::NORM: snprintf(buf, n, "N(%g, %g)", pri->a, pri->b); break;
    }
}

double prior_apply(const prior* pri, double u) {
    switch (pri->kind) {
        case prior::DELTA: return pri->a;
        case prior::UNIF: return pri->a + u * (pri->b - pri->a);
        case prior::NORM: return pri->a + pri->b * normal_quantile_as26223(u);
    }
    return NAN;
}

// bounds of the support; the normal prior is cut at +-7 sigma as in
// src/prior/norm.c:84-96
double prior_lower(const prior* pri) {
    switch (pri->kind) {
        case prior::DELTA: return pri->a;
        case prior::UNIF: return pri->a;
        case prior::NORM: return pri->a - 7 * pri->b;
    }
    return NAN;
}

double prior_upper(const prior* pri) {
    switch (pri->kind) {
        case prior::DELTA: return pri->a;
        case prior::UNIF: return pri->b;
        case prior::NORM: return pri->a + 7 * pri->b;
    }
    return NAN;
}

int prior_pseudo(const prior* pri) { return pri && pri->kind == prior::DELTA; }

}  // extern "C"
