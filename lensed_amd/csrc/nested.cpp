// Nested sampling driver (the caller of the likelihood path).
//
// Lensed hands its likelihood to MultiNest's run() (src/lensed.c:1236-1288),
// which calls loglike() once per candidate and dumper() every updInt
// iterations (src/nested.c:17,133).  MultiNest is an external Fortran library
// that is not linked here; this file provides a nested sampler with
// the same entry point and callbacks, and a batched variant that proposes K
// points at a time so one launch evaluates them all.
//
// Algorithm (Skilling 2004; ellipsoidal bounds as in Mukherjee, Parkinson &
// Liddle 2006 and Feroz, Hobson & Bridges 2009): nlive points uniform in the
// unit cube; at each iteration the lowest-likelihood point dies with
// prior-volume shrinkage X_i = exp(-i / nlive) and is replaced by a draw from
// the prior restricted to L > L_min.  Draws come from a union of ellipsoids
// bounding the live points (covariance-shaped, scaled to contain them all,
// enlarged), rejected where L <= L_min or outside the cube (wrapped for
// periodic parameters).  Proposals
// are drawn in batches; a batch is consumed in order against the rising
// threshold, which keeps every accepted point a uniform draw from the
// constrained prior as long as the ellipsoid still bounds it.  The evidence
// accumulates Z += L_i (X_{i-1} - X_i); the run stops when the live points
// can raise ln Z by at most tol (MultiNest's tolerance), and the live points
// are then added with the remaining volume.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lensed.h"
#include "context.h"

namespace lsd {
void set_error(const std::string& msg);
}

namespace {

struct Point {
    std::vector<double> u;     // unit cube, ndim
    std::vector<double> phys;  // physical parameters, npar (sampler order)
    double logl;
};

double logaddexp(double a, double b) {
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    double m = std::max(a, b);
    return m + std::log(std::exp(a - m) + std::exp(b - m));
}

// Cholesky of a symmetric positive-definite n x n matrix (row-major, lower)
bool cholesky(std::vector<double>& a, int n) {
    for (int j = 0; j < n; ++j) {
        double s = a[j * n + j];
        for (int k = 0; k < j; ++k) s -= a[j * n + k] * a[j * n + k];
        if (s <= 0) return false;
        a[j * n + j] = std::sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double t = a[i * n + j];
            for (int k = 0; k < j; ++k) t -= a[i * n + k] * a[j * n + k];
            a[i * n + j] = t / a[j * n + j];
        }
        for (int k = j + 1; k < n; ++k) a[j * n + k] = 0;
    }
    return true;
}

double log_unit_ball(int n) { return 0.5 * n * std::log(M_PI) - std::lgamma(0.5 * n + 1); }

// An ellipsoid {x : |L^-1 (x - mean)| <= scale} with L the Cholesky factor
// of the covariance of the points it was fitted to.
struct Ellipsoid {
    int n = 0;
    std::vector<double> mean, chol;
    double scale = 1;
    double logvol = 0;  // log volume

    // Mahalanobis radius^2 of x (in units of the covariance, before scale)
    double radius2(const double* x) const {
        double r2 = 0;
        double y[64];
        for (int i = 0; i < n; ++i) {
            double t = x[i] - mean[i];
            for (int k = 0; k < i; ++k) t -= chol[i * n + k] * y[k];
            y[i] = t / chol[i * n + i];
            r2 += y[i] * y[i];
        }
        return r2;
    }

    bool contains(const double* x) const { return radius2(x) <= scale * scale; }

    // bound the m points xs[m][n], then multiply the volume by `volfac`;
    // with boot > 0 the bound is further expanded by the bootstrap (fit on a
    // resample, measure how far the points left out stick out, keep the
    // largest ratio over `boot` resamples), which a piece of a region split
    // off from the rest needs, since its points' hull undercovers its share
    bool fit(const std::vector<const double*>& xs, int dim, double volfac, int boot = 0,
             std::mt19937_64* rng = nullptr) {
        n = dim;
        int m = (int)xs.size();
        mean.assign(n, 0.0);
        for (const double* x : xs)
            for (int i = 0; i < n; ++i) mean[i] += x[i] / m;
        std::vector<double> cov((size_t)n * n, 0.0);
        for (const double* x : xs)
            for (int i = 0; i < n; ++i)
                for (int j = 0; j <= i; ++j) cov[i * n + j] += (x[i] - mean[i]) * (x[j] - mean[j]) / (m - 1);
        for (int i = 0; i < n; ++i) {
            cov[i * n + i] += 1e-12;
            for (int j = 0; j < i; ++j) cov[j * n + i] = cov[i * n + j];
        }
        chol = cov;
        if (m <= n || !cholesky(chol, n)) return false;
        double rmax = 0;
        for (const double* x : xs) rmax = std::max(rmax, radius2(x));
        double expand = 1;
        for (int b = 0; b < boot && rng; ++b) {
            std::uniform_int_distribution<int> pick(0, m - 1);
            std::vector<char> in(m, 0);
            std::vector<const double*> sub(m);
            for (int k = 0; k < m; ++k) {
                int j = pick(*rng);
                sub[k] = xs[j];
                in[j] = 1;
            }
            Ellipsoid e;
            if (!e.fit(sub, n, 1.0)) continue;
            double far = 0;
            for (int k = 0; k < m; ++k)
                if (!in[k]) far = std::max(far, std::sqrt(e.radius2(xs[k])) / e.scale);
            expand = std::max(expand, far);
        }
        scale = std::sqrt(rmax) * expand * std::pow(volfac, 1.0 / n);
        logvol = log_unit_ball(n) + n * std::log(scale);
        for (int i = 0; i < n; ++i) logvol += std::log(chol[i * n + i]);
        return true;
    }

    // the sphere around the whole unit cube
    void cube(int dim) {
        n = dim;
        mean.assign(n, 0.5);
        chol.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) chol[i * n + i] = 1;
        scale = 0.5 * std::sqrt((double)n) * 1.0001;
        logvol = log_unit_ball(n) + n * std::log(scale);
    }

    void sample(std::mt19937_64& rng, double* out) const {
        std::normal_distribution<double> g(0, 1);
        std::uniform_real_distribution<double> u(0, 1);
        double z[64];
        double r2 = 0;
        for (int i = 0; i < n; ++i) {
            z[i] = g(rng);
            r2 += z[i] * z[i];
        }
        double rad = std::pow(u(rng), 1.0 / n) / std::sqrt(r2);
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = 0; k <= i; ++k) s += chol[i * n + k] * z[k];
            out[i] = mean[i] + scale * rad * s;
        }
    }
};

// The union of ellipsoids bounding the live points (Feroz, Hobson & Bridges
// 2009, MultiNest's bound): the set is split recursively by 2-means while
// the two halves' ellipsoids take less than half the volume of their parent,
// or while the parent is more than twice the volume its points are expected
// to occupy (their share m/N of the remaining prior volume X, enlarged).
// Periodic coordinates are first rotated so that the circular mean of the
// points sits at 0.5, which keeps a mode across the 0/1 seam in one piece.
bool inside(const double* u, int n, const std::vector<int>& wrap, double* out);

struct Bound {
    static constexpr int kBoot = 5;  // bootstrap resamples of a split-off piece
    static constexpr int kProbe = 1024;  // draws that measure the share of the bound inside the cube
    int n = 0;
    int boot = kBoot;           // 0 in constant efficiency mode
    std::vector<double> shift;  // x = frac(u + shift) for periodic coordinates
    std::vector<Ellipsoid> ells;
    std::vector<double> cum;    // cumulative volume fractions
    double logvol_union = 0;    // log of the summed volumes (bounds the union's)
    double logvol_cube = 0;     // log of the volume of the union's part inside the cube (estimated)

    void to_x(const double* u, double* x) const {
        for (int i = 0; i < n; ++i) {
            double v = u[i] + shift[i];
            x[i] = shift[i] != 0 ? v - std::floor(v) : v;
        }
    }

    static double logsum(const std::vector<Ellipsoid>& es) {
        double top = -INFINITY, acc = 0;
        for (const auto& e : es) top = std::max(top, e.logvol);
        for (const auto& e : es) acc += std::exp(e.logvol - top);
        return top + std::log(acc);
    }

    // the leaves that bound xs: the parent itself, or the pieces of a
    // recursive split when they take less volume than the parent
    std::vector<Ellipsoid> split(const std::vector<const double*>& xs, const Ellipsoid& parent, double volfac,
                                 int nlive, double logx, std::mt19937_64& rng, int depth) {
        const int m = (int)xs.size(), minpts = 2 * (n + 1);
        if (m < 2 * minpts || depth > 8) return {parent};
        // 2-means, started from a random point's farthest point and that one's farthest
        std::vector<double> ca(n), cb(n);
        {
            auto far = [&](const double* c) {
                int best = 0;
                double dmax = -1;
                for (int k = 0; k < m; ++k) {
                    double d = 0;
                    for (int i = 0; i < n; ++i) d += (xs[k][i] - c[i]) * (xs[k][i] - c[i]);
                    if (d > dmax) dmax = d, best = k;
                }
                return best;
            };
            int a = far(xs[std::uniform_int_distribution<int>(0, m - 1)(rng)]);
            int b = far(xs[a]);
            ca.assign(xs[a], xs[a] + n);
            cb.assign(xs[b], xs[b] + n);
        }
        std::vector<char> side(m, 0);
        for (int it = 0; it < 20; ++it) {
            bool changed = false;
            for (int k = 0; k < m; ++k) {
                double da = 0, db = 0;
                for (int i = 0; i < n; ++i) {
                    da += (xs[k][i] - ca[i]) * (xs[k][i] - ca[i]);
                    db += (xs[k][i] - cb[i]) * (xs[k][i] - cb[i]);
                }
                char s = db < da;
                changed |= s != side[k];
                side[k] = s;
            }
            std::fill(ca.begin(), ca.end(), 0.0);
            std::fill(cb.begin(), cb.end(), 0.0);
            int na = 0, nb = 0;
            for (int k = 0; k < m; ++k) {
                auto& c = side[k] ? cb : ca;
                for (int i = 0; i < n; ++i) c[i] += xs[k][i];
                (side[k] ? nb : na) += 1;
            }
            if (!na || !nb) break;
            for (int i = 0; i < n; ++i) ca[i] /= na, cb[i] /= nb;
            if (!changed && it) break;
        }
        std::vector<const double*> xa, xb;
        for (int k = 0; k < m; ++k) (side[k] ? xb : xa).push_back(xs[k]);
        if ((int)xa.size() < minpts || (int)xb.size() < minpts) return {parent};
        Ellipsoid ea, eb;
        if (!ea.fit(xa, n, volfac, boot, &rng) || !eb.fit(xb, n, volfac, boot, &rng)) return {parent};
        // Look deeper when the halves halve the volume, or when the parent is
        // more than twice the volume its points are expected to occupy (their
        // share m/N of the remaining prior volume X, enlarged) -- several
        // modes or a curved region, whose first cut may not gain yet.
        double both = logsum({ea, eb});
        double expected = logx + std::log(volfac * m / nlive);
        if (!(both < parent.logvol + std::log(0.5)) && !(parent.logvol > expected + std::log(2.0))) return {parent};
        std::vector<Ellipsoid> leaves = split(xa, ea, volfac, nlive, logx, rng, depth + 1);
        std::vector<Ellipsoid> more = split(xb, eb, volfac, nlive, logx, rng, depth + 1);
        leaves.insert(leaves.end(), more.begin(), more.end());
        if (logsum(leaves) < parent.logvol + std::log(0.8)) return leaves;
        return {parent};
    }

    void fit(const std::vector<Point>& pts, double volfac, double logx, const std::vector<int>& wrap, bool multi,
             std::mt19937_64& rng) {
        n = (int)pts[0].u.size();
        int m = (int)pts.size();
        shift.assign(n, 0.0);
        for (int i = 0; i < n && i < (int)wrap.size(); ++i) {
            if (!wrap[i]) continue;
            double cs = 0, sn = 0;
            for (const auto& p : pts) {
                cs += std::cos(2 * M_PI * p.u[i]);
                sn += std::sin(2 * M_PI * p.u[i]);
            }
            double c = std::atan2(sn, cs) / (2 * M_PI);
            shift[i] = 0.5 - (c - std::floor(c));
            if (shift[i] == 0) shift[i] = 1;  // still marks the coordinate periodic
        }
        std::vector<double> xbuf((size_t)m * n);
        std::vector<const double*> xs(m);
        for (int k = 0; k < m; ++k) {
            to_x(pts[k].u.data(), &xbuf[(size_t)k * n]);
            xs[k] = &xbuf[(size_t)k * n];
        }
        ells.clear();
        Ellipsoid root;
        if (!root.fit(xs, n, volfac)) {  // degenerate live set
            std::fill(shift.begin(), shift.end(), 0.0);
            root.cube(n);
            ells.push_back(root);
        } else if (multi) {
            ells = split(xs, root, volfac, m, logx, rng, 0);
        } else {
            ells.push_back(root);
        }
        cum.assign(ells.size(), 0.0);
        double top = -INFINITY, acc = 0;
        for (const auto& e : ells) top = std::max(top, e.logvol);
        for (size_t k = 0; k < ells.size(); ++k) cum[k] = acc += std::exp(ells[k].logvol - top);
        for (auto& c : cum) c /= acc;
        logvol_union = top + std::log(acc);
        // Where the union is larger than the cube, the share of its draws
        // that land inside the cube gives the volume it actually proposes
        // from: in many dimensions an ellipsoid many times the cube's volume
        // can still cover a small part of it, and drawing from it then beats
        // drawing from the prior.
        logvol_cube = logvol_union;
        if (logvol_union >= 0) {
            double x[64], c[64];
            int hit = 0;
            for (int k = 0; k < kProbe; ++k) {
                sample(rng, x);
                hit += inside(x, n, wrap, c);
            }
            logvol_cube = logvol_union + std::log(std::max(hit, 1) / (double)kProbe);
        }
    }

    // a uniform draw from the union (in unit-cube coordinates, unwrapped)
    void sample(std::mt19937_64& rng, double* out) const {
        std::uniform_real_distribution<double> u(0, 1);
        double x[64];
        for (;;) {
            size_t k = std::lower_bound(cum.begin(), cum.end(), u(rng)) - cum.begin();
            if (k >= ells.size()) k = ells.size() - 1;
            ells[k].sample(rng, x);
            int q = 0;  // the union's overlap: accept with 1 / (ellipsoids holding x)
            for (const auto& e : ells) q += e.contains(x);
            if (q <= 1 || u(rng) * q < 1) break;
        }
        for (int i = 0; i < n; ++i) out[i] = x[i] - shift[i];
    }
};

struct Options {
    int nlive = 300, batch = 1, seed = 0, updint = 1000, maxiter = 0, ceff = 0, resume = 0, outfile = 0;
    int feedback = 0;
    double tol = 0.1, efr = 0.8;
    std::vector<int> wrap;
    std::string root;
};

typedef void (*BatchFn)(const double* cubes, int n, double* phys, double* lnew, void* ctx);

struct Result {
    double logz = -INFINITY, logz_err = 0, info = 0, max_logl = -INFINITY;
    long nevals = 0, niter = 0;
    std::vector<double> mean, sigma, ml;
    std::vector<Point> dead;
    std::vector<double> dead_logw;  // log prior-volume widths
};

bool inside(const double* u, int n, const std::vector<int>& wrap, double* out) {
    for (int i = 0; i < n; ++i) {
        double v = u[i];
        if (i < (int)wrap.size() && wrap[i]) v -= std::floor(v);
        if (v < 0 || v > 1) return false;
        out[i] = v;
    }
    return true;
}

void write_checkpoint(const std::string& path, const std::vector<Point>& live, const Result& r, int ndim, int npar) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return;
    fprintf(f, "lensed-nested 1 %d %d %zu %zu %ld %ld\n", ndim, npar, live.size(), r.dead.size(), r.niter, r.nevals);
    auto put = [&](const Point& p, double lw) {
        fprintf(f, "%.17g %.17g", p.logl, lw);
        for (double v : p.u) fprintf(f, " %.17g", v);
        for (double v : p.phys) fprintf(f, " %.17g", v);
        fprintf(f, "\n");
    };
    for (const auto& p : live) put(p, 0);
    for (size_t i = 0; i < r.dead.size(); ++i) put(r.dead[i], r.dead_logw[i]);
    fprintf(f, "%.17g %.17g %.17g\n", r.logz, r.info, r.max_logl);
    fclose(f);
}

bool read_checkpoint(const std::string& path, std::vector<Point>& live, Result& r, int ndim, int npar) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    int nd, np;
    size_t nl, ndd;
    bool ok = fscanf(f, "lensed-nested 1 %d %d %zu %zu %ld %ld", &nd, &np, &nl, &ndd, &r.niter, &r.nevals) == 6 &&
              nd == ndim && np == npar;
    auto get = [&](Point& p, double& lw) {
        p.u.resize(ndim);
        p.phys.resize(npar);
        bool good = fscanf(f, "%lf %lf", &p.logl, &lw) == 2;
        for (int i = 0; i < ndim && good; ++i) good = fscanf(f, "%lf", &p.u[i]) == 1;
        for (int i = 0; i < npar && good; ++i) good = fscanf(f, "%lf", &p.phys[i]) == 1;
        return good;
    };
    live.assign(ok ? nl : 0, Point());
    double lw;
    for (size_t i = 0; ok && i < nl; ++i) ok = get(live[i], lw);
    r.dead.assign(ok ? ndd : 0, Point());
    r.dead_logw.assign(ok ? ndd : 0, 0.0);
    for (size_t i = 0; ok && i < ndd; ++i) ok = get(r.dead[i], r.dead_logw[i]);
    ok = ok && fscanf(f, "%lf %lf %lf", &r.logz, &r.info, &r.max_logl) == 3;
    fclose(f);
    if (!ok) {  // unusable checkpoint: start afresh
        r = Result();
        live.clear();
    }
    return ok;
}

// posterior weights of dead + final live points -> mean, sigma, ML
void summarise(Result& r, const std::vector<Point>& live, double logx, int npar) {
    std::vector<const Point*> pts;
    std::vector<double> lw;
    for (size_t i = 0; i < r.dead.size(); ++i) {
        pts.push_back(&r.dead[i]);
        lw.push_back(r.dead_logw[i] + r.dead[i].logl - r.logz);
    }
    double lwl = logx - std::log((double)live.size());
    for (const auto& p : live) {
        pts.push_back(&p);
        lw.push_back(lwl + p.logl - r.logz);
    }
    r.mean.assign(npar, 0.0);
    r.sigma.assign(npar, 0.0);
    r.ml.assign(npar, 0.0);
    double wsum = 0, best = -INFINITY;
    for (size_t i = 0; i < pts.size(); ++i) {
        double w = std::exp(lw[i]);
        wsum += w;
        for (int k = 0; k < npar; ++k) r.mean[k] += w * pts[i]->phys[k];
        if (pts[i]->logl > best) {
            best = pts[i]->logl;
            r.ml = pts[i]->phys;
        }
    }
    for (int k = 0; k < npar; ++k) r.mean[k] /= wsum;
    for (size_t i = 0; i < pts.size(); ++i) {
        double w = std::exp(lw[i]) / wsum;
        for (int k = 0; k < npar; ++k) r.sigma[k] += w * (pts[i]->phys[k] - r.mean[k]) * (pts[i]->phys[k] - r.mean[k]);
    }
    for (int k = 0; k < npar; ++k) r.sigma[k] = std::sqrt(r.sigma[k]);
    r.max_logl = best;
}

typedef void (*DumperFn)(int*, int*, int*, double**, double**, double**, double*, double*, double*, double*, void*);

// After a batch is evaluated: the candidates whose logL came from an
// approximation (the EPL emulator) and that may still be accepted -- logL
// above the current threshold, which only rises while the batch is used up
// -- get their exact logL (phys in sampler order; ll replaced in place), so
// every point that joins the live set, and so every posterior sample,
// carries the exact value.
using RecheckFn = std::function<void(const double* phys, int n, double lmin, double* ll)>;

void call_dumper(DumperFn dumper, void* ctx, Result& r, const std::vector<Point>& live, double logx, int npar) {
    if (!dumper) return;
    summarise(r, live, logx, npar);
    int ns = (int)r.dead.size(), nl = (int)live.size(), np = npar;
    // MultiNest's arrays, Fortran (column-major) layout
    std::vector<double> physlive((size_t)nl * (npar + 1)), post((size_t)std::max(ns, 1) * (npar + 2)),
        cons((size_t)4 * npar);
    for (int i = 0; i < nl; ++i) {
        for (int k = 0; k < npar; ++k) physlive[(size_t)k * nl + i] = live[i].phys[k];
        physlive[(size_t)npar * nl + i] = live[i].logl;
    }
    for (int i = 0; i < ns; ++i) {
        for (int k = 0; k < npar; ++k) post[(size_t)k * ns + i] = r.dead[i].phys[k];
        post[(size_t)npar * ns + i] = r.dead[i].logl;
        post[(size_t)(npar + 1) * ns + i] = std::exp(r.dead_logw[i] + r.dead[i].logl - r.logz);
    }
    for (int k = 0; k < npar; ++k) {
        cons[k] = r.mean[k];
        cons[npar + k] = r.sigma[k];
        cons[2 * npar + k] = r.ml[k];
        cons[3 * npar + k] = r.ml[k];  // MAP = ML for the uniform cube prior
    }
    double* pl = physlive.data();
    double* po = post.data();
    double* co = cons.data();
    double maxl = r.max_logl, logz = r.logz, ins = r.logz, err = r.logz_err;
    dumper(&ns, &nl, &np, &pl, &po, &co, &maxl, &logz, &ins, &err, ctx);
}

// Constant efficiency mode (MultiNest's ceff): the enlargement is steered
// once per bound, from that bound's acceptance, towards acceptance efr, by
// at most a factor 2 a step; as in MultiNest the bound may then shrink below
// the one that encloses the live points, which trades evidence accuracy for
// speed.  Shrinking raises the acceptance where the constrained region is
// compact (and its outermost live points stick out into thin tails), but
// lowers it where the region is a thin or curved sheet that falls out of a
// shrunken ellipsoid.  So when a step meant to raise the acceptance more
// than halved it instead, or a limit stops the steps, the direction turns
// round; the enlargement then stays near the one that accepts best.
struct Steer {
    double efr;
    double acc_prev = -1;
    bool raise_prev = false;  // the last step was meant to raise the acceptance
    int dir = -1;             // the step that should raise it: shrink (-1) or grow (+1)

    double operator()(double volfac, long acc_n, long prop_n, bool prior_draws) {
        const double acc = (double)acc_n / prop_n;
        const double expect = acc_prev * prop_n;
        if (prior_draws)
            dir = -1;  // the bound held the whole cube: only shrinking changes the draws
        else if (raise_prev && acc_n + 2 * std::sqrt(expect) < 0.5 * expect)
            dir = -dir;
        double f;
        bool raise = acc < efr;
        if (!raise) {
            // (no growing past a bound that already holds the whole cube)
            f = prior_draws ? 1.0 : std::min(acc / efr, 2.0);
            dir = -1;
        } else {
            const double g = std::max(acc / efr, 0.5);
            f = dir < 0 ? g : 1 / g;
        }
        const double next = std::min(1e3, std::max(1e-2, volfac * f));
        if (raise && next == volfac) {
            dir = -dir;
            raise = false;
        }
        acc_prev = acc;
        raise_prev = raise;
        return next;
    }
};

int sample(int ndim, int npar, BatchFn fn, void* ctx, const Options& o, DumperFn dumper, void* dctx, Result& r,
           const RecheckFn& recheck = nullptr) {
    if (ndim <= 0 || ndim > 64 || o.nlive <= ndim) {
        lsd::set_error("nested sampling needs 0 < ndim < nlive and ndim <= 64");
        return -1;
    }
    std::mt19937_64 rng((unsigned long long)o.seed * 6364136223846793005ull + 1442695040888963407ull);
    std::uniform_real_distribution<double> uni(0, 1);
    std::vector<Point> live;
    std::string ckpt = o.root.empty() ? "" : o.root + "resume.dat";
    bool resumed = o.resume && !ckpt.empty() && read_checkpoint(ckpt, live, r, ndim, npar);
    if (resumed) rng.seed(rng() ^ (unsigned long long)r.niter);  // a fresh stream after the checkpoint
    if (!resumed) {
        live.assign(o.nlive, Point());
        std::vector<double> cubes((size_t)o.nlive * ndim), phys((size_t)o.nlive * npar), ll(o.nlive);
        for (auto& c : cubes) c = uni(rng);
        fn(cubes.data(), o.nlive, phys.data(), ll.data(), ctx);
        if (recheck) recheck(phys.data(), o.nlive, -INFINITY, ll.data());
        r.nevals += o.nlive;
        for (int i = 0; i < o.nlive; ++i) {
            live[i].u.assign(cubes.begin() + (size_t)i * ndim, cubes.begin() + (size_t)(i + 1) * ndim);
            live[i].phys.assign(phys.begin() + (size_t)i * npar, phys.begin() + (size_t)(i + 1) * npar);
            live[i].logl = ll[i];
        }
    }
    const int n = o.nlive;
    const double shrink = -1.0 / n;  // log X_i - log X_{i-1}
    double logx = shrink * r.niter;
    // enlargement: without ceff, efr is MultiNest's shrinking factor and the
    // bounding volume grows by 1/efr (by at least half); with ceff, efr is
    // the target acceptance and Steer adapts the enlargement to it
    double volfac = o.ceff ? 1.5 : std::max(1.5, 1.0 / o.efr);
    Bound bound;
    if (o.ceff) bound.boot = 0;
    Steer steer{o.efr};
    const double refit_prop = std::max(100.0, 4.0 * std::max(1, n / 10) / o.efr);
    long fb_acc = 0, fb_prop = 0;
    std::vector<double> cubes, phys, ll;
    size_t next = 0, have = 0;
    int since_fit = n;
    long fit_acc = 0, fit_prop = 0;  // proposals drawn from the current bound
    long b_acc = 0, b_prop = 0;      // from the current batch
    double acc_last = 1.0;
    while (true) {
        int worst = 0;
        for (int i = 1; i < n; ++i)
            if (live[i].logl < live[worst].logl) worst = i;
        double lmin = live[worst].logl;
        // termination: the live points can raise Z by at most a fraction tol
        double lmax = -INFINITY;
        for (const auto& p : live) lmax = std::max(lmax, p.logl);
        // (MultiNest's criterion: ln(Z + dZ_live) - ln Z < tol)
        if (r.niter > n && lmax + logx < r.logz + std::log(std::expm1(o.tol))) break;
        if (o.maxiter > 0 && r.niter >= o.maxiter) break;
        // the dead point: weight width X_{i-1} - X_i
        double logw = logx + std::log1p(-std::exp(shrink));
        double logz_new = logaddexp(r.logz, logw + lmin);
        r.info = std::exp(logw + lmin - logz_new) * lmin +
                 (r.logz == -INFINITY ? 0 : std::exp(r.logz - logz_new) * (r.info + r.logz)) - logz_new;
        r.logz = logz_new;
        r.dead.push_back(live[worst]);
        r.dead_logw.push_back(logw);
        logx += shrink;
        r.niter += 1;
        // replacement from the constrained prior
        bool done = false;
        while (!done) {
            if (next >= have) {
                // The bound is refitted after n/10 replacements; in constant
                // efficiency mode also after 4 x the proposals those should take
                // at acceptance efr, so that a bound far too large or small is
                // not kept while its few replacements trickle in.
                if (bound.ells.empty() || since_fit >= std::max(1, n / 10) || (o.ceff && fit_prop >= refit_prop)) {
                    if (o.ceff && fit_prop > 0) volfac = steer(volfac, fit_acc, fit_prop, bound.logvol_cube >= 0);
                    bound.fit(live, volfac, logx, o.wrap, true, rng);
                    since_fit = 0;
                    fit_acc = fit_prop = 0;
                }
                // A batch is drawn from one bound, which grows stale as the
                // live points contract while the batch is used up (its
                // acceptance falls, and in constant efficiency mode the
                // enlargement would shrink to make up for it).  So a large
                // batch is sized to what the next refit needs, 1.5 x n/10
                // replacements at the last batch's acceptance, in powers of
                // two from 256 (a few launch shapes) up to o.batch.
                if (b_prop > 0) acc_last = (double)b_acc / b_prop;
                b_acc = b_prop = 0;
                int b = std::max(1, o.batch);
                if (b > 256) {
                    const double want = 1.5 * std::max(1, n / 10) / std::max(acc_last, 1e-4);
                    int c = 256;
                    while (c < b && c < want) c *= 2;
                    b = std::min(b, c);
                }
                cubes.resize((size_t)b * ndim);
                phys.resize((size_t)b * npar);
                ll.resize(b);
                // While the bound's part inside the cube is larger than the
                // cube itself would be, i.e. early on (Bound::logvol_cube), the
                // prior is the better proposal; either is a uniform draw over a
                // superset of the constrained region, which is all the
                // acceptance test needs.
                const bool use_cube = bound.logvol_cube >= 0;
                // kStreams independent streams seeded from the main one fill
                // the batch (candidate k from stream k % kStreams), on threads
                // for large batches; the draws do not depend on the thread count
                constexpr int kStreams = 8;
                unsigned long long seeds[kStreams];
                for (auto& sd : seeds) sd = rng();
                auto fill = [&](int st) {
                    std::mt19937_64 g(seeds[st]);
                    std::uniform_real_distribution<double> u01(0, 1);
                    double x[64];
                    for (int k = st; k < b; k += kStreams) {
                        double* c = &cubes[(size_t)k * ndim];
                        bool got = false;
                        for (int tries = 0; !use_cube && !got && tries < 1000; ++tries) {
                            bound.sample(g, x);
                            got = inside(x, ndim, o.wrap, c);
                        }
                        if (!got)
                            for (int i = 0; i < ndim; ++i) c[i] = u01(g);
                    }
                };
                if (b >= 512) {
                    std::vector<std::thread> pool;
                    for (int st = 1; st < kStreams; ++st) pool.emplace_back(fill, st);
                    fill(0);
                    for (auto& t : pool) t.join();
                } else {
                    for (int st = 0; st < kStreams; ++st) fill(st);
                }
                fn(cubes.data(), b, phys.data(), ll.data(), ctx);
                if (recheck) recheck(phys.data(), b, lmin, ll.data());
                r.nevals += b;
                next = 0;
                have = b;
            }
            size_t k = next++;
            fit_prop += 1;
            fb_prop += 1;
            b_prop += 1;
            if (ll[k] > lmin) {
                live[worst].u.assign(cubes.begin() + k * ndim, cubes.begin() + (k + 1) * ndim);
                live[worst].phys.assign(phys.begin() + k * npar, phys.begin() + (k + 1) * npar);
                live[worst].logl = ll[k];
                fit_acc += 1;
                fb_acc += 1;
                b_acc += 1;
                since_fit += 1;
                done = true;
            }
        }
        if (o.feedback && r.niter % (o.updint > 0 ? o.updint : 1000) == 0)
        {
            fprintf(stderr,
                    "nested: iteration %ld  ln Z %.4f  ln X %.2f  evaluations %ld  ellipsoids %zu  ln V %.2f  "
                    "acceptance %.2e  enlargement %.3g%s\n",
                    r.niter, r.logz, logx, r.nevals, bound.ells.size(), bound.logvol_cube,
                    fb_prop ? (double)fb_acc / fb_prop : 0.0, volfac,
                    bound.logvol_cube >= 0 ? "  (prior draws)" : "");
            fb_acc = fb_prop = 0;
        }
        if (o.updint > 0 && r.niter % o.updint == 0) {
            call_dumper(dumper, dctx, r, live, logx, npar);
            if (!ckpt.empty()) write_checkpoint(ckpt, live, r, ndim, npar);
        }
    }
    // the checkpoint keeps the state of the iteration (a resumed run goes on
    // from here), before the live points are added to the evidence
    if (!ckpt.empty()) write_checkpoint(ckpt, live, r, ndim, npar);
    // remaining live points share the last volume
    for (const auto& p : live) {
        double lw = logx - std::log((double)n);
        double lz = logaddexp(r.logz, lw + p.logl);
        r.info = std::exp(lw + p.logl - lz) * p.logl + std::exp(r.logz - lz) * (r.info + r.logz) - lz;
        r.logz = lz;
    }
    r.logz_err = std::sqrt(std::max(r.info, 0.0) / n);
    summarise(r, live, logx, npar);
    if (o.outfile && !o.root.empty()) {
        FILE* f = fopen((o.root + ".txt").c_str(), "w");
        if (f) {  // MultiNest's <root>.txt: weight, -2 logL, parameters
            for (size_t i = 0; i < r.dead.size(); ++i) {
                fprintf(f, "%.10e %.10e", std::exp(r.dead_logw[i] + r.dead[i].logl - r.logz), -2 * r.dead[i].logl);
                for (double v : r.dead[i].phys) fprintf(f, " %.10e", v);
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }
    call_dumper(dumper, dctx, r, live, logx, npar);
    return 0;
}

// the MultiNest callback evaluated one candidate at a time
struct SerialCtx {
    void (*loglike)(double*, int*, int*, double*, void*);
    void* ctx;
    int ndim, npar;
};

void serial_batch(const double* cubes, int n, double* phys, double* lnew, void* c) {
    SerialCtx* s = (SerialCtx*)c;
    std::vector<double> cube(s->npar);
    for (int i = 0; i < n; ++i) {
        std::fill(cube.begin(), cube.end(), 0.0);
        std::copy(cubes + (size_t)i * s->ndim, cubes + (size_t)(i + 1) * s->ndim, cube.begin());
        s->loglike(cube.data(), &s->ndim, &s->npar, &lnew[i], s->ctx);
        std::copy(cube.begin(), cube.end(), phys + (size_t)i * s->npar);
    }
}

struct BatchCtx {
    lensed* l;
    size_t s;
    int ndim;
};

void lensed_batch(const double* cubes, int n, double* phys, double* lnew, void* c) {
    BatchCtx* b = (BatchCtx*)c;
    if (lensed_loglike_batch(b->l, b->s, cubes, (size_t)n, phys, lnew))
        for (int i = 0; i < n; ++i) lnew[i] = -DBL_MAX;
}

Options to_options(const lensed_ns_options* in, int npar) {
    Options o;
    if (!in) return o;
    o.nlive = in->nlive > 0 ? in->nlive : 300;
    o.tol = in->tol > 0 ? in->tol : 0.1;
    o.efr = in->efr > 0 ? in->efr : 0.8;
    o.ceff = in->ceff;
    o.batch = in->batch > 0 ? in->batch : 1;
    o.seed = in->seed;
    o.updint = in->updint;
    o.maxiter = in->maxiter;
    o.resume = in->resume;
    o.outfile = in->outfile;
    o.feedback = in->feedback;
    if (in->wrap) o.wrap.assign(in->wrap, in->wrap + npar);
    if (in->root) o.root = in->root;
    return o;
}

void fill_result(const Result& r, lensed_ns_result* out, double* mean, double* sigma, double* ml, int npar) {
    if (out) {
        out->logz = r.logz;
        out->logz_err = r.logz_err;
        out->info = r.info;
        out->max_loglike = r.max_logl;
        out->nevals = r.nevals;
        out->niter = r.niter;
    }
    for (int k = 0; k < npar; ++k) {
        if (mean) mean[k] = r.mean[k];
        if (sigma) sigma[k] = r.sigma[k];
        if (ml) ml[k] = r.ml[k];
    }
}

}  // namespace

extern "C" {

int lensed_nested_sample(int ndim, int npar, lensed_batch_fn fn, void* ctx, const lensed_ns_options* opts,
                         double* mean, double* sigma, double* ml, lensed_ns_result* result) {
    Options o = to_options(opts, npar);
    Result r;
    if (sample(ndim, npar, (BatchFn)fn, ctx, o, nullptr, nullptr, r)) return -1;
    fill_result(r, result, mean, sigma, ml, npar);
    return 0;
}

int lensed_nested_run(lensed* l, size_t s, const lensed_ns_options* opts, double* mean, double* sigma, double* ml,
                      lensed_ns_result* result) {
    int ndim = (int)lensed_ndims(l, s), npar = (int)lensed_npars(l, s);
    BatchCtx b{l, s, ndim};
    Options o = to_options(opts, npar);
    if (opts && opts->batch < 0) {  // the batch shape from the autotuner
        size_t k = 0;
        if (lensed_autotune_batch(l, s, &k, nullptr, nullptr, nullptr)) return -1;
        o.batch = (int)k;
    }
    if (!opts || !opts->wrap) {  // the parameters' own periodic flags (sampler order), as Lensed passes them
        const auto& sys = l->sys[s];
        o.wrap.resize(npar);
        for (int i = 0; i < npar; ++i) o.wrap[i] = sys.pars[sys.pmap[i]].wrap;
    }
    Result r;
    // accepted points of an emulated launch get their exact logL
    RecheckFn recheck = [l, s](const double* phys, int n, double lmin, double* ll) {
        if (lsd::recheck_emulated(l, s, phys, n, lmin, ll))
            for (int i = 0; i < n; ++i) ll[i] = -DBL_MAX;
    };
    if (sample(ndim, npar, lensed_batch, &b, o, s == 0 ? dumper : nullptr, l, r, recheck)) return -1;
    fill_result(r, result, mean, sigma, ml, npar);
    return 0;
}

void run(int IS, int mmodal, int ceff, int nlive, double tol, double efr, int ndims, int nPar, int nClsPar,
         int maxModes, int updInt, double Ztol, char root[], int seed, int* pWrap, int fb, int resume, int outfile,
         int initMPI, double logZero, int maxiter, void (*LogLike)(double*, int*, int*, double*, void*),
         void (*Dumper)(int*, int*, int*, double**, double**, double**, double*, double*, double*, double*, void*),
         void* context) {
    (void)IS; (void)mmodal; (void)nClsPar; (void)maxModes; (void)Ztol; (void)initMPI; (void)logZero;
    Options o;
    o.nlive = nlive;
    o.tol = tol;
    o.efr = efr;
    o.ceff = ceff;
    o.updint = updInt;
    o.seed = seed < 0 ? (int)std::random_device()() : seed;
    o.maxiter = maxiter;
    o.resume = resume;
    o.outfile = outfile;
    o.feedback = fb;
    o.batch = 1;
    if (pWrap) o.wrap.assign(pWrap, pWrap + nPar);
    if (root) o.root = root;
    SerialCtx sc{LogLike, context, ndims, nPar};
    Result r;
    if (sample(ndims, nPar, serial_batch, &sc, o, Dumper, context, r))
        fprintf(stderr, "lensed: nested sampling failed\n");
}

}  // extern "C"
