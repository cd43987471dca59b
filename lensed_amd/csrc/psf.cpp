// PSF factorisation for the separable convolution path.
//
// The reference convolves directly (kernel/lensed.cl:96-98).  When the PSF
// is numerically a sum of a few outer products, psf = sum_k c_k r_k^T, the
// same clamp-to-edge convolution is a sum of row passes followed by column
// passes (the clamped extension of the image is separable in x and y, and
// convolution is linear), at (PW + PH) instead of PW*PH multiply-adds per
// pixel and rank.  The rank is the smallest whose truncation error is below
// FP32 rounding of the PSF itself (1e-7 of its Frobenius norm), so the two
// paths agree to FP32 rounding; otherwise the direct path is kept.
#include "psf.h"

#include <algorithm>
#include <cmath>

namespace lsd {

// one-sided Jacobi SVD of an m x n matrix (row-major), m, n <= 64:
// a = U diag(s) V^T, singular values sorted descending
static void jacobi_svd(int m, int n, const std::vector<double>& a, std::vector<double>& s,
                       std::vector<double>& u, std::vector<double>& v) {
    std::vector<double> w(a);  // columns of w converge to U * s
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < m; ++i) {
                    double x = w[(size_t)i * n + p], y = w[(size_t)i * n + q];
                    alpha += x * x;
                    beta += y * y;
                    gamma += x * y;
                }
                if (gamma == 0.0) continue;
                off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
                double zeta = (beta - alpha) / (2 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                double c = 1 / std::sqrt(1 + t * t), sn = c * t;
                for (int i = 0; i < m; ++i) {
                    double x = w[(size_t)i * n + p], y = w[(size_t)i * n + q];
                    w[(size_t)i * n + p] = c * x - sn * y;
                    w[(size_t)i * n + q] = sn * x + c * y;
                }
                for (int i = 0; i < n; ++i) {
                    double x = v[(size_t)i * n + p], y = v[(size_t)i * n + q];
                    v[(size_t)i * n + p] = c * x - sn * y;
                    v[(size_t)i * n + q] = sn * x + c * y;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<int> order(n);
    std::vector<double> norms(n);
    for (int j = 0; j < n; ++j) {
        double s2 = 0;
        for (int i = 0; i < m; ++i) s2 += w[(size_t)i * n + j] * w[(size_t)i * n + j];
        norms[j] = std::sqrt(s2);
        order[j] = j;
    }
    std::sort(order.begin(), order.end(), [&](int x, int y) { return norms[x] > norms[y]; });
    s.assign(n, 0.0);
    u.assign((size_t)m * n, 0.0);
    std::vector<double> vs((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        int j = order[k];
        s[k] = norms[j];
        for (int i = 0; i < m; ++i) u[(size_t)i * n + k] = norms[j] > 0 ? w[(size_t)i * n + j] / norms[j] : 0.0;
        for (int i = 0; i < n; ++i) vs[(size_t)i * n + k] = v[(size_t)i * n + j];
    }
    v.swap(vs);
}

PsfFactors factor_psf(const float* psf, int pw, int ph) {
    PsfFactors out;
    if (pw <= 0 || ph <= 0) return out;
    std::vector<double> a((size_t)pw * ph);
    double fro = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        a[i] = psf[i];
        fro += a[i] * a[i];
    }
    fro = std::sqrt(fro);
    if (fro == 0) return out;
    std::vector<double> s, u, v;
    jacobi_svd(ph, pw, a, s, u, v);  // a is ph rows x pw columns
    int n = pw;
    std::vector<double> tail(n + 1, 0.0);
    for (int k = n - 1; k >= 0; --k) tail[k] = tail[k + 1] + s[k] * s[k];
    int rank = n;
    for (int k = 1; k <= n; ++k)
        if (std::sqrt(tail[k]) <= 1e-7 * fro) {
            rank = k;
            break;
        }
    // worth it only if clearly fewer multiply-adds than the direct sum
    if (rank > kMaxRank || 2 * rank * (pw + ph) > pw * ph) return out;
    out.rank = rank;
    for (int k = 0; k < rank; ++k) {
        double sq = std::sqrt(s[k]);
        // column factor (length ph) then row factor (length pw)
        double sign = 0;
        for (int i = 0; i < ph; ++i) sign += u[(size_t)i * n + k];
        sign = sign < 0 ? -1.0 : 1.0;
        for (int i = 0; i < ph; ++i) out.factors.push_back((float)(sign * sq * u[(size_t)i * n + k]));
        for (int i = 0; i < pw; ++i) out.factors.push_back((float)(sign * sq * v[(size_t)i * n + k]));
    }
    return out;
}

}  // namespace lsd
