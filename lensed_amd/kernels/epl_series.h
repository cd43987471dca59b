// The EPL angular factor in FP64: Omega(psi) = sum_{k=0..10} C_k e^{i(2k+1) psi}
// with C_0 = 1, C_k = C_{k-1} (-f (2k - T)/(2k + T)), f = (1-q)/(1+q), T = 2-t
// (the tenth-order series of objects/epl.cl), and its psi derivative.
// Table nodes are taken in the diamond angle p of epl_table.h.
#pragma once

__device__ inline void lensed_epl_omega(double q, double t, double psi, double& ox, double& oy, double& dx,
                                        double& dy)
{
    const double f = (1 - q) / (1 + q), T = 2 - t;
    double c = 1.0;
    ox = oy = dx = dy = 0.0;
    for (int k = 0; k <= 10; ++k) {
        if (k > 0) c *= -f * (2 * k - T) / (2 * k + T);
        double m = 2 * k + 1, sn, cs;
        sincos(m * psi, &sn, &cs);
        ox += c * cs;
        oy += c * sn;
        dx -= c * m * sn;
        dy += c * m * cs;
    }
}

// The tabulated factor at diamond angle p = sin / (cos + sin) of psi in the
// first quadrant: W(p) = g(p)^h Omega(psi(p)) with g = p^2 + (1-p)^2
// (|u|^2 = (|u_x| + |u_y|)^2 g) and h = (1-t)/2, so that the deflection's
// |u|^(1-t) Omega is (|u_x| + |u_y|)^(1-t) W.  psi = atan2(p, 1-p), dpsi/dp
// = 1/g; the derivative is per unit of the interval coordinate f = p NT - i.
__device__ inline void lensed_epl_node(double q, double t, double p, double& wx, double& wy, double& dx,
                                       double& dy)
{
    const double g = p * p + (1 - p) * (1 - p), h = 0.5 * (1 - t);
    const double gh = pow(g, h);
    double ox, oy, gx, gy;
    lensed_epl_omega(q, t, atan2(p, 1 - p), ox, oy, gx, gy);
    const double lg = h * (4 * p - 2) / g;  // d log(g^h) / dp
    wx = gh * ox;
    wy = gh * oy;
    dx = gh * (lg * ox + gx / g) / LENSED_EPL_NT;
    dy = gh * (lg * oy + gy / g) / LENSED_EPL_NT;
}

// cubic Hermite coefficients (a0, a1, a2, a3) of one interval from the
// values and f-derivatives at its ends
__device__ inline float4 lensed_hermite(double v0, double d0, double v1, double d1)
{
    return make_float4((float)v0, (float)d0, (float)(3 * (v1 - v0) - 2 * d0 - d1), (float)(2 * (v0 - v1) + d0 + d1));
}
