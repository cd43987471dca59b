// The EPL angular factor in FP64: Omega(psi) = sum_{k=0..10} C_k e^{i(2k+1) psi}
// with C_0 = 1, C_k = C_{k-1} (-f (2k - T)/(2k + T)), f = (1-q)/(1+q), T = 2-t
// (the tenth-order series of objects/epl.cl), and its psi derivative.
// Table nodes are taken in the octant variable tau of epl_table.h.
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

// octant `half` (0: psi = atan tau, 1: psi = pi/2 - atan tau); the
// derivative is per unit of the interval coordinate f = tau * NT - i
__device__ inline void lensed_epl_node(double q, double t, int half, double tau, double& ox, double& oy,
                                       double& dx, double& dy)
{
    double a = atan(tau);
    double psi = half ? 1.57079632679489661923 - a : a;
    double gx, gy;
    lensed_epl_omega(q, t, psi, ox, oy, gx, gy);
    double dpsi = (half ? -1.0 : 1.0) / (1 + tau * tau) / LENSED_EPL_NT;
    dx = gx * dpsi;
    dy = gy * dpsi;
}

// cubic Hermite coefficients (a0, a1, a2, a3) of one interval from the
// values and f-derivatives at its ends
__device__ inline float4 lensed_hermite(double v0, double d0, double v1, double d1)
{
    return make_float4((float)v0, (float)d0, (float)(3 * (v1 - v0) - 2 * d0 - d1), (float)(2 * (v0 - v1) + d0 + d1));
}
