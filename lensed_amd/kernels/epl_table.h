// Tabulated angular factor of the elliptical power-law lens.
//
// The EPL deflection (objects/epl.cl, epl_plus_shear.cl) is
// n |u|^(1-t) Omega(phi) with Omega the tenth-order series in e^{i phi}.
// Omega has real series coefficients, so Omega_x is odd in u_x and even in
// u_y and Omega_y the reverse: Omega(u) = (sgn u_x Ox(psi), sgn u_y Oy(psi))
// with psi the first-quadrant angle of (|u_x|, |u_y|).  The quadrant is
// covered by two octants in the variable tau = min(|u_x|, |u_y|) /
// max(|u_x|, |u_y|) in [0, 1]: half 0 is psi = atan(tau), half 1 is
// psi = pi/2 - atan(tau).  Each half has LENSED_EPL_NT intervals; per
// interval the table holds the cubic Hermite polynomial in the local
// coordinate f in [0, 1] for each component, as two float4 (a0, a1, a2, a3):
//     O(f) = a0 + f (a1 + f (a2 + f a3)).
// The table of a candidate is built from the series (or by the emulator)
// and staged in LDS by the render kernel.
#pragma once

// LENSED_EPL_NT (intervals per octant) and LENSED_EPL_TABLE_F4 (float4 per
// candidate) come from desc.h.

#ifdef LCL_GPU
#pragma clang force_cuda_host_device begin
#endif
namespace lcl {

inline float2 lensed_epl_omega_tab(float2 u, float uu, const float4* tab) {
    float au = fabs(u.x), av = fabs(u.y);
    float lo = fmin(au, av), hi = fmax(au, av);
    float tau = uu > 0 ? lo / hi : 0.0f;
    float s = tau * (float)LENSED_EPL_NT;
    int i = min((int)s, LENSED_EPL_NT - 1);
    float f = s - (float)i;
    int half = av > au ? 1 : 0;
    const float4 cx = tab[2 * (half * LENSED_EPL_NT + i)];
    const float4 cy = tab[2 * (half * LENSED_EPL_NT + i) + 1];
    float ox = cx.x + f * (cx.y + f * (cx.z + f * cx.w));
    float oy = cy.x + f * (cy.y + f * (cy.z + f * cy.w));
    return float2(copysign(ox, u.x), copysign(oy, u.y));
}

// the epl / epl_plus_shear deflection with the tabulated angular factor;
// D is the object's data struct (fields c, fwd, bwd, [shear], hexp, norm)
template <class D>
inline float2 lensed_epl_table_deflection(const D* self, float2 x, const float4* tab) {
    float2 d = x - self->c;
    float2 u = mv22(self->fwd, d);
    float uu = dot(u, u);
    float2 om = lensed_epl_omega_tab(u, uu, tab);
    return mv22(self->bwd, (self->norm * powr(uu, self->hexp)) * om);
}

template <class D>
inline float2 lensed_epl_table_deflection_shear(const D* self, float2 x, const float4* tab) {
    float2 d = x - self->c;
    float2 u = mv22(self->fwd, d);
    float uu = dot(u, u);
    float2 om = lensed_epl_omega_tab(u, uu, tab);
    return mv22(self->bwd, (self->norm * powr(uu, self->hexp)) * om) + mv22(self->shear, d);
}

}  // namespace lcl
#ifdef LCL_GPU
#pragma clang force_cuda_host_device end
#endif
