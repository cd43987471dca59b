// Tabulated angular factor of the elliptical power-law lens.
//
// The EPL deflection (objects/epl.cl, epl_plus_shear.cl) is
// n |u|^(1-t) Omega(phi) with Omega the tenth-order series in e^{i phi}.
// Omega has real series coefficients, so Omega_x is odd in u_x and even in
// u_y and Omega_y the reverse: Omega(u) = (sgn u_x Ox(psi), sgn u_y Oy(psi))
// with psi the first-quadrant angle of (|u_x|, |u_y|).  The quadrant is
// indexed by the diamond angle p = |u_y| / (|u_x| + |u_y|) in [0, 1] (one
// division, no octant choice), and with s = |u_x| + |u_y| the deflection's
// |u|^(1-t) Omega = s^(1-t) W(p), W = g^h Omega, g = p^2 + (1-p)^2,
// h = (1-t)/2: the radial power needs log2(s) only, the rest is in the
// table.  LENSED_EPL_NT intervals; per interval the table holds the cubic
// Hermite polynomial in the local coordinate f in [0, 1) for each
// component, as float4 (a0, a1, a2, a3):
//     W(f) = a0 + f (a1 + f (a2 + f a3)),
// and one more row holds the value at p = 1 (a1 = a2 = a3 = 0), so the row
// index floor(p NT) needs no clamp.  Rows are stored per component,
// LENSED_EPL_AT in desc.h.  The table of a candidate is built from the
// series (or by the emulator, epl_series.h / emulator.hip) and staged in
// LDS by the render kernel.
//
// The lens-frame position u = F (x - c) and the ray position with the
// lens's linear (external shear) deflection already taken off, x - S (x - c),
// are linear in the ray position x = p + q_n (pixel centre plus quadrature
// node), so the render kernel forms both as a pixel part plus a node part,
// the node parts computed once per node and candidate (lensed_epl_pre*).
// The generated ray code then subtracts only the tabulated EPL deflection
// (and any other lens's) from the second pair: the shear costs nothing per
// ray.
#pragma once

// LENSED_EPL_NT (intervals), LENSED_EPL_ROWS and LENSED_EPL_TABLE_F4
// (float4 per candidate) come from desc.h.

#ifdef LCL_GPU
#pragma clang force_cuda_host_device begin
#endif
namespace lcl {

inline float lensed_fract(float s) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fractf(s);
#else
    return s - floorf(s);
#endif
}

// W at the diamond angle of u, with the signs of u (s = |u_x| + |u_y|)
inline float2 lensed_epl_w_tab(float2 u, float s, const float4* tab) {
    // p NT in [0, NT]; at u = 0 it is 0 * inf = NaN, which fmin turns into
    // NT (the end row: any finite W, since the power factor s^(1-t) decides)
    float x = fminf(fabs(u.y) * ((float)LENSED_EPL_NT / s), (float)LENSED_EPL_NT);
    int i = (int)x;
    float f = lensed_fract(x);
    const float4 cx = tab[LENSED_EPL_AT(0, i)];
    const float4 cy = tab[LENSED_EPL_AT(1, i)];
    float wx = cx.x + f * (cx.y + f * (cx.z + f * cx.w));
    float wy = cy.x + f * (cy.y + f * (cy.z + f * cy.w));
    return float2(copysign(wx, u.x), copysign(wy, u.y));
}

// the pair of the offset r = x - c at position x: (F r, x) without shear,
// (F r, x - S r) with it; both are linear in (r, x), so the pixel part
// (r = p - c, x = p) and the node part (r = x = q_n) add up to the ray's.
// D is the object's data struct (fields c, fwd, bwd, [shear], hexp; bwd
// carries the normalisation).
template <class D>
inline float4 lensed_epl_pre(const D* self, float2 r, float2 x) {
    float2 u = mv22(self->fwd, r);
    return float4(u.x, u.y, x.x, x.y);
}

template <class D>
inline float4 lensed_epl_pre_shear(const D* self, float2 r, float2 x) {
    float2 u = mv22(self->fwd, r), g = x - mv22(self->shear, r);
    return float4(u.x, u.y, g.x, g.y);
}

// the EPL deflection (without the shear) from the pair, in the lens frame
// (v, with a = bwd v) and in the image frame: s^(2 hexp) W with
// s = |u_x| + |u_y| (at u = 0: 0 for t < 1, else not finite, as the series)
template <class D>
inline float2 lensed_epl_table_vector_pre(const D* self, float4 pre, const float4* tab) {
    float2 u = float2(pre.x, pre.y);
    float s = fabs(u.x) + fabs(u.y);
    return exp2((2.0f * self->hexp) * log2(s)) * lensed_epl_w_tab(u, s, tab);
}

template <class D>
inline float2 lensed_epl_table_deflection_pre(const D* self, float4 pre, const float4* tab) {
    return mv22(self->bwd, lensed_epl_table_vector_pre(self, pre, tab));
}

// Lens-frame form of the ray-shooting step for one source behind the lens
// that describes itself in a frame u = T (y - c_s) (frame_map T linear):
// with y = x - S (x - c) - bwd v the source sees
//     u = T (x - S (x - c) - c_s) - (T bwd) v,
// the first term formed from pixel and node parts like the pair (pre.zw),
// the columns c0, c1 of T bwd the same for every ray of a candidate.

// The guard of the lens-frame form: each component of v is clamped to
// [-1e10, 1e10], a NaN going to -1e10 (fmax returns its other operand).
// The only non-finite v of finite parameters is at u = 0 with t >= 1 (where
// the image-frame guard rejects the ray); such a ray, and any deflected by
// more than ~1e10 |bwd| pixels, ends ~1e10 pixels out, where every framed
// source's brightness is 0 in FP32 -- as at the 1e10 displacement the
// image-frame form gives a rejected ray.  At u = 0 with t < 1, v = 0
// exactly, as in the reference.
inline float2 lensed_clamp_deflection(float2 v) {
    return float2(fminf(fmaxf(v.x, -1E10f), 1E10f), fminf(fmaxf(v.y, -1E10f), 1E10f));
}

// (n0, n1 = -c0, -c1: the negation stays with the per-candidate columns)
inline float2 lensed_frame_trace(float4 pre, float2 n0, float2 n1, float2 v) {
    return float2(fma(v.y, n1.x, fma(v.x, n0.x, pre.z)), fma(v.y, n1.y, fma(v.x, n0.y, pre.w)));
}

// ------------------------------------------------------------------ culling
// A pixel whose rays all land where the source's brightness is exactly 0
// need not be traced: its value is 0 either way, bit for bit.  For the
// pixel centre x_c (the pixel part of the pair, `base`) and any of its rays
// x = x_c + q, |q| <= rho, with y = x - S (x - c) - bwd v(u), u = F (x - c),
// v = |u|^(1-t) Omega(psi):
//     |y(x) - y(x_c)| <= |I - S| rho + |bwd| sup|dv/du| |F| rho,
//     |dv/du| <= |u|^(-t) (|1-t| A + B),  A = sum |C_k|, B = sum (2k+1) |C_k|
// (the radial and the angular derivative of the series), the sup taken over
// |u| within |F| rho of |u_c| (a pixel whose rays can reach u = 0 is never
// skipped).  Norms are Frobenius norms, which bound the operator norms.  In
// the source's frame u_s = T (y - c_s) every ray is then within |T| times
// that of the centre ray's u_s, and the pixel is skipped when even the
// nearest is beyond the source's cutoff (frame_cutoff) by a margin of 1/16
// of a frame unit: the FP32 rounding of the traced positions and of the
// bound, and the table's interpolation error, come to ~1e-5 of a unit.
template <class D>
inline float lensed_epl_slope_bound(const D* self) {
    float a = 1.0f, b = 1.0f;
    for (int k = 0; k < 10; ++k) {
        const float ck = fabsf(self->coef[k]);
        a += ck;
        b += (float)(2 * k + 3) * ck;
    }
    return fabsf(2.0f * self->hexp) * a + b;
}

inline float lensed_frob(mat22 m) { return sqrtf(m.x * m.x + m.y * m.y + m.z * m.z + m.w * m.w); }

// the candidate's constants, once (set_params): cc[0] = (n0, n1), cc[1] =
// (|F|, |bwd| (|1-t| A + B) |T|, |I - S| |T|, cutoff + 1/16), cc[2].x = t
template <class D>
inline void lensed_frame_cull_consts(const D* self, float2 n0, float2 n1, float tnorm, float lnorm, float cutoff,
                                     float4* cc) {
    cc[0] = float4(n0.x, n0.y, n1.x, n1.y);
    cc[1] = float4(lensed_frob(self->fwd), lensed_frob(self->bwd) * lensed_epl_slope_bound(self) * tnorm,
                   lnorm * tnorm, cutoff + 0.0625f);
    cc[2] = float4(1.0f - 2.0f * self->hexp, 0.0f, 0.0f, 0.0f);
}

// per pixel: whether all its rays provably miss the source
template <class D>
inline bool lensed_frame_cull(const D* self, float4 base, const float4* tab, const float4* cc, float rho) {
    const float4 a = cc[0], b = cc[1];
    const float t = cc[2].x;
    const float2 us = lensed_frame_trace(base, float2(a.x, a.y), float2(a.z, a.w),
                                         lensed_clamp_deflection(lensed_epl_table_vector_pre(self, base, tab)));
    const float rc = sqrtf(base.x * base.x + base.y * base.y);
    const float fr = b.x * rho;
    const float rlo = rc - fr;
    const float pw = exp2(-t * log2(t > 0.0f ? rlo : rc + fr));  // sup of |u|^(-t) over [rlo, rc + fr]
    const float d = rho * (b.z + b.y * pw * b.x);                 // |T| times the bound on |y - y_c|
    return rlo > 0.0f && sqrtf(us.x * us.x + us.y * us.y) - d > b.w;
}

}  // namespace lcl
#ifdef LCL_GPU
#pragma clang force_cuda_host_device end
#endif
