/*
 * Type-generic body of the oracle; oracle.c includes it once with
 * REAL = double (suffix _d) and once with REAL = float (suffix _f).
 * Math calls go through <tgmath.h>, so the float pass evaluates in float.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.c).
 */

typedef struct { REAL x, y; } SFX(v2);

static SFX(v2) SFX(mk)(REAL x, REAL y) { SFX(v2) r; r.x = x; r.y = y; return r; }

/* mv22 of kernel/constants.cl:35-38 with mat22 = (m0, m1, m2, m3) */
static SFX(v2) SFX(mv)(const REAL m[4], SFX(v2) v) { return SFX(mk)(m[0] * v.x + m[1] * v.y, m[2] * v.x + m[3] * v.y); }

typedef struct {
    int kind;
    REAL cx, cy;          /* centre */
    REAL m[4], w[4], g[4];/* rotation, inverse rotation, shear */
    REAL a0, a1, a2, a3;  /* object constants, meaning per kind */
} SFX(obj);

static const REAL SFX(DEG2RAD) = (REAL)0.0174532925199432957692369076848861271344287188854;
static const REAL SFX(LN10) = (REAL)2.3025850929940456840179914546843642076011014886288;
static const REAL SFX(LNPI) = (REAL)1.1447298858494001741434273513530587116472948129153;
static const REAL SFX(PI_) = (REAL)3.1415926535897932384626433832795028841971693993751;

static void SFX(rot)(SFX(obj)* o, REAL pa) {
    REAL c = cos(pa * SFX(DEG2RAD)), s = sin(pa * SFX(DEG2RAD));
    o->m[0] = c; o->m[1] = s; o->m[2] = -s; o->m[3] = c;
    o->w[0] = c; o->w[1] = -s; o->w[2] = s; o->w[3] = c;
}

static void SFX(shear)(SFX(obj)* o, REAL g1, REAL g2) {
    o->g[0] = g1; o->g[1] = g2; o->g[2] = g2; o->g[3] = -g1;
}

/* Sersic b(n): objects/sersic.cl:32-33 (MacArthur, Courteau & Holtzman 2003) */
static REAL SFX(sersic_b)(REAL n) {
    if (n > (REAL)0.36)
        return 2 * n - (REAL)1 / 3 + (REAL)4 / (405 * n) + (REAL)46 / (25515 * (n * n)) +
               (REAL)131 / (1148175 * (n * n * n)) - (REAL)2194697 / ((REAL)30690717750.0 * (n * n * n * n));
    return (REAL)0.01945 - (REAL)0.8902 * n + (REAL)10.95 * (n * n) - (REAL)19.67 * (n * n * n) +
           (REAL)13.43 * (n * n * n * n);
}

/* set() of each object: objects/<name>.cl, the set function */
static void SFX(set)(SFX(obj)* o, int kind, const REAL* p) {
    memset(o, 0, sizeof *o);
    o->kind = kind;
    switch (kind) {
    case ORC_SIS: case ORC_NSIS: case ORC_POINT_MASS:  /* sis.cl:25-32, nsis.cl:30-40, point_mass.cl:25-32 */
        o->cx = p[0]; o->cy = p[1]; o->a0 = p[2]; o->a1 = kind == ORC_NSIS ? p[3] : 0;
        if (kind == ORC_POINT_MASS) o->a0 = p[2] * p[2];
        break;
    case ORC_SIS_SHEAR:  /* sis_plus_shear.cl:30-40 */
        o->cx = p[0]; o->cy = p[1]; o->a0 = p[2]; SFX(shear)(o, p[3], p[4]);
        break;
    case ORC_SIE: case ORC_SIE_SHEAR: {  /* sie.cl:46-64, sie_plus_shear.cl:49-70 */
        REAL q = p[3];
        o->cx = p[0]; o->cy = p[1]; SFX(rot)(o, p[4]);
        if (kind == ORC_SIE_SHEAR) SFX(shear)(o, p[5], p[6]);
        o->a0 = q * q; o->a1 = sqrt(1 - q * q); o->a2 = p[2] * sqrt(q) / sqrt(1 - q * q);
        break;
    }
    case ORC_NSIE: {  /* nsie.cl:48-69 */
        REAL q = p[4];
        o->cx = p[0]; o->cy = p[1]; SFX(rot)(o, p[5]);
        o->a3 = p[3]; o->a0 = q * q; o->a1 = sqrt(1 - q * q); o->a2 = p[2] * sqrt(q) / sqrt(1 - q * q);
        break;
    }
    case ORC_EPL: case ORC_EPL_SHEAR: {  /* epl.cl:70-93, epl_plus_shear.cl:78-104 */
        REAL r = p[2], t = p[3], q = p[4];
        REAL c = cos(p[5] * SFX(DEG2RAD)), s = sin(p[5] * SFX(DEG2RAD));
        REAL k = 1 / r / sqrt(q);
        o->cx = p[0]; o->cy = p[1];
        o->m[0] = k * q * c; o->m[1] = k * q * s; o->m[2] = -k * s; o->m[3] = k * c;
        o->w[0] = c; o->w[1] = -s; o->w[2] = s; o->w[3] = c;
        if (kind == ORC_EPL_SHEAR) SFX(shear)(o, p[6], p[7]);
        o->a0 = t; o->a1 = (1 - q) / (1 + q); o->a2 = 2 * r * sqrt(q) / (1 + q);
        break;
    }
    case ORC_SERSIC: case ORC_SERSIC_OLD: {  /* sersic.cl:29-47, sersic-old.cl:29-45 */
        REAL r = p[2], mag = p[3], n = p[4], q = p[5];
        REAL b = kind == ORC_SERSIC ? SFX(sersic_b)(n) : (REAL)1.9992 * n - (REAL)0.3271;
        REAL c = cos(p[6] * SFX(DEG2RAD)), s = sin(p[6] * SFX(DEG2RAD)), sq = sqrt(q);
        o->cx = p[0]; o->cy = p[1];
        o->m[0] = q * c / sq; o->m[1] = q * s / sq; o->m[2] = -s / sq; o->m[3] = c / sq;
        o->a0 = -(REAL)0.4 * mag * SFX(LN10) + 2 * n * log(b) - SFX(LNPI) - 2 * log(r) - log(tgamma(2 * n + 1));
        o->a1 = log(b) - log(r) / n;
        o->a2 = (REAL)0.5 / n;
        break;
    }
    case ORC_GAUSS: case ORC_EXPONENTIAL: case ORC_DEVAUC: {  /* gauss.cl:28-41, exponential.cl:27-43, devauc.cl:32-48 */
        REAL r = p[2], mag = p[3], q = p[4];
        REAL c = cos(p[5] * SFX(DEG2RAD)), s = sin(p[5] * SFX(DEG2RAD));
        REAL lum = exp(-(REAL)0.4 * mag * SFX(LN10));
        o->cx = p[0]; o->cy = p[1];
        o->m[0] = q * c; o->m[1] = q * s; o->m[2] = -s; o->m[3] = c;
        if (kind == ORC_GAUSS) { o->a0 = r * r; o->a1 = lum * (REAL)0.5 / SFX(PI_) / (r * r) / q; }
        if (kind == ORC_EXPONENTIAL) { o->a0 = r; o->a1 = lum * (REAL)0.5 / SFX(PI_) / r / r / q; }
        /* de Vaucouleurs: b^8/8! with b = 7.6692494425008039044 */
        if (kind == ORC_DEVAUC) { o->a0 = r; o->a1 = lum / SFX(PI_) / r / r / q * (REAL)296.826303766893; }
        break;
    }
    case ORC_SKY:  /* sky.cl:21-25 */
        o->a0 = p[0]; o->a1 = p[1]; o->a2 = p[2];
        break;
    }
}

/* EPL angular series, epl.cl:47-61: tenth order in e^{2 i phi} */
static SFX(v2) SFX(epl_series)(REAL t, REAL f, REAL phi, int nterms) {
    const REAL T = 2 - t;
    REAL c = cos(phi), s = sin(phi), c2 = cos(2 * phi), s2 = sin(2 * phi);
    SFX(v2) A = SFX(mk)(c, s), a = A;
    for (int k = 1; k <= nterms; ++k) {
        REAL fac = -f * (2 * k - T) / (2 * k + T);
        SFX(v2) RA = SFX(mk)(c2 * A.x - s2 * A.y, s2 * A.x + c2 * A.y);
        A = SFX(mk)(fac * RA.x, fac * RA.y);
        a.x += A.x; a.y += A.y;
    }
    return a;
}

static SFX(v2) SFX(deflection)(const SFX(obj)* o, SFX(v2) x) {
    SFX(v2) d = SFX(mk)(x.x - o->cx, x.y - o->cy);
    switch (o->kind) {
    case ORC_SIS: case ORC_SIS_SHEAR: {  /* sis.cl:19-23, sis_plus_shear.cl:21-28 */
        REAL l = sqrt(d.x * d.x + d.y * d.y);
        SFX(v2) a = SFX(mk)(o->a0 * d.x / l, o->a0 * d.y / l);
        if (o->kind == ORC_SIS_SHEAR) { SFX(v2) g = SFX(mv)(o->g, d); a.x += g.x; a.y += g.y; }
        return a;
    }
    case ORC_NSIS: {  /* nsis.cl:21-28 */
        REAL k = o->a0 / (o->a1 + sqrt(d.x * d.x + d.y * d.y));
        return SFX(mk)(k * d.x, k * d.y);
    }
    case ORC_POINT_MASS: {  /* point_mass.cl:18-23 */
        REAL k = o->a0 / (d.x * d.x + d.y * d.y);
        return SFX(mk)(k * d.x, k * d.y);
    }
    case ORC_SIE: case ORC_SIE_SHEAR: {  /* sie.cl:27-44, sie_plus_shear.cl:30-47 */
        SFX(v2) y = SFX(mv)(o->m, d);
        REAL r = o->a1 / sqrt(o->a0 * y.x * y.x + y.y * y.y);
        SFX(v2) a = SFX(mv)(o->w, SFX(mk)(o->a2 * atan(y.x * r), o->a2 * atanh(y.y * r)));
        if (o->kind == ORC_SIE_SHEAR) { SFX(v2) g = SFX(mv)(o->g, d); a.x += g.x; a.y += g.y; }
        return a;
    }
    case ORC_NSIE: {  /* nsie.cl:29-46 */
        SFX(v2) y = SFX(mv)(o->m, d);
        REAL r = sqrt(o->a0 * y.x * y.x + y.y * y.y);
        return SFX(mv)(o->w, SFX(mk)(o->a2 * atan(y.x * o->a1 / (o->a3 + r)),
                                     o->a2 * atanh(y.y * o->a1 / (o->a3 * o->a0 + r))));
    }
    case ORC_EPL: case ORC_EPL_SHEAR: {  /* epl.cl:25-68, epl_plus_shear.cl:28-76 */
        SFX(v2) u = SFX(mv)(o->m, d);
        REAL r = sqrt(u.x * u.x + u.y * u.y);
        REAL phi = atan2(u.y, u.x);
        SFX(v2) a = SFX(epl_series)(o->a0, o->a1, phi, 10);
        REAL rad = o->a2 * pow(r, 1 - o->a0);
        a = SFX(mv)(o->w, SFX(mk)(rad * a.x, rad * a.y));
        if (o->kind == ORC_EPL_SHEAR) { SFX(v2) g = SFX(mv)(o->g, d); a.x += g.x; a.y += g.y; }
        return a;
    }
    }
    return SFX(mk)(0, 0);
}

static REAL SFX(brightness)(const SFX(obj)* o, SFX(v2) x) {
    SFX(v2) d = SFX(mk)(x.x - o->cx, x.y - o->cy);
    switch (o->kind) {
    case ORC_SERSIC: case ORC_SERSIC_OLD: {  /* sersic.cl:23-27 */
        SFX(v2) y = SFX(mv)(o->m, d);
        return exp(o->a0 - exp(o->a1 + o->a2 * log(y.x * y.x + y.y * y.y)));
    }
    case ORC_GAUSS: {  /* gauss.cl:21-26 */
        SFX(v2) y = SFX(mv)(o->m, d);
        return o->a1 * exp(-(REAL)0.5 * (y.x * y.x + y.y * y.y) / o->a0);
    }
    case ORC_EXPONENTIAL: {  /* exponential.cl:21-25 */
        SFX(v2) y = SFX(mv)(o->m, d);
        return o->a1 * exp(-sqrt(y.x * y.x + y.y * y.y) / o->a0);
    }
    case ORC_DEVAUC: {  /* devauc.cl:26-30 */
        SFX(v2) y = SFX(mv)(o->m, d);
        return o->a1 * exp(-(REAL)7.6692494425008039044 * sqrt(sqrt(sqrt(y.x * y.x + y.y * y.y) / o->a0)));
    }
    }
    return 0;
}

static REAL SFX(foreground)(const SFX(obj)* o, SFX(v2) x) {  /* sky.cl:16-19 */
    return o->a0 + o->a1 * (x.x - 1) + o->a2 * (x.y - 1);
}

static int SFX(is_lens)(int k) { return orc_type(k) == 'L'; }

/* the generated compute() of src/kernel.c:65-111 */
static REAL SFX(compute)(int nobj, const SFX(obj)* objs, SFX(v2) x) {
    SFX(v2) y = x, a = SFX(mk)(0, 0);
    REAL f = 0;
    int trigger = 0, open = 0;
    for (int i = 0; i < nobj; ++i) {
        int ty = orc_type(objs[i].kind);
        if (ty != trigger && ty != 'F') {
            if (trigger == 'L' && open) {
                if (a.x * a.x + a.y * a.y < HUGE_VAL) { y.x -= a.x; y.y -= a.y; }
                else { y.x -= (REAL)1e10; y.y -= (REAL)1e10; }
                open = 0;
            }
            trigger = ty;
        }
        if (ty == 'L') {
            if (!open) { a = SFX(mk)(0, 0); open = 1; }
            SFX(v2) d = SFX(deflection)(&objs[i], y);
            a.x += d.x; a.y += d.y;
        } else if (ty == 'S') {
            f += SFX(brightness)(&objs[i], y);
        } else {
            f += SFX(foreground)(&objs[i], x);
        }
    }
    if (open) {
        if (a.x * a.x + a.y * a.y < HUGE_VAL) { y.x -= a.x; y.y -= a.y; }
        else { y.x -= (REAL)1e10; y.y -= (REAL)1e10; }
    }
    (void)SFX(is_lens);
    return f;
}

/* the generated set_params() of src/kernel.c:113-150,401-656, with
 * image-plane positions traced back through the lenses of earlier planes */
static void SFX(set_all)(int nobj, const int* kinds, const int* ipp, const REAL* params, SFX(obj)* objs) {
    int trigger = 0, plane = 0;
    for (int i = 0, p = 0; i < nobj; ++i) {
        int ty = orc_type(kinds[i]), np = orc_npars(kinds[i]);
        REAL args[16];
        if (ty != trigger && ty != 'F') {
            if (trigger == 'L' && ty == 'S') plane = i;
            trigger = ty;
        }
        for (int j = 0; j < np; ++j) args[j] = params[p + j];
        if (ipp && ipp[i]) {
            /* positions are the first two parameters of every positioned object */
            SFX(v2) x = SFX(mk)(params[p], params[p + 1]), a = SFX(mk)(0, 0);
            int trig2 = 0;
            for (int k = 0; k < plane; ++k) {
                int tk = orc_type(objs[k].kind);
                if (tk != trig2 && tk != 'F') {
                    if (trig2 == 'L') {
                        if (a.x * a.x + a.y * a.y < HUGE_VAL) { x.x -= a.x; x.y -= a.y; }
                        else { x.x -= (REAL)1e10; x.y -= (REAL)1e10; }
                        a = SFX(mk)(0, 0);
                    }
                    trig2 = tk;
                }
                if (tk == 'L') { SFX(v2) d = SFX(deflection)(&objs[k], x); a.x += d.x; a.y += d.y; }
            }
            if (trig2 == 'L') {
                if (a.x * a.x + a.y * a.y < HUGE_VAL) { x.x -= a.x; x.y -= a.y; }
                else { x.x -= (REAL)1e10; x.y -= (REAL)1e10; }
            }
            args[0] = x.x; args[1] = x.y;
        }
        SFX(set)(&objs[i], kinds[i], args);
        p += np;
    }
}

/* render, kernel/lensed.cl:9-38: value and error of the quadrature at every
 * pixel centre x = (rx + sx*i, ry + sy*j); nodes qq (already scaled) and
 * weights ww = (value weight, error weight) as 2*nq arrays */
void SFX(orc_render)(int nobj, const int* kinds, const int* ipp, const REAL* params, int W, int H,
                     const REAL* pcs, int nq, const REAL* qq, const REAL* ww, REAL* value, REAL* error) {
    SFX(obj) objs[64];
    SFX(set_all)(nobj, kinds, ipp, params, objs);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int k = 0; k < W * H; ++k) {
        REAL cx = pcs[0] + pcs[2] * (REAL)(k % W), cy = pcs[1] + pcs[3] * (REAL)(k / W);
        REAL f = 0, e = 0;
        for (int n = 0; n < nq; ++n) {
            REAL v = SFX(compute)(nobj, objs, SFX(mk)(cx + qq[2 * n], cy + qq[2 * n + 1]));
            f += ww[2 * n] * v;
            e += ww[2 * n + 1] * v;
        }
        value[k] = f;
        if (error) error[k] = e;
    }
}

/* deflection of object list entry `which` at a batch of points */
void SFX(orc_deflect)(int nobj, const int* kinds, const REAL* params, int which, int npts, const REAL* xy, REAL* a) {
    SFX(obj) objs[64];
    SFX(set_all)(nobj, kinds, NULL, params, objs);
    for (int i = 0; i < npts; ++i) {
        SFX(v2) d = SFX(deflection)(&objs[which], SFX(mk)(xy[2 * i], xy[2 * i + 1]));
        a[2 * i] = d.x; a[2 * i + 1] = d.y;
    }
}

/* convolve, kernel/lensed.cl:56-103: flipped PSF, clamp-to-edge borders */
void SFX(orc_convolve)(int W, int H, const REAL* in, int PW, int PH, const REAL* psf, REAL* out) {
    int cx = PW - 1 - PW / 2, cy = PH - 1 - PH / 2;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            REAL s = 0;
            for (int j = 0; j < PH; ++j) {
                int yy = y + cy - j;
                yy = yy < 0 ? 0 : yy >= H ? H - 1 : yy;
                for (int i = 0; i < PW; ++i) {
                    int xx = x + cx - i;
                    xx = xx < 0 ? 0 : xx >= W ? W - 1 : xx;
                    s += psf[j * PW + i] * in[yy * W + xx];
                }
            }
            out[y * W + x] = s;
        }
}

/* EPL angular factor Omega(phi) of the tenth-order series */
void SFX(orc_epl_omega)(int n, const REAL* q, const REAL* t, const REAL* phi, int nterms, REAL* out) {
    for (int i = 0; i < n; ++i) {
        SFX(v2) a = SFX(epl_series)(t[i], (1 - q[i]) / (1 + q[i]), phi[i], nterms);
        out[2 * i] = a.x; out[2 * i + 1] = a.y;
    }
}
