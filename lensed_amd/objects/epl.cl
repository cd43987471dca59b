// Elliptical power-law lens (Tessore & Metcalf 2015).
// In elliptical coordinates u (scaled by the Einstein radius) with polar angle
// phi, the deflection is  n |u|^(1-t) * Omega(phi),  where the angular factor
// Omega = sum_k A_k is the hypergeometric series with A_0 = e^{i phi} and
// A_k = -f (2k - T)/(2k + T) e^{2 i phi} A_{k-1},  T = 2 - t, f = (1-q)/(1+q),
// summed to tenth order.  cos(phi), sin(phi) are taken directly from u/|u|,
// and the sum is evaluated as e^{i phi} P(e^{2 i phi}) with the real
// polynomial P(z) = 1 + sum_k C_k z^k, C_k the running products of the
// recurrence factors, in Horner form.
type = LENS;

params
{
    { "x",  POSITION_X              },
    { "y",  POSITION_Y              },
    { "r",  RADIUS                  },
    { "t",  PARAMETER, { 0.f, 2.f } },
    { "q",  AXIS_RATIO              },
    { "pa", POS_ANGLE               }
};

data
{
    float2 c;       // centre
    mat22 fwd;      // image frame -> elliptical frame, scaled by 1/(r sqrt(q))
    mat22 bwd;      // elliptical frame -> image frame: rotation times 2 r sqrt(q)/(1 + q)
    float coef[10]; // C_k = prod_{j<=k} -f (2j - T)/(2j + T), k = 1..10
    float hexp;     // (1 - t)/2, exponent applied to |u|^2
};

static float2 deflection(local data* this, float2 x)
{
    float2 u = mv22(this->fwd, x - this->c);
    float uu = dot(u, u);
    float ir = rsqrt(uu);
    float2 e = (float2)(uu > 0 ? u.x*ir : 1.0f, uu > 0 ? u.y*ir : 0.0f);
    float2 z = (float2)(e.x*e.x - e.y*e.y, 2*e.x*e.y);
    float2 p = (float2)(this->coef[9], 0);
    for(int k = 8; k >= 0; --k)
        p = (float2)(this->coef[k] + z.x*p.x - z.y*p.y, z.y*p.x + z.x*p.y);
    p = (float2)(1 + z.x*p.x - z.y*p.y, z.y*p.x + z.x*p.y);
    float2 sum = (float2)(e.x*p.x - e.y*p.y, e.y*p.x + e.x*p.y);
    return mv22(this->bwd, powr(uu, this->hexp)*sum);
}

static void set(local data* this, float x, float y, float r, float t, float q, float pa)
{
    float cs;
    float sn = sincos(pa*DEG2RAD, &cs);
    float T = 2 - t;
    float f = (1 - q)/(1 + q);
    this->c = (float2)(x, y);
    this->fwd = (1/r/sqrt(q))*(mat22)(q*cs, q*sn, -sn, cs);
    this->bwd = (2*r*sqrt(q)/(1 + q))*(mat22)(cs, -sn, sn, cs);
    float c = 1;
    for(int k = 1; k <= 10; ++k)
    {
        c *= -f*(2*k - T)/(2*k + T);
        this->coef[k-1] = c;
    }
    this->hexp = 0.5f*(1 - t);
}
