// Elliptical power-law lens (Tessore & Metcalf 2015) plus external shear.
// Same angular series as the epl object; the shear (g1, g2) acts on the
// offset from the lens centre.
type = LENS;

params
{
    { "x",  POSITION_X              },
    { "y",  POSITION_Y              },
    { "r",  RADIUS                  },
    { "t",  PARAMETER, { 0.f, 2.f } },
    { "q",  AXIS_RATIO              },
    { "pa", POS_ANGLE               },
    { "g1", PARAMETER               },
    { "g2", PARAMETER               }
};

data
{
    float2 c;       // centre
    mat22 fwd;      // image frame -> elliptical frame, scaled by 1/(r sqrt(q))
    mat22 bwd;      // elliptical frame -> image frame: rotation times 2 r sqrt(q)/(1 + q)
    mat22 shear;    // shear matrix
    float coef[10]; // C_k = prod_{j<=k} -f (2j - T)/(2j + T), k = 1..10
    float hexp;     // (1 - t)/2, exponent applied to |u|^2
};

static float2 deflection(local data* this, float2 x)
{
    float2 d = x - this->c;
    float2 u = mv22(this->fwd, d);
    float uu = dot(u, u);
    float ir = rsqrt(uu);
    float2 e = (float2)(uu > 0 ? u.x*ir : 1.0f, uu > 0 ? u.y*ir : 0.0f);
    float2 z = (float2)(e.x*e.x - e.y*e.y, 2*e.x*e.y);
    float2 p = (float2)(this->coef[9], 0);
    for(int k = 8; k >= 0; --k)
        p = (float2)(this->coef[k] + z.x*p.x - z.y*p.y, z.y*p.x + z.x*p.y);
    p = (float2)(1 + z.x*p.x - z.y*p.y, z.y*p.x + z.x*p.y);
    float2 sum = (float2)(e.x*p.x - e.y*p.y, e.y*p.x + e.x*p.y);
    return mv22(this->bwd, powr(uu, this->hexp)*sum) + mv22(this->shear, d);
}

static void set(local data* this, float x, float y, float r, float t, float q, float pa, float g1, float g2)
{
    float cs;
    float sn = sincos(pa*DEG2RAD, &cs);
    float T = 2 - t;
    float f = (1 - q)/(1 + q);
    this->c = (float2)(x, y);
    this->fwd = (1/r/sqrt(q))*(mat22)(q*cs, q*sn, -sn, cs);
    this->bwd = (2*r*sqrt(q)/(1 + q))*(mat22)(cs, -sn, sn, cs);
    this->shear = (mat22)(g1, g2, g2, -g1);
    float c = 1;
    for(int k = 1; k <= 10; ++k)
    {
        c *= -f*(2*k - T)/(2*k + T);
        this->coef[k-1] = c;
    }
    this->hexp = 0.5f*(1 - t);
}
