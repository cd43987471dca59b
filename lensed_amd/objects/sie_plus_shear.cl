// Singular isothermal ellipsoid plus external shear (g1, g2).
type = LENS;

params
{
    { "x",  POSITION_X },
    { "y",  POSITION_Y },
    { "r",  RADIUS     },
    { "q",  AXIS_RATIO },
    { "pa", POS_ANGLE  },
    { "g1", PARAMETER  },
    { "g2", PARAMETER  }
};

data
{
    float2 c;    // centre
    mat22 fwd;   // image frame -> axis frame
    mat22 bwd;   // axis frame -> image frame
    mat22 shear; // shear matrix
    float qq;    // q^2
    float ecc;   // sqrt(1 - q^2)
    float amp;   // r sqrt(q)/sqrt(1 - q^2)
};

static float2 deflection(local data* this, float2 x)
{
    float2 d = x - this->c;
    float2 u = mv22(this->fwd, d);
    float s = this->ecc/sqrt(this->qq*u.x*u.x + u.y*u.y);
    return mv22(this->bwd, this->amp*(float2)(atan(u.x*s), atanh(u.y*s))) + mv22(this->shear, d);
}

static void set(local data* this, float x, float y, float r, float q, float pa, float g1, float g2)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->fwd = (mat22)(cs, sn, -sn, cs);
    this->bwd = (mat22)(cs, -sn, sn, cs);
    this->shear = (mat22)(g1, g2, g2, -g1);
    this->qq = q*q;
    this->ecc = sqrt(1 - q*q);
    this->amp = r*sqrt(q)/sqrt(1 - q*q);
}
