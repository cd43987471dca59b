// Singular isothermal ellipsoid (Kormann, Schneider & Bartelmann 1994).
// In the frame aligned with the major axis, with e = sqrt(1 - q^2):
//   alpha = r sqrt(q)/e * ( atan(e u1/w), atanh(e u2/w) ),  w = sqrt(q^2 u1^2 + u2^2)
type = LENS;

params
{
    { "x",  POSITION_X },
    { "y",  POSITION_Y },
    { "r",  RADIUS     },
    { "q",  AXIS_RATIO },
    { "pa", POS_ANGLE  }
};

data
{
    float2 c;  // centre
    mat22 fwd; // image frame -> axis frame
    mat22 bwd; // axis frame -> image frame
    float qq;  // q^2
    float ecc; // sqrt(1 - q^2)
    float amp; // r sqrt(q)/sqrt(1 - q^2)
};

static float2 deflection(local data* this, float2 x)
{
    float2 u = mv22(this->fwd, x - this->c);
    float s = this->ecc/sqrt(this->qq*u.x*u.x + u.y*u.y);
    return mv22(this->bwd, this->amp*(float2)(atan(u.x*s), atanh(u.y*s)));
}

static void set(local data* this, float x, float y, float r, float q, float pa)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->fwd = (mat22)(cs, sn, -sn, cs);
    this->bwd = (mat22)(cs, -sn, sn, cs);
    this->qq = q*q;
    this->ecc = sqrt(1 - q*q);
    this->amp = r*sqrt(q)/sqrt(1 - q*q);
}
