// Non-singular isothermal ellipsoid with core radius rc
// (Kormann, Schneider & Bartelmann 1994, cored form).
type = LENS;

params
{
    { "x",  POSITION_X },
    { "y",  POSITION_Y },
    { "r",  RADIUS     },
    { "rc", RADIUS     },
    { "q",  AXIS_RATIO },
    { "pa", POS_ANGLE  }
};

data
{
    float2 c;   // centre
    mat22 fwd;  // image frame -> axis frame
    mat22 bwd;  // axis frame -> image frame
    float core; // core radius
    float qq;   // q^2
    float ecc;  // sqrt(1 - q^2)
    float amp;  // r sqrt(q)/sqrt(1 - q^2)
};

static float2 deflection(local data* this, float2 x)
{
    float2 u = mv22(this->fwd, x - this->c);
    float w = sqrt(this->qq*u.x*u.x + u.y*u.y);
    float ax = atan(this->ecc*u.x/(this->core + w));
    float ay = atanh(this->ecc*u.y/(this->core*this->qq + w));
    return mv22(this->bwd, this->amp*(float2)(ax, ay));
}

static void set(local data* this, float x, float y, float r, float rc, float q, float pa)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->fwd = (mat22)(cs, sn, -sn, cs);
    this->bwd = (mat22)(cs, -sn, sn, cs);
    this->core = rc;
    this->qq = q*q;
    this->ecc = sqrt(1 - q*q);
    this->amp = r*sqrt(q)/sqrt(1 - q*q);
}
