// Elliptical Gaussian source with standard deviation sigma along the major axis.
type = SOURCE;

params
{
    { "x",     POSITION_X },
    { "y",     POSITION_Y },
    { "sigma", RADIUS     },
    { "mag",   MAGNITUDE  },
    { "q",     AXIS_RATIO },
    { "pa",    POS_ANGLE  }
};

data
{
    float2 c;   // centre
    mat22 tr;   // rotation and axis scaling
    float k2;   // -log2(e)/(2 sigma^2)
    float peak; // central surface brightness L/(2 pi sigma^2 q)
};

static float brightness(local data* this, float2 x)
{
    float2 u = mv22(this->tr, x - this->c);
    return this->peak*exp2(this->k2*dot(u, u));
}

static void set(local data* this, float x, float y, float sigma, float mag, float q, float pa)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->tr = (mat22)(q*cs, q*sn, -sn, cs);
    this->k2 = -0.72134752044448170368f/(sigma*sigma);
    this->peak = exp(-0.4f*mag*LOG_10)*0.5f/PI/(sigma*sigma)/q;
}
