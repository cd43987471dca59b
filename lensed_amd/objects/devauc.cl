// de Vaucouleurs (Sersic n = 4) profile, I = I0 exp(-b (R/r)^{1/4}) with
// b = 7.66924944... (the root of Gamma(8) = 2 gamma(8, b)); with r the scale
// parameter, I0 = L b^8 / (8! pi r^2 q).
#define DEVAUC_BN 7.6692494425008039044f
#define DEVAUC_BN8_OVER_8FAC 296.826303766893f

type = SOURCE;

params
{
    { "x",   POSITION_X },
    { "y",   POSITION_Y },
    { "r",   RADIUS     },
    { "mag", MAGNITUDE  },
    { "q",   AXIS_RATIO },
    { "pa",  POS_ANGLE  }
};

data
{
    float2 c;   // centre
    mat22 tr;   // rotation and axis scaling
    float rs;   // scale radius
    float peak; // central surface brightness
};

static float brightness(local data* this, float2 x)
{
    float s = length(mv22(this->tr, x - this->c))/this->rs;
    return this->peak*exp(-DEVAUC_BN*sqrt(sqrt(s)));
}

static void set(local data* this, float x, float y, float r, float mag, float q, float pa)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->tr = (mat22)(q*cs, q*sn, -sn, cs);
    this->rs = r;
    this->peak = exp(-0.4f*mag*LOG_10)/PI/r/r/q*DEVAUC_BN8_OVER_8FAC;
}
