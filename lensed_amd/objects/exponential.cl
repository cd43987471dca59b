// Elliptical exponential disk with scale length rs.
type = SOURCE;

params
{
    { "x",   POSITION_X },
    { "y",   POSITION_Y },
    { "rs",  RADIUS     },
    { "mag", MAGNITUDE  },
    { "q",   AXIS_RATIO },
    { "pa",  POS_ANGLE  }
};

data
{
    float2 c;   // centre
    mat22 tr;   // rotation and axis scaling
    float rs;   // scale length
    float peak; // central surface brightness L/(2 pi rs^2 q)
};

static float brightness(local data* this, float2 x)
{
    return this->peak*exp(-length(mv22(this->tr, x - this->c))/this->rs);
}

static void set(local data* this, float x, float y, float rs, float mag, float q, float pa)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->tr = (mat22)(q*cs, q*sn, -sn, cs);
    this->rs = rs;
    this->peak = exp(-0.4f*mag*LOG_10)*0.5f/PI/rs/rs/q;
}
