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

// in its own frame u = tr (x - c): brightness(x) = frame_profile(frame_map(x - frame_centre))
static float2 frame_centre(local data* this)
{
    return this->c;
}

static float2 frame_map(local data* this, float2 d)
{
    return mv22(this->tr, d);
}

static float frame_profile(local data* this, float2 u)
{
    return this->peak*exp(-length(u)/this->rs);
}

// |u| beyond which frame_profile is below 2^-127 (127 ln 2 = 88.03): a
// subnormal, which the kernels' FP32 arithmetic (subnormals flushed, as the
// reference's -cl-denorms-are-zero) takes as 0
static float frame_cutoff(local data* this)
{
    return this->rs*fmax(log(this->peak) + 88.03f, 0.0f);
}

static float brightness(local data* this, float2 x)
{
    return frame_profile(this, frame_map(this, x - frame_centre(this)));
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
