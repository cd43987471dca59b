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
    mat22 tr;   // rotation and axis scaling, times sqrt(log2(e)/2)/sigma
    float lpk;  // log2 of the central surface brightness L/(2 pi sigma^2 q)
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

// peak exp(-R^2/(2 sigma^2)) as exp2(log2(peak) - |u|^2)
static float frame_profile(local data* this, float2 u)
{
    return exp2(fma(-u.x, u.x, fma(-u.y, u.y, this->lpk)));
}

// |u| beyond which frame_profile is below 2^-127: a subnormal, which the
// kernels' FP32 arithmetic (subnormals flushed, as the reference's
// -cl-denorms-are-zero) takes as 0, so the ray adds exactly nothing
static float frame_cutoff(local data* this)
{
    return sqrt(fmax(this->lpk + 127, 0.0f));
}

static float brightness(local data* this, float2 x)
{
    return frame_profile(this, frame_map(this, x - frame_centre(this)));
}

static void set(local data* this, float x, float y, float sigma, float mag, float q, float pa)
{
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->tr = (0.84932180028801904272f/sigma)*(mat22)(q*cs, q*sn, -sn, cs);
    this->lpk = -0.4f*mag*3.32192809488736234787f - log2(2*PI*sigma*sigma*q);
}
