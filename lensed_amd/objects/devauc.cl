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
    float s = length(u)/this->rs;
    return this->peak*exp(-DEVAUC_BN*sqrt(sqrt(s)));
}

// |u| beyond which frame_profile is below 2^-127 (127 ln 2 = 88.03): a
// subnormal, which the kernels' FP32 arithmetic (subnormals flushed, as the
// reference's -cl-denorms-are-zero) takes as 0
static float frame_cutoff(local data* this)
{
    float a = fmax(log(this->peak) + 88.03f, 0.0f)/DEVAUC_BN;
    return this->rs*(a*a)*(a*a);
}

static float brightness(local data* this, float2 x)
{
    return frame_profile(this, frame_map(this, x - frame_centre(this)));
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
