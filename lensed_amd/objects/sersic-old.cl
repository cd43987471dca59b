// Sersic profile with the linear b(n) = 1.9992 n - 0.3271 (valid for
// 0.5 < n < 8), the approximation Lensed used before its MacArthur-series
// update.  Otherwise identical to the sersic object.
type = SOURCE;

params
{
    { "x",   POSITION_X },
    { "y",   POSITION_Y },
    { "r",   RADIUS     },
    { "mag", MAGNITUDE  },
    { "n",   PARAMETER, { 0.5f, 8.0f } },
    { "q",   AXIS_RATIO },
    { "pa",  POS_ANGLE  }
};

data
{
    float2 c;     // centre
    mat22 tr;     // rotation and axis scaling
    float lnorm;  // log2 of the central surface brightness
    float lscale; // (log(b) - log(r)/n) log2(e)
    float hinv;   // 1/(2n)
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
    return exp2(this->lnorm - 1.44269504088896340736f*exp2(this->lscale + this->hinv*log2(dot(u, u))));
}

// |u| beyond which frame_profile is below 2^-127 (its exponent below -127):
// a subnormal, which the kernels' FP32 arithmetic (subnormals flushed, as
// the reference's -cl-denorms-are-zero) takes as 0
static float frame_cutoff(local data* this)
{
    float e = (this->lnorm + 127)/1.44269504088896340736f;
    return e > 0 ? exp2(0.5f*(log2(e) - this->lscale)/this->hinv) : 0.0f;
}

static float brightness(local data* this, float2 x)
{
    return frame_profile(this, frame_map(this, x - frame_centre(this)));
}

static void set(local data* this, float x, float y, float r, float mag, float n, float q, float pa)
{
    float b = 1.9992f*n - 0.3271f;
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->tr = (mat22)(q*cs, q*sn, -sn, cs)/sqrt(q);
    this->lnorm = (-0.4f*mag*LOG_10 + 2*n*log(b) - LOG_PI - 2*log(r) - log(tgamma(2*n + 1)))*1.44269504088896340736f;
    this->lscale = (log(b) - log(r)/n)*1.44269504088896340736f;
    this->hinv = 0.5f/n;
}
