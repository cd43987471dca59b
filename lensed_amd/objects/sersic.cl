// Sersic profile with effective radius r, index n and total magnitude mag.
// I(R) = L/(2 pi n r^2 q) b^{2n}/Gamma(2n) exp(-b (R/r)^{1/n}) written as
// exp(lnorm - exp(lscale + log(R^2)/(2n))), evaluated in base 2.  b(n) from MacArthur, Courteau &
// Holtzman (2003): the Ciotti & Bertin (1999) asymptotic series for n > 0.36,
// their polynomial fit below.
type = SOURCE;

params
{
    { "x",   POSITION_X },
    { "y",   POSITION_Y },
    { "r",   RADIUS     },
    { "mag", MAGNITUDE  },
    { "n",   PARAMETER, POS_BOUND },
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
    float b;
    if(n > 0.36f)
        b = 2.0f*n - 1.0f/3 + 4.0f/(405*n) + 46.0f/(25515*(n*n)) + 131.0f/(1148175*(n*n*n))
          - 2194697.0f/(30690717750*(n*n*n*n));
    else
        b = 0.01945f - 0.8902f*n + 10.95f*(n*n) - 19.67f*(n*n*n) + 13.43f*(n*n*n*n);
    float cs = cos(pa*DEG2RAD);
    float sn = sin(pa*DEG2RAD);
    this->c = (float2)(x, y);
    this->tr = (mat22)(q*cs, q*sn, -sn, cs)/sqrt(q);
    this->lnorm = (-0.4f*mag*LOG_10 + 2*n*log(b) - LOG_PI - 2*log(r) - log(tgamma(2*n + 1)))*1.44269504088896340736f;
    this->lscale = (log(b) - log(r)/n)*1.44269504088896340736f;
    this->hinv = 0.5f/n;
}
