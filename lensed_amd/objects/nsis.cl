// Non-singular isothermal sphere with core radius rc:
// alpha = r (x - c) / (rc + |x - c|).
type = LENS;

params
{
    { "x",  POSITION_X },
    { "y",  POSITION_Y },
    { "r",  RADIUS     },
    { "rc", RADIUS     }
};

data
{
    float2 c;   // centre
    float rein; // Einstein radius
    float core; // core radius
};

static float2 deflection(local data* this, float2 x)
{
    float2 d = x - this->c;
    return (this->rein/(this->core + length(d)))*d;
}

static void set(local data* this, float x, float y, float r, float rc)
{
    this->c = (float2)(x, y);
    this->rein = r;
    this->core = rc;
}
