// Point mass: alpha = r^2 (x - c)/|x - c|^2.
type = LENS;

params
{
    { "x", POSITION_X },
    { "y", POSITION_Y },
    { "r", RADIUS     }
};

data
{
    float2 c;  // centre
    float re2; // squared Einstein radius
};

static float2 deflection(local data* this, float2 x)
{
    float2 d = x - this->c;
    return (this->re2/dot(d, d))*d;
}

static void set(local data* this, float x, float y, float r)
{
    this->c = (float2)(x, y);
    this->re2 = r*r;
}
