// Singular isothermal sphere: alpha = r * (x - c)/|x - c|.
type = LENS;

params
{
    { "x", POSITION_X },
    { "y", POSITION_Y },
    { "r", RADIUS     }
};

data
{
    float2 c;   // centre
    float rein; // Einstein radius
};

static float2 deflection(local data* this, float2 x)
{
    return this->rein*normalize(x - this->c);
}

static void set(local data* this, float x, float y, float r)
{
    this->c = (float2)(x, y);
    this->rein = r;
}
