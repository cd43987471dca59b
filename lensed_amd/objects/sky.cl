// Flat sky background with an optional linear gradient, measured from the
// centre (1, 1) of the first pixel.  The gradient defaults to zero.
type = FOREGROUND;

params
{
    { "bg" },
    { "dx", PARAMETER, UNBOUNDED, -0.0f },
    { "dy", PARAMETER, UNBOUNDED, -0.0f }
};

data
{
    float level; // background level
    float2 grad; // gradient (d/dx, d/dy)
};

static float foreground(local data* this, float2 x)
{
    return this->level + dot(this->grad, x - (float2)(1, 1));
}

static void set(local data* this, float bg, float dx, float dy)
{
    this->level = bg;
    this->grad = (float2)(dx, dy);
}
