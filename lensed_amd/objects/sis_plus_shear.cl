// Singular isothermal sphere with external shear (g1, g2).
type = LENS;

params
{
    { "x",  POSITION_X },
    { "y",  POSITION_Y },
    { "r",  RADIUS     },
    { "g1", PARAMETER  },
    { "g2", PARAMETER  }
};

data
{
    float2 c;    // centre
    mat22 shear; // symmetric traceless shear matrix
    float rein;  // Einstein radius
};

static float2 deflection(local data* this, float2 x)
{
    float2 d = x - this->c;
    return this->rein*normalize(d) + mv22(this->shear, d);
}

static void set(local data* this, float x, float y, float r, float g1, float g2)
{
    this->c = (float2)(x, y);
    this->rein = r;
    this->shear = (mat22)(g1, g2, g2, -g1);
}
