// Definition of the opaque prior type (src/input.h:73), shared by the
// prior registry and the context, which keeps its own copies.
#pragma once

struct prior {
    enum Kind { DELTA, UNIF, NORM } kind;
    double a, b;
};
