// Per-pixel cubature rules (src/quadrature.c:21-43).  The node tables are
// regenerated from the mathematics of each rule by gen_quad.py.
#include <cstring>

#include "../../include/lensed.h"
#include "quad_tables.h"

static const lensed_quad::Rule* find_rule(const char* name) {
    if (!name) name = "g3k7";
    for (const lensed_quad::Rule* r = lensed_quad::RULES; r->name; ++r)
        if (strcmp(r->name, name) == 0) return r;
    return nullptr;
}

extern "C" {

int lensed_quad_size(const char* rule) {
    const lensed_quad::Rule* r = find_rule(rule);
    return r ? r->size : -1;
}

// nodes scaled by the pixel scale in double, then rounded to float, as the
// node buffers of the OpenCL program are filled
int lensed_quad_rule(const char* rule, float* qq, float* ww, double sx, double sy) {
    const lensed_quad::Rule* r = find_rule(rule);
    if (!r) return -1;
    for (int i = 0; i < r->size; ++i) {
        if (qq) {
            qq[2 * i] = (float)(sx * r->pts[i][0]);
            qq[2 * i + 1] = (float)(sy * r->pts[i][1]);
        }
        if (ww) {
            ww[2 * i] = (float)r->wht[i];
            ww[2 * i + 1] = (float)r->err[i];
        }
    }
    return r->size;
}

int lensed_quad_table(const char* rule, double* pts, double* wht, double* err) {
    const lensed_quad::Rule* r = find_rule(rule);
    if (!r) return -1;
    for (int i = 0; i < r->size; ++i) {
        if (pts) {
            pts[2 * i] = r->pts[i][0];
            pts[2 * i + 1] = r->pts[i][1];
        }
        if (wht) wht[i] = r->wht[i];
        if (err) err[i] = r->err[i];
    }
    return r->size;
}

const char* lensed_quad_name(int index) {
    int n = 0;
    while (lensed_quad::RULES[n].name) ++n;
    return (index >= 0 && index < n) ? lensed_quad::RULES[index].name : nullptr;
}

}  // extern "C"
