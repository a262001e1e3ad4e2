// ref_wrap.cpp -- exposes the reference's own trig lookup tables
// (include/so_linterp.h, instantiated as at src/Projection.cxx:34-40)
// through extern "C" entry points, so that golden vectors can be produced
// from the reference source compiled in this container.  The reference
// header is included from /root/reference at build time (see Makefile);
// nothing of it is copied here.  Test infrastructure only.
#include <cstdint>
#include <cstring>
#include "so_linterp.h"

static asinTable g_asin(1 << 14);
static atan2Table g_atan2(1 << 14);

extern "C" {
void ref_tables(double *asin_vals, double *atan2_vals) {
    memcpy(asin_vals, g_asin.vals.data(), sizeof(double) * g_asin.n);
    memcpy(atan2_vals, g_atan2.vals.data(), sizeof(double) * g_atan2.n);
}
void ref_asin_lookup(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; i++) y[i] = g_asin.get(x[i]);
}
void ref_atan2_lookup(const double *y, const double *x, double *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = g_atan2.get(y[i], x[i]);
}
}
