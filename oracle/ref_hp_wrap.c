/* ref_hp_wrap.c -- array entry points over the reference's HEALPix routines
 * (src/healpix_bare.c, compiled from /root/reference by oracle/Makefile).
 * Used only to write golden vectors and pin the oracle.  Test infrastructure. */
#include <stdint.h>
#include "healpix_bare.h"

void ref_ang2pix(int64_t nside, int nest, const double *theta, const double *phi,
                 int64_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        t_ang a = {theta[i], phi[i]};
        out[i] = nest ? ang2nest(nside, a) : ang2ring(nside, a);
    }
}
void ref_pix2ang(int64_t nside, int nest, const int64_t *pix, double *theta,
                 double *phi, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        t_ang a = nest ? nest2ang(nside, pix[i]) : ring2ang(nside, pix[i]);
        theta[i] = a.theta;
        phi[i] = a.phi;
    }
}
void ref_nest2ring(int64_t nside, const int64_t *pix, int64_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = nest2ring(nside, pix[i]);
}
