// so3g_oracle.cpp -- CPU oracle for the projection hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain-C++ restatement of the
// arithmetic of so3g's ProjectionEngine (src/Projection.cxx), its trig
// lookup tables (include/so_linterp.h) and the HEALPix ang2pix routines
// (src/healpix_bare.c).  It exists so that the HIP kernels can be checked
// against an independent CPU computation on the same inputs.  Only tests,
// the smoke check and the CPU-baseline leg of the benchmark may load it;
// nothing in the product path calls it.
//
// Build: g++ -O3 -ffp-contract=off -fopenmp -shared -fPIC (no -march), which
// matches the reference's x86-64 Release build (SSE2 doubles, no FMA).
// Every expression below keeps the reference's left-to-right evaluation
// order; reordering any of them changes bits.
//
// Pinning (what checks each part independently of this file):
//   * trig tables / lookups  -> golden vectors from include/so_linterp.h
//                               compiled by oracle/_ref (tests/golden/)
//   * ang2nest / ang2ring    -> golden vectors from src/healpix_bare.c
//                               compiled by oracle/_ref, plus brute-force
//                               pixel-centre round trips for small nside
//   * quaternion product     -> Hamilton-product identities, rotation
//                               composition tests
//   * projections            -> FITS-II closed forms (Calabretta & Greisen
//                               2002) evaluated in numpy, tolerance tests
//   * pixelizors             -> closed-form index arithmetic incl. off-map,
//                               stale-slot and truncation cases
//   * to_map / to_weights    -> dense brute-force P^T P on tiny inputs,
//                               the one-detector weights matrix in
//                               docs/proj.rst:319-324, det_weights=0 => 0
//   * from_map               -> polarisation-purity closed form
//                               (test/test_proj_zen.py:107-133)

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#pragma STDC FP_CONTRACT OFF

namespace {

const double kPi = 3.141592653589793238462643383279502884197;

// --------------------------------------------------------------------------
// Linear-interpolation trig tables: include/so_linterp.h:8-71,
// instantiated with 1<<14 entries at src/Projection.cxx:34-40.
// --------------------------------------------------------------------------
struct Table {
    int n = 0;
    double dx = 0.;
    std::vector<double> v;
    // so_linterp.h:20-29
    double raw(double x) const {
        double d = x / dx;
        int idx = (int)d;
        if (x < 0) return v[0];
        if (idx >= n - 1) return v[n - 1];
        d -= idx;
        return v[idx] * (1 - d) + v[idx + 1] * d;
    }
};

struct Tables {
    Table as, at;
    Tables() {
        const int n = 1 << 14;
        as.n = at.n = n;
        as.v.resize(n);
        at.v.resize(n);
        // so_linterp.h:35-39 and :51-56
        as.dx = 1. / (n - 1);
        at.dx = 1. / (n - 1);
        for (int i = 0; i < n; i++) as.v[i] = asin(i * as.dx);
        for (int i = 0; i < n; i++) at.v[i] = atan2(i * at.dx, 1.);
    }
    // so_linterp.h:41-45
    double asin_lu(double x) const {
        if (x < 0) return -as.raw(-x);
        return as.raw(x);
    }
    // so_linterp.h:58-70
    double atan2_lu(double y, double x) const {
        if (y < 0) return -atan2_lu(-y, x);
        if (x < 0) return kPi - atan2_lu(y, -x);
        if (y == 0) return at.v[0];
        if (x < y) return kPi / 2 - at.raw(x / y);
        return at.raw(y / x);
    }
};

const Tables &tables() {
    static Tables t;
    return t;
}

// --------------------------------------------------------------------------
// HEALPix ang2pix: src/healpix_bare.c:46-78 (loc2hpc), :109-114 (ang2loc),
// :152-161 (spread_bits), :178-179 (hpd2nest), :187-211 (hpd2ring),
// :271-275 (loc2hpd), :303-306 (ang2ring / ang2nest).
// Algorithm: Gorski et al. 2005, ApJ 622, 759.
// --------------------------------------------------------------------------
const int kJrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
const int kJpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

struct HpLoc { double z, s, phi; };
struct HpCont { double x, y; int32_t f; };
struct HpDisc { int64_t x, y; int32_t f; };

HpLoc hp_ang2loc(double theta, double phi) {
    double cth = cos(theta), sth = sin(theta);
    if (sth < 0) { sth = -sth; phi += kPi; }
    return HpLoc{cth, sth, phi};
}

HpCont hp_loc2hpc(HpLoc loc) {
    double za = fabs(loc.z);
    double x = loc.phi * (1. / (2. * kPi));
    if (x < 0.)
        x += (int64_t)x + 1.;
    else if (x >= 1.)
        x -= (int64_t)x;
    double tt = 4. * x;
    if (za <= 2. / 3.) {
        double t1 = 0.5 + tt;
        double t2 = loc.z * 0.75;
        double jp = t1 - t2;
        double jm = t1 + t2;
        int ifp = (int)jp;
        int ifm = (int)jm;
        int32_t face = (ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : (ifm + 8));
        return HpCont{jm - ifm, 1 + ifp - jp, face};
    }
    int64_t ntt = (int64_t)tt;
    if (ntt >= 4) ntt = 3;
    double tp = tt - ntt;
    double tmp = loc.s / sqrt((1. + za) * (1. / 3.));
    double jp = tp * tmp;
    double jm = (1.0 - tp) * tmp;
    if (jp > 1.) jp = 1.;
    if (jm > 1.) jm = 1.;
    if (loc.z >= 0) return HpCont{1. - jm, 1. - jp, (int32_t)ntt};
    return HpCont{jp, jm, (int32_t)ntt + 8};
}

HpDisc hp_loc2hpd(int64_t nside, HpLoc loc) {
    HpCont c = hp_loc2hpc(loc);
    return HpDisc{(int64_t)(c.x * nside), (int64_t)(c.y * nside), c.f};
}

int64_t hp_spread(int64_t v) {
    int64_t r = v & 0xffffffff;
    r = (r ^ (r << 16)) & 0x0000ffff0000ffff;
    r = (r ^ (r << 8)) & 0x00ff00ff00ff00ff;
    r = (r ^ (r << 4)) & 0x0f0f0f0f0f0f0f0f;
    r = (r ^ (r << 2)) & 0x3333333333333333;
    r = (r ^ (r << 1)) & 0x5555555555555555;
    return r;
}

int64_t hp_hpd2nest(int64_t nside, HpDisc h) {
    return (h.f * nside * nside) + hp_spread(h.x) + (hp_spread(h.y) << 1);
}

int64_t hp_hpd2ring(int64_t ns, HpDisc h) {
    int64_t nl4 = 4 * ns;
    int64_t jr = (kJrll[h.f] * ns) - h.x - h.y - 1;
    if (jr < ns) {
        int64_t jp = (kJpll[h.f] * jr + h.x - h.y + 1) / 2;
        jp = (jp > nl4) ? jp - nl4 : ((jp < 1) ? jp + nl4 : jp);
        return 2 * jr * (jr - 1) + jp - 1;
    } else if (jr > 3 * ns) {
        jr = nl4 - jr;
        int64_t jp = (kJpll[h.f] * jr + h.x - h.y + 1) / 2;
        jp = (jp > nl4) ? jp - nl4 : ((jp < 1) ? jp + nl4 : jp);
        return 12 * ns * ns - 2 * (jr + 1) * jr + jp - 1;
    }
    int64_t jp = (kJpll[h.f] * ns + h.x - h.y + 1 + ((jr - ns) & 1)) / 2;
    jp = (jp > nl4) ? jp - nl4 : ((jp < 1) ? jp + nl4 : jp);
    return 2 * ns * (ns - 1) + (jr - ns) * nl4 + jp - 1;
}

int64_t hp_ang2nest(int64_t nside, double theta, double phi) {
    return hp_hpd2nest(nside, hp_loc2hpd(nside, hp_ang2loc(theta, phi)));
}
int64_t hp_ang2ring(int64_t nside, double theta, double phi) {
    return hp_hpd2ring(nside, hp_loc2hpd(nside, hp_ang2loc(theta, phi)));
}

// --------------------------------------------------------------------------
// Pointing: src/Projection.cxx:238-507.
// --------------------------------------------------------------------------
enum { PROJ_FLAT = 0, PROJ_QUAT, PROJ_CAR, PROJ_CEA, PROJ_ARC, PROJ_SIN,
       PROJ_TAN, PROJ_ZEA };

// Hamilton product with boost::math::quaternion<double>::operator*= term
// order (used at Projection.cxx:256, 298, 351, 386, 422, 458, 493).
inline void qmul(const double *l, const double *r, double *o) {
    const double a = l[0], b = l[1], c = l[2], d = l[3];
    const double ar = r[0], br = r[1], cr = r[2], dr = r[3];
    o[0] = a * ar - b * br - c * cr - d * dr;
    o[1] = a * br + b * ar + c * dr - d * cr;
    o[2] = a * cr - b * dr + c * ar + d * br;
    o[3] = a * dr + b * cr - c * br + d * ar;
}

void get_coords(int proj, const double *qb, const double *qd, double *out) {
    const Tables &T = tables();
    if (proj == PROJ_FLAT) {  // Projection.cxx:262-276
        out[0] = qb[0] + qd[0];
        out[1] = qb[1] + qd[1];
        const double c2 = qb[2];
        out[2] = qb[2] * qd[2] - qb[3] * qd[3];
        out[3] = qb[3] * qd[2] + c2 * qd[3];
        return;
    }
    double q[4];
    qmul(qb, qd, q);
    const double a = q[0], b = q[1], c = q[2], d = q[3];
    switch (proj) {
    case PROJ_QUAT:  // :241-257
        out[0] = a; out[1] = b; out[2] = c; out[3] = d;
        return;
    case PROJ_ARC: {  // :284-326
        const double ct2 = a * a + d * d;
        const double st2 = b * b + c * c;
        const double sc = c * a + d * b;
        const double ss = a * b - c * d;
        const double hst = sqrt(sc * sc + ss * ss);
        double R;
        if (hst < 1e-8)
            R = 2 + 1.33333333333 * hst * hst;
        else
            R = T.asin_lu(hst * 2) / hst;
        if (ct2 < st2) R = kPi / hst - R;
        out[0] = ss * R;
        out[1] = sc * R;
        out[2] = (a * a - d * d) / ct2;
        out[3] = (2 * a * d) / ct2;
        return;
    }
    case PROJ_SIN: {  // :337-364
        const double ct2 = a * a + d * d;
        out[0] = 2 * (a * b - c * d);
        out[1] = 2 * (c * a + d * b);
        out[2] = (a * a - d * d) / ct2;
        out[3] = (2 * a * d) / ct2;
        return;
    }
    case PROJ_TAN: {  // :372-400
        const double ct2 = a * a + d * d;
        const double ct = 2 * ct2 - 1;
        out[0] = (a * b - c * d) * 2 / ct;
        out[1] = (a * c + b * d) * 2 / ct;
        out[2] = (a * a - d * d) / ct2;
        out[3] = (2 * a * d) / ct2;
        return;
    }
    case PROJ_ZEA: {  // :408-437
        const double ct2 = a * a + d * d;
        const double c2 = sqrt(ct2);
        out[0] = (a * b - c * d) * 2 / c2;
        out[1] = (a * c + b * d) * 2 / c2;
        out[2] = (a * a - d * d) / ct2;
        out[3] = (2 * a * d) / ct2;
        return;
    }
    case PROJ_CEA: {  // :444-472
        const double ct = a * a - b * b - c * c + d * d;
        const double h = 0.5 * sqrt(1 - ct * ct);
        out[0] = atan2(c * d - a * b, c * a + d * b);
        out[1] = ct;
        out[2] = (a * c - b * d) / h;
        out[3] = (c * d + a * b) / h;
        return;
    }
    default: {  // PROJ_CAR, :479-507
        const double ct = a * a - b * b - c * c + d * d;
        const double h = 0.5 * sqrt(1 - ct * ct);
        out[0] = T.atan2_lu(c * d - a * b, c * a + d * b);
        out[1] = T.asin_lu(ct);
        out[2] = (a * c - b * d) / h;
        out[3] = (c * d + a * b) / h;
        return;
    }
    }
}

// --------------------------------------------------------------------------
// Geometry / pixelizors: Projection.cxx:510-1166.
// --------------------------------------------------------------------------
enum { PIX_FLAT = 0, PIX_HEALPIX = 1 };

}  // namespace

extern "C" {

// Flat description of one engine: the template parameters of
// ProjectionEngine<C,P,S> plus the pixelizor constructor arguments.
typedef struct {
    int32_t proj;          // PROJ_*
    int32_t pixelizor;     // PIX_FLAT / PIX_HEALPIX
    int32_t tiled;         // 0/1
    int32_t interp;        // 0 nearest, 1 bilinear
    int32_t ncomp;         // 1 T, 2 QU, 3 TQU
    int32_t ny, nx;        // flat: naxis
    double cdelt_y, cdelt_x, crpix_y, crpix_x;
    int32_t tile_y, tile_x;
    int32_t nside, nest, nside_tile;
} oracle_geom;

}  // extern "C"

namespace {

struct Pix {
    const oracle_geom *g;
    int n_tx() const { return (g->nx + g->tile_x - 1) / g->tile_x; }
    int index_count() const {
        if (g->pixelizor == PIX_HEALPIX) return g->tiled ? 2 : 1;
        return g->tiled ? 3 : 2;
    }
    int npix_hp() const { return 12 * g->nside * g->nside; }
    int npt_hp() const {
        int ntiles = 12 * g->nside_tile * g->nside_tile;
        return npix_hp() / ntiles;
    }
    int hp_index(const double *coords, bool nest) const {
        double phi = coords[0];
        double theta = kPi / 2 - coords[1];
        int64_t ix = nest ? hp_ang2nest(g->nside, theta, phi)
                          : hp_ang2ring(g->nside, theta, phi);
        return (int)ix;
    }

    // GetPixel (used by pixels() and pointing_matrix()).  Off-map
    // samples write only slot 0 (= -1), leaving the other slots as
    // they were.  Projection.cxx:558-564, 704-712, 846-863, 1030-1051.
    void get_pixel(const double *coords, int *pi) const {
        if (g->pixelizor == PIX_HEALPIX) {
            if (!g->tiled) {
                pi[0] = hp_index(coords, g->nest != 0);
            } else {
                int ix = hp_index(coords, true);
                int npt = npt_hp();
                pi[0] = ix / npt;
                pi[1] = ix % npt;
            }
            return;
        }
        double ix = coords[0] / g->cdelt_x + g->crpix_x - 1 + 0.5;
        if (ix < 0 || ix >= g->nx) { pi[0] = -1; return; }
        double iy = coords[1] / g->cdelt_y + g->crpix_y - 1 + 0.5;
        if (iy < 0 || iy >= g->ny) { pi[0] = -1; return; }
        if (!g->tiled) {
            pi[0] = int(iy);
            pi[1] = int(ix);
        } else {
            int sy = int(iy) / g->tile_y;
            int sx = int(ix) / g->tile_x;
            pi[0] = sy * n_tx() + sx;
            pi[1] = int(iy) - sy * g->tile_y;
            pi[2] = int(ix) - sx * g->tile_x;
        }
    }

    // GetPixels (used by from_map/to_map/to_weight_map and the ranges
    // helpers).  Returns the number of (pixel, weight) points.
    // Projection.cxx:566-574, 714-724, 913-952, 1125-1166.
    int get_pixels(const double *coords, int pi[4][3], float w[4]) const {
        if (g->pixelizor == PIX_HEALPIX) {
            get_pixel(coords, pi[0]);
            w[0] = 1;
            return 1;
        }
        if (!g->tiled) {
            if (!g->interp) {
                double x = coords[0] / g->cdelt_x + g->crpix_x - 1 + 0.5;
                double y = coords[1] / g->cdelt_y + g->crpix_y - 1 + 0.5;
                if (x < 0 || x >= g->nx) return 0;
                if (y < 0 || y >= g->ny) return 0;
                pi[0][0] = int(y) - int(y < 0);
                pi[0][1] = int(x) - int(x < 0);
                w[0] = 1;
                return 1;
            }
            double x = coords[0] / g->cdelt_x + g->crpix_x - 1 + 0.5;
            double y = coords[1] / g->cdelt_y + g->crpix_y - 1 + 0.5;
            int x1 = int(x) - int(x < 0);
            int y1 = int(y) - int(y < 0);
            double wx[2] = {x - x1, 1 - (x - x1)};
            double wy[2] = {y - y1, 1 - (y - y1)};
            int n = 0;
            for (int iy = y1; iy < y1 + 2; iy++) {
                if (iy < 0 || iy >= g->ny) continue;
                for (int ix = x1; ix < x1 + 2; ix++) {
                    if (ix < 0 || ix >= g->nx) continue;
                    pi[n][0] = iy;
                    pi[n][1] = ix;
                    w[n] = wy[iy - y1] * wx[ix - x1];
                    n++;
                }
            }
            return n;
        }
        if (!g->interp) {
            int ix = int(coords[0] / g->cdelt_x + g->crpix_x - 1 + 0.5);
            if (ix < 0 || ix >= g->nx) return 0;
            int iy = int(coords[1] / g->cdelt_y + g->crpix_y - 1 + 0.5);
            if (iy < 0 || iy >= g->ny) return 0;
            int sy = iy / g->tile_y;
            int sx = ix / g->tile_x;
            pi[0][0] = sy * n_tx() + sx;
            pi[0][1] = iy - sy * g->tile_y;
            pi[0][2] = ix - sx * g->tile_x;
            w[0] = 1;
            return 1;
        }
        double x = coords[0] / g->cdelt_x + g->crpix_x - 1 + 0.5;
        double y = coords[1] / g->cdelt_y + g->crpix_y - 1 + 0.5;
        int x1 = int(x);
        int y1 = int(y);
        double wx[2] = {x - x1, 1 - (x - x1)};
        double wy[2] = {y - y1, 1 - (y - y1)};
        int n = 0;
        for (int iy = y1; iy < y1 + 2; iy++) {
            if (iy < 0 || iy >= g->ny) continue;
            int sy = iy / g->tile_y;
            for (int ix = x1; ix < x1 + 2; ix++) {
                if (ix < 0 || ix >= g->nx) continue;
                int sx = ix / g->tile_x;
                pi[n][0] = sy * n_tx() + sx;
                pi[n][1] = iy - sy * g->tile_y;
                pi[n][2] = ix - sx * g->tile_x;
                w[n] = wy[iy - y1] * wx[ix - x1];
                n++;
            }
        }
        return n;
    }

    // Element offset of pixel pi within its (tile) map, for a map whose
    // leading axes hold `lead` component slots.  Returns the map base in
    // *base; nullptr if the tile is inactive.
    // Non-tiled flat maps: (lead, ny, nx); tiled flat: per tile
    // (lead, th, tw) with edge tiles clipped (Projection.cxx:1018-1019);
    // HEALPix: (lead, npix) or per tile (lead, npix_per_tile).
    int64_t plane(int tile) const {
        if (g->pixelizor == PIX_HEALPIX) return g->tiled ? npt_hp() : npix_hp();
        if (!g->tiled) return (int64_t)g->ny * g->nx;
        int ty = tile / n_tx(), tx = tile % n_tx();
        int th = std::min(g->tile_y, g->ny - ty * g->tile_y);
        int tw = std::min(g->tile_x, g->nx - tx * g->tile_x);
        return (int64_t)th * tw;
    }
    int64_t offset(const int *pi) const {
        if (g->pixelizor == PIX_HEALPIX) return g->tiled ? pi[1] : pi[0];
        if (!g->tiled) return (int64_t)pi[0] * g->nx + pi[1];
        int tx = pi[0] % n_tx();
        int tw = std::min(g->tile_x, g->nx - tx * g->tile_x);
        return (int64_t)pi[1] * tw + pi[2];
    }
    double *map_base(double *map, double *const *tiles, const int *pi) const {
        if (!g->tiled) return map;
        return tiles[pi[0]];
    }
};

// spin_proj_factors: Projection.cxx:1184-1209.
inline void spin_factors(int ncomp, const double *coords, float T, float P, float *pf) {
    if (ncomp == 1) { pf[0] = T; return; }
    const double c = coords[2];
    const double s = coords[3];
    if (ncomp == 2) {
        pf[0] = P * (c * c - s * s);
        pf[1] = P * (2 * c * s);
        return;
    }
    pf[0] = T;
    pf[1] = P * (c * c - s * s);
    pf[2] = P * (2 * c * s);
}

inline int wmul(int a, int b) { return (int)(int32_t)((uint32_t)a * (uint32_t)b); }
inline int wadd(int a, int b) { return (int)(int32_t)((uint32_t)a + (uint32_t)b); }

struct Segs {
    // Either an explicit (det, t0, t1) list or the full range of every
    // detector (derive_ranges(None), Projection.cxx:1875-1879).
    const int32_t *seg;
    int64_t nseg;
    int64_t n_det, n_t;
};

template <class F>
void for_each_segment(const Segs &S, F f) {
    if (S.seg == nullptr) {
        for (int64_t d = 0; d < S.n_det; d++) f((int)d, 0, (int)S.n_t);
        return;
    }
    for (int64_t i = 0; i < S.nseg; i++)
        f(S.seg[3 * i], S.seg[3 * i + 1], S.seg[3 * i + 2]);
}

}  // namespace

extern "C" {

// Error codes shared with the tests: 0 ok, 1 inactive tile hit.

void oracle_tables(double *asin_vals, double *atan2_vals) {
    const Tables &T = tables();
    memcpy(asin_vals, T.as.v.data(), sizeof(double) * T.as.n);
    memcpy(atan2_vals, T.at.v.data(), sizeof(double) * T.at.n);
}

void oracle_asin_lookup(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; i++) y[i] = tables().asin_lu(x[i]);
}

void oracle_atan2_lookup(const double *yv, const double *xv, double *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = tables().atan2_lu(yv[i], xv[i]);
}

void oracle_ang2pix(int64_t nside, int nest, const double *theta, const double *phi,
                    int64_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        out[i] = nest ? hp_ang2nest(nside, theta[i], phi[i])
                      : hp_ang2ring(nside, theta[i], phi[i]);
}

void oracle_qmul(const double *l, const double *r, double *o, int64_t n) {
    for (int64_t i = 0; i < n; i++) qmul(l + 4 * i, r + 4 * i, o + 4 * i);
}

// ProjectionEngine::coords, Projection.cxx:1316-1349.  out: (n_det, n_t, 4).
void oracle_coords(const oracle_geom *g, const double *bore, int64_t n_t,
                   const double *det, int64_t n_det, double *out) {
#pragma omp parallel for
    for (int64_t d = 0; d < n_det; d++)
        for (int64_t t = 0; t < n_t; t++)
            get_coords(g->proj, bore + 4 * t, det + 4 * d, out + 4 * (d * n_t + t));
}

// ProjectionEngine::pixels, Projection.cxx:1351-1385.  out: (n_det, n_t, k).
// The per-detector index buffer starts as {-1, 0, ...} (:1372).
void oracle_pixels(const oracle_geom *g, const double *bore, int64_t n_t,
                   const double *det, int64_t n_det, int32_t *out) {
    Pix P{g};
    const int k = P.index_count();
#pragma omp parallel for
    for (int64_t d = 0; d < n_det; d++) {
        int pi[3] = {-1, 0, 0};
        for (int64_t t = 0; t < n_t; t++) {
            double co[4];
            get_coords(g->proj, bore + 4 * t, det + 4 * d, co);
            P.get_pixel(co, pi);
            for (int j = 0; j < k; j++) out[(d * n_t + t) * k + j] = pi[j];
        }
    }
}

// ProjectionEngine::pointing_matrix, Projection.cxx:1391-1438.
void oracle_pointing_matrix(const oracle_geom *g, const double *bore, int64_t n_t,
                            const double *det, int64_t n_det, const float *resp,
                            int32_t *pix, float *spin) {
    Pix P{g};
    const int k = P.index_count();
    const int nc = g->ncomp;
#pragma omp parallel for
    for (int64_t d = 0; d < n_det; d++) {
        int pi[3] = {-1, 0, 0};
        const float T = resp[2 * d], Pp = resp[2 * d + 1];
        for (int64_t t = 0; t < n_t; t++) {
            double co[4];
            float pf[3];
            get_coords(g->proj, bore + 4 * t, det + 4 * d, co);
            P.get_pixel(co, pi);
            spin_factors(nc, co, T, Pp, pf);
            for (int j = 0; j < k; j++) pix[(d * n_t + t) * k + j] = pi[j];
            for (int j = 0; j < nc; j++) spin[(d * n_t + t) * nc + j] = pf[j];
        }
    }
}

// ProjectionEngine::from_map, Projection.cxx:1728-1770.
// signal: (n_det, n_t), accumulated in place.
int oracle_from_map(const oracle_geom *g, const double *map, double *const *tiles,
                    const double *bore, int64_t n_t, const double *det, int64_t n_det,
                    const float *resp, float *signal) {
    Pix P{g};
    const int nc = g->ncomp;
    int err = 0;
#pragma omp parallel for
    for (int64_t d = 0; d < n_det; d++) {
        const float T = resp[2 * d], Pp = resp[2 * d + 1];
        int pi[4][3];
        float w[4];
        float pf[3];
        for (int64_t t = 0; t < n_t; t++) {
            double co[4];
            get_coords(g->proj, bore + 4 * t, det + 4 * d, co);
            spin_factors(nc, co, T, Pp, pf);
            float *sig = signal + d * n_t + t;
            int np = P.get_pixels(co, pi, w);
            for (int ip = 0; ip < np; ip++) {
                const double *base = P.map_base(const_cast<double *>(map), tiles, pi[ip]);
                if (base == nullptr) {
#pragma omp atomic write
                    err = 1;
                    continue;
                }
                int64_t off = P.offset(pi[ip]);
                int64_t pl = P.plane(pi[ip][0]);
                for (int c = 0; c < nc; c++)
                    *sig += base[c * pl + off] * pf[c] * w[ip];
            }
        }
    }
    return err;
}

// to_map_single_thread (Projection.cxx:1772-1811) over a segment list.
int oracle_to_map(const oracle_geom *g, double *map, double *const *tiles,
                  const double *bore, int64_t n_t, const double *det, int64_t n_det,
                  const float *resp, const float *signal, const float *det_weights,
                  const int32_t *seg, int64_t nseg) {
    Pix P{g};
    const int nc = g->ncomp;
    int err = 0;
    Segs S{seg, nseg, n_det, n_t};
    for_each_segment(S, [&](int d, int t0, int t1) {
        float det_wt = 1.;
        if (det_weights) {
            det_wt = det_weights[d];
            if (det_wt == 0.) return;
        }
        const float T = resp[2 * d], Pp = resp[2 * d + 1];
        int pi[4][3];
        float w[4] = {0, 0, 0, 0};
        float pf[3];
        for (int t = t0; t < t1; t++) {
            double co[4];
            get_coords(g->proj, bore + 4 * (int64_t)t, det + 4 * (int64_t)d, co);
            spin_factors(nc, co, T, Pp, pf);
            const float sig = signal[(int64_t)d * n_t + t];
            int np = P.get_pixels(co, pi, w);
            for (int ip = 0; ip < np; ip++) {
                double *base = P.map_base(map, tiles, pi[ip]);
                if (base == nullptr) { err = 1; continue; }
                int64_t off = P.offset(pi[ip]);
                int64_t pl = P.plane(pi[ip][0]);
                for (int c = 0; c < nc; c++)
                    base[c * pl + off] += sig * pf[c] * w[ip] * det_wt;
            }
        }
    });
    return err;
}

// to_weight_map_single_thread (Projection.cxx:1813-1854).  Weight maps
// have leading axes (ncomp, ncomp); only the upper triangle is filled.
int oracle_to_weights(const oracle_geom *g, double *map, double *const *tiles,
                      const double *bore, int64_t n_t, const double *det, int64_t n_det,
                      const float *resp, const float *det_weights,
                      const int32_t *seg, int64_t nseg) {
    Pix P{g};
    const int nc = g->ncomp;
    int err = 0;
    Segs S{seg, nseg, n_det, n_t};
    for_each_segment(S, [&](int d, int t0, int t1) {
        float det_wt = 1.;
        if (det_weights) {
            det_wt = det_weights[d];
            if (det_wt == 0.) return;
        }
        const float T = resp[2 * d], Pp = resp[2 * d + 1];
        int pi[4][3];
        float w[4] = {0, 0, 0, 0};
        float pf[3];
        for (int t = t0; t < t1; t++) {
            double co[4];
            get_coords(g->proj, bore + 4 * (int64_t)t, det + 4 * (int64_t)d, co);
            spin_factors(nc, co, T, Pp, pf);
            int np = P.get_pixels(co, pi, w);
            for (int ip = 0; ip < np; ip++) {
                double *base = P.map_base(map, tiles, pi[ip]);
                if (base == nullptr) { err = 1; continue; }
                int64_t off = P.offset(pi[ip]);
                int64_t pl = P.plane(pi[ip][0]);
                for (int i = 0; i < nc; i++)
                    for (int j = i; j < nc; j++)
                        base[(i * nc + j) * pl + off] +=
                            pf[i] * pf[j] * w[ip] * w[ip] * det_wt;
            }
        }
    });
    return err;
}

// Per-sample domain used by pixel_ranges / tile_ranges
// (Projection.cxx:1489-1506, 1655-1670): the domain of every point of the
// sample if they agree, n_domain if they disagree, -1 if no point.
// mode 0: stripe (pixelizor stripe(), :601-603, 776-780, 894-898, 1106-1108)
// mode 1: value of a (1, ...) thread map (same layout as a T map)
// mode 2: tile -> domain table (tile_dom[n_tile])
void oracle_sample_domains(const oracle_geom *g, const double *bore, int64_t n_t,
                           const double *det, int64_t n_det, int mode, int n_domain,
                           const double *tmap, double *const *ttiles,
                           const int32_t *tile_dom, int32_t *out) {
    Pix P{g};
#pragma omp parallel for
    for (int64_t d = 0; d < n_det; d++) {
        int pi[4][3] = {{-1, 0, 0}};
        float w[4];
        for (int64_t t = 0; t < n_t; t++) {
            double co[4];
            get_coords(g->proj, bore + 4 * t, det + 4 * d, co);
            int np = P.get_pixels(co, pi, w);
            int dom = -1;
            for (int ip = 0; ip < np; ip++) {
                int pd;
                if (mode == 1) {
                    const double *base = P.map_base(const_cast<double *>(tmap), ttiles, pi[ip]);
                    if (base == nullptr) { pd = -2; dom = -2; break; }
                    pd = (int)base[P.offset(pi[ip])];
                } else if (mode == 2) {
                    pd = tile_dom[pi[ip][0]];
                } else {
                    // 32-bit int arithmetic with two's-complement wrap,
                    // as the reference's int expressions compile on x86-64.
                    if (g->pixelizor == PIX_HEALPIX) {
                        if (!g->tiled) {
                            pd = wmul(pi[ip][0], n_domain) / P.npix_hp();
                        } else {
                            int npt = P.npt_hp();
                            int ntiles = 12 * g->nside_tile * g->nside_tile;
                            int ipix = wadd(wmul(pi[ip][0], npt), pi[ip][1]);
                            pd = wmul(ipix, n_domain) / wmul(npt, ntiles);
                        }
                    } else if (!g->tiled) {
                        pd = (pi[ip][0] < 0) ? -1 : wmul(pi[ip][1], n_domain) / g->nx;
                    } else {
                        pd = pi[ip][0] % n_domain;
                    }
                }
                if (ip > 0 && pd != dom) { dom = n_domain; break; }
                dom = pd;
            }
            out[d * n_t + t] = dom;
        }
    }
}

// tile_hits, Projection.cxx:1543-1601.  hits: n_tile.
void oracle_tile_hits(const oracle_geom *g, const double *bore, int64_t n_t,
                      const double *det, int64_t n_det, int32_t *hits) {
    Pix P{g};
    for (int64_t d = 0; d < n_det; d++) {
        int pi[4][3] = {{-1, 0, 0}};
        float w[4];
        for (int64_t t = 0; t < n_t; t++) {
            double co[4];
            get_coords(g->proj, bore + 4 * t, det + 4 * d, co);
            int np = P.get_pixels(co, pi, w);
            for (int ip = 0; ip < np; ip++) hits[pi[ip][0]]++;
        }
    }
}

// ---------------------------------------------------------------------------
// ProjEng_Precomp (Projection.cxx:2012-2253).  pix: (n_det, n_t, k) with k
// the index count; spin: (n_det, n_t, nc).  tiled selects the map layout
// (flat pixelizor only, as in the reference).  For tiled maps the tile
// widths come from (ny, nx, tile_y, tile_x) in g.
// ---------------------------------------------------------------------------
int oracle_precomp_from_map(const oracle_geom *g, const double *map, double *const *tiles,
                            const int32_t *pix, int k, const float *spin, int nc,
                            int64_t n_det, int64_t n_t, float *signal) {
    Pix P{g};
    int err = 0;
#pragma omp parallel for
    for (int64_t d = 0; d < n_det; d++) {
        for (int64_t t = 0; t < n_t; t++) {
            const int32_t *po = pix + (d * n_t + t) * k;
            if (po[0] < 0) continue;
            const double *base = P.map_base(const_cast<double *>(map), tiles, po);
            if (base == nullptr) {
#pragma omp atomic write
                err = 1;
                continue;
            }
            int64_t off = P.offset(po), pl = P.plane(po[0]);
            float s = 0;
            for (int c = 0; c < nc; c++) {
                float pf = spin[(d * n_t + t) * nc + c];
                s += pf * base[c * pl + off];
            }
            signal[d * n_t + t] += s;
        }
    }
    return err;
}

int oracle_precomp_to_map(const oracle_geom *g, double *map, double *const *tiles,
                          const int32_t *pix, int k, const float *spin, int nc,
                          int64_t n_det, int64_t n_t, const float *signal,
                          const float *det_weights, const int32_t *seg, int64_t nseg) {
    Pix P{g};
    int err = 0;
    Segs S{seg, nseg, n_det, n_t};
    for_each_segment(S, [&](int d, int t0, int t1) {
        float det_wt = 1.;
        if (det_weights) {
            det_wt = det_weights[d];
            if (det_wt == 0.) return;
        }
        for (int t = t0; t < t1; t++) {
            int64_t i = (int64_t)d * n_t + t;
            const int32_t *po = pix + i * k;
            if (po[0] < 0) continue;
            const float sig = signal[i];
            const float *pf = spin + i * nc;
            double *base = P.map_base(map, tiles, po);
            if (base == nullptr) { err = 1; continue; }
            int64_t off = P.offset(po), pl = P.plane(po[0]);
            for (int c = 0; c < nc; c++) base[c * pl + off] += det_wt * sig * pf[c];
        }
    });
    return err;
}

int oracle_precomp_to_weights(const oracle_geom *g, double *map, double *const *tiles,
                              const int32_t *pix, int k, const float *spin, int nc,
                              int64_t n_det, int64_t n_t, const float *det_weights,
                              const int32_t *seg, int64_t nseg) {
    Pix P{g};
    int err = 0;
    Segs S{seg, nseg, n_det, n_t};
    for_each_segment(S, [&](int d, int t0, int t1) {
        float det_wt = 1.;
        if (det_weights) {
            det_wt = det_weights[d];
            if (det_wt == 0.) return;
        }
        for (int t = t0; t < t1; t++) {
            int64_t i = (int64_t)d * n_t + t;
            const int32_t *po = pix + i * k;
            if (po[0] < 0) continue;
            const float *pf = spin + i * nc;
            double *base = P.map_base(map, tiles, po);
            if (base == nullptr) { err = 1; continue; }
            int64_t off = P.offset(po), pl = P.plane(po[0]);
            for (int a = 0; a < nc; a++)
                for (int b = a; b < nc; b++)
                    base[(a * nc + b) * pl + off] += pf[a] * pf[b] * det_wt;
        }
    });
    return err;
}

// ---------------------------------------------------------------------------
// Multi-threaded to_map for the CPU baseline: the reference's OpenMP
// schedule (Projection.cxx:1957-1963) with the thread ranges supplied as
// per-thread segment lists; each thread's segments must be pixel-disjoint
// from the others'.  seg_off[n_thread+1] indexes into seg.
// ---------------------------------------------------------------------------
int oracle_to_map_threaded(const oracle_geom *g, double *map, double *const *tiles,
                           const double *bore, int64_t n_t, const double *det,
                           int64_t n_det, const float *resp, const float *signal,
                           const float *det_weights, const int32_t *seg,
                           const int64_t *seg_off, int n_thread) {
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int th = 0; th < n_thread; th++) {
        int e = oracle_to_map(g, map, tiles, bore, n_t, det, n_det, resp, signal,
                              det_weights, seg + 3 * seg_off[th],
                              seg_off[th + 1] - seg_off[th]);
        if (e) {
#pragma omp atomic write
            err = e;
        }
    }
    return err;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
