// ptg.h -- device-side pointing, pixelization and spin-projection code for
// the MI355X projection engine.
//
// Every floating-point expression that feeds a pixel index, a coordinate or
// a spin factor is written in the same order as so3g's CPU code
// (src/Projection.cxx) and is compiled without FMA contraction
// (-ffp-contract=off plus the pragma below), so that the GPU reproduces
// the CPU path's float64 results bit for bit: x86-64 doubles with SSE2 and
// no FMA, IEEE-rounded +, -, *, / and sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "cr_atan2.h"

#pragma clang fp contract(off)

namespace so3g {

constexpr double kPi = 3.141592653589793238462643383279502884197;
constexpr int kTrigN = 1 << 14;  // TRIG_TABLE_SIZE, Projection.cxx:34

enum { PROJ_FLAT = 0, PROJ_QUAT, PROJ_CAR, PROJ_CEA, PROJ_ARC, PROJ_SIN,
       PROJ_TAN, PROJ_ZEA };

// Pixelizor kinds (template parameter of the kernels).
enum { PK_FLAT = 0, PK_FLAT_TILED = 1, PK_HP = 2, PK_HP_TILED = 3 };

// Device view of one engine's geometry, passed by value.
struct Geo {
    int32_t ncomp, interp;
    int32_t ny, nx;
    double cdelt_y, cdelt_x, crpix_y, crpix_x;
    double icdelt_y, icdelt_x;   // RN(1 / cdelt)
    int32_t ty, tx, n_tx;        // flat tiles
    int32_t nside, nest, npt;    // HEALPix (npt = pixels per tile)
    int64_t npix;                // pixels per plane, non-tiled
};

// Map view: either one (lead..., npix) array or per-tile arrays.
template <typename T>
struct MapView {
    T *map;                      // non-tiled base
    T *const *tiles;             // tiled: per-tile bases (nullptr = inactive)
    const int32_t *tile_dims;    // tiled: (th, tw) per tile
};

// ---------------------------------------------------------------------------
// a / b, correctly rounded, from yb = RN(1 / b): q = RN(a yb) lies within an
// ulp of a / b, the remainder r = a - b q is exact (one FMA), and RN(q + r yb)
// is the correctly rounded quotient (Markstein's theorem; Muller et al.,
// Handbook of Floating-Point Arithmetic, ch. 4: division with a correctly
// rounded reciprocal).  One multiply and two FMAs instead of a full
// division; the same bits as a / b.  Outside the theorem's range (|a| below
// 2^-900, where the remainder may underflow, a zero or non-finite a, or a
// quotient near overflow) it divides.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ double div_rn(double a, double b, double yb) {
    const double q = a * yb;
    const double r = fma(-b, q, a);
    double qq = fma(r, yb, q);
    if (!(fabs(a) >= 0x1p-900 && fabs(qq) < 0x1p1000)) {
#ifdef __HIP_DEVICE_COMPILE__
        // an identity lane move (DPP quad_perm [0,1,2,3]) of a: convergent,
        // so the division that depends on it stays in this rarely taken
        // branch (the compiler would otherwise compute it for every call
        // and select), while an unused quotient is still removed
        const uint64_t u = __builtin_bit_cast(uint64_t, a);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0xE4, 0xF, 0xF, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0xE4, 0xF, 0xF, false);
        a = __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
#endif
        qq = a / b;
    }
    return qq;
}

// MK: quotients by div_rn (true) or by plain division (false: pixels(), where
// the extra branch of each div_rn measured slower, 4.9 vs 5.1 ms on 7.2e8
// CAR samples, while the kernels that also use the spin terms gain).
template <bool MK>
__device__ __forceinline__ double div_q(double a, double b, double yb) {
    if constexpr (MK) return div_rn(a, b, yb);
    else return a / b;
}

// ---------------------------------------------------------------------------
// Linear-interpolation trig tables.  Lookup arithmetic as in
// include/so_linterp.h:20-29 (raw), :41-45 (asin), :58-70 (atan2).
// ---------------------------------------------------------------------------
template <bool MK = true>
__device__ __forceinline__ double lut_raw(const double *__restrict__ v, double x) {
    constexpr double dx = 1. / (kTrigN - 1);
    constexpr double idx_scale = 1. / dx;   // RN(1 / dx)
    double d = div_q<MK>(x, dx, idx_scale);    // x / dx
    // Inputs are bounded (|x| <= ~1), so d fits an int.
    int idx = (int)d;
    if (x < 0) return v[0];
    if (idx >= kTrigN - 1) return v[kTrigN - 1];
    d -= idx;
    return v[idx] * (1 - d) + v[idx + 1] * d;
}

template <bool MK = true>
__device__ __forceinline__ double asin_lu(const double *__restrict__ v, double x) {
    if (x < 0) return -lut_raw<MK>(v, -x);
    return lut_raw<MK>(v, x);
}

// The table's recursive atan2 folding, unrolled: a negative y negates the
// result, a negative x reflects through pi, then the octant choice.
template <bool MK = true>
__device__ __forceinline__ double atan2_octant(const double *__restrict__ v, double y,
                                               double x) {
    if (y == 0) return v[0];
    if (x < y) return kPi / 2 - lut_raw<MK>(v, x / y);
    return lut_raw<MK>(v, y / x);
}

template <bool MK = true>
__device__ __forceinline__ double atan2_lu(const double *__restrict__ v, double y,
                                           double x) {
    const bool neg = y < 0;
    if (neg) y = -y;
    double r = (x < 0) ? kPi - atan2_octant<MK>(v, y, -x) : atan2_octant<MK>(v, y, x);
    return neg ? -r : r;
}

// ---------------------------------------------------------------------------
// Quaternion product q = l * r with the term order of
// boost::math::quaternion<double>::operator*= (Projection.cxx:47, 493).
// ---------------------------------------------------------------------------
struct Q4 { double a, b, c, d; };

__device__ __forceinline__ Q4 qmul(const Q4 &l, const Q4 &r) {
    Q4 o;
    o.a = l.a * r.a - l.b * r.b - l.c * r.c - l.d * r.d;
    o.b = l.a * r.b + l.b * r.a + l.c * r.d - l.d * r.c;
    o.c = l.a * r.c - l.b * r.d + l.c * r.a + l.d * r.b;
    o.d = l.a * r.d + l.b * r.c - l.c * r.b + l.d * r.a;
    return o;
}

struct Tabs {
    const double *asin_v;
    const double *atan2_v;
};

// Pointer<C>::GetCoords for each coordinate system, Projection.cxx:238-507.
template <int PROJ, bool MK = true>
__device__ __forceinline__ void get_coords(const Q4 &qb, const Q4 &qd, const Tabs &T,
                                           double *co) {
    if constexpr (PROJ == PROJ_FLAT) {
        co[0] = qb.a + qd.a;
        co[1] = qb.b + qd.b;
        co[2] = qb.c * qd.c - qb.d * qd.d;
        co[3] = qb.d * qd.c + qb.c * qd.d;
    } else {
        const Q4 q = qmul(qb, qd);
        const double a = q.a, b = q.b, c = q.c, d = q.d;
        if constexpr (PROJ == PROJ_QUAT) {
            co[0] = a; co[1] = b; co[2] = c; co[3] = d;
        } else if constexpr (PROJ == PROJ_CAR) {
            const double ct = a * a - b * b - c * c + d * d;
            const double h = 0.5 * sqrt(1 - ct * ct);
            co[0] = atan2_lu<MK>(T.atan2_v, c * d - a * b, c * a + d * b);
            co[1] = asin_lu<MK>(T.asin_v, ct);
            const double ih = 1. / h;   // one division for the pair of quotients
            co[2] = div_q<MK>(a * c - b * d, h, ih);
            co[3] = div_q<MK>(c * d + a * b, h, ih);
        } else if constexpr (PROJ == PROJ_CEA) {
            const double ct = a * a - b * b - c * c + d * d;
            const double h = 0.5 * sqrt(1 - ct * ct);
            // the C library's atan2 in the reference (Projection.cxx:468);
            // here correctly rounded (cr_atan2.h)
            co[0] = cr_atan2(c * d - a * b, c * a + d * b);
            co[1] = ct;
            const double ih = 1. / h;
            co[2] = div_q<MK>(a * c - b * d, h, ih);
            co[3] = div_q<MK>(c * d + a * b, h, ih);
        } else if constexpr (PROJ == PROJ_ARC) {
            const double ct2 = a * a + d * d;
            const double st2 = b * b + c * c;
            const double sc = c * a + d * b;
            const double ss = a * b - c * d;
            const double hst = sqrt(sc * sc + ss * ss);
            double R;
            if (hst < 1e-8)
                R = 2 + 1.33333333333 * hst * hst;
            else
                R = asin_lu<MK>(T.asin_v, hst * 2) / hst;
            if (ct2 < st2) R = kPi / hst - R;
            co[0] = ss * R;
            co[1] = sc * R;
            const double ict2 = 1. / ct2;
            co[2] = div_q<MK>(a * a - d * d, ct2, ict2);
            co[3] = div_q<MK>(2 * a * d, ct2, ict2);
        } else if constexpr (PROJ == PROJ_SIN) {
            const double ct2 = a * a + d * d;
            co[0] = 2 * (a * b - c * d);
            co[1] = 2 * (c * a + d * b);
            const double ict2 = 1. / ct2;
            co[2] = div_q<MK>(a * a - d * d, ct2, ict2);
            co[3] = div_q<MK>(2 * a * d, ct2, ict2);
        } else if constexpr (PROJ == PROJ_TAN) {
            const double ct2 = a * a + d * d;
            const double ct = 2 * ct2 - 1;
            const double ict = 1. / ct;
            co[0] = div_q<MK>((a * b - c * d) * 2, ct, ict);
            co[1] = div_q<MK>((a * c + b * d) * 2, ct, ict);
            const double ict2 = 1. / ct2;
            co[2] = div_q<MK>(a * a - d * d, ct2, ict2);
            co[3] = div_q<MK>(2 * a * d, ct2, ict2);
        } else {  // PROJ_ZEA
            const double ct2 = a * a + d * d;
            const double c2 = sqrt(ct2);
            const double ic2 = 1. / c2;
            co[0] = div_q<MK>((a * b - c * d) * 2, c2, ic2);
            co[1] = div_q<MK>((a * c + b * d) * 2, c2, ic2);
            const double ict2 = 1. / ct2;
            co[2] = div_q<MK>(a * a - d * d, ct2, ict2);
            co[3] = div_q<MK>(2 * a * d, ct2, ict2);
        }
    }
}

// ---------------------------------------------------------------------------
// HEALPix ang2pix (Gorski et al. 2005), with the operation sequence of the
// healpix_bare routines so3g compiles in (src/healpix_bare.c:46-78,
// 109-114, 152-161, 178-211, 271-275, 303-306).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t hp_spread(int64_t v) {
    int64_t r = v & 0xffffffff;
    r = (r ^ (r << 16)) & 0x0000ffff0000ffff;
    r = (r ^ (r << 8)) & 0x00ff00ff00ff00ff;
    r = (r ^ (r << 4)) & 0x0f0f0f0f0f0f0f0f;
    r = (r ^ (r << 2)) & 0x3333333333333333;
    r = (r ^ (r << 1)) & 0x5555555555555555;
    return r;
}

// The same interleave of a value below 2^16 in 32-bit arithmetic.
__device__ __forceinline__ uint32_t hp_spread16(uint32_t v) {
    uint32_t r = v & 0xffffu;
    r = (r ^ (r << 8)) & 0x00ff00ffu;
    r = (r ^ (r << 4)) & 0x0f0f0f0fu;
    r = (r ^ (r << 2)) & 0x33333333u;
    r = (r ^ (r << 1)) & 0x55555555u;
    return r;
}

__device__ __forceinline__ int64_t hp_ang2pix(int64_t ns, bool nest, double theta,
                                              double phi) {
    // sin and cos from one argument reduction: the same reduction and
    // polynomials as separate sin() and cos() calls, so the same bits
    double z, s;
    sincos(theta, &s, &z);
    if (s < 0) { s = -s; phi += kPi; }
    // Continuous face coordinates.
    const double za = fabs(z);
    // phi is a longitude, so these truncations fit 32 bits (the reference
    // truncates to 64 bits: the same values)
    double x = phi * (1. / (2. * kPi));
    if (x < 0.)
        x += (int)x + 1.;
    else if (x >= 1.)
        x -= (int)x;
    const double tt = 4. * x;
    double fx, fy;
    int32_t face;
    if (za <= 2. / 3.) {
        const double t1 = 0.5 + tt;
        const double t2 = z * 0.75;
        const double jp = t1 - t2;
        const double jm = t1 + t2;
        const int ifp = (int)jp;
        const int ifm = (int)jm;
        fx = jm - ifm;
        fy = 1 + ifp - jp;
        face = (ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : (ifm + 8));
    } else {
        int ntt = (int)tt;
        if (ntt >= 4) ntt = 3;
        const double tp = tt - ntt;
        const double tmp = s / sqrt((1. + za) * (1. / 3.));
        double jp = tp * tmp;
        double jm = (1.0 - tp) * tmp;
        if (jp > 1.) jp = 1.;
        if (jm > 1.) jm = 1.;
        if (z >= 0) { fx = 1. - jm; fy = 1. - jp; face = (int32_t)ntt; }
        else { fx = jp; fy = jm; face = (int32_t)ntt + 8; }
    }
    // 0 <= fx, fy <= 1 and nside <= 2^29: 32-bit truncations
    const int64_t ix = (int32_t)(fx * ns);
    const int64_t iy = (int32_t)(fy * ns);
    if (nest) {
        if (ns <= 65536)   // face coordinates below 2^16: 32-bit interleave
            return (face * ns * ns) +
                   (int64_t)(hp_spread16((uint32_t)ix) | (hp_spread16((uint32_t)iy) << 1));
        return (face * ns * ns) + hp_spread(ix) + (hp_spread(iy) << 1);
    }
    // Ring index.  Face row jrll = {2,2,2,2,3,3,3,3,4,4,4,4} and face
    // column jpll = {1,3,5,7,0,2,4,6,1,3,5,7} (odd for the polar faces,
    // even for the equatorial ones).
    const int jrll = (face >> 2) + 2;
    const int jp_base = ((face >> 2) == 1) ? ((face & 3) << 1) : (((face & 3) << 1) + 1);
    const int64_t nl4 = 4 * ns;
    int64_t jr = (jrll * ns) - ix - iy - 1;
    if (jr < ns) {
        int64_t jp = (jp_base * jr + ix - iy + 1) / 2;
        jp = (jp > nl4) ? jp - nl4 : ((jp < 1) ? jp + nl4 : jp);
        return 2 * jr * (jr - 1) + jp - 1;
    } else if (jr > 3 * ns) {
        jr = nl4 - jr;
        int64_t jp = (jp_base * jr + ix - iy + 1) / 2;
        jp = (jp > nl4) ? jp - nl4 : ((jp < 1) ? jp + nl4 : jp);
        return 12 * ns * ns - 2 * (jr + 1) * jr + jp - 1;
    }
    int64_t jp = (jp_base * ns + ix - iy + 1 + ((jr - ns) & 1)) / 2;
    jp = (jp > nl4) ? jp - nl4 : ((jp < 1) ? jp + nl4 : jp);
    return 2 * ns * (ns - 1) + (jr - ns) * nl4 + jp - 1;
}

// ---------------------------------------------------------------------------
// Pixelizors.
// ---------------------------------------------------------------------------

// Pixel index tuple + map element offset of one interpolation point.
struct Pt {
    int32_t i0, i1, i2;   // index tuple (unused slots ignored)
    float w;              // interpolation weight
};

// co / cdelt through the geometry's RN(1 / cdelt) (div_rn: the same bits)
template <bool MK = true>
__device__ __forceinline__ double flat_x(const Geo &g, const double *co) {
    return div_q<MK>(co[0], g.cdelt_x, g.icdelt_x) + g.crpix_x - 1 + 0.5;
}
template <bool MK = true>
__device__ __forceinline__ double flat_y(const Geo &g, const double *co) {
    return div_q<MK>(co[1], g.cdelt_y, g.icdelt_y) + g.crpix_y - 1 + 0.5;
}

template <int PK>
constexpr int index_count() {
    return PK == PK_FLAT ? 2 : PK == PK_FLAT_TILED ? 3 : PK == PK_HP ? 1 : 2;
}

// GetPixel (pixels(), pointing_matrix()): Projection.cxx:558-564, 704-712,
// 846-863, 1030-1051.  Returns false for an off-map sample; then only slot
// 0 is meaningful (-1).
template <int PK, bool MK = true>
__device__ __forceinline__ bool get_pixel(const Geo &g, const double *co, int32_t *pi) {
    if constexpr (PK == PK_HP || PK == PK_HP_TILED) {
        const double phi = co[0];
        const double theta = kPi / 2 - co[1];
        const int ix = (int)hp_ang2pix(g.nside, PK == PK_HP_TILED || g.nest, theta, phi);
        if constexpr (PK == PK_HP) {
            pi[0] = ix;
        } else {
            pi[0] = ix / g.npt;
            pi[1] = ix % g.npt;
        }
        return true;
    } else {
        const double x = flat_x<MK>(g, co);
        if (x < 0 || x >= g.nx) { pi[0] = -1; return false; }
        const double y = flat_y<MK>(g, co);
        if (y < 0 || y >= g.ny) { pi[0] = -1; return false; }
        const int iy = int(y), ix = int(x);
        if constexpr (PK == PK_FLAT) {
            pi[0] = iy;
            pi[1] = ix;
        } else {
            const int sy = iy / g.ty, sx = ix / g.tx;
            pi[0] = sy * g.n_tx + sx;
            pi[1] = iy - sy * g.ty;
            pi[2] = ix - sx * g.tx;
        }
        return true;
    }
}

// GetPixels (projection operations): Projection.cxx:566-574, 714-724,
// 913-952, 1125-1166.  Fills up to 4 points, returns their count.
template <int PK, int INTERP>
__device__ __forceinline__ int get_pixels(const Geo &g, const double *co, Pt *p) {
    if constexpr (PK == PK_HP || PK == PK_HP_TILED) {
        int32_t pi[2];
        get_pixel<PK>(g, co, pi);
        p[0].i0 = pi[0];
        p[0].i1 = pi[1];
        p[0].w = 1;
        return 1;
    } else if constexpr (INTERP == 0) {
        if constexpr (PK == PK_FLAT) {
            const double x = flat_x(g, co);
            const double y = flat_y(g, co);
            if (x < 0 || x >= g.nx) return 0;
            if (y < 0 || y >= g.ny) return 0;
            p[0].i0 = int(y);
            p[0].i1 = int(x);
            p[0].w = 1;
            return 1;
        } else {
            // Tiled nearest neighbour truncates before the bounds test.
            const int ix = int(flat_x(g, co));
            if (ix < 0 || ix >= g.nx) return 0;
            const int iy = int(flat_y(g, co));
            if (iy < 0 || iy >= g.ny) return 0;
            const int sy = iy / g.ty, sx = ix / g.tx;
            p[0].i0 = sy * g.n_tx + sx;
            p[0].i1 = iy - sy * g.ty;
            p[0].i2 = ix - sx * g.tx;
            p[0].w = 1;
            return 1;
        }
    } else {
        const double x = flat_x(g, co);
        const double y = flat_y(g, co);
        int x1, y1;
        if constexpr (PK == PK_FLAT) {
            x1 = int(x) - int(x < 0);
            y1 = int(y) - int(y < 0);
        } else {
            x1 = int(x);
            y1 = int(y);
        }
        const double wx[2] = {x - x1, 1 - (x - x1)};
        const double wy[2] = {y - y1, 1 - (y - y1)};
        int n = 0;
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const int iy = y1 + dy;
            if (iy < 0 || iy >= g.ny) continue;
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
                const int ix = x1 + dx;
                if (ix < 0 || ix >= g.nx) continue;
                if constexpr (PK == PK_FLAT) {
                    p[n].i0 = iy;
                    p[n].i1 = ix;
                } else {
                    const int sy = iy / g.ty, sx = ix / g.tx;
                    p[n].i0 = sy * g.n_tx + sx;
                    p[n].i1 = iy - sy * g.ty;
                    p[n].i2 = ix - sx * g.tx;
                }
                p[n].w = wy[dy] * wx[dx];
                n++;
            }
        }
        return n;
    }
}

// Address of slot `lead` (component, or i*ncomp+j for weights) of the
// map element for point p; nullptr for an inactive tile.
template <int PK, typename T>
__device__ __forceinline__ T *elem(const Geo &g, const MapView<T> &m, const Pt &p,
                                   int64_t lead) {
    if constexpr (PK == PK_FLAT) {
        return m.map + lead * g.npix + ((int64_t)p.i0 * g.nx + p.i1);
    } else if constexpr (PK == PK_HP) {
        return m.map + lead * g.npix + p.i0;
    } else if constexpr (PK == PK_HP_TILED) {
        T *b = m.tiles[p.i0];
        if (b == nullptr) return nullptr;
        return b + lead * g.npt + p.i1;
    } else {
        T *b = m.tiles[p.i0];
        if (b == nullptr) return nullptr;
        const int th = m.tile_dims[2 * p.i0], tw = m.tile_dims[2 * p.i0 + 1];
        return b + lead * ((int64_t)th * tw) + ((int64_t)p.i1 * tw + p.i2);
    }
}

// spin_proj_factors<S>, Projection.cxx:1184-1209: computed in float64
// (the float response is promoted), stored as float32.
__device__ __forceinline__ void spin_factors(int ncomp, const double *co, float T,
                                             float P, float *pf) {
    if (ncomp == 1) { pf[0] = T; return; }
    const double c = co[2], s = co[3];
    const float q = P * (c * c - s * s);
    const float u = P * (2 * c * s);
    if (ncomp == 2) { pf[0] = q; pf[1] = u; return; }
    pf[0] = T; pf[1] = q; pf[2] = u;
}

__device__ __forceinline__ Q4 load_q(const double *__restrict__ p) {
    const double2 lo = *reinterpret_cast<const double2 *>(p);
    const double2 hi = *reinterpret_cast<const double2 *>(p + 2);
    return Q4{lo.x, lo.y, hi.x, hi.y};
}

}  // namespace so3g
