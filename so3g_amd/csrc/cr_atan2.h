// cr_atan2.h -- atan2 in double-double arithmetic, rounded once to double.
//
// so3g's CEA pointer takes the longitude from the C library's atan2
// (Projection.cxx:468; glibc on the reference's x86-64 build).  glibc's
// atan2 returns the correctly rounded result for these arguments, so the
// device computes atan2 to about 2^-100 relative accuracy and rounds once:
// the nearest double, the same bits.  ROCm's own atan2 (1-2 ulp) differed in
// a quarter of the benchmark's samples.
//
// Method (double-double arithmetic as in Hida, Li & Bailey, "Algorithms for
// quad-double precision floating point arithmetic", 2001: error-free sums
// and FMA products):
//   t = min(|x|, |y|) / max(|x|, |y|) in [0, 1] (double-double);
//   k = nearest integer to 128 t, c = k / 128;
//   atan(t) = atan(c) + atan(u),  u = (t - c) / (1 + t c),  |u| <= 1/256;
//   atan(u) = u (1 - u^2/3 + u^4/5 - ... ), the u^2 term in double-double,
//   u^4 .. u^10 in double (they are below 5e-11);
//   octant and quadrant by pi/2 - a and pi - a; the sign of y last.
// atan(k / 128) is tabulated as a double-double pair (hi + lo, |lo| <= ulp(hi)/2),
// computed with 80-digit decimal arithmetic (argument halving to |x| < 0.05,
// then the Taylor series).  Zeros, infinities and NaNs take the library's
// atan2, whose results there are exact constants.
#pragma once

#include <hip/hip_runtime.h>

// the error-free transformations below must not be contracted into FMAs
#pragma clang fp contract(off)

namespace so3g {

struct DD {
    double hi, lo;
};

__host__ __device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}

__host__ __device__ __forceinline__ DD dd_fast_two_sum(double a, double b) {   // |a| >= |b|
    const double s = a + b;
    return DD{s, b - (s - a)};
}

__host__ __device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}

__host__ __device__ __forceinline__ DD dd_neg(DD a) { return DD{-a.hi, -a.lo}; }

__host__ __device__ __forceinline__ DD dd_mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return dd_fast_two_sum(p, e);
}

__host__ __device__ __forceinline__ DD dd_mul_d(DD a, double b) {
    const double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e += a.lo * b;
    return dd_fast_two_sum(p, e);
}

__host__ __device__ __forceinline__ DD dd_div(DD a, DD b) {
    const double q1 = a.hi / b.hi;
    DD r = dd_add(a, dd_neg(dd_mul_d(b, q1)));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_neg(dd_mul_d(b, q2)));
    const double q3 = r.hi / b.hi;
    return dd_add(dd_fast_two_sum(q1, q2), DD{q3, 0.});
}

// atan(k / 128), k = 0 .. 128, as double-double pairs.
__host__ __device__ __forceinline__ DD atan_k128(int k) {
    constexpr double tab[129][2] = {
    {0x0.0p+0, 0x0.0p+0},
    {0x1.fffd555bbba97p-8, 0x1.68062351fbbe6p-63},
    {0x1.fff555bbb729bp-7, -0x1.220c39d4dff50p-61},
    {0x1.7fee0184a5c36p-6, -0x1.43189fc0a354bp-60},
    {0x1.ffd55bba97625p-6, -0x1.5ec431444912cp-60},
    {0x1.3fd65f169c9d9p-5, 0x1.7230a716461b5p-61},
    {0x1.7fb818430da2ap-5, -0x1.86ef8f794f105p-63},
    {0x1.bf8ddf139c444p-5, -0x1.89fe34b2a7fa8p-59},
    {0x1.ff55bb72cfdeap-5, -0x1.c934d86d23f1dp-60},
    {0x1.1f86dbf082d59p-4, -0x1.095dc7732ef81p-59},
    {0x1.3f59f0e7c559dp-4, 0x1.ac4ce285df847p-58},
    {0x1.5f2324fd2d7b2p-4, 0x1.8a8da4401318ep-58},
    {0x1.7ee182602f10fp-4, -0x1.cfb654c0c3d98p-58},
    {0x1.9e94153cfdcf1p-4, 0x1.a332e1d69c47ep-58},
    {0x1.be39ebe6f07c3p-4, 0x1.f7b8f29a05987p-58},
    {0x1.ddd21701eba6ep-4, 0x1.94effcd76fe58p-58},
    {0x1.fd5ba9aac2f6ep-4, -0x1.cd37686760c17p-59},
    {0x1.0e6adccf40882p-3, -0x1.d71a31bb98d0dp-57},
    {0x1.1e1fafb043727p-3, -0x1.b485914dacf8cp-59},
    {0x1.2dcbdb2fba1ffp-3, 0x1.8f28705561534p-58},
    {0x1.3d6eee8c6626cp-3, 0x1.61a3b0ce9281bp-57},
    {0x1.4d087a9da4f17p-3, 0x1.1f323f1adf158p-57},
    {0x1.5c9811e3ec26ap-3, -0x1.054ab2c010f3dp-58},
    {0x1.6c1d4898933d9p-3, -0x1.2954a7603c427p-58},
    {0x1.7b97b4bce5b02p-3, 0x1.347b0b4f881cap-58},
    {0x1.8b06ee2879c29p-3, -0x1.118cd30308c4fp-57},
    {0x1.9a6a8e96c8626p-3, 0x1.cf601e7b4348ep-59},
    {0x1.a9c231b403279p-3, 0x1.0e8bbe89cca85p-57},
    {0x1.b90d7529260a2p-3, 0x1.17b10d2e0e5abp-61},
    {0x1.c84bf8a742e6ep-3, -0x1.95bdd0682ea26p-58},
    {0x1.d77d5df205736p-3, 0x1.c648d1534597ep-57},
    {0x1.e6a148e96ec4dp-3, 0x1.866b22029f765p-57},
    {0x1.f5b75f92c80ddp-3, 0x1.8ab6e3cf7afbdp-57},
    {0x1.025fa510665b6p-2, -0x1.672df6832fa48p-56},
    {0x1.09dc597d86362p-2, 0x1.62e47390cb865p-56},
    {0x1.1151a362431cap-2, -0x1.4dc8dc9077b9fp-56},
    {0x1.18bf5a30bf178p-2, 0x1.30ca4748b1bf9p-57},
    {0x1.2025567e47c96p-2, -0x1.1832328f4290ep-57},
    {0x1.278372057ef46p-2, -0x1.077cdd36dfc81p-56},
    {0x1.2ed987a823cfep-2, 0x1.b91258ea012cap-57},
    {0x1.362773707ebccp-2, -0x1.963a544b672d8p-57},
    {0x1.3d6d129271134p-2, 0x1.137ca41cc958ap-56},
    {0x1.44aa436c2af0ap-2, -0x1.5d5e43c55b3bap-56},
    {0x1.4bdee586890e7p-2, -0x1.e4dc77c22a757p-57},
    {0x1.530ad9951cd4ap-2, -0x1.2566480884082p-57},
    {0x1.5a2e0175e0f4ep-2, 0x1.13b7a8f82e457p-56},
    {0x1.614840309cfe2p-2, -0x1.a725715711f00p-56},
    {0x1.685979f5fa6fep-2, -0x1.257814d1ada9cp-59},
    {0x1.6f61941e4def1p-2, -0x1.c63aae6f6e918p-56},
    {0x1.7660752817502p-2, -0x1.dd11791cc7600p-59},
    {0x1.7d5604b63b3f7p-2, 0x1.69c885c2b249ap-56},
    {0x1.84422b8df95d7p-2, 0x1.d76a0299b41b6p-56},
    {0x1.8b24d394a1b25p-2, 0x1.b6d0ba3748fa8p-56},
    {0x1.91fde7cd0c662p-2, 0x1.1074188054b53p-56},
    {0x1.98cd5454d6b18p-2, 0x1.9e6c988fd0a77p-56},
    {0x1.9f93066168002p-2, -0x1.c827047c9439ap-56},
    {0x1.a64eec3cc23fdp-2, -0x1.24dec1b50b7ffp-56},
    {0x1.ad00f5422058bp-2, 0x1.fc4c33891d2e8p-56},
    {0x1.b3a911da65c6cp-2, 0x1.ae187b1ca5040p-56},
    {0x1.ba473378624a5p-2, 0x1.519a1b46e4affp-56},
    {0x1.c0db4c94ec9f0p-2, -0x1.cc1ce70934c34p-56},
    {0x1.c76550aad71f9p-2, -0x1.74b8bff7043e4p-56},
    {0x1.cde53432c1351p-2, -0x1.a2cfa4418f1adp-56},
    {0x1.d45aec9ec862bp-2, 0x1.89421163ef92dp-57},
    {0x1.dac670561bb4fp-2, 0x1.a2b7f222f65e2p-56},
    {0x1.e127b6b0744b0p-2, -0x1.2b0986398d4abp-58},
    {0x1.e77eb7f175a34p-2, 0x1.0e53dc1bf3435p-56},
    {0x1.edcb6d43f8435p-2, -0x1.fc976330884e4p-58},
    {0x1.f40dd0b541418p-2, -0x1.a3992dc382a23p-57},
    {0x1.fa45dd3029259p-2, -0x1.ca563dc28d8b5p-56},
    {0x1.0039c73c1a40cp-1, -0x1.b32c949c9d593p-55},
    {0x1.034b709250488p-1, 0x1.8f9b38d855410p-56},
    {0x1.0657e94db30d0p-1, -0x1.d5b495f6349e6p-56},
    {0x1.095f30861a590p-1, -0x1.121b20a15a9f3p-56},
    {0x1.0c6145b5b43dap-1, 0x1.974fa13b5404fp-58},
    {0x1.0f5e28b67e295p-1, 0x1.311b17ec990d0p-65},
    {0x1.1255d9bfbd2a9p-1, -0x1.2bdaee1c0ee35p-58},
    {0x1.154859637646ap-1, -0x1.4ba7c548bf3c3p-55},
    {0x1.1835a88be7c13p-1, 0x1.c621cec00c301p-55},
    {0x1.1b1dc87904285p-1, -0x1.21e8c8aef8f29p-57},
    {0x1.1e00babdefeb4p-1, -0x1.928df287a668fp-58},
    {0x1.20de813e823b2p-1, -0x1.791d753ebb744p-55},
    {0x1.23b71e2cc9e6ap-1, 0x1.c421c9f38224ep-57},
    {0x1.268a940696da6p-1, 0x1.d1348a04c73ccp-58},
    {0x1.2958e59308e31p-1, -0x1.09e73b0c6c087p-56},
    {0x1.2c2215e024466p-1, -0x1.4b810da3a4be1p-59},
    {0x1.2ee628406cbcap-1, 0x1.c5d5e9ff0cf8dp-55},
    {0x1.31a52048874bep-1, 0x1.40cab87a7ac24p-55},
    {0x1.345f01cce37bbp-1, 0x1.1021137c71102p-55},
    {0x1.3713d0df6c504p-1, -0x1.4f789e031606dp-58},
    {0x1.39c391cd4171ap-1, -0x1.2304331d8bf46p-55},
    {0x1.3c6e491c78dc5p-1, -0x1.e145094fd0ba7p-55},
    {0x1.3f13fb89e96f4p-1, 0x1.ecf8b492644f0p-56},
    {0x1.41b4ae06fea41p-1, 0x1.3d60a53277652p-57},
    {0x1.445065b795b56p-1, -0x1.f76d0163f79c8p-56},
    {0x1.46e727efe4716p-1, -0x1.39b9b1b844cc9p-57},
    {0x1.4978fa3269ee1p-1, 0x1.2419a87f2a458p-56},
    {0x1.4c05e22de94e5p-1, -0x1.c0ac1f09f2edfp-55},
    {0x1.4e8de5bb6ec04p-1, 0x1.4a33dbeb3796cp-55},
    {0x1.51110adc5ed81p-1, 0x1.23dcd6832a63ep-56},
    {0x1.538f57b89061fp-1, -0x1.1bb74abda520cp-55},
    {0x1.5608d29c70c34p-1, 0x1.9939cf0de8088p-55},
    {0x1.587d81f732fbbp-1, -0x1.5e5c9d8c5a950p-56},
    {0x1.5aed6c5909517p-1, 0x1.7312f714a9436p-55},
    {0x1.5d58987169b18p-1, 0x1.0028e4bc5e7cap-57},
    {0x1.5fbf0d0d5cc4ap-1, -0x1.b4cfd000b7158p-58},
    {0x1.6220d115d7b8ep-1, -0x1.2b785350ee8c1p-57},
    {0x1.647deb8e20b90p-1, -0x1.eca04023a51cfp-58},
    {0x1.66d663923e087p-1, -0x1.6ea6febe8bbbap-56},
    {0x1.692a40556fb6ap-1, 0x1.d94b95a8ea2ccp-55},
    {0x1.6b798920b3d99p-1, -0x1.a80386188c50ep-55},
    {0x1.6dc44551553afp-1, -0x1.bf8863573828ep-58},
    {0x1.700a7c5784634p-1, -0x1.8c34d25aadef6p-56},
    {0x1.724c35b4fae7bp-1, 0x1.948b32db3499bp-58},
    {0x1.748978fba8e0fp-1, 0x1.7b2a6165884a1p-59},
    {0x1.76c24dcc6c6c0p-1, 0x1.1952551adc83dp-55},
    {0x1.78f6bbd5d315ep-1, 0x1.406a089803740p-55},
    {0x1.7b26cad2e50fep-1, -0x1.ce80df30411fbp-55},
    {0x1.7d528289fa093p-1, 0x1.560821e2f3aa9p-55},
    {0x1.7f79eacb97898p-1, 0x1.fd5ca80ead221p-55},
    {0x1.819d0b7158a4dp-1, -0x1.bf76229d3b917p-56},
    {0x1.83bbec5cdee22p-1, 0x1.3107104ffc6c3p-57},
    {0x1.85d69576cc2c5p-1, 0x1.6b66e7fc8b8c3p-57},
    {0x1.87ed0eadc5a2ap-1, 0x1.0af5ad957f4bcp-56},
    {0x1.89ff5ff57f1f8p-1, -0x1.55b9a5e177a1bp-55},
    {0x1.8c0d9145cf49dp-1, 0x1.bea4076dc4333p-55},
    {0x1.8e17aa99cc05ep-1, -0x1.ec182ab042f61p-56},
    {0x1.901db3eeef187p-1, 0x1.68665e5603c8fp-55},
    {0x1.921fb54442d18p-1, 0x1.1a62633145c07p-55},
    };
    return DD{tab[k][0], tab[k][1]};
}

__host__ __device__ __forceinline__ double cr_atan2(double y, double x) {
    if (!(isfinite(x) && isfinite(y)) || x == 0. || y == 0.) return atan2(y, x);
    const double ax = fabs(x), ay = fabs(y);
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    const double th = num / den;
    const DD t = dd_fast_two_sum(th, fma(-th, den, num) / den);   // num / den
    const int k = (int)(t.hi * 128. + 0.5);
    const double c = (double)k * (1. / 128.);
    const DD u = dd_div(dd_add(t, DD{-c, 0.}), dd_add(DD{1., 0.}, dd_mul_d(t, c)));
    const DD u2 = dd_mul(u, u);
    const double p = u2.hi;
    const double tail = p * p * (1. / 5. - p * (1. / 7. - p * (1. / 9. - p * (1. / 11.))));
    const DD third{0x1.5555555555555p-2, 0x1.5555555555555p-56};
    const DD s = dd_add(dd_neg(dd_mul(u2, third)), DD{tail, 0.});   // -u^2/3 + u^4/5 - ...
    DD a = dd_add(atan_k128(k), dd_add(u, dd_mul(u, s)));
    if (swap) a = dd_add(DD{0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54}, dd_neg(a));
    if (x < 0.) a = dd_add(DD{0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53}, dd_neg(a));
    const double r = a.hi + a.lo;
    return y < 0. ? -r : r;
}

}  // namespace so3g
