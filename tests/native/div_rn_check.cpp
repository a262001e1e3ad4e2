// Host check of so3g::div_rn (so3g_amd/csrc/ptg.h) against IEEE division:
// random quotients over the divisors the pointing code uses (the trig
// table step, map cdelt, the spin denominators) and over random exponents,
// exact and near-halfway quotients, zeros, tiny and huge operands.  Prints
// the number of checks and of quotients whose bits differ.
#include "../../so3g_amd/csrc/ptg.h"

#include <cstdio>
#include <cstring>
#include <random>

static uint64_t bits(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 2000000;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1, 1);
    long checks = 0, bad = 0;
    auto check = [&](double a, double b) {
        const double x = so3g::div_rn(a, b, 1. / b), r = a / b;
        checks++;
        if (bits(x) != bits(r) && !(x != x && r != r)) {
            if (bad < 5) printf("a=%a b=%a got %a want %a\n", a, b, x, r);
            bad++;
        }
    };
    const double dx = 1. / 16383;
    for (long i = 0; i < n; i++) {
        check(U(rng), dx);
        check(3.2 * U(rng), (i & 1 ? -1 : 1) * std::pow(10., -5 + 1.5 * (U(rng) + 1)));
        check(U(rng), 0.5 * std::sqrt(std::fabs(U(rng))));
        check(U(rng) * std::ldexp(1., (int)(60 * U(rng))), U(rng) * std::ldexp(1., (int)(60 * U(rng))));
        const double b = 1 + std::fabs(U(rng)), q = 1 + std::fabs(U(rng)), a = b * q;
        check(a, b);
        check(std::nextafter(a, 0.), b);
        check(std::nextafter(a, 4.), b);
    }
    for (double a : {0., -0., 1e-300, -1e-300, 0x1p-1000, 4.9e-324, 1e300, -1e300})
        for (double b : {dx, -3e-4, 0.25, 1e-300, 0x1p-1070})
            check(a, b);
    printf("%ld %ld\n", checks, bad);
    return bad != 0;
}
