// Host check of div_r (harm_phys.hpp) against IEEE division, bit for bit:
// random quotients over a wide exponent range, quotients within 2^-60 ulp of a
// rounding midpoint (the hard cases of correct rounding), significands near 1
// and 2 for divisor and dividend, mixed signs and signed zeros.
// Usage: fdiv_check <n> ; prints "checked N mismatches M".
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../harm_cool_amd/csrc/harm_phys.hpp"

using harm::div_r;

static uint64_t bits(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
static double from_bits(uint64_t u)
{
    double x;
    memcpy(&x, &u, 8);
    return x;
}
static uint64_t s_state = 0x9E3779B97F4A7C15ull;
static uint64_t next()  // splitmix64
{
    uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unit() { return (next() >> 11) * 0x1.0p-53; }
static double with_exp(uint64_t mant, int e) { return from_bits(((uint64_t)(1023 + e) << 52) | (mant & ((1ull << 52) - 1))); }

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    long bad = 0, checked = 0;
    for (long t = 0; t < n; t++) {
        double a, b;
        switch (t % 4) {
        case 0:
            a = with_exp(next(), (int)(next() % 400) - 200);
            b = with_exp(next(), (int)(next() % 400) - 200);
            break;
        case 1: {
            // a = RN(m b) for a midpoint m = (2K+1) 2^-53 in [1, 2): a/b lies next to m
            b = with_exp(next(), (int)(next() % 60) - 30);
            const uint64_t M = (1ull << 53) | (next() & ((1ull << 53) - 1)) | 1ull;
            const long double m = (long double)M / (long double)(1ull << 53);
            a = (double)(m * (long double)b);
            break;
        }
        case 2: {
            const uint64_t mb = (next() & 1) ? (next() & 0xfffff) : ((1ull << 52) - 1 - (next() & 0xfffff));
            a = with_exp((1ull << 52) - 1 - (next() & 0xffffff), (int)(next() % 40) - 20);
            b = with_exp(mb, (int)(next() % 40) - 20);
            break;
        }
        default:
            a = -with_exp(next(), (int)(next() % 100) - 50);
            b = with_exp(next(), (int)(next() % 100) - 50);
            if (next() & 1) b = -b;
            if (next() & 1) a = unit() < 0.5 ? 0. : -0.;
            break;
        }
        const double rb = 1. / b;
        checked++;
        if (bits(div_r(a, b, rb)) != bits(a / b)) {
            if (bad < 5) printf("mismatch a=%a b=%a want %a got %a\n", a, b, a / b, div_r(a, b, rb));
            bad++;
        }
    }
    printf("checked %ld mismatches %ld\n", checked, bad);
    return bad != 0;
}
