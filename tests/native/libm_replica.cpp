// Host-side check that libm_glibc.h's exp and sincos return glibc's numbers:
// random arguments over each branch of the two functions, compared bit for
// bit with the C library this program links.  glibc's sincos is called
// through a pointer so the compiler cannot replace it.  Built with
// -ffp-contract=off and run by tests/test_libm_replica.py.
#include "../../mctomo_amd/csrc/libm_glibc.h"
#include <cmath>
#include <cstdio>
#include <cstdlib>

extern "C" void sincos(double, double *, double *);
static void (*volatile lib_sincos)(double, double *, double *) = sincos;
static double (*volatile lib_exp)(double) = ::exp;

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    // branch edges: tiny, polynomial (< 0.126), table (< 0.855), pi/2 - |x|
    // (< 2.426), reduced; both signs
    const double sc_ranges[][2] = {{1e-9, 0.126}, {0.126, 0.855}, {0.855, 2.426}, {2.426, 60.0},
                                   {60.0, 4000.0}, {-300.0, 300.0}};
    const double ex_ranges[][2] = {{-130.0, 0.0}, {-32.0, 0.0}, {-1e-6, 1e-6}, {-1.0, 1.0}, {0.0, 300.0}};
    unsigned long long s = 88172645463325252ULL;
    auto next = [&]() {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        return (s >> 11) * 0x1p-53;
    };
    long bad = 0;
    for (const auto &r : sc_ranges) {
        long b = 0;
        for (long i = 0; i < n; i++) {
            const double x = r[0] + (r[1] - r[0]) * next();
            double gs, gc, ms, mc;
            lib_sincos(x, &gs, &gc);
            mct_libm::libm_sincos(x, &ms, &mc);
            if (gs != ms || gc != mc) {
                if (b < 3) printf("sincos(%a): glibc %a %a, here %a %a\n", x, gs, gc, ms, mc);
                b++;
            }
        }
        printf("sincos [%g, %g]: %ld arguments, mismatches %ld\n", r[0], r[1], n, b);
        bad += b;
    }
    for (const auto &r : ex_ranges) {
        long b = 0;
        for (long i = 0; i < n; i++) {
            const double x = r[0] + (r[1] - r[0]) * next();
            const double g = lib_exp(x), m = mct_libm::libm_exp(x);
            if (g != m) {
                if (b < 3) printf("exp(%a): glibc %a, here %a\n", x, g, m);
                b++;
            }
        }
        printf("exp [%g, %g]: %ld arguments, mismatches %ld\n", r[0], r[1], n, b);
        bad += b;
    }
    // special arguments: zero, the smallest normals, branch boundaries
    const double sp[] = {0.0, -0.0, 0x1p-1022, 0x1p-60, 0x1p-54, 0x1p-27, 0x1p-26, 0.126, 0.85546875,
                         2.426265, 1.0e8, 1.0, -1.0, 16.0, -32.0, -60.0};
    for (double x : sp) {
        double gs, gc, ms, mc;
        lib_sincos(x, &gs, &gc);
        mct_libm::libm_sincos(x, &ms, &mc);
        if (gs != ms || gc != mc) { printf("sincos(%a) special: glibc %a %a, here %a %a\n", x, gs, gc, ms, mc); bad++; }
        if (fabs(x) < 512 && lib_exp(x) != mct_libm::libm_exp(x)) {
            printf("exp(%a) special: glibc %a, here %a\n", x, lib_exp(x), mct_libm::libm_exp(x));
            bad++;
        }
    }
    printf("total mismatches %ld\n", bad);
    return bad != 0;
}
