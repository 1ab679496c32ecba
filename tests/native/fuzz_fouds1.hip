// Host-side check that the first-order update with each root solved once per
// node (ff_cands0f) returns the same number as the reference-shaped update
// (ff_cands<0>) on random stencils: near-equal neighbours, far (1e38) and
// out-of-grid neighbours, several node spacings and slownesses.  Built with
// hipcc and run on the CPU by tests/test_fm2d_update_equivalence.py.
#include "../../mctomo_amd/csrc/fm2d_fast.hip"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

// the launchers fm2d_fast.hip refers to live in fm2d.hip; not used here
int mct_launch_fm2d_gridder(const mct_fm2d_args &, hipStream_t) { return 0; }
int mct_launch_fm2d_solve(const mct_fm2d_args &, hipStream_t) { return 0; }
int mct_launch_fm2d_rays(const mct_fm2d_args &, hipStream_t) { return 0; }
int mct_launch_fm2d_wave(const mct_fm2d_args &, int32_t *, hipStream_t) { return 0; }
int mct_fill_i32(int32_t *, long, int32_t, hipStream_t) { return 0; }

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 rng(20261202);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double dns[] = {0.5, 0.25, 1.0, 0.125, 0.7, 1.0 / 3.0};
    long bad = 0, far = 0;
    for (long it = 0; it < n; it++) {
        const double dnx = dns[rng() % 6], dnz = dns[rng() % 6];
        const double slown = 1.0 / (1.5 + 4.0 * U(rng));
        const double base = 100.0 * U(rng);
        ff_sten S;
        for (int s = 0; s < 2; s++) {
            S.ix1[s] = (rng() % 8) != 0;
            S.iz1[s] = (rng() % 8) != 0;
            S.ix2[s] = S.iz2[s] = false;
            const double spread = (rng() % 3 == 0) ? 1e-9 : (rng() % 2 ? 0.3 : 3.0);
            S.tx1[s] = (rng() % 6 == 0) ? FF_BIG : base + spread * (2.0 * U(rng) - 1.0);
            S.tz1[s] = (rng() % 6 == 0) ? FF_BIG : base + spread * (2.0 * U(rng) - 1.0);
            if (rng() % 20 == 0) S.tz1[s] = S.tx1[s];
            if (!S.ix1[s]) S.tx1[s] = FF_BIG;
            if (!S.iz1[s]) S.tz1[s] = FF_BIG;
            S.tx2[s] = S.tz2[s] = FF_BIG;
        }
        const double a = ff_cands<0>(S, dnx, dnz, slown);
        const double b = ff_cands0f(S, ff_k0_make(dnx, dnz), slown);
        if (a >= FF_THR0) far++;
        if (memcmp(&a, &b, sizeof a) != 0) {
            if (bad < 5) printf("mismatch %.17g %.17g\n", a, b);
            bad++;
        }
    }
    printf("cases %ld no-update %ld mismatches %ld\n", n, far, bad);
    return bad != 0;
}
