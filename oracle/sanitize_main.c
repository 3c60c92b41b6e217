/*
 * sanitize_main.c -- calls every entry point of the CPU oracle on small
 * inputs, edge cases included, so that an AddressSanitizer +
 * UndefinedBehaviorSanitizer build (`make -C oracle sanitize`) checks the
 * oracle's memory accesses and arithmetic.  TEST INFRASTRUCTURE ONLY, like
 * the oracle itself.  Exit status 0 means every call returned; the sanitizers
 * abort on the first finding.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qsim_oracle.h"

static orc_cplx *field(size_t n, unsigned seed)
{
    orc_cplx *a = (orc_cplx *)malloc((n ? n : 1) * sizeof(orc_cplx));
    unsigned s = seed * 2654435761u + 1u;
    for (size_t i = 0; i < n; ++i) {
        s = s * 1664525u + 1013904223u;
        a[i].re = (double)(s >> 8) / (double)(1u << 24) - 0.5;
        s = s * 1664525u + 1013904223u;
        a[i].im = (double)(s >> 8) / (double)(1u << 24) - 0.5;
    }
    return a;
}

static int calls = 0;
#define RUN(expr) do { (void)(expr); ++calls; } while (0)

static void ffts(void)
{
    static const size_t n1[] = {1, 2, 3, 5, 7, 16, 30, 617, 1234};
    for (size_t k = 0; k < sizeof n1 / sizeof n1[0]; ++k) {
        orc_cplx *a = field(n1[k], 1), *b = field(n1[k], 2);
        for (int inv = 0; inv < 2; ++inv) {
            RUN(orc_fft(1, &n1[k], inv, a, b, 1));
            RUN(orc_fft(1, &n1[k], inv, a, a, 2));   /* in place */
        }
        if (n1[k] <= 30) RUN(orc_dft_bruteforce(n1[k], 0, a, b));
        free(a);
        free(b);
    }
    const size_t d2[2] = {12, 23}, d3[3] = {12, 23, 34};
    orc_cplx *a = field(12 * 23 * 34, 3), *b = field(12 * 23 * 34, 4);
    RUN(orc_fft(2, d2, 0, a, b, 2));
    RUN(orc_fft(2, d2, 1, b, b, 1));
    RUN(orc_fft(3, d3, 0, a, b, 2));
    RUN(orc_fft(3, d3, 1, b, a, 1));
    RUN(orc_norm2(a, 0));
    RUN(orc_norm2(a, 12 * 23 * 34));
    free(a);
    free(b);
}

static void splits(void)
{
    const size_t n = 64;
    orc_cplx *V = field(n, 5), *psi = field(n, 6);
    for (int method = 2; method <= 4; method += 2)
        for (int boundary = 0; boundary < 2; ++boundary)
            for (int cv = 0; cv < 2; ++cv)
                RUN(orc_split1d_run(n, -10, 10, 0.05, cv ? -0.01 : 0.0, 1, 1, method, boundary,
                                    V, cv, psi, 3, 2, cv, 1 + cv));
    RUN(orc_norm2_range_1d(psi, n, -10, 10, -20, 20));   /* range beyond the grid */
    RUN(orc_norm2_range_1d(psi, n, -10, 10, 5, -5));     /* empty range */
    RUN(orc_norm2_range_1d(psi, n, -10, 10, 0, 10));
    {
        double vr[n], out[4];
        for (size_t i = 0; i < n; ++i) vr[i] = V[i].re;
        for (int boundary = 0; boundary < 2; ++boundary)
            RUN(orc_observables(1, n, 1, -10, 10, 0, 1, 1.3, 0.7, boundary, psi, vr, out));
    }
    free(V);
    free(psi);
    const size_t nx = 16, ny = 32;
    V = field(nx * ny, 7);
    psi = field(nx * ny, 8);
    for (int method = 2; method <= 4; method += 2)
        for (int cv = 0; cv < 2; ++cv)
            RUN(orc_split2d_run(nx, ny, -8, 8, -4, 4, 0.05, 0.0, 1, 1, method, V, cv, psi, 2, 3,
                                cv, 2));
    RUN(orc_split2d_time(nx, ny, -8, 8, -4, 4, 0.05, 1, 1, V, 1, psi, 2, 2, 1));
    {
        double vr[nx * ny], out[4];
        for (size_t i = 0; i < nx * ny; ++i) vr[i] = V[i].re;
        RUN(orc_observables(2, nx, ny, -8, 8, -4, 4, 1.3, 0.7, 0, psi, vr, out));
    }
    free(V);
    free(psi);
}

static void born(void)
{
    const double dir2[2] = {1, 0}, dir3[3] = {0, 0, 1};
    {
        const size_t n[2] = {32, 24};
        const double lo[2] = {-80, -80}, hi[2] = {80, 80};
        const size_t N = n[0] * n[1];
        orc_cplx *V = field(N, 9), *dp = field(N, 10), *ve = field(N, 11), *p0 = field(N, 12);
        for (int pc = 0; pc < 2; ++pc)
            for (int pre = 0; pre < 4; ++pre) {
                double eps = 0.0, dn = 0.0;
                memset(dp, 0, N * sizeof(orc_cplx));
                RUN(orc_born_run(2, n, lo, hi, 0.5, dir2, 1, 1, pc, pre, 0.5, 2, V, dp,
                                 pre ? ve : NULL, &eps, &dn, 2, 1 + pc));
            }
        RUN(orc_plane_wave(2, n, lo, hi, 0.5, dir2, 1, 1, p0));
        RUN(orc_xsection2d(n, lo, hi, 0.5, 1, 1, p0, dp, V, 1, 0));
        RUN(orc_total_xsection2d(n, lo, hi, 0.5, 1, 1, p0, dp, V, 16));
        free(V);
        free(dp);
        free(ve);
        free(p0);
    }
    {
        const size_t n[3] = {16, 12, 10};
        const double lo[3] = {-8, -8, -8}, hi[3] = {8, 8, 8};
        const size_t N = n[0] * n[1] * n[2];
        orc_cplx *V = field(N, 13), *dp = field(N, 14), *ve = field(N, 15), *p0 = field(N, 16);
        orc_cplx amp;
        for (int pc = 0; pc < 2; ++pc)
            for (int pre = 0; pre < 4; ++pre) {
                double eps = 0.0, dn = 0.0;
                memset(dp, 0, N * sizeof(orc_cplx));
                RUN(orc_born_run(3, n, lo, hi, 2.0, dir3, 1, 1, pc, pre, 0.5, 2, V, dp,
                                 pre ? ve : NULL, &eps, &dn, 2, 1 + pc));
            }
        RUN(orc_plane_wave(3, n, lo, hi, 2.0, dir3, 1, 1, p0));
        RUN(orc_xsection3d(n, lo, hi, 2.0, 1, 1, p0, dp, V, 0, 0, 1, &amp));
        RUN(orc_xsection3d(n, lo, hi, 2.0, 1, 1, p0, dp, V, 0.6, 0, 0.8, NULL));
        RUN(orc_total_xsection3d(n, lo, hi, 2.0, 1, 1, p0, dp, V, 6, 5));
        RUN(orc_born_time(3, n, lo, hi, 2.0, dir3, 1, 1, 1, 1, 0.5, 2, V, 2, 1));
        RUN(orc_born_time(3, n, lo, hi, 2.0, dir3, 1, 1, 0, 0, 1.0, 2, V, 1, 2));
        free(V);
        free(dp);
        free(ve);
        free(p0);
    }
}

int main(void)
{
    ffts();
    splits();
    born();
    printf("oracle sanitizer driver: %d calls returned\n", calls);
    return 0;
}
