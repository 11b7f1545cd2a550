/*
 * dna_irt_math.h — FP64 special functions and the counter-based RNG of the IRT path.
 *
 * Special functions (the methods named by public/irt-worker.js):
 *   dna_erfcinv  Abramowitz & Stegun 26.2.23 rational approximation of the normal
 *                quantile, rescaled to erfc^-1 (irt-worker.js:31-39); |error| <= 4.5e-4
 *                in the quantile, as published.
 *   dna_erfcx    exp(x^2) erfc(x): Chebyshev expansion in t = 2/(2+x) with the
 *                erfc coefficients of Numerical Recipes 3rd ed. (section 6.2.2) for
 *                0 <= x < 6, the Laplace continued fraction (30 terms) for x >= 6, and
 *                2 exp(x^2) - erfcx(-x) for x < 0 (irt-worker.js:53-88).
 *
 * RNG.  The reference draws Math.random() (unseeded).  Here every random number is a
 * pure function of (seed, instance, primary, particle pair, draw index): a SplitMix64
 * finaliser applied to a per-pair state plus a Weyl increment.  Results therefore do
 * not depend on thread scheduling or evaluation order, which is what lets the GPU
 * kernel and the CPU restatement draw identical numbers.
 *   track key  tk = mix(mix(seed + G*(instance+1)) + G*(pid+1))
 *   pair state ps = tk ^ (a << 32 | b)          (a < b particle indices, or a = b = k
 *                                                for the per-particle displacement stream)
 *   draw n     u  = (mix(ps + G*(n+1)) >> 11) * 2^-53, in [0, 1)
 * Draws 0..5 of a pair stream are the Brownian synchronisation normals, draws 6.. the
 * reaction-time sampling.
 */
#ifndef DNA_IRT_MATH_H
#define DNA_IRT_MATH_H

#include "dna_fmath.h"

#define DNA_GOLDEN 0x9E3779B97F4A7C15ull

DNA_FN uint64_t dna_mix64(uint64_t z)
{
    z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27; z *= 0x94d049bb133111ebull;
    z ^= z >> 31;
    return z;
}

DNA_FN uint64_t dna_track_key(uint64_t seed, uint64_t instance, uint64_t pid)
{
    return dna_mix64(dna_mix64(seed + DNA_GOLDEN * (instance + 1ull)) + DNA_GOLDEN * (pid + 1ull));
}

DNA_FN uint64_t dna_pair_state(uint64_t tk, uint32_t a, uint32_t b)
{
    return tk ^ (((uint64_t)a << 32) | (uint64_t)b);
}

DNA_FN double dna_u01(uint64_t ps, uint32_t n)
{
    uint64_t x = dna_mix64(ps + DNA_GOLDEN * (uint64_t)(n + 1u));
    return (double)(x >> 11) * 1.1102230246251565404e-16;   /* 2^-53 */
}

/* erfc^-1 by A&S 26.2.23 (irt-worker.js:31-39) */
DNA_FN double dna_erfcinv(double x)
{
    if (x <= 0.0 || x >= 2.0) return 0.0;
    double p = x > 1.0 ? 2.0 - x : x;
    double t = sqrt(-2.0 * dna_log(p / 2.0));
    double num = 2.515517 + 0.802853 * t + 0.010328 * t * t;
    double den = 1.0 + 1.432788 * t + 0.189269 * t * t + 0.001308 * t * t * t;
    double y = t - num / den;
    y *= 0.7071067811865475;
    return x > 1.0 ? -y : y;
}

/* exp(x^2) erfc(x) for x >= 0 */
DNA_FN double dna_erfcx_pos(double x)
{
    if (x >= 6.0) {
        double f = 0.0;
        for (int n = 30; n >= 1; n--) f = (0.5 * n) / (x + f);
        return 0.5641895835477563 / (x + f);          /* 1/sqrt(pi) */
    }
    /* Numerical Recipes erfc Chebyshev coefficients, c[0..27] */
    const double c[28] = {
        -1.3026537197817094, 6.4196979235649026e-1, 1.9476473204185836e-2,
        -9.561514786808631e-3, -9.46595344482036e-4, 3.66839497852761e-4,
        4.2523324806907e-5, -2.0278578112534e-5, -1.624290004647e-6,
        1.303655835580e-6, 1.5626441722e-8, -8.5238095915e-8, 6.529054439e-9,
        5.059343495e-9, -9.91364156e-10, -2.27365122e-10, 9.6467911e-11,
        2.394038e-12, -6.886027e-12, 8.94487e-13, 3.13092e-13, -1.12708e-13,
        3.81e-16, 7.106e-15, -1.523e-15, -9.4e-17, 1.21e-16, -2.8e-17 };
    double t = 2.0 / (2.0 + x);
    double ty = 4.0 * t - 2.0;
    double d = 0.0, dd = 0.0;
    for (int j = 27; j > 0; j--) {
        double tmp = d;
        d = ty * d - dd + c[j];
        dd = tmp;
    }
    return t * dna_exp(0.5 * (c[0] + ty * d) - dd);
}

DNA_FN double dna_erfcx(double x)
{
    if (x < 0.0) return 2.0 * dna_exp(x * x) - dna_erfcx_pos(-x);
    return dna_erfcx_pos(x);
}

/* ---- Karamitros 2011 reaction scheme (irt-worker.js:191-286) ---- */
/* species: 0 OH, 1 e-aq, 2 H, 3 H3O+, 4 H2O2, 5 OH- */
#define DNA_N_SPECIES 6
#define DNA_T_END_NS 1000.0
#define DNA_PI 3.141592653589793

#endif
