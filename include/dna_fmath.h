/*
 * dna_fmath.h — deterministic elementary functions for the Geant4-DNA chain.
 *
 * The transport path computes in FP32 and the IRT path in FP64.  Every
 * transcendental on those paths (log, exp, sin/cos, pow, sinh/cosh) is evaluated by
 * the routines below, built only from IEEE-754 operations that are correctly
 * rounded on both the host CPU and gfx950: +, -, *, /, sqrt, fma, rint, floor and
 * integer bit manipulation.  Code that includes this header must be compiled with
 * floating-point contraction disabled (-ffp-contract=off) so that no mul+add is
 * fused behind the source's back.  Under those two conditions a CPU build and a
 * gfx950 build of the same expression give bit-identical results, which is what
 * makes per-track bit-exact comparison of the HIP kernels with the CPU restatement
 * possible.
 *
 * Methods (all standard, written here from the mathematics):
 *   log:  x = 2^e * m, m in [sqrt(2)/2, sqrt(2)); log(m) = f - (f^2/2 - s(f^2/2 + R(s^2)))
 *         with f = m-1, s = f/(2+f) and R the atanh series 2/3 z + 2/5 z^2 + ...
 *         (the reduction used by the classic fdlibm/Cephes family).
 *   exp:  Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor polynomial for e^r,
 *         exact scaling by 2^k.
 *   sin/cos: three-part Cody-Waite reduction by pi/2, Taylor polynomials on |r|<=pi/4.
 * Accuracy (checked against libm in double/long double by the tests): FP32 log/exp/sin/cos
 * within 2 ulp over the argument ranges the path uses, FP64 within 2 ulp.
 *
 * The header is plain C99 and C++; under hipcc every function is __host__ __device__.
 */
#ifndef DNA_FMATH_H
#define DNA_FMATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#define DNA_FN __host__ __device__ static inline
#else
#define DNA_FN static inline
#endif

/* ---------------------------------------------------------------- bit casts */
DNA_FN uint32_t dna_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
DNA_FN float dna_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
DNA_FN uint64_t dna_d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
DNA_FN double dna_u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

/* ================================================================== FP32 */

/* natural log, x > 0 finite; x below FLT_MIN is treated through the same reduction
 * (callers clamp their arguments to >= 1e-30, so subnormals never reach it). */
DNA_FN float dna_logf(float x)
{
    const float ln2_hi = 6.93145751953125e-01f;   /* 0x3f317200, 15 significant bits */
    const float ln2_lo = 1.428606765330187e-06f;   /* ln2 - ln2_hi, rounded          */
    uint32_t ix = dna_f2u(x);
    int e = (int)(ix >> 23) - 127;
    float m = dna_u2f((ix & 0x007fffffu) | 0x3f800000u);   /* [1,2) */
    if (m > 1.41421353816986083984f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;                       /* exact */
    float s = f / (2.0f + f);
    float z = s * s;
    float R = fmaf(z, 0.2222222238779068f, 0.2857142984867096f);   /* 2/9, 2/7 */
    R = fmaf(z, R, 0.4000000059604645f);                             /* 2/5 */
    R = fmaf(z, R, 0.6666666865348816f);                             /* 2/3 */
    R = z * R;
    float hfsq = 0.5f * f * f;
    float ef = (float)e;
    /* log(m) = f - (hfsq - s*(hfsq+R)) */
    float corr = fmaf(s, hfsq + R, ef * ln2_lo);
    return fmaf(ef, ln2_hi, f - (hfsq - corr));
}

/* 2^k as a float, for -149 <= k <= 127 (two-step for the subnormal range). */
DNA_FN float dna_scalbnf(float y, int k)
{
    if (k > 127) { y = y * 1.7014118346046923e38f; k -= 127; }   /* 2^127 */
    if (k < -126) { y = y * 1.1754943508222875e-38f; k += 126;    /* 2^-126 */
                    if (k < -126) { y = y * 1.1754943508222875e-38f; k += 126; } }
    return y * dna_u2f((uint32_t)(k + 127) << 23);
}

DNA_FN float dna_expf(float x)
{
    if (x > 88.72283935546875f) return HUGE_VALF;
    if (x < -103.97208404541015625f) return 0.0f;
    const float ln2_hi = 6.93145751953125e-01f;
    const float ln2_lo = 1.428606765330187e-06f;
    float kf = rintf(x * 1.44269502162933349609f);
    float r = fmaf(kf, -ln2_hi, x);
    r = fmaf(kf, -ln2_lo, r);                          /* |r| <= 0.3466 */
    float p = fmaf(r, 1.98412701e-04f, 1.38888892e-03f);   /* 1/5040, 1/720 */
    p = fmaf(r, p, 8.33333377e-03f);                        /* 1/120 */
    p = fmaf(r, p, 4.16666679e-02f);                        /* 1/24  */
    p = fmaf(r, p, 1.66666672e-01f);                        /* 1/6   */
    p = fmaf(r, p, 0.5f);
    float y = 1.0f + fmaf(r * r, p, r);
    return dna_scalbnf(y, (int)kf);
}

/* sin and cos together; |x| <= 1e5 (the path uses [0, 2pi] and helix phases < 6e3). */
DNA_FN void dna_sincosf(float x, float *sn, float *cs)
{
    const float p1 = 1.5707963705062866211f;       /* float(pi/2)           */
    const float p2 = -4.3711388286737928865e-08f;  /* pi/2 - p1, rounded    */
    const float p3 = -1.7151245100058819e-15f;     /* remainder, rounded    */
    float kf = rintf(x * 0.63661974668502807617f);
    float r = fmaf(kf, -p1, x);
    r = fmaf(kf, -p2, r);
    r = fmaf(kf, -p3, r);
    float z = r * r;
    /* sin r = r + r^3 (-1/6 + z (1/120 + z (-1/5040 + z/362880))) */
    float ps = fmaf(z, 2.75573188e-06f, -1.98412701e-04f);
    ps = fmaf(z, ps, 8.33333377e-03f);
    ps = fmaf(z, ps, -1.66666672e-01f);
    float s = fmaf(r * z, ps, r);
    /* cos r = 1 - z/2 + z^2 (1/24 + z (-1/720 + z (1/40320 - z/3628800))) */
    float pc = fmaf(z, -2.75573188e-07f, 2.48015876e-05f);
    pc = fmaf(z, pc, -1.38888892e-03f);
    pc = fmaf(z, pc, 4.16666679e-02f);
    float c = fmaf(z * z, pc, 1.0f - 0.5f * z);
    int q = ((int)kf) & 3;
    float so, co;
    if (q == 0) { so = s; co = c; }
    else if (q == 1) { so = c; co = -s; }
    else if (q == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *sn = so; *cs = co;
}

DNA_FN float dna_cosf(float x) { float s, c; dna_sincosf(x, &s, &c); return c; }
DNA_FN float dna_sinf(float x) { float s, c; dna_sincosf(x, &s, &c); return s; }

/* x^y for x > 0 as exp(y log x) (used only for the W^1.3 track-length fit). */
DNA_FN float dna_powf(float x, float y) { return dna_expf(y * dna_logf(x)); }

/* ================================================================== FP64 */

DNA_FN double dna_scalbn(double y, int k)
{
    if (k > 1023) { y = y * 8.98846567431158e307; k -= 1023; }        /* 2^1023 */
    if (k < -1022) { y = y * 2.2250738585072014e-308; k += 1022;       /* 2^-1022 */
                     if (k < -1022) { y = y * 2.2250738585072014e-308; k += 1022; } }
    return y * dna_u2d((uint64_t)(k + 1023) << 52);
}

DNA_FN double dna_exp(double x)
{
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double ln2_hi = 6.93147180369123816490e-01;   /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10;   /* ln2 - ln2_hi       */
    double kd = rint(x * 1.44269504088896338700e+00);
    double r = fma(kd, -ln2_hi, x);
    r = fma(kd, -ln2_lo, r);                            /* |r| <= 0.3466 */
    /* Taylor to r^13: coefficients 1/n! for n = 13 .. 2 */
    double p = fma(r, 1.6059043836821613e-10, 2.08767569878680989792e-09);
    p = fma(r, p, 2.50521083854417187751e-08);
    p = fma(r, p, 2.75573192239858906526e-07);
    p = fma(r, p, 2.75573192239858906526e-06);
    p = fma(r, p, 2.48015873015873015873e-05);
    p = fma(r, p, 1.98412698412698412698e-04);
    p = fma(r, p, 1.38888888888888888889e-03);
    p = fma(r, p, 8.33333333333333333333e-03);
    p = fma(r, p, 4.16666666666666666667e-02);
    p = fma(r, p, 1.66666666666666666667e-01);
    p = fma(r, p, 0.5);
    double y = 1.0 + fma(r * r, p, r);
    return dna_scalbn(y, (int)kd);
}

/* Series coefficients 2/(2k+1), k = 11..1, of dna_log.  In device code they are read from
   constant memory with scalar loads at the point of use: as literals, the compiler hoists
   them out of loops into vector registers and, under register pressure, spills them to
   scratch memory, so every call inside a loop reloaded them from memory.  Same values on
   both sides, so the results do not change. */
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ double dna_log_series[11] = {
#else
static const double dna_log_series[11] = {
#endif
    0.08695652173913043478, 0.09523809523809523810, 0.10526315789473684211, 0.11764705882352941176,
    0.13333333333333333333, 0.15384615384615384615, 0.18181818181818181818, 0.22222222222222222222,
    0.28571428571428571429, 0.40000000000000000000, 0.66666666666666666667};

DNA_FN double dna_log(double x)
{
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    uint64_t ix = dna_d2u(x);
    int e = (int)(ix >> 52) - 1023;
    double m = dna_u2d((ix & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (e == -1023) {                     /* subnormal input: renormalise */
        double xs = x * 18014398509481984.0;   /* 2^54 */
        ix = dna_d2u(xs);
        e = (int)(ix >> 52) - 1023 - 54;
        m = dna_u2d((ix & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    }
    if (m > 1.41421356237309514547) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    /* R = sum_{k=1..11} 2/(2k+1) z^k */
    const double *c = dna_log_series;
    double R = fma(z, c[0], c[1]);           /* 2/23, 2/21 */
    R = fma(z, R, c[2]);                     /* 2/19 */
    R = fma(z, R, c[3]);                     /* 2/17 */
    R = fma(z, R, c[4]);                     /* 2/15 */
    R = fma(z, R, c[5]);                     /* 2/13 */
    R = fma(z, R, c[6]);                     /* 2/11 */
    R = fma(z, R, c[7]);                     /* 2/9  */
    R = fma(z, R, c[8]);                     /* 2/7  */
    R = fma(z, R, c[9]);                     /* 2/5  */
    R = fma(z, R, c[10]);                    /* 2/3  */
    R = z * R;
    double hfsq = 0.5 * f * f;
    double ed = (double)e;
    double corr = fma(s, hfsq + R, ed * ln2_lo);
    return fma(ed, ln2_hi, f - (hfsq - corr));
}

DNA_FN void dna_sincos(double x, double *sn, double *cs)
{
    const double p1 = 1.57079632673412561417e+00;   /* 0x3ff921fb54400000 */
    const double p2 = 6.077100506506192e-11;        /* pi/2 - p1, rounded  */
    const double p3 = 3.5215598651832e-27;          /* remainder, rounded  */
    double kd = rint(x * 6.36619772367581382433e-01);
    double r = fma(kd, -p1, x);
    r = fma(kd, -p2, r);
    r = fma(kd, -p3, r);
    double z = r * r;
    /* sin: r + r z (-1/3! + z(1/5! + ... + z^7/17!)) */
    double ps = fma(z, 2.81145725434552076319e-15, -7.64716373181981647590e-13);
    ps = fma(z, ps, 1.60590438368216145994e-10);
    ps = fma(z, ps, -2.50521083854417187751e-08);
    ps = fma(z, ps, 2.75573192239858906526e-06);
    ps = fma(z, ps, -1.98412698412698412698e-04);
    ps = fma(z, ps, 8.33333333333333333333e-03);
    ps = fma(z, ps, -1.66666666666666666667e-01);
    double s = fma(r * z, ps, r);
    /* cos: 1 - z/2 + z^2 (1/4! - z/6! + ... + z^6/16!) */
    double pc = fma(z, 4.77947733238738529744e-14, -1.14707455977297247139e-11);
    pc = fma(z, pc, 2.08767569878680989792e-09);
    pc = fma(z, pc, -2.75573192239858906526e-07);
    pc = fma(z, pc, 2.48015873015873015873e-05);
    pc = fma(z, pc, -1.38888888888888888889e-03);
    pc = fma(z, pc, 4.16666666666666666667e-02);
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    double c = w + (((1.0 - w) - hz) + z * z * pc);
    int q = ((int)kd) & 3;
    double so, co;
    if (q == 0) { so = s; co = c; }
    else if (q == 1) { so = c; co = -s; }
    else if (q == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *sn = so; *cs = co;
}

DNA_FN double dna_cos(double x) { double s, c; dna_sincos(x, &s, &c); return c; }

/* sinh and cosh of the same argument (series below 0.5 to avoid cancellation). */
DNA_FN void dna_sinhcosh(double x, double *sh, double *ch)
{
    double ax = fabs(x);
    double ep = dna_exp(ax);
    double em = 1.0 / ep;
    double c = 0.5 * (ep + em);
    double s;
    if (ax < 0.5) {
        double z = x * x;
        double p = fma(z, 7.64716373181981647590e-13, 1.60590438368216145994e-10);  /* 1/15!, 1/13! */
        p = fma(z, p, 2.50521083854417187751e-08);   /* 1/11! */
        p = fma(z, p, 2.75573192239858906526e-06);   /* 1/9!  */
        p = fma(z, p, 1.98412698412698412698e-04);   /* 1/7!  */
        p = fma(z, p, 8.33333333333333333333e-03);   /* 1/5!  */
        p = fma(z, p, 1.66666666666666666667e-01);   /* 1/3!  */
        s = fma(x * z, p, x);
    } else {
        s = 0.5 * (ep - em);
        if (x < 0) s = -s;
    }
    *sh = s; *ch = c;
}

#endif /* DNA_FMATH_H */
