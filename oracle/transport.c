/*
 * transport.c — CPU restatement of Phase A (src/shaders/primary.wgsl) and Phase B
 * (src/shaders/secondary.wgsl) with the lookups of src/shaders/helpers.wgsl and the
 * host seeding of src/gpu/buffers.ts:91-104.  TEST INFRASTRUCTURE (see oracle.h).
 *
 * FP32 throughout; transcendentals from include/dna_fmath.h; compiled with
 * -ffp-contract=off so every expression rounds exactly as written.
 *
 * Output order (canonical): for primary 0..n-1 in turn, the primary's own radical
 * entries in emission order, then those of each of its secondaries in the order the
 * primary emitted them.  The reference appends with atomics, so its order is
 * nondeterministic; this order is one of its possible orders up to interleaving.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"
#include "../include/dna_fmath.h"

static dna_tables T;
static int T_loaded = 0;

int oracle_load_tables(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    size_t got = fread(&T, 1, sizeof(T), f);
    int extra = fgetc(f);
    fclose(f);
    if (got != sizeof(T) || extra != EOF || sizeof(T) != DNA_TABLES_BYTES) return 2;
    T_loaded = 1;
    return 0;
}
const dna_tables *oracle_tables(void) { return T_loaded ? &T : NULL; }
void oracle_free(void *p) { free(p); }

/* WGSL abstract-float constants as they land in f32 (helpers.wgsl:1-2, primary.wgsl) */
#define PI_F     3.14159265f
#define TWO_PI_F 6.2831853f
#define NW_F     33.4f
static const float SB[5] = {10.79f, 13.39f, 16.05f, 32.30f, 539.0f};   /* helpers.wgsl:11 */
static const float EXL[5] = {8.22f, 10.00f, 11.24f, 12.61f, 13.77f};    /* helpers.wgsl:13 */
static const float DEA_E[14] = {0.0f, 4.2f, 4.3f, 5.8f, 6.0f, 6.2f, 6.4f, 7.0f, 7.5f, 7.7f,
                                8.4f, 10.2f, 11.2f, 12.3f};             /* helpers.wgsl:175 */
static const float DEA_S[14] = {0.0f, 0.0f, 0.0f, 2.5e-5f, 2.8e-4f, 6.0e-4f, 6.7e-4f, 3.7e-4f,
                                1.09e-4f, 7.0e-5f, 1.4e-4f, 2.0e-5f, 6.0e-5f, 0.0f};

static inline float fmaxf_(float a, float b) { return a > b ? a : b; }   /* WGSL max */
static inline float fminf_(float a, float b) { return a < b ? a : b; }
static inline float clampf_(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }

/* ------------------------------------------------------------ helpers.wgsl */

/* xs_all, helpers.wgsl:15-27 */
static void xs_all(float E, float xs[3])
{
    if (E <= T.XE[0]) { xs[0] = 0.0f; xs[1] = 0.0f; xs[2] = T.XL[0]; return; }
    if (E >= T.XE[DNA_XN - 1]) { xs[0] = T.XI[DNA_XN-1]; xs[1] = T.XC[DNA_XN-1]; xs[2] = T.XL[DNA_XN-1]; return; }
    float t = (dna_logf(E) - T.LOG_XE0) * T.INV_LOG_XE_STEP;
    uint32_t i = (uint32_t)clampf_(floorf(t), 0.0f, (float)(DNA_XN - 2));
    float f = t - (float)i;
    uint32_t j = i + 1;
    xs[0] = fmaxf_(0.0f, T.XI[i] + (T.XI[j] - T.XI[i]) * f);
    xs[1] = fmaxf_(0.0f, T.XC[i] + (T.XC[j] - T.XC[i]) * f);
    xs[2] = fmaxf_(0.0f, T.XL[i] + (T.XL[j] - T.XL[i]) * f);
}

/* xs_shell_fracs, helpers.wgsl:29-43 */
static void xs_shell_fracs(float E, float sf[5])
{
    int k;
    if (E <= T.XE[0]) { for (k = 0; k < 5; k++) sf[k] = T.XSF[k][0]; return; }
    if (E >= T.XE[DNA_XN - 1]) { for (k = 0; k < 5; k++) sf[k] = T.XSF[k][DNA_XN-1]; return; }
    float t = (dna_logf(E) - T.LOG_XE0) * T.INV_LOG_XE_STEP;
    uint32_t i = (uint32_t)clampf_(floorf(t), 0.0f, (float)(DNA_XN - 2));
    float f = t - (float)i;
    for (k = 0; k < 5; k++) sf[k] = fmaxf_(0.0f, T.XSF[k][i] + (T.XSF[k][i+1] - T.XSF[k][i]) * f);
}

/* xs_exc_fracs, helpers.wgsl:45-64 (index clamped at 85, f = 0 there) */
static void xs_exc_fracs(float E, float ef[5])
{
    int k;
    if (E <= T.XE[0]) { for (k = 0; k < 5; k++) ef[k] = T.XEF[k][0]; return; }
    if (E >= T.XE[DNA_XN - 1]) { for (k = 0; k < 5; k++) ef[k] = T.XEF[k][85]; return; }
    float t = (dna_logf(E) - T.LOG_XE0) * T.INV_LOG_XE_STEP;
    uint32_t i = (uint32_t)clampf_(floorf(t), 0.0f, (float)(DNA_XN - 2));
    if (i > 85u) i = 85u;
    float f = (i == 85u) ? 0.0f : t - (float)i;
    uint32_t j = (i + 1u < 85u) ? i + 1u : 85u;
    for (k = 0; k < 5; k++) ef[k] = fmaxf_(0.0f, T.XEF[k][i] + (T.XEF[k][j] - T.XEF[k][i]) * f);
}

/* xs_el_cos, helpers.wgsl:66-82 */
static float xs_el_cos(float E, float r)
{
    float rr = clampf_(r, 0.0f, 1.0f);
    float jf = rr * (float)(DNA_N_AC - 1);
    uint32_t jlo = (uint32_t)floorf(jf);
    uint32_t jhi = (jlo + 1u < DNA_N_AC - 1u) ? jlo + 1u : DNA_N_AC - 1u;
    float fj = jf - (float)jlo;
    float tE = (dna_logf(fmaxf_(E, T.XAE[0])) - T.LOG_XAE0) * T.INV_LOG_XAE_STEP;
    uint32_t iE = (uint32_t)clampf_(floorf(tE), 0.0f, (float)(DNA_N_AE - 2));
    float fE = clampf_(tE - (float)iE, 0.0f, 1.0f);
    uint32_t b0 = iE * DNA_N_AC, b1 = (iE + 1u) * DNA_N_AC;
    float c00 = T.XAC[b0 + jlo], c01 = T.XAC[b0 + jhi];
    float c10 = T.XAC[b1 + jlo], c11 = T.XAC[b1 + jhi];
    float cE0 = c00 + (c01 - c00) * fj;
    float cE1 = c10 + (c11 - c10) * fj;
    return clampf_(cE0 + (cE1 - cE0) * fE, -1.0f, 1.0f);
}

/* sample_W_sec + bsearch_cdf, helpers.wgsl:88-117: nearest XWE row, then the linear
 * scan for the last c <= r that stops at the first c > r, then linear interpolation. */
static float sample_W_sec(float E, uint32_t shell, float r)
{
    uint32_t lo = 0, hi = DNA_N_WE - 1;
    while (hi - lo > 1u) { uint32_t mid = (lo + hi) / 2u; if (T.XWE[mid] <= E) lo = mid; else hi = mid; }
    uint32_t iE = (fabsf(T.XWE[hi] - E) < fabsf(T.XWE[lo] - E)) ? hi : lo;
    uint32_t base = iE * DNA_N_WR;
    const float *C = T.XWC[shell], *W = T.XWT[shell];
    uint32_t jlo = 0;
    for (uint32_t j = 0; j < DNA_N_WR - 1u; j++) {
        if (C[base + j] <= r) jlo = j; else break;
    }
    uint32_t jhi = (jlo + 1u < DNA_N_WR - 1u) ? jlo + 1u : DNA_N_WR - 1u;
    float c0 = C[base + jlo], c1 = C[base + jhi];
    float e0 = W[base + jlo], e1 = W[base + jhi];
    if (c1 <= c0) return e0;
    float f = clampf_((r - c0) / (c1 - c0), 0.0f, 1.0f);
    return e0 + (e1 - e0) * f;
}

/* xs_vib_total, helpers.wgsl:135-144 */
static float xs_vib_total(float E)
{
    if (E < 2.0f || E > 100.0f) return 0.0f;
    for (uint32_t i = 0; i < DNA_N_VIB - 1u; i++) {
        if (E >= T.XVE[i] && E < T.XVE[i+1]) {
            float f = (E - T.XVE[i]) / (T.XVE[i+1] - T.XVE[i]);
            return fmaxf_(0.0f, T.XVS[i] + (T.XVS[i+1] - T.XVS[i]) * f);
        }
    }
    return 0.0f;
}

/* sample_vib_mode, helpers.wgsl:148-170 */
static uint32_t sample_vib_mode(float E, float r)
{
    uint32_t bin = 0;
    for (uint32_t i = 0; i < DNA_N_VIB - 1u; i++)
        if (E >= T.XVE[i] && E < T.XVE[i+1]) { bin = i; break; }
    uint32_t base = bin * 9u;
    float cs[9], cum = 0.0f;
    for (uint32_t k = 0; k < 9; k++) {
        float frac = (T.VIB_LEV[k] > E) ? 0.0f : T.XVMF[base + k];
        cum = cum + frac;
        cs[k] = cum;
    }
    if (cum <= 0.0f) return 0;
    float thr = r * cum;
    for (uint32_t k = 0; k < 9; k++) if (thr < cs[k]) return k;
    return 8;
}

/* xs_dea, helpers.wgsl:177-186 */
static float xs_dea(float E)
{
    if (E < 4.0f || E > 13.0f) return 0.0f;
    for (int i = 0; i < 13; i++) {
        if (E >= DEA_E[i] && E < DEA_E[i+1]) {
            float f = (E - DEA_E[i]) / (DEA_E[i+1] - DEA_E[i]);
            return fmaxf_(0.0f, DEA_S[i] + (DEA_S[i+1] - DEA_S[i]) * f);
        }
    }
    return 0.0f;
}

/* xoshiro128++ step and uniform, helpers.wgsl:188-190.  rf = f32(x>>1) / f32(2147483647)
 * and f32(2147483647) == 2^31, so rf = round_to_f32(x>>1) * 2^-31 (can be exactly 1). */
static inline uint32_t rl(uint32_t x, uint32_t k) { return (x << k) | (x >> (32u - k)); }
static inline uint32_t rn(uint32_t s[4])
{
    uint32_t r = rl(s[0] + s[3], 7u) + s[0];
    uint32_t t = s[1] << 9u;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rl(s[3], 11u);
    return r;
}
static inline float rf(uint32_t s[4]) { return (float)(rn(s) >> 1u) / 2147483648.0f; }

/* Meesungnoen 2002 thermalisation sigma, primary.wgsl:90-98 (Horner in f32) */
static float meesungnoen_sigma(float k)
{
    if (k <= 0.1f) return 0.0f;
    float r = -4.06217193e-08f;
    r = r*k + 3.06848412e-06f; r = r*k - 9.93217814e-05f; r = r*k + 1.80172797e-03f;
    r = r*k - 2.01135480e-02f; r = r*k + 1.42939448e-01f; r = r*k - 6.48348714e-01f;
    r = r*k + 1.85227848e+00f; r = r*k - 3.36450378e+00f; r = r*k + 4.37785068e+00f;
    r = r*k - 4.20557339e+00f; r = r*k + 3.81679083e+00f; r = r*k - 2.34069784e-01f;
    return fmaxf_(r, 0.0f) * 0.6267f;
}

/* host seeding, buffers.ts:91-104.  The first product is formed in JS doubles and then
 * ToUint32'd; it is reproduced here in double for exactness. */
void oracle_seed_primary(uint32_t np, uint32_t seed, uint32_t *st)
{
    for (uint32_t i = 0; i < np; i++) {
        double v = (double)(i + 1) * 2654435761.0 + (double)seed * 1013904223.0;
        double m = fmod(v, 4294967296.0);
        uint32_t s = (uint32_t)m;
        for (uint32_t j = 0; j < 4; j++) {
            s ^= s >> 16;
            s *= 0x45d9f3bu;
            s ^= s >> 16;
            st[4*i + j] = s + j * 0x9E3779B9u;
        }
    }
}

void oracle_rng_draws(const uint32_t *state4, uint32_t n, uint32_t *raw, float *rfo)
{
    uint32_t s[4] = {state4[0], state4[1], state4[2], state4[3]};
    uint32_t s2[4] = {state4[0], state4[1], state4[2], state4[3]};
    for (uint32_t i = 0; i < n; i++) { if (raw) raw[i] = rn(s); if (rfo) rfo[i] = rf(s2); }
}

/* ------------------------------------------------------------- emission sink */

typedef struct {
    float *rad; uint64_t n, cap;            /* radical entries (x,y,z,w)        */
    uint32_t cnt[8];                        /* counters slots                   */
    uint32_t sst[8];                        /* sec_stats                        */
    uint32_t dbg[8];
    uint32_t *dose;                         /* shared grid (atomic adds)        */
} sink;

static void emit(sink *k, float x, float y, float z, float w)
{
    if (k->n == k->cap) {
        k->cap = k->cap ? 2 * k->cap : 1024;
        k->rad = (float *)realloc(k->rad, k->cap * 16);
    }
    float *e = k->rad + 4 * k->n++;
    e[0] = x; e[1] = y; e[2] = z; e[3] = w;
}

typedef struct { float box; uint32_t vc; } world;

/* deposit, primary.wgsl:102-112 / secondary.wgsl:31-41 (100 units per eV, truncating) */
static void deposit(sink *k, const world *W, float px, float py, float pz, float dep)
{
    if (dep <= 0.0f) return;
    float vs = 2.0f * W->box / (float)W->vc;
    int ix = (int)floorf((px + W->box) / vs);
    int iy = (int)floorf((py + W->box) / vs);
    int iz = (int)floorf((pz + W->box) / vs);
    int n = (int)W->vc;
    if (ix < 0 || ix >= n || iy < 0 || iy >= n || iz < 0 || iz >= n) return;
    uint32_t vi = (uint32_t)((iz * n + iy) * n + ix);
    uint32_t add = (uint32_t)(dep * 100.0f);
#pragma omp atomic
    k->dose[vi] += add;
}

/* dna_near, primary.wgsl:41-86 (WGSL round() is round-half-to-even -> rintf) */
static uint32_t dna_near(const dna_params *p, float px, float py, float pz)
{
    if (p->dna_enable == 0u) return 0u;
    const float r_damage = 0.29f;
    uint32_t N = p->dna_grid_n;
    float spacing = p->dna_spacing;
    float grid_off = -((float)(N - 1u) * spacing) * 0.5f;
    float inv_s = 1.0f / spacing;
    int fi = (int)rintf((py - grid_off) * inv_s);
    int fj = (int)rintf((pz - grid_off) * inv_s);
    if (fi < 0 || fi >= (int)N || fj < 0 || fj >= (int)N) return 0u;
    float fy = grid_off + (float)fi * spacing;
    float fz = grid_off + (float)fj * spacing;
    float y_rel = py - fy, z_rel = pz - fz;
    float r2 = y_rel * y_rel + z_rel * z_rel;
    float R_reach = p->dna_r_bb + r_damage;
    if (r2 > R_reach * R_reach) return 0u;
    int bp_est = (int)rintf((px - p->dna_x0) / p->dna_rise);
    int b0 = bp_est - 1 > 0 ? bp_est - 1 : 0;
    int b1 = bp_est + 1;
    const float d_phase = 0.598398600f;           /* f32(2*PI/10.5) */
    float r_bb = p->dna_r_bb;
    for (int b = b0; b <= b1; b++) {
        float bx = p->dna_x0 + (float)b * p->dna_rise;
        float phi = (float)b * d_phase;
        float s0y = r_bb * dna_cosf(phi), s0z = r_bb * dna_sinf(phi);
        float dx = px - bx, dy0 = y_rel - s0y, dz0 = z_rel - s0z;
        if (dx*dx + dy0*dy0 + dz0*dz0 < r_damage * r_damage) return 1u;
        float s1y = r_bb * dna_cosf(phi + PI_F), s1z = r_bb * dna_sinf(phi + PI_F);
        float dy1 = y_rel - s1y, dz1 = z_rel - s1z;
        if (dx*dx + dy1*dy1 + dz1*dz1 < r_damage * r_damage) return 1u;
    }
    return 0u;
}

/* Gaussian 3-vector from 4 draws: (r cos, r sin, r' cos') with r = sqrt(-2 ln u)*sig.
 * Pattern of primary.wgsl:216-220 (and every displacement that repeats it). */
static void gauss3(uint32_t s[4], float sig, float d[3])
{
    float u1 = fmaxf_(rf(s), 1e-30f); float u2 = rf(s);
    float r = sqrtf(-2.0f * dna_logf(u1)) * sig;
    float sn, cs; dna_sincosf(TWO_PI_F * u2, &sn, &cs);
    d[0] = r * cs; d[1] = r * sn;
    float u3 = fmaxf_(rf(s), 1e-30f); float u4 = rf(s);
    d[2] = sqrtf(-2.0f * dna_logf(u3)) * sig * dna_cosf(TWO_PI_F * u4);
}

/* H2Ovib dissociation channels after electron-hole recombination
 * (primary.wgsl:247-276): 13.65% 2OH+H2, 35.75% OH+H, 15.6% 2H, else relax. */
static void h2ovib(sink *k, uint32_t s[4], float x, float y, float z, float pid)
{
    float r_vd = rf(s);
    if (r_vd < 0.1365f) {
        k->cnt[0] += 2; k->cnt[5] += 1;
        emit(k, x, y, z, 0.0f + pid); emit(k, x, y, z, 0.0f + pid); emit(k, x, y, z, 7.0f + pid);
    } else if (r_vd < 0.494f) {
        k->cnt[0] += 1; k->cnt[2] += 1;
        emit(k, x, y, z, 0.0f + pid); emit(k, x, y, z, 2.0f + pid);
    } else if (r_vd < 0.650f) {
        k->cnt[2] += 2;
        emit(k, x, y, z, 2.0f + pid); emit(k, x, y, z, 2.0f + pid);
    }
}

/* H2O+ / e-aq pair with Onsager recombination (primary.wgsl:229-288, 473-536, 560-619;
 * secondary.wgsl:192-255): mother at m = site + gauss(1.1547), e-aq at e0 + gauss(sig_e);
 * P = 1 - exp(-0.711 / |e - m|).  e0 is the ionisation site (ionisation) or the mother
 * (autoionisation); eaq_code is 5 (pre-thermalised) or 1.  The mother displacement is
 * drawn by the caller (its position in the draw sequence differs between call sites). */
static void hole_pair(sink *k, uint32_t s[4], const float m[3], const float e0[3],
                      float sig_e, float eaq_code, float pid)
{
    float e[3] = {e0[0], e0[1], e0[2]};
    if (sig_e > 0.0f) {
        float d[3]; gauss3(s, sig_e, d);
        e[0] += d[0]; e[1] += d[1]; e[2] += d[2];
    }
    float rx = e[0] - m[0], ry = e[1] - m[1], rz = e[2] - m[2];
    float r_sep = fmaxf_(sqrtf(rx*rx + ry*ry + rz*rz), 1e-6f);
    float p_recomb = 1.0f - dna_expf(-0.711f / r_sep);
    float r_recomb = rf(s);
    if (r_recomb < p_recomb) {
        h2ovib(k, s, m[0], m[1], m[2], pid);
    } else {
        k->cnt[0] += 1; k->cnt[1] += 1; k->cnt[3] += 1;
        emit(k, m[0], m[1], m[2], 0.0f + pid);
        emit(k, e[0], e[1], e[2], eaq_code + pid);
        emit(k, m[0], m[1], m[2], 3.0f + pid);
    }
}

/* mother displacement, RMS 2 nm (primary.wgsl:215-221) */
static void mother(uint32_t s[4], float px, float py, float pz, float m[3])
{
    float d[3]; gauss3(s, 1.1547005f, d);
    m[0] = px + d[0]; m[1] = py + d[1]; m[2] = pz + d[2];
}

/* electronic excitation products, primary.wgsl:431-622 / secondary.wgsl:274-419 */
static void excitation_products(sink *k, uint32_t s[4], uint32_t lvl, float px, float py,
                                float pz, float pid)
{
    float r_ch = rf(s);
    if (lvl == 0u) {
        if (r_ch < 0.65f) {
            k->cnt[0] += 1; k->cnt[2] += 1;
            emit(k, px, py, pz, 0.0f + pid); emit(k, px, py, pz, 2.0f + pid);
        }
    } else if (lvl == 1u) {
        if (r_ch < 0.175f) {
        } else if (r_ch < 0.2075f) {
            k->cnt[0] += 2; k->cnt[5] += 1;
            emit(k, px, py, pz, 0.0f + pid); emit(k, px, py, pz, 0.0f + pid); emit(k, px, py, pz, 7.0f + pid);
        } else if (r_ch < 0.7075f) {
            float m[3]; mother(s, px, py, pz, m);
            hole_pair(k, s, m, m, meesungnoen_sigma(1.7f), 1.0f, pid);
        } else if (r_ch < 0.961f) {
            k->cnt[0] += 1; k->cnt[2] += 1;
            emit(k, px, py, pz, 0.0f + pid); emit(k, px, py, pz, 2.0f + pid);
        } else {
            k->cnt[2] += 2;
            emit(k, px, py, pz, 2.0f + pid); emit(k, px, py, pz, 2.0f + pid);
        }
    } else {
        if (r_ch < 0.50f) {
            float m[3]; mother(s, px, py, pz, m);
            hole_pair(k, s, m, m, meesungnoen_sigma(1.7f), 1.0f, pid);
        }
    }
}

/* elastic deflection of the flight direction, primary.wgsl:164-170 */
static void rotate(float *dx, float *dy, float *dz, float cos_t, float sin_t, float cp, float sp)
{
    if (fabsf(*dz) > 0.99999f) {
        float sw = (*dz > 0.0f) ? 1.0f : -1.0f;
        *dx = sin_t * cp; *dy = sin_t * sp * sw; *dz = cos_t * sw;
    } else {
        float q = sqrtf(1.0f - *dz * *dz);
        float inv = 1.0f / q;
        float nx = *dx * cos_t + sin_t * (*dx * *dz * cp - *dy * sp) * inv;
        float ny = *dy * cos_t + sin_t * (*dy * *dz * cp + *dx * sp) * inv;
        float nz = *dz * cos_t - q * sin_t * cp;
        float len = sqrtf(nx*nx + ny*ny + nz*nz);
        *dx = nx / len; *dy = ny / len; *dz = nz / len;
    }
}

typedef struct { float x, y, z, E, dx, dy, dz; uint32_t s[4]; } secrec;

/* sample the shell (primary.wgsl:184-193) */
static uint32_t pick_shell(uint32_t s[4], float E, float *bind)
{
    float sf[5]; xs_shell_fracs(E, sf);
    float fsum = fmaxf_(sf[0] + sf[1] + sf[2] + sf[3] + sf[4], 1e-12f);
    float r_sh = rf(s) * fsum;
    uint32_t sh;
    if (r_sh < sf[0]) sh = 0;
    else if (r_sh < sf[0] + sf[1]) sh = 1;
    else if (r_sh < sf[0] + sf[1] + sf[2]) sh = 2;
    else if (r_sh < sf[0] + sf[1] + sf[2] + sf[3]) sh = 3;
    else sh = 4;
    *bind = SB[sh];
    return sh;
}

static uint32_t pick_level(uint32_t s[4], float E)
{
    float ef[5]; xs_exc_fracs(E, ef);
    float efsum = fmaxf_(ef[0] + ef[1] + ef[2] + ef[3] + ef[4], 1e-12f);
    float r_ex = rf(s) * efsum;
    if (r_ex < ef[0]) return 0;
    if (r_ex < ef[0] + ef[1]) return 1;
    if (r_ex < ef[0] + ef[1] + ef[2]) return 2;
    if (r_ex < ef[0] + ef[1] + ef[2] + ef[3]) return 3;
    return 4;
}

/* Phase A: one primary, primary.wgsl:114-663.  Secondaries go to *secs in emission order. */
static void run_primary(const dna_params *p, const world *W, uint32_t idx, float E0,
                        uint32_t s[4], sink *k, secrec **secs, uint32_t *nsec, uint32_t *capsec,
                        dna_result *res)
{
    float E = E0;
    float pid = (float)idx * 8.0f;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    float sh = p->start_half;
    if (sh > 0.0f) {
        px = (rf(s) * 2.0f - 1.0f) * sh;
        py = (rf(s) * 2.0f - 1.0f) * sh;
        pz = (rf(s) * 2.0f - 1.0f) * sh;
    }
    float spx = px, spy = py, spz = pz;
    float dx, dy, dz;
    {
        float u1 = fmaxf_(rf(s), 1e-10f);
        float u2 = rf(s);
        float cos_i = 1.0f - 2.0f * u1;
        float sin_i = sqrtf(fmaxf_(0.0f, 1.0f - cos_i * cos_i));
        float sn, cs; dna_sincosf(TWO_PI_F * u2, &sn, &cs);
        dx = sin_i * cs; dy = sin_i * sn; dz = cos_i;
        if (p->dir_mode == 1u) { dx = 0.0f; dy = 0.0f; dz = 1.0f; }
    }
    float total_path = 0.0f, productive_path = 0.0f, max_r = 0.0f;
    uint32_t n_ions = 0, n_exc = 0, escaped = 0;
    float box = p->box, ce = p->ce;

    for (uint32_t step = 0; step < p->ms; step++) {
        if (E < ce) break;
        if (fabsf(px) >= box || fabsf(py) >= box || fabsf(pz) >= box) { escaped = 1; break; }
        int productive = E > 12.0f;
        float xs[3]; xs_all(E, xs);
        float s_vib = xs_vib_total(E), s_dea = xs_dea(E);
        float s_tot = xs[0] + xs[1] + xs[2] + s_vib + s_dea;
        if (s_tot <= 0.0f) break;
        float lambda = 1.0f / (NW_F * s_tot);
        float dist = -dna_logf(fmaxf_(rf(s), 1e-10f)) * lambda;
        float r_el = rf(s);
        float cos_el = xs_el_cos(E, r_el);
        float sin_el = sqrtf(fmaxf_(0.0f, 1.0f - cos_el * cos_el));
        float sp, cp; dna_sincosf(TWO_PI_F * rf(s), &sp, &cp);
        rotate(&dx, &dy, &dz, cos_el, sin_el, cp, sp);
        px += dx * dist; py += dy * dist; pz += dz * dist;
        total_path += dist;
        if (productive) productive_path += dist;
        float rdx = px - spx, rdy = py - spy, rdz = pz - spz;
        float r = sqrtf(rdx*rdx + rdy*rdy + rdz*rdz);
        if (r > max_r) max_r = r;
        if (fabsf(px) >= box || fabsf(py) >= box || fabsf(pz) >= box) { escaped = 1; break; }

        float r_type = rf(s) * s_tot;
        if (r_type < xs[0] && E > SB[0]) {
            float bind; uint32_t shell = pick_shell(s, E, &bind);
            if (E > bind) {
                float r_w = rf(s);
                float Wmax = (E + bind) * 0.5f;
                float Wt = clampf_(sample_W_sec(E, shell, r_w), bind, Wmax);
                float W_sec = fmaxf_(Wt - bind, 0.0f);
                E -= Wt;
                n_ions++;
                k->dbg[4] += (uint32_t)(Wt * 10.0f);
                deposit(k, W, px, py, pz, bind);
                if (W_sec <= ce) deposit(k, W, px, py, pz, W_sec);
                if (dna_near(p, px, py, pz)) k->cnt[4]++;
                float m[3]; mother(s, px, py, pz, m);
                if (W_sec <= ce) {
                    float site[3] = {px, py, pz};
                    hole_pair(k, s, m, site, meesungnoen_sigma(W_sec), 5.0f, pid);
                } else {
                    /* tracked secondary: track-length separation model, primary.wgsl:289-353 */
                    float r_track = 0.04f * dna_powf(W_sec, 1.3f);
                    float r_tail = meesungnoen_sigma(fminf_(W_sec, 7.4f));
                    float r_sep = fmaxf_(r_track + r_tail, 1e-3f);
                    float p_recomb = 1.0f - dna_expf(-0.711f / r_sep);
                    float r_recomb = rf(s);
                    if (r_recomb < p_recomb) {
                        h2ovib(k, s, m[0], m[1], m[2], pid);
                    } else {
                        k->cnt[0] += 1; k->cnt[3] += 1;
                        emit(k, m[0], m[1], m[2], 0.0f + pid);
                        emit(k, m[0], m[1], m[2], 3.0f + pid);
                    }
                }
                if (W_sec > ce) {
                    /* secondary emission with G4DNABornAngle, primary.wgsl:362-413 */
                    k->cnt[6]++;
                    float E_before = E + bind + W_sec;
                    float cos_s;
                    if (W_sec < 50.0f) cos_s = 2.0f * rf(s) - 1.0f;
                    else if (W_sec <= 200.0f) {
                        if (rf(s) <= 0.1f) cos_s = 2.0f * rf(s) - 1.0f;
                        else cos_s = rf(s) * 0.70710678f;
                    } else {
                        float sin2 = clampf_((1.0f - W_sec / E_before) / (1.0f + W_sec / 1022000.0f), 0.0f, 1.0f);
                        cos_s = sqrtf(fmaxf_(0.0f, 1.0f - sin2));
                    }
                    float sin_s = sqrtf(fmaxf_(0.0f, 1.0f - cos_s * cos_s));
                    float sps, cps; dna_sincosf(TWO_PI_F * rf(s), &sps, &cps);
                    float sdx = dx, sdy = dy, sdz = dz;
                    rotate(&sdx, &sdy, &sdz, cos_s, sin_s, cps, sps);
                    float p_inc = sqrtf(E_before), p_sec = sqrtf(W_sec);
                    float pfx = p_inc * dx - p_sec * sdx;
                    float pfy = p_inc * dy - p_sec * sdy;
                    float pfz = p_inc * dz - p_sec * sdz;
                    float pfl = sqrtf(pfx*pfx + pfy*pfy + pfz*pfz);
                    if (pfl > 1e-10f) { dx = pfx / pfl; dy = pfy / pfl; dz = pfz / pfl; }
                    if (*nsec == *capsec) {
                        *capsec = *capsec ? 2 * *capsec : 64;
                        *secs = (secrec *)realloc(*secs, *capsec * sizeof(secrec));
                    }
                    secrec *q = &(*secs)[(*nsec)++];
                    q->x = px; q->y = py; q->z = pz; q->E = W_sec;
                    q->dx = sdx; q->dy = sdy; q->dz = sdz;
                    q->s[0] = rn(s); q->s[1] = rn(s); q->s[2] = rn(s); q->s[3] = rn(s);
                }
            }
        } else if (r_type < xs[0] + xs[1]) {
            uint32_t lvl = pick_level(s, E);
            float exc_E = EXL[lvl];
            float dep = fminf_(exc_E, E);
            E -= dep;
            deposit(k, W, px, py, pz, dep);
            n_exc++;
            excitation_products(k, s, lvl, px, py, pz, pid);
            if (dna_near(p, px, py, pz)) k->cnt[4]++;
        } else if (r_type < xs[0] + xs[1] + s_vib) {
            uint32_t mode = sample_vib_mode(E, rf(s));
            float dep = fminf_(T.VIB_LEV[mode], E);
            E -= dep;
            deposit(k, W, px, py, pz, dep);
        } else if (r_type < xs[0] + xs[1] + s_vib + s_dea) {
            deposit(k, W, px, py, pz, E);
            k->cnt[0] += 1; k->cnt[5] += 1;
            emit(k, px, py, pz, 0.0f + pid); emit(k, px, py, pz, 6.0f + pid); emit(k, px, py, pz, 7.0f + pid);
            E = 0.0f;
        }
    }
    if (escaped == 0u && E > 0.0f) deposit(k, W, px, py, pz, E);
    res->path = total_path; res->prod = productive_path; res->fE = E;
    res->ni = n_ions; res->nx = n_exc; res->esc = escaped; res->mx = max_r; res->_pad = 0;
    k->dbg[0] += n_ions; k->dbg[1] += n_exc; k->dbg[2] += escaped; k->dbg[3] += 1;
}

/* Phase B for one secondary: the per-dispatch body of secondary.wgsl:83-450 applied
 * up to max_sec_steps times (the host issues MAX_SEC_STEPS dispatches, dispatch.ts:137). */
static void run_secondary(const dna_params *p, const world *W, float pid, secrec *q, sink *k)
{
    float px = q->x, py = q->y, pz = q->z, E = q->E, dx = q->dx, dy = q->dy, dz = q->dz;
    uint32_t s[4] = {q->s[0], q->s[1], q->s[2], q->s[3]};
    float box = p->box, ce = p->ce;
    for (uint32_t call = 0; call < p->max_sec_steps; call++) {
        if (E < ce) {
            deposit(k, W, px, py, pz, E);
            k->cnt[1] += 1;
            float sig = meesungnoen_sigma(E);
            float e[3] = {px, py, pz};
            if (sig > 0.0f) { float d[3]; gauss3(s, sig, d); e[0] += d[0]; e[1] += d[1]; e[2] += d[2]; }
            emit(k, e[0], e[1], e[2], 5.0f + pid);
            k->sst[1]++;
            return;
        }
        if (fabsf(px) >= box || fabsf(py) >= box || fabsf(pz) >= box) { k->sst[2]++; return; }
        float xs[3]; xs_all(E, xs);
        float s_vib = xs_vib_total(E), s_dea = xs_dea(E);
        float s_tot = xs[0] + xs[1] + xs[2] + s_vib + s_dea;
        if (s_tot <= 0.0f) return;
        float lambda = 1.0f / (NW_F * s_tot);
        float dist = -dna_logf(fmaxf_(rf(s), 1e-10f)) * lambda;
        float r_el = rf(s);
        float cos_el = xs_el_cos(E, r_el);
        float sin_el = sqrtf(fmaxf_(0.0f, 1.0f - cos_el * cos_el));
        float sp, cp; dna_sincosf(TWO_PI_F * rf(s), &sp, &cp);
        rotate(&dx, &dy, &dz, cos_el, sin_el, cp, sp);
        px += dx * dist; py += dy * dist; pz += dz * dist;
        k->sst[3]++;
        float r_type = rf(s) * s_tot;
        if (r_type < xs[0] && E > SB[0]) {
            float bind; uint32_t shell = pick_shell(s, E, &bind);
            if (E > bind) {
                float r_w = rf(s);
                float Wmax = (E + bind) * 0.5f;
                float Wt = clampf_(sample_W_sec(E, shell, r_w), bind, Wmax);
                float W_sec = fmaxf_(Wt - bind, 0.0f);
                E -= Wt;
                k->sst[4]++;
                deposit(k, W, px, py, pz, Wt);
                float m[3]; mother(s, px, py, pz, m);
                float site[3] = {px, py, pz};
                hole_pair(k, s, m, site, meesungnoen_sigma(W_sec), 5.0f, pid);
                if (dna_near(p, px, py, pz)) k->cnt[4]++;
            }
        } else if (r_type < xs[0] + xs[1]) {
            uint32_t lvl = pick_level(s, E);
            float dep = fminf_(EXL[lvl], E);
            E -= dep;
            deposit(k, W, px, py, pz, dep);
            excitation_products(k, s, lvl, px, py, pz, pid);
            if (dna_near(p, px, py, pz)) k->cnt[4]++;
        } else if (r_type < xs[0] + xs[1] + s_vib) {
            uint32_t mode = sample_vib_mode(E, rf(s));
            float dep = fminf_(T.VIB_LEV[mode], E);
            E -= dep;
            deposit(k, W, px, py, pz, dep);
        } else if (r_type < xs[0] + xs[1] + s_vib + s_dea) {
            deposit(k, W, px, py, pz, E);
            k->cnt[0] += 1; k->cnt[5] += 1;
            emit(k, px, py, pz, 0.0f + pid); emit(k, px, py, pz, 6.0f + pid); emit(k, px, py, pz, 7.0f + pid);
            E = 0.0f;
        }
    }
}

int oracle_transport_range(const dna_params *p, const float *energies, uint32_t first,
                           uint32_t count, int n_threads, oracle_transport_out *out)
{
    if (!T_loaded) return 1;
    memset(out, 0, sizeof(*out));
    world W = {p->box, p->vc};
    size_t nvox = (size_t)p->vc * p->vc * p->vc;
    out->dose = (uint32_t *)calloc(nvox, 4);
    out->results = (dna_result *)calloc(count ? count : 1, sizeof(dna_result));
    uint32_t *seeds = (uint32_t *)malloc(4 * sizeof(uint32_t) * (size_t)(first + count + 1));
    oracle_seed_primary(first + count, p->seed, seeds);
    sink *sinks = (sink *)calloc(count ? count : 1, sizeof(sink));
    uint32_t *nsecs = (uint32_t *)calloc(count ? count : 1, 4);
    uint32_t *nkeep = (uint32_t *)calloc(count ? count : 1, 4);
    secrec **secs = (secrec **)calloc(count ? count : 1, sizeof(secrec *));
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#endif
    /* primaries first, then the secondaries: with a cap (p->max_sec != 0) only the first
     * max_sec secondaries of the range in canonical (primary, emission) order are
     * transported, as the reference drops the secondaries beyond MAX_SEC
     * (primary.wgsl:364, dispatch.ts:124-126) */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
#endif
    for (int64_t c = 0; c < (int64_t)count; c++) {
        uint32_t i = first + (uint32_t)c;
        sink *k = &sinks[c];
        k->dose = out->dose;
        uint32_t s[4] = {seeds[4*i], seeds[4*i+1], seeds[4*i+2], seeds[4*i+3]};
        uint32_t capsec = 0;
        float E0 = energies ? energies[i] : p->be;
        run_primary(p, &W, i, E0, s, k, &secs[c], &nsecs[c], &capsec, &out->results[c]);
    }
    uint64_t sec_raw = 0;
    for (uint32_t c = 0; c < count; c++) {
        uint64_t keep = nsecs[c];
        if (p->max_sec) keep = sec_raw >= p->max_sec ? 0 : (p->max_sec - sec_raw < keep ? p->max_sec - sec_raw : keep);
        nkeep[c] = (uint32_t)keep;
        sec_raw += nsecs[c];
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
#endif
    for (int64_t c = 0; c < (int64_t)count; c++) {
        uint32_t i = first + (uint32_t)c;
        float pid = (float)((uint32_t)((float)(i + 1u)) - 1u) * 8.0f;   /* secondary.wgsl:88 */
        for (uint32_t q = 0; q < nkeep[c]; q++) run_secondary(p, &W, pid, &secs[c][q], &sinks[c]);
        free(secs[c]);
    }
    uint64_t tot = 0;
    for (uint32_t c = 0; c < count; c++) tot += sinks[c].n;
    /* with p->max_rad != 0 the first max_rad entries in canonical order are kept
     * (MAX_RAD, dispatch.ts:190-192) */
    const uint64_t kept = (p->max_rad && tot > p->max_rad) ? p->max_rad : tot;
    out->rad_buf = (float *)malloc(16 * (kept ? kept : 1));
    uint64_t o = 0;
    for (uint32_t c = 0; c < count; c++) {
        sink *k = &sinks[c];
        uint64_t m = k->n;
        if (o + m > kept) m = kept - o;
        if (m) memcpy(out->rad_buf + 4 * o, k->rad, 16 * m);
        o += m;
        for (int j = 0; j < 8; j++) { out->counters[j] += k->cnt[j]; out->sec_stats[j] += k->sst[j]; out->dbg[j] += k->dbg[j]; }
        out->sec_n += nkeep[c];
        free(k->rad);
    }
    out->rad_n = kept;
    out->rad_n_raw = tot;
    out->sec_n_raw = sec_raw;
    out->counters[7] = (uint32_t)tot;
    out->counters[6] = (uint32_t)sec_raw;
    uint64_t ds = 0;
    for (size_t v = 0; v < nvox; v++) ds += out->dose[v];
    out->dose_sum = ds;
    free(sinks); free(nsecs); free(nkeep); free(secs); free(seeds);
    return 0;
}

int oracle_transport(const dna_params *p, const float *energies, int n_threads,
                     oracle_transport_out *out)
{
    return oracle_transport_range(p, energies, 0, p->n, n_threads, out);
}

/* ------------------------------------------------------------ golden probes */

void oracle_xs(int which, const float *E, float *o, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        if (which == 0) xs_all(E[i], o + 3 * i);
        else if (which == 1) xs_shell_fracs(E[i], o + 5 * i);
        else if (which == 2) xs_exc_fracs(E[i], o + 5 * i);
        else if (which == 3) o[i] = xs_vib_total(E[i]);
        else if (which == 4) o[i] = xs_dea(E[i]);
    }
}
float oracle_xs_el_cos(float E, float r) { return xs_el_cos(E, r); }
float oracle_sample_W(float E, uint32_t shell, float r) { return sample_W_sec(E, shell, r); }
uint32_t oracle_sample_vib_mode(float E, float r) { return sample_vib_mode(E, r); }
float oracle_meesungnoen_sigma(float k) { return meesungnoen_sigma(k); }

void oracle_fmath32(int fn, const float *x, float *y, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        switch (fn) {
        case 0: y[i] = dna_logf(x[i]); break;
        case 1: y[i] = dna_expf(x[i]); break;
        case 2: y[i] = dna_sinf(x[i]); break;
        case 3: y[i] = dna_cosf(x[i]); break;
        case 4: y[i] = dna_powf(x[i], 1.3f); break;
        default: y[i] = 0.0f;
        }
    }
}
