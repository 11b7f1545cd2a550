/*
 * irt.c — CPU restatement of the IRT chemistry worker, public/irt-worker.js:178-699
 * (Karamitros 2011 nine-reaction scheme, per-primary O(n^2) pair IRT, min-heap event
 * loop with product re-pairing).  TEST INFRASTRUCTURE (see oracle.h).
 *
 * FP64 arithmetic as in the worker.  Two deliberate, documented canonicalisations:
 *  - Math.random() is replaced by the counter-based streams of include/dna_irt_math.h
 *    (one stream per particle pair / per particle), so results are reproducible;
 *  - events are ordered by (t, a, b) with a < b, a strict total order, where the
 *    worker's heap breaks ties of equal t arbitrarily (irt-worker.js:314-334).
 * Positions are stored as float after displacement (differences and distances in
 * double), matching the GPU kernel's storage.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"
#include "../include/dna_irt_math.h"

/* reaction table, irt-worker.js:196-226 */
static const double IRT_D[6] = {2.2, 4.9, 7.0, 9.46, 2.3, 5.3};
static const double VDW_R[6] = {0.22, 0.50, 0.19, 0.25, 0.21, 0.33};
static const double CHARGE[6] = {0, -1, 0, 1, 0, -1};
typedef struct { int a, b; double k; int np; int prods[2]; int type; } rxn_def;
static const rxn_def RXN[9] = {
    {0, 0, 0.55e10, 1, {4, 0}, 1}, {0, 1, 2.95e10, 1, {5, 0}, 1}, {0, 2, 1.55e10, 0, {0, 0}, 1},
    {1, 1, 0.636e10, 2, {5, 5}, 0}, {1, 2, 2.50e10, 1, {5, 0}, 0}, {1, 3, 2.11e10, 1, {2, 0}, 1},
    {2, 2, 0.503e10, 0, {0, 0}, 0}, {1, 4, 1.10e10, 2, {5, 0}, 1}, {3, 5, 1.13e11, 0, {0, 0}, 0},
};
static double rSigma[9], rSigmaEff[9], rRc[9], rKdif[9], rKact[9], rProb[9], rKobs[9];
static int rxnMap[36];
static double R_CUT2;
static double SIG_OH_PROD, SIG_H_EXC, SIG_EAQ_AUTO;
static int params_ready = 0;

/* parameter precomputation, irt-worker.js:240-286, 349-365 */
static void init_params(void)
{
    if (params_ready) return;
    const double ONSAGER = 1.44 / (0.02527 * 80.1);
    const double Rs = 0.29;
    const double K_CONV = 1e24 / (6.022e23 * 1e9);
    for (int r = 0; r < 9; r++) {
        int a = RXN[r].a, b = RXN[r].b;
        double D_sum = IRT_D[a] + IRT_D[b];
        int self = (a == b);
        double rc = CHARGE[a] * CHARGE[b] * ONSAGER;
        rRc[r] = rc;
        double k_nm = RXN[r].k * K_CONV;
        rKobs[r] = k_nm;
        if (RXN[r].type == 0) {
            double sig = k_nm / (4 * DNA_PI * D_sum);
            if (self) sig *= 2;
            rSigma[r] = sig; rSigmaEff[r] = sig; rKdif[r] = 0; rKact[r] = 0; rProb[r] = 1;
        } else {
            double sc = VDW_R[a] + VDW_R[b];
            rSigma[r] = sc;
            double se, kd;
            if (rc == 0) { se = sc; kd = 4 * DNA_PI * D_sum * sc; }
            else { se = -rc / (1 - dna_exp(rc / sc)); kd = 4 * DNA_PI * D_sum * se; }
            if (self) kd /= 2;
            rSigmaEff[r] = se; rKdif[r] = kd;
            double ka = kd * k_nm / (kd - k_nm);
            rKact[r] = ka;
            rProb[r] = Rs / (Rs + (kd / ka) * (se + Rs));
        }
    }
    for (int i = 0; i < 36; i++) rxnMap[i] = -1;
    for (int r = 0; r < 9; r++) { rxnMap[RXN[r].a * 6 + RXN[r].b] = r; rxnMap[RXN[r].b * 6 + RXN[r].a] = r; }
    double R_CUT = 1.45 + 2 * sqrt(8 * 9.46 * 1000);
    R_CUT2 = R_CUT * R_CUT;
    /* Meesungnoen 2002 at 1.7 eV as a power sum, irt-worker.js:16-28 */
    static const double g[13] = {-4.06217193e-08, 3.06848412e-06, -9.93217814e-05, 1.80172797e-03,
        -2.01135480e-02, 1.42939448e-01, -6.48348714e-01, 1.85227848e+00, -3.36450378e+00,
        4.37785068e+00, -4.20557339e+00, 3.81679083e+00, -2.34069784e-01};
    double rm = 0, pw[13];
    pw[0] = 1.0; for (int i = 1; i < 13; i++) pw[i] = pw[i-1] * 1.7;
    for (int i = 12; i >= 0; i--) rm += g[12 - i] * pw[i];
    SIG_EAQ_AUTO = rm * 0.62665706865775006;
    SIG_OH_PROD = 0.8 / sqrt(3.0);
    SIG_H_EXC = (17.0 / 18.0) * 2.4 / sqrt(3.0);
    params_ready = 1;
}

void oracle_irt_params(double *o)
{
    init_params();
    for (int r = 0; r < 9; r++) {
        o[6*r+0] = rSigma[r]; o[6*r+1] = rSigmaEff[r]; o[6*r+2] = rRc[r];
        o[6*r+3] = rKdif[r]; o[6*r+4] = rKact[r]; o[6*r+5] = rProb[r];
    }
}

/* SamplePDC, irt-worker.js:91-113.  Draws 7+2t and 8+2t of the pair stream. */
static double sample_pdc(double a, double b, uint64_t ps)
{
    double p = 2.0 * sqrt(2.0 * b / a);
    double q = 2.0 / sqrt(2.0 * b / a);
    double M1 = 1.0 / (a * a), M2 = 3.0 * b / a;
    double M = M1 > M2 ? M1 : M2;
    for (int trial = 0; trial < 10000; trial++) {
        double U = dna_u01(ps, 7u + 2u * (uint32_t)trial);
        double X, v;
        if (U < p / (p + q * M)) { v = U * (p + q * M) / 2; X = v * v; }
        else { v = 2 / ((1 - U) * (p + q * M) / M); X = v * v; }
        U = dna_u01(ps, 8u + 2u * (uint32_t)trial);
        double sqX = sqrt(X);
        double lam = dna_exp(-b * b / X) * (1.0 - a * sqrt(DNA_PI * X) * dna_erfcx(b / sqX + a * sqX));
        if ((X <= 2.0 * b / a && U <= lam) || (X >= 2.0 * b / a && U * M / X <= lam)) return X;
    }
    return -1.0;
}

/* sampleIRT, irt-worker.js:116-184.  The first uniform is draw 6 of the pair stream. */
static double sample_irt(double r0, int ri, double D, uint64_t ps)
{
    if (RXN[ri].type == 0) {
        double sigma = rSigma[ri], rc = rRc[ri];
        if (sigma <= 0 || D <= 0) return -1;
        if (r0 <= sigma) return 0;
        double r0e = r0;
        if (rc != 0) r0e = -rc / (1 - dna_exp(rc / r0));
        if (r0e <= sigma) return 0;
        double Winf = sigma / r0e;
        double U = dna_u01(ps, 6u);
        if (U <= 0 || U >= Winf) return -1;
        double ei = dna_erfcinv(r0e * U / sigma);
        if (fabs(ei) < 1e-10) return -1;
        double dr = r0e - sigma;
        return 0.25 * dr * dr / (D * ei * ei);
    }
    double sc = rSigma[ri], rc = rRc[ri], kobs = rKobs[ri], kdif = rKdif[ri], kact = rKact[ri];
    double sig_chk = sc, r0_chk = r0;
    if (rc != 0) { r0_chk = -rc / (1 - dna_exp(rc / r0)); sig_chk = rSigmaEff[ri]; }
    if (sig_chk > r0_chk) return (rProb[ri] > dna_u01(ps, 6u)) ? 0.0 : -1.0;
    double Winf = sig_chk / r0_chk * kobs / kdif;
    if (!(Winf > dna_u01(ps, 6u))) return -1;
    double a, b;
    if (rc == 0) { a = (1.0 / sc) * kact / kobs; b = (r0 - sc) / 2; }
    else {
        double v = kact / (4 * DNA_PI * sc * sc * dna_exp(-rc / sc));
        double alpha = v + rc * D / (sc * sc * (1 - dna_exp(-rc / sc)));
        double sh, ch; dna_sinhcosh(rc / (2 * sc), &sh, &ch);
        a = 4 * sc * sc * alpha / (D * rc * rc) * sh * sh;
        double sh0, ch0; dna_sinhcosh(rc / (2 * r0), &sh0, &ch0);
        b = rc / 4 * (ch0 / sh0 - ch / sh);
    }
    double X = sample_pdc(a, b, ps);
    if (X > 0) return X / D;
    return -1;
}

double oracle_irt_sample(double r0, int ri, uint64_t seed, uint32_t track, uint32_t a, uint32_t b)
{
    init_params();
    uint64_t tk = dna_track_key(seed, 0, track);
    int s0 = RXN[ri].a, s1 = RXN[ri].b;
    return sample_irt(r0, ri, IRT_D[s0] + IRT_D[s1], dna_pair_state(tk, a, b));
}

/* ------------------------------------------------------------ event heap */
typedef struct { double t; uint32_t a, b; } ev;
typedef struct { ev *v; int64_t n, cap; } heap;
static int ev_less(const ev *x, const ev *y)
{
    if (x->t != y->t) return x->t < y->t;
    if (x->a != y->a) return x->a < y->a;
    return x->b < y->b;
}
static void hpush(heap *h, double t, uint32_t a, uint32_t b)
{
    if (h->n == h->cap) { h->cap = h->cap ? 2 * h->cap : 256; h->v = (ev *)realloc(h->v, h->cap * sizeof(ev)); }
    int64_t c = h->n++;
    h->v[c].t = t; h->v[c].a = a; h->v[c].b = b;
    while (c > 0) {
        int64_t p = (c - 1) >> 1;
        if (!ev_less(&h->v[c], &h->v[p])) break;
        ev tmp = h->v[p]; h->v[p] = h->v[c]; h->v[c] = tmp; c = p;
    }
}
static ev hpop(heap *h)
{
    ev top = h->v[0];
    h->n--;
    if (h->n > 0) {
        h->v[0] = h->v[h->n];
        int64_t p = 0;
        for (;;) {
            int64_t s = p, l = 2 * p + 1, r = 2 * p + 2;
            if (l < h->n && ev_less(&h->v[l], &h->v[s])) s = l;
            if (r < h->n && ev_less(&h->v[r], &h->v[s])) s = r;
            if (s == p) break;
            ev tmp = h->v[p]; h->v[p] = h->v[s]; h->v[s] = tmp; p = s;
        }
    }
    return top;
}

/* ------------------------------------------------------------ one track */
typedef struct {
    float *x, *y, *z; int8_t *sp; uint8_t *alive; double *tb, *td; int64_t cap;
    heap h;
    uint64_t n_all, n_samp, n_rep;   /* pairs with a channel, within R_CUT, re-pairings */
} work;

static void work_reserve(work *w, int64_t cap)
{
    if (cap <= w->cap) return;
    w->x = (float *)realloc(w->x, cap * 4); w->y = (float *)realloc(w->y, cap * 4);
    w->z = (float *)realloc(w->z, cap * 4); w->sp = (int8_t *)realloc(w->sp, cap);
    w->alive = (uint8_t *)realloc(w->alive, cap); w->tb = (double *)realloc(w->tb, cap * 8);
    w->td = (double *)realloc(w->td, cap * 8);
    w->cap = cap;
}

static inline double clf6(uint64_t ps, uint32_t *n)
{
    double s = dna_u01(ps, (*n)++);
    for (int i = 1; i < 6; i++) s += dna_u01(ps, (*n)++);
    return (s - 3) * 1.4142135623730951;
}

/* pairWithAlive, irt-worker.js:430-465 */
static void pair_with_alive(work *w, uint64_t tk, uint32_t idx, uint32_t n_total, double t_new)
{
    int si = w->sp[idx];
    double Di = IRT_D[si];
    for (uint32_t jj = 0; jj < n_total; jj++) {
        if (jj == idx || !w->alive[jj]) continue;
        int sj = w->sp[jj];
        int ri = rxnMap[si * 6 + sj];
        if (ri < 0) continue;
        w->n_rep++;
        double Dj = IRT_D[sj];
        uint32_t a = idx < jj ? idx : jj, b = idx < jj ? jj : idx;
        uint64_t ps = dna_pair_state(tk, a, b);
        double dt = t_new - w->tb[jj];
        double xj = w->x[jj], yj = w->y[jj], zj = w->z[jj];
        if (dt > 0) {
            double sig = sqrt(2.0 * Dj * dt);
            double g0 = sqrt(-2 * dna_log(dna_u01(ps, 0) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 1));
            double g1 = sqrt(-2 * dna_log(dna_u01(ps, 2) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 3));
            double g2 = sqrt(-2 * dna_log(dna_u01(ps, 4) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 5));
            xj += g0 * sig; yj += g1 * sig; zj += g2 * sig;
        }
        double ddx = (double)w->x[idx] - xj, ddy = (double)w->y[idx] - yj, ddz = (double)w->z[idx] - zj;
        double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (r2 > R_CUT2) continue;
        double t = sample_irt(sqrt(r2), ri, Di + Dj, ps);
        if (t >= 0) {
            double ta = t_new + t;
            if (ta < DNA_T_END_NS) hpush(&w->h, ta, a, b);
        }
    }
}

static void snapshot(const work *w, uint32_t n_total, uint64_t h2, uint64_t *tally)
{
    uint64_t oh = 0, eq = 0, hh = 0, h2o2 = 0;
    for (uint32_t k = 0; k < n_total; k++) {
        if (!w->alive[k]) continue;
        int s = w->sp[k];
        if (s == 0) oh++; else if (s == 1) eq++; else if (s == 2) hh++; else if (s == 4) h2o2++;
    }
    tally[0] += oh; tally[1] += eq; tally[2] += hh; tally[3] += h2o2; tally[4] += h2;
}

/* Phase 3 body for one primary, irt-worker.js:470-660 */
static uint32_t run_track(work *w, const float *rad, const uint64_t *idx, uint32_t n, uint64_t tk,
                          uint64_t init_h2, int64_t CAP, const double *tc, int nc,
                          uint64_t (*tally)[5], uint64_t *rxn_counts, uint64_t *n_reacted, uint64_t *n_events)
{
    work_reserve(w, CAP > n ? CAP : n);
    for (uint32_t k = 0; k < n; k++) {
        const float *e = rad + 4 * idx[k];
        double x = e[0], y = e[1], z = e[2];
        int s = (int)fmod(floor((double)e[3] + 0.5), 8.0);
        w->alive[k] = 1; w->tb[k] = 0.0; w->td[k] = INFINITY;
        uint64_t ps = dna_pair_state(tk, k, k);
        uint32_t dn = 0;
        double sig = 0.0;
        if (s == 5) { w->sp[k] = 1; }
        else if (s == 0 || s == 3) {
            w->sp[k] = (int8_t)s;
            if (dna_u01(ps, dn++) < 0.5) sig = SIG_OH_PROD;
        } else if (s == 1) { w->sp[k] = 1; sig = SIG_EAQ_AUTO; }
        else if (s == 2) { w->sp[k] = 2; sig = SIG_H_EXC; }
        else if (s == 6) { w->sp[k] = 5; }
        else { w->alive[k] = 0; w->sp[k] = (int8_t)s; }
        if (sig != 0.0) {
            x += sig * clf6(ps, &dn);
            y += sig * clf6(ps, &dn);
            z += sig * clf6(ps, &dn);
        }
        w->x[k] = (float)x; w->y[k] = (float)y; w->z[k] = (float)z;
    }
    w->h.n = 0;
    /* initial pairs, irt-worker.js:541-558 */
    for (uint32_t i = 0; i < n; i++) {
        if (!w->alive[i]) continue;
        int si = w->sp[i];
        double Di = IRT_D[si];
        for (uint32_t j = i + 1; j < n; j++) {
            if (!w->alive[j]) continue;
            int sj = w->sp[j];
            int ri = rxnMap[si * 6 + sj];
            if (ri < 0) continue;
            w->n_all++;
            double ddx = (double)w->x[i] - (double)w->x[j];
            double ddy = (double)w->y[i] - (double)w->y[j];
            double ddz = (double)w->z[i] - (double)w->z[j];
            double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (r2 > R_CUT2) continue;
            w->n_samp++;
            double t = sample_irt(sqrt(r2), ri, Di + IRT_D[sj], dna_pair_state(tk, i, j));
            if (t >= 0 && t < DNA_T_END_NS) hpush(&w->h, t, i, j);
        }
    }
    *n_events += (uint64_t)w->h.n;
    uint64_t h2 = init_h2;
    int cp = 0;
    uint32_t n_total = n;
    while (w->h.n > 0) {
        ev e = hpop(&w->h);
        while (cp < nc && e.t >= tc[cp]) { snapshot(w, n_total, h2, tally[cp]); cp++; }
        if (!w->alive[e.a] || !w->alive[e.b]) continue;
        int si = w->sp[e.a], sj = w->sp[e.b];
        int ri = rxnMap[(si < sj ? si : sj) * 6 + (si < sj ? sj : si)];
        if (ri < 0) continue;
        w->alive[e.a] = 0; w->alive[e.b] = 0;
        w->td[e.a] = e.t; w->td[e.b] = e.t;
        (*n_reacted)++; rxn_counts[ri]++;
        if (ri == 3 || ri == 4 || ri == 6) h2++;
        int np = RXN[ri].np;
        if (np == 0) continue;
        double sqDi = sqrt(IRT_D[si]), sqDj = sqrt(IRT_D[sj]);
        double wsum = sqDi + sqDj;
        double mx = (sqDj * w->x[e.a] + sqDi * w->x[e.b]) / wsum;
        double my = (sqDj * w->y[e.a] + sqDi * w->y[e.b]) / wsum;
        double mz = (sqDj * w->z[e.a] + sqDi * w->z[e.b]) / wsum;
        for (int pp = 0; pp < np; pp++) {
            int64_t id = n_total;
            if (id >= CAP) break;
            work_reserve(w, id + 1);
            w->sp[id] = (int8_t)RXN[ri].prods[pp];
            w->alive[id] = 1;
            w->tb[id] = e.t; w->td[id] = INFINITY;
            if (np == 1) { w->x[id] = (float)mx; w->y[id] = (float)my; w->z[id] = (float)mz; }
            else if (pp == 0) { w->x[id] = w->x[e.a]; w->y[id] = w->y[e.a]; w->z[id] = w->z[e.a]; }
            else { w->x[id] = w->x[e.b]; w->y[id] = w->y[e.b]; w->z[id] = w->z[e.b]; }
            n_total++;
            pair_with_alive(w, tk, (uint32_t)id, n_total, e.t);
        }
    }
    while (cp < nc) { snapshot(w, n_total, h2, tally[cp]); cp++; }
    return n_total;
}

/* The OH alive at the end of a track (1 us), in particle order, each moved by one Brownian
 * step over t_end - t_birth with draws 32..37 of its own stream (tk, k, k): Box-Muller as in
 * the re-pairing step (irt-worker.js:443-456).  The worker keeps no positions at 1 us
 * (chem_pos_final = null, worker.ts:69-70); this is the definition the indirect-SSB
 * scoring (ssb-dsb.ts:43-121) is given here. */
static void collect_oh(const work *w, uint32_t n_total, uint64_t tk, float **buf, uint64_t *n)
{
    uint64_t cnt = 0;
    for (uint32_t k = 0; k < n_total; k++) if (w->alive[k] && w->sp[k] == 0) cnt++;
    *buf = (float *)malloc(16 * (cnt ? cnt : 1));
    *n = cnt;
    uint64_t o = 0;
    for (uint32_t k = 0; k < n_total; k++) {
        if (!w->alive[k] || w->sp[k] != 0) continue;
        double x = w->x[k], y = w->y[k], z = w->z[k];
        double dt = DNA_T_END_NS - w->tb[k];
        if (dt > 0) {
            uint64_t ps = dna_pair_state(tk, k, k);
            double sig = sqrt(2.0 * IRT_D[0] * dt);
            double g0 = sqrt(-2 * dna_log(dna_u01(ps, 32) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 33));
            double g1 = sqrt(-2 * dna_log(dna_u01(ps, 34) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 35));
            double g2 = sqrt(-2 * dna_log(dna_u01(ps, 36) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 37));
            x += g0 * sig; y += g1 * sig; z += g2 * sig;
        }
        float *q = *buf + 4 * o++;
        q[0] = (float)x; q[1] = (float)y; q[2] = (float)z; q[3] = 0.0f;
    }
}

/* Every particle slot of a track at the end: (x, y, z, pid * 8 + rad_buf code) with the
 * species mapped back to the rad_buf codes (OH- is 6), birth and death time (infinite if it
 * never reacted).  6 doubles per slot: x, y, z, w, t_birth, t_death. */
static void collect_particles(const work *w, uint32_t n_total, uint64_t pid, double **buf, uint64_t *n)
{
    static const int code[6] = {0, 1, 2, 3, 4, 6};
    *buf = (double *)malloc(6 * 8 * (n_total ? n_total : 1));
    *n = n_total;
    for (uint32_t k = 0; k < n_total; k++) {
        double *q = *buf + 6 * k;
        int s = w->sp[k];
        q[0] = w->x[k]; q[1] = w->y[k]; q[2] = w->z[k];
        q[3] = (double)(pid * 8 + (uint64_t)(s >= 0 && s < 6 ? code[s] : s));
        q[4] = w->tb[k]; q[5] = w->td[k];
    }
}

typedef struct { int64_t pid; uint64_t first, n, h2; } trk;

static int cmp_u64pair(const void *a, const void *b)
{
    const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

/* One worker invocation (irt-worker.js:370-699) over rad[0..rad_n); instance keys the RNG. */
static int irt_instance(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                        const double *tc, int nc, uint64_t seed, uint64_t instance, int n_threads,
                        dna_chem_out *out, float **oh_out, uint64_t *oh_n, double **pt_out, uint64_t *pt_n)
{
    init_params();
    memset(out, 0, sizeof(*out));
    out->n_ckpt = nc;
    for (int c = 0; c < nc; c++) out->t_ckpt[c] = tc[c];
    out->chem_n = rad_n;
    /* Phase 1: group by pid in order of first appearance (Map insertion order).  Sort
     * (pid, index) pairs; particle order within a track stays the rad_buf order. */
    uint64_t *pairs = (uint64_t *)malloc(16 * (rad_n ? rad_n : 1));
    uint64_t np = 0;
    int64_t *h2count_pid = NULL;
    for (uint64_t i = 0; i < rad_n; i++) {
        double w = floor((double)rad[4 * i + 3] + 0.5);
        double spd = fmod(w, 8.0);
        if (spd < 0 || spd == 4) continue;
        pairs[2 * np] = (uint64_t)floor(w / 8.0); pairs[2 * np + 1] = i | ((spd == 7) ? (1ull << 63) : 0);
        np++;
    }
    (void)h2count_pid;
    qsort(pairs, np, 16, cmp_u64pair);
    /* tracks: runs of equal pid; markers (bit 63) count as initial H2 */
    trk *tr = (trk *)malloc(sizeof(trk) * (np ? np : 1));
    uint64_t *idx = (uint64_t *)malloc(8 * (np ? np : 1));
    uint64_t ntr = 0, nidx = 0, maxN = 0;
    for (uint64_t k = 0; k < np;) {
        uint64_t pid = pairs[2 * k], first = nidx, h2 = 0;
        while (k < np && pairs[2 * k] == pid) {
            uint64_t v = pairs[2 * k + 1];
            if (v >> 63) h2++; else idx[nidx++] = v & ~(1ull << 63);
            k++;
        }
        uint64_t n = nidx - first;
        if (n == 0) continue;              /* pid with markers only: not in priMap */
        tr[ntr].pid = (int64_t)pid; tr[ntr].first = first; tr[ntr].n = n; tr[ntr].h2 = h2; ntr++;
        if (n > maxN) maxN = n;
    }
    out->n_tracks = ntr;
    int64_t CAP = (int64_t)maxN * 2 + 512;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    uint64_t (*tally)[DNA_MAX_CKPT][5] = calloc((size_t)n_threads, sizeof(*tally));
    uint64_t (*rxc)[DNA_N_RXN] = calloc((size_t)n_threads, sizeof(*rxc));
    uint64_t *nre = calloc((size_t)n_threads, 8), *nev = calloc((size_t)n_threads, 8);
    uint64_t (*work3)[3] = calloc((size_t)n_threads, sizeof(*work3));
    float **oh_tr = oh_out ? (float **)calloc(ntr ? ntr : 1, sizeof(float *)) : NULL;
    uint64_t *oh_tn = oh_out ? (uint64_t *)calloc(ntr ? ntr : 1, 8) : NULL;
    double **pt_tr = pt_out ? (double **)calloc(ntr ? ntr : 1, sizeof(double *)) : NULL;
    uint64_t *pt_tn = pt_out ? (uint64_t *)calloc(ntr ? ntr : 1, 8) : NULL;
#ifdef _OPENMP
#pragma omp parallel num_threads(n_threads)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        work w; memset(&w, 0, sizeof(w));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t t = 0; t < (int64_t)ntr; t++) {
            uint64_t tk = dna_track_key(seed, instance, (uint64_t)tr[t].pid);
            uint32_t nt = run_track(&w, rad, idx + tr[t].first, (uint32_t)tr[t].n, tk, tr[t].h2, CAP, tc, nc,
                                    tally[tid], rxc[tid], &nre[tid], &nev[tid]);
            if (oh_tr) collect_oh(&w, nt, tk, &oh_tr[t], &oh_tn[t]);
            if (pt_tr) collect_particles(&w, nt, (uint64_t)tr[t].pid, &pt_tr[t], &pt_tn[t]);
        }
        work3[tid][0] = w.n_all; work3[tid][1] = w.n_samp; work3[tid][2] = w.n_rep;
        free(w.x); free(w.y); free(w.z); free(w.sp); free(w.alive); free(w.tb); free(w.td); free(w.h.v);
    }
    for (int th = 0; th < n_threads; th++) {
        for (int c = 0; c < nc; c++) for (int s = 0; s < 5; s++) out->tally[c][s] += tally[th][c][s];
        for (int r = 0; r < 9; r++) out->rxn_counts[r] += rxc[th][r];
        out->n_reacted += nre[th]; out->n_events += nev[th];
        out->n_pair_all += work3[th][0]; out->n_pair_sampled += work3[th][1]; out->n_repair += work3[th][2];
    }
    out->n_pair_cand = out->n_pair_all;   /* the worker visits every pair */
    free(work3);
    double per100 = (double)n_therm * (double)E_eV / 100.0;
    for (int c = 0; c < nc; c++) for (int s = 0; s < 5; s++)
        out->G[c][s] = per100 > 0 ? (double)out->tally[c][s] / per100 : 0.0;
    if (oh_out) {                         /* concatenate in track order */
        uint64_t tot = 0;
        for (uint64_t t = 0; t < ntr; t++) tot += oh_tn[t];
        *oh_out = (float *)malloc(16 * (tot ? tot : 1));
        uint64_t o = 0;
        for (uint64_t t = 0; t < ntr; t++) {
            memcpy(*oh_out + 4 * o, oh_tr[t], 16 * oh_tn[t]);
            o += oh_tn[t];
            free(oh_tr[t]);
        }
        *oh_n = tot;
        free(oh_tr); free(oh_tn);
    }
    if (pt_out) {
        uint64_t tot = 0;
        for (uint64_t t = 0; t < ntr; t++) tot += pt_tn[t];
        *pt_out = (double *)malloc(6 * 8 * (tot ? tot : 1));
        uint64_t o = 0;
        for (uint64_t t = 0; t < ntr; t++) {
            memcpy(*pt_out + 6 * o, pt_tr[t], 6 * 8 * pt_tn[t]);
            o += pt_tn[t];
            free(pt_tr[t]);
        }
        *pt_n = tot;
        free(pt_tr); free(pt_tn);
    }
    free(tally); free(rxc); free(nre); free(nev); free(tr); free(idx); free(pairs);
    return 0;
}

int oracle_irt(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
               const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads, dna_chem_out *out)
{
    if (n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) return 1;
    return irt_instance(rad, rad_n, n_therm, E_eV, t_ckpt, n_ckpt, seed, 0, n_threads, out, NULL, NULL, NULL, NULL);
}

/* oracle_irt, then the OH alive at 1 us (collect_oh) of every track in track order;
 * *oh_out is malloc'ed (oracle_free). */
int oracle_irt_oh(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                  const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads, dna_chem_out *out,
                  float **oh_out, uint64_t *oh_n)
{
    if (n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) return 1;
    return irt_instance(rad, rad_n, n_therm, E_eV, t_ckpt, n_ckpt, seed, 0, n_threads, out, oh_out, oh_n, NULL, NULL);
}

/* oracle_irt plus every particle slot of every track at the end, in track order:
 * [n][6] doubles (x, y, z, pid * 8 + code, t_birth, t_death); *pt_out is malloc'ed */
int oracle_irt_particles(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                         const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads, dna_chem_out *out,
                         double **pt_out, uint64_t *pt_n)
{
    if (n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) return 1;
    return irt_instance(rad, rad_n, n_therm, E_eV, t_ckpt, n_ckpt, seed, 0, n_threads, out, NULL, NULL, pt_out, pt_n);
}

int oracle_irt_batch(const float *rad, const uint64_t *off, uint32_t n_inst, const uint32_t *n_therm,
                     const float *E, const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads,
                     dna_chem_out *out)
{
    if (n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) return 1;
    for (uint32_t i = 0; i < n_inst; i++) {
        int rc = irt_instance(rad + 4 * off[i], off[i + 1] - off[i], n_therm[i], E[i], t_ckpt, n_ckpt,
                              seed, i, n_threads, &out[i], NULL, NULL, NULL, NULL);
        if (rc) return rc;
    }
    return 0;
}

/* FP64 elementary-function probe for accuracy tests (codes of oracle.h) */
void oracle_fmath64(int fn, const double *x, double *y, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        double s, c;
        switch (fn) {
        case 10: y[i] = dna_log(x[i]); break;
        case 11: y[i] = dna_exp(x[i]); break;
        case 12: y[i] = dna_cos(x[i]); break;
        case 13: dna_sinhcosh(x[i], &s, &c); y[i] = s; break;
        case 14: dna_sinhcosh(x[i], &s, &c); y[i] = c; break;
        case 15: y[i] = dna_erfcx(x[i]); break;
        case 16: y[i] = dna_erfcinv(x[i]); break;
        default: y[i] = 0.0;
        }
    }
}
