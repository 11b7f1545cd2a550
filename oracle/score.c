/*
 * score.c — CPU restatement of the SSB/DSB scoring of src/scoring/ssb-dsb.ts and of the
 * spatial-hash cell assignment of src/shaders/chemistry.wgsl.  TEST INFRASTRUCTURE.
 *
 * JS semantics reproduced: Float32Array inputs promoted to double, Math.round(x) =
 * floor(x + 0.5), the LCG of makeSsbRng (ssb-dsb.ts:421-427) as exact u32 arithmetic.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle.h"

static inline double js_round(double x) { return floor(x + 0.5); }

/* scoreDirectSSB_events, ssb-dsb.ts:135-217 */
static void direct_events(const dna_dna_target *d, const float *rad, uint64_t rad_n,
                          uint32_t *lcg, uint8_t *hits, dna_score_out *o)
{
    const double r_direct = 0.29, r2d = r_direct * r_direct, p_direct = 0.15;
    double x_half = (d->n_bp_per - 1) * d->rise * 0.5;
    double rise_inv = 1 / d->rise;
    double grid_off = -((d->grid_N - 1) * d->spacing_nm) * 0.5;
    double inv_spacing = 1 / d->spacing_nm;
    uint64_t i = 0;
    while (i < rad_n) {
        float xf = rad[4*i], yf = rad[4*i+1], zf = rad[4*i+2];
        double x = xf, y = yf, z = zf;
        uint64_t skip = 1;
        if (i + 1 < rad_n && rad[4*(i+1)] == xf && rad[4*(i+1)+1] == yf && rad[4*(i+1)+2] == zf) skip = 2;
        i += skip;
        if (x < -x_half - r_direct || x > x_half + r_direct) continue;
        double fi = js_round((y - grid_off) * inv_spacing);
        double fj = js_round((z - grid_off) * inv_spacing);
        if (fi < 0 || fi >= d->grid_N || fj < 0 || fj >= d->grid_N) continue;
        int64_t fiber = (int64_t)fi * d->grid_N + (int64_t)fj;
        double y_rel = y - d->fy[fiber], z_rel = z - d->fz[fiber];
        double r2 = y_rel * y_rel + z_rel * z_rel;
        double outer = d->r_bb + r_direct;
        if (r2 > outer * outer) continue;
        o->candidates++;
        double bp_est = js_round((x + x_half) * rise_inv);
        int64_t b0 = (int64_t)(bp_est - 2 > 0 ? bp_est - 2 : 0);
        int64_t b1 = (int64_t)(bp_est + 2 < d->n_bp_per - 1 ? bp_est + 2 : d->n_bp_per - 1);
        double best = INFINITY; int64_t best_bp = -1; int best_strand = -1;
        for (int64_t b = b0; b <= b1; b++) {
            double dx = x - (d->x0 + b * d->rise);
            double dy0 = y_rel - d->rbb0[2*b], dz0 = z_rel - d->rbb0[2*b+1];
            double d20 = dx * dx + dy0 * dy0 + dz0 * dz0;
            if (d20 < best) { best = d20; best_bp = b; best_strand = 0; }
            double dy1 = y_rel - d->rbb1[2*b], dz1 = z_rel - d->rbb1[2*b+1];
            double d21 = dx * dx + dy1 * dy1 + dz1 * dz1;
            if (d21 < best) { best = d21; best_bp = b; best_strand = 1; }
        }
        if (best < r2d) {
            o->in_reach++;
            *lcg = *lcg * 1664525u + 1013904223u;
            double u = (double)*lcg / 4294967296.0;
            if (u < p_direct) {
                uint64_t idx = (uint64_t)(fiber * d->n_bp_per + best_bp) + (uint64_t)best_strand * d->n_bp;
                if (hits[idx] == 0) { hits[idx] = 1; o->ssb_count++; }
            }
        }
    }
}

/* scoreIndirectSSB, ssb-dsb.ts:43-121, on a list of alive OH positions: the same geometry
 * with the OH reaction radius 0.29 nm and p = 0.4 (constants.ts:122-123), no duplicate
 * skipping, one LCG draw per in-reach OH in list order, hits in its own array */
static void indirect_oh(const dna_dna_target *d, const float *oh, uint64_t n, uint32_t *lcg, uint8_t *hits,
                        dna_indirect_out *o)
{
    const double r_damage = 0.29, r2d = r_damage * r_damage, p_ssb = 0.4;
    double x_half = (d->n_bp_per - 1) * d->rise * 0.5;
    double rise_inv = 1 / d->rise;
    double grid_off = -((d->grid_N - 1) * d->spacing_nm) * 0.5;
    double inv_spacing = 1 / d->spacing_nm;
    for (uint64_t i = 0; i < n; i++) {
        double x = oh[4*i], y = oh[4*i+1], z = oh[4*i+2];
        if (x < -x_half - r_damage || x > x_half + r_damage) continue;
        double fi = js_round((y - grid_off) * inv_spacing);
        double fj = js_round((z - grid_off) * inv_spacing);
        if (fi < 0 || fi >= d->grid_N || fj < 0 || fj >= d->grid_N) continue;
        int64_t fiber = (int64_t)fi * d->grid_N + (int64_t)fj;
        double y_rel = y - d->fy[fiber], z_rel = z - d->fz[fiber];
        double r2 = y_rel * y_rel + z_rel * z_rel;
        double outer = d->r_bb + r_damage;
        if (r2 > outer * outer) continue;
        o->candidates++;
        double bp_est = js_round((x + x_half) * rise_inv);
        int64_t b0 = (int64_t)(bp_est - 2 > 0 ? bp_est - 2 : 0);
        int64_t b1 = (int64_t)(bp_est + 2 < d->n_bp_per - 1 ? bp_est + 2 : d->n_bp_per - 1);
        double best = INFINITY; int64_t best_bp = -1; int best_strand = -1;
        for (int64_t b = b0; b <= b1; b++) {
            double dx = x - (d->x0 + b * d->rise);
            double dy0 = y_rel - d->rbb0[2*b], dz0 = z_rel - d->rbb0[2*b+1];
            double d20 = dx * dx + dy0 * dy0 + dz0 * dz0;
            if (d20 < best) { best = d20; best_bp = b; best_strand = 0; }
            double dy1 = y_rel - d->rbb1[2*b], dz1 = z_rel - d->rbb1[2*b+1];
            double d21 = dx * dx + dy1 * dy1 + dz1 * dz1;
            if (d21 < best) { best = d21; best_bp = b; best_strand = 1; }
        }
        if (best < r2d) {
            o->in_reach++;
            *lcg = *lcg * 1664525u + 1013904223u;
            double u = (double)*lcg / 4294967296.0;
            if (u < p_ssb) {
                uint64_t idx = (uint64_t)(fiber * d->n_bp_per + best_bp) + (uint64_t)best_strand * d->n_bp;
                if (hits[idx] == 0) { hits[idx] = 1; if (best_strand == 0) o->ssb0++; else o->ssb1++; }
            }
        }
    }
}

/* clusterDSB, ssb-dsb.ts:348-407 */
static void cluster(const dna_dna_target *d, const uint8_t *hits, dna_score_out *o)
{
    const int W = 10;
    uint64_t n = d->n_bp, n_per = d->n_bp_per;
    uint32_t *s0 = (uint32_t *)malloc(4 * n_per), *s1 = (uint32_t *)malloc(4 * n_per);
    uint8_t *used = (uint8_t *)malloc(n_per);
    for (uint64_t fi = 0; fi < d->n_fibers; fi++) {
        uint64_t base = fi * n_per;
        int any = 0;
        for (uint64_t b = 0; b < n_per && !any; b++) if (hits[base + b] || hits[base + b + n]) any = 1;
        if (!any) continue;
        uint64_t k0 = 0, k1 = 0;
        for (uint64_t b = 0; b < n_per; b++) {
            if (hits[base + b] == 1) s0[k0++] = (uint32_t)b;
            if (hits[base + b + n] == 1) s1[k1++] = (uint32_t)b;
        }
        o->ssb0 += k0; o->ssb1 += k1;
        memset(used, 0, k1);
        uint64_t jlo = 0;
        for (uint64_t a = 0; a < k0; a++) {
            int64_t b0 = s0[a];
            while (jlo < k1 && (int64_t)s1[jlo] < b0 - W) jlo++;
            for (uint64_t j = jlo; j < k1 && (int64_t)s1[j] <= b0 + W; j++)
                if (!used[j]) { used[j] = 1; o->dsb++; break; }
        }
        if (k0 > 0 && k1 > 0 && n_per > 0) {
            double p_pair = (double)(2 * W + 1) / (double)n_per;
            if (p_pair > 1) p_pair = 1;
            o->dsb_expected += 1 - pow(1 - p_pair, (double)(k0 * k1));
        }
    }
    free(s0); free(s1); free(used);
}

int oracle_score(const dna_dna_target *tgt, const float *rad, uint64_t rad_n,
                 uint32_t *lcg_state, uint8_t *hits_out, dna_score_out *out)
{
    memset(out, 0, sizeof(*out));
    uint8_t *hits = (uint8_t *)calloc(2 * (size_t)tgt->n_bp, 1);
    /* indirect hits are empty on the default worker path (app.ts:178-180), so
     * combineHits(direct, indirect) == direct */
    direct_events(tgt, rad, rad_n, lcg_state, hits, out);
    cluster(tgt, hits, out);
    if (hits_out) memcpy(hits_out, hits, 2 * (size_t)tgt->n_bp);
    free(hits);
    return 0;
}

/* scoreDamageAt10keV with chemistry positions (app.ts:176-186): indirect SSB from the OH
 * list first, then direct from rad_buf with the same LCG, combineHits (ssb-dsb.ts:413-418)
 * as the OR of the two hit arrays, clusterDSB on the combination */
int oracle_score_oh(const dna_dna_target *tgt, const float *rad, uint64_t rad_n, const float *oh, uint64_t n_oh,
                    uint32_t *lcg_state, uint8_t *hits_out, dna_score_out *out, dna_indirect_out *ind)
{
    memset(out, 0, sizeof(*out));
    memset(ind, 0, sizeof(*ind));
    size_t nh = 2 * (size_t)tgt->n_bp;
    uint8_t *hi = (uint8_t *)calloc(nh, 1), *hd = (uint8_t *)calloc(nh, 1);
    indirect_oh(tgt, oh, n_oh, lcg_state, hi, ind);
    direct_events(tgt, rad, rad_n, lcg_state, hd, out);
    for (size_t k = 0; k < nh; k++) hd[k] |= hi[k];
    cluster(tgt, hd, out);
    if (hits_out) memcpy(hits_out, hd, nh);
    free(hi); free(hd);
    return 0;
}

/* Spatial hash, chemistry.wgsl:66-80 (hash3) and 110-122 (build_hash): cell index
 * floor(p / cell) with an f32 division, Teschner key with wrapping i32 products,
 * reinterpreted as u32, modulo 2^bits.  Buckets are emitted as a counting sort. */
void oracle_hash_cells(const float *pos4, uint64_t n, float cell_nm, uint32_t bits,
                       uint32_t *key, uint32_t *start, uint32_t *order)
{
    uint64_t nb = 1ull << bits;
    memset(start, 0, 4 * (nb + 1));
    for (uint64_t i = 0; i < n; i++) {
        int32_t ix = (int32_t)floorf(pos4[4*i] / cell_nm);
        int32_t iy = (int32_t)floorf(pos4[4*i+1] / cell_nm);
        int32_t iz = (int32_t)floorf(pos4[4*i+2] / cell_nm);
        uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
        key[i] = h & (uint32_t)(nb - 1);
        start[key[i] + 1]++;
    }
    for (uint64_t b = 0; b < nb; b++) start[b + 1] += start[b];
    uint32_t *fill = (uint32_t *)malloc(4 * nb);
    memcpy(fill, start, 4 * nb);
    for (uint64_t i = 0; i < n; i++) order[fill[key[i]]++] = (uint32_t)i;
    free(fill);
}

/* clusterDSB alone on a caller-supplied hit array (test hook) */
void oracle_cluster(const dna_dna_target *tgt, const uint8_t *hits, dna_score_out *out)
{
    memset(out, 0, sizeof(*out));
    cluster(tgt, hits, out);
}
