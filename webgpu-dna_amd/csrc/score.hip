// score.hip — SSB/DSB scoring (src/scoring/ssb-dsb.ts) and spatial-hash binning
// (src/shaders/chemistry.wgsl:66-80,110-122) on gfx950.
//
// scoreDirectSSB_events (ssb-dsb.ts:135-217) is sequential in two ways; both are made
// parallel without changing the result:
//  * adjacent-duplicate skipping: inside a maximal run of equal (x,y,z) entries starting
//    at s, the loop visits exactly s, s+2, s+4, ...; each entry finds its run start by
//    walking back, so visited(i) = (i - s) even;
//  * the LCG (makeSsbRng, ssb-dsb.ts:421-427) is consumed once per in-reach event in
//    index order: an exclusive scan gives each in-reach event its rank r, and the draw
//    is state_{r+1} = A^(r+1) s0 + C(r+1) mod 2^32, composed from a table of the
//    affine maps for 2^k steps.
// Geometry uses doubles on the f32 inputs with Math.round = floor(x + 0.5), as JS does.
// Hits are a bitmap; a hit counts when atomicOr finds its bit clear (the number of
// distinct bits, independent of order).  clusterDSB (ssb-dsb.ts:348-407) runs one
// workgroup per fiber.
#include "dna_internal.h"
#include "dna_ctx.h"
#include <vector>
#include <cmath>
#include <mutex>
#include <cstring>

namespace dna {

struct ScoreGeom {
    double x_half, rise_inv, grid_off, inv_spacing, r_bb, x0, rise;
    uint32_t grid_N, n_bp_per, n_bp, n_fibers;
};

__constant__ uint32_t c_lcg_a[32], c_lcg_c[32];   // affine map of 2^k LCG steps

__device__ __forceinline__ double js_round(double x) { return floor(x + 0.5); }

__device__ __forceinline__ bool same_xyz(float4 a, float4 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// per entry: 1 = candidate, 2 = in reach (then *hit = bit index)
__device__ int classify(const ScoreGeom &G, const float *fy, const float *fz, const float *rbb0,
                        const float *rbb1, float4 e, uint64_t *hit)
{
    const double r_direct = 0.29, r2d = r_direct * r_direct;
    double x = e.x, y = e.y, z = e.z;
    if (x < -G.x_half - r_direct || x > G.x_half + r_direct) return 0;
    double fi = js_round((y - G.grid_off) * G.inv_spacing);
    double fj = js_round((z - G.grid_off) * G.inv_spacing);
    if (fi < 0 || fi >= G.grid_N || fj < 0 || fj >= G.grid_N) return 0;
    int64_t fiber = (int64_t)fi * G.grid_N + (int64_t)fj;
    double y_rel = y - fy[fiber], z_rel = z - fz[fiber];
    double r2 = y_rel * y_rel + z_rel * z_rel;
    double outer = G.r_bb + r_direct;
    if (r2 > outer * outer) return 0;
    double bp_est = js_round((x + G.x_half) * G.rise_inv);
    int64_t b0 = (int64_t)(bp_est - 2 > 0 ? bp_est - 2 : 0);
    int64_t b1 = (int64_t)(bp_est + 2 < G.n_bp_per - 1 ? bp_est + 2 : G.n_bp_per - 1);
    double best = 1e300; int64_t best_bp = -1; int best_strand = -1;
    for (int64_t b = b0; b <= b1; b++) {
        double dx = x - (G.x0 + b * G.rise);
        double dy0 = y_rel - rbb0[2 * b], dz0 = z_rel - rbb0[2 * b + 1];
        double d20 = dx * dx + dy0 * dy0 + dz0 * dz0;
        if (d20 < best) { best = d20; best_bp = b; best_strand = 0; }
        double dy1 = y_rel - rbb1[2 * b], dz1 = z_rel - rbb1[2 * b + 1];
        double d21 = dx * dx + dy1 * dy1 + dz1 * dz1;
        if (d21 < best) { best = d21; best_bp = b; best_strand = 1; }
    }
    if (!(best < r2d)) return 1;
    *hit = (uint64_t)(fiber * G.n_bp_per + best_bp) + (uint64_t)best_strand * G.n_bp;
    return 2;
}

// pass 1: visited/classify; flag[i] = 1 if an in-reach event; hitidx[i] = bit; counts.
// dedupe: the direct scorer's adjacent-duplicate skipping (the indirect scorer visits all).
// Entries [i0, n) of rad; flag and hitidx are indexed from i0 (a chunk of the input), while
// the duplicate walk-back reads the whole input before i0.
__global__ void k_ssb_classify(const float4 *rad, uint64_t i0, uint64_t n, ScoreGeom G, const float *fy, const float *fz,
                               const float *rbb0, const float *rbb1, uint32_t *flag, uint64_t *hitidx,
                               unsigned long long *cnt /* [0] candidates, [1] in reach */, int dedupe)
{
    uint64_t i = i0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    unsigned long long cand = 0, reach = 0;
    if (i < n) {
        float4 e = rad[i];
        uint64_t s = i;
        if (dedupe) while (s > 0 && same_xyz(rad[s - 1], rad[s])) s--;
        uint32_t f = 0;
        if (((i - s) & 1ull) == 0) {
            uint64_t h = 0;
            int c = classify(G, fy, fz, rbb0, rbb1, e, &h);
            if (c >= 1) cand = 1;
            if (c == 2) { reach = 1; f = 1; hitidx[i - i0] = h; }
        }
        flag[i - i0] = f;
    }
    for (int o = 32; o > 0; o >>= 1) { cand += __shfl_xor(cand, o, 64); reach += __shfl_xor(reach, o, 64); }
    if ((threadIdx.x & 63) == 0) {
        if (cand) atomicAdd(&cnt[0], cand);
        if (reach) atomicAdd(&cnt[1], reach);
    }
}

__device__ __forceinline__ uint32_t lcg_jump(uint32_t s0, uint32_t m)
{
    // apply m LCG steps: compose the 2^k maps for the set bits of m
    uint32_t s = s0;
    for (int k = 0; k < 32; k++) if ((m >> k) & 1u) s = c_lcg_a[k] * s + c_lcg_c[k];
    return s;
}

// pass 2: draws for in-reach events (rank from the exclusive scan of flag, plus *carry, the
// in-reach events of earlier chunks, when carry is set); a hit counts when its bit was
// clear, in cnt[2] (direct) or per strand in cnt[2], cnt[3] (indirect: per_strand, bits at
// or above n_bp belong to strand 1)
__global__ void k_ssb_hits(const uint32_t *flag, const uint32_t *rank, const uint64_t *hitidx, uint64_t n,
                           uint32_t s0, double p, uint64_t n_bp, int per_strand, uint32_t *bits,
                           unsigned long long *cnt, const uint32_t *carry)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    unsigned long long add0 = 0, add1 = 0;
    if (i < n && flag[i]) {
        const uint32_t before = carry ? *carry : 0u;
        uint32_t s = lcg_jump(s0, before + rank[i] + 1u);
        double u = (double)s / 4294967296.0;
        if (u < p) {
            uint64_t h = hitidx[i];
            uint32_t bit = 1u << (h & 31u);
            uint32_t old = atomicOr(&bits[h >> 5], bit);
            if (!(old & bit)) { if (per_strand && h >= n_bp) add1 = 1; else add0 = 1; }
        }
    }
    for (int o = 32; o > 0; o >>= 1) { add0 += __shfl_xor(add0, o, 64); add1 += __shfl_xor(add1, o, 64); }
    if ((threadIdx.x & 63) == 0) {
        if (add0) atomicAdd(&cnt[2], add0);
        if (add1) atomicAdd(&cnt[3], add1);
    }
}

__global__ void k_or_bits(uint32_t *dst, const uint32_t *src, uint64_t nwords)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < nwords) dst[i] |= src[i];
}

__device__ __forceinline__ bool getbit(const uint32_t *bits, uint64_t i) { return (bits[i >> 5] >> (i & 31u)) & 1u; }

// clusterDSB per fiber: strand hit lists in bp order, greedy pairing within +-10 bp
#define CL_BLOCK 256
#define CL_MAXK 4096
__global__ void __launch_bounds__(CL_BLOCK) k_cluster(const uint32_t *bits, ScoreGeom G, uint32_t *fiber_k,
                                                      unsigned long long *cnt /* [3] dsb, [4] ssb0, [5] ssb1 */,
                                                      uint32_t *overflow)
{
    __shared__ uint32_t s0[CL_MAXK], s1[CL_MAXK];
    __shared__ uint32_t c0w[CL_BLOCK], c1w[CL_BLOCK];
    __shared__ uint8_t used[CL_MAXK];
    const uint32_t fi = blockIdx.x, tid = threadIdx.x, np = G.n_bp_per;
    const uint64_t base = (uint64_t)fi * np;
    uint32_t per = (np + CL_BLOCK - 1) / CL_BLOCK;
    uint32_t lo = tid * per, hi = lo + per < np ? lo + per : np;
    uint32_t a = 0, b = 0;
    for (uint32_t q = lo; q < hi; q++) { a += getbit(bits, base + q); b += getbit(bits, base + q + G.n_bp); }
    c0w[tid] = a; c1w[tid] = b;
    __syncthreads();
    if (tid == 0) {
        uint32_t r0 = 0, r1 = 0;
        for (int t = 0; t < CL_BLOCK; t++) { uint32_t x = c0w[t], y = c1w[t]; c0w[t] = r0; c1w[t] = r1; r0 += x; r1 += y; }
        fiber_k[2 * fi] = r0; fiber_k[2 * fi + 1] = r1;
    }
    __syncthreads();
    uint32_t k0 = fiber_k[2 * fi], k1 = fiber_k[2 * fi + 1];
    if (k0 > CL_MAXK || k1 > CL_MAXK) { if (tid == 0) atomicOr(overflow, 1u); return; }
    uint32_t o0 = c0w[tid], o1 = c1w[tid];
    for (uint32_t q = lo; q < hi; q++) {
        if (getbit(bits, base + q)) s0[o0++] = q;
        if (getbit(bits, base + q + G.n_bp)) s1[o1++] = q;
    }
    for (uint32_t j = tid; j < k1; j += CL_BLOCK) used[j] = 0;
    __syncthreads();
    if (tid == 0) {
        const int64_t W = 10;
        unsigned long long dsb = 0;
        uint32_t jlo = 0;
        for (uint32_t i = 0; i < k0; i++) {
            int64_t b0 = s0[i];
            while (jlo < k1 && (int64_t)s1[jlo] < b0 - W) jlo++;
            for (uint32_t j = jlo; j < k1 && (int64_t)s1[j] <= b0 + W; j++)
                if (!used[j]) { used[j] = 1; dsb++; break; }
        }
        if (dsb) atomicAdd(&cnt[3], dsb);
        if (k0) atomicAdd(&cnt[4], (unsigned long long)k0);
        if (k1) atomicAdd(&cnt[5], (unsigned long long)k1);
    }
}

__global__ void k_bits_to_u8(const uint32_t *bits, uint64_t n, uint8_t *out)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint8_t)getbit(bits, i);
}

// ------------------------------------------------------------ spatial hash
__global__ void k_hash_key(const float4 *pos, uint64_t n, float cell, uint32_t mask, uint32_t *key, uint32_t *count)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    int32_t ix = (int32_t)floorf(p.x / cell);
    int32_t iy = (int32_t)floorf(p.y / cell);
    int32_t iz = (int32_t)floorf(p.z / cell);
    uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    h &= mask;
    key[i] = h;
    atomicAdd(&count[h], 1u);
}

__global__ void k_hash_scatter(const uint32_t *key, uint64_t n, const uint32_t *start, uint32_t *cursor, uint32_t *order)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = key[i];
    order[start[k] + atomicAdd(&cursor[k], 1u)] = (uint32_t)i;
}

// buckets are tiny: each lane sorts one bucket by insertion (stable order = by index)
__global__ void k_hash_sort_buckets(const uint32_t *start, uint32_t nb, uint32_t *order)
{
    uint32_t bkt = blockIdx.x * blockDim.x + threadIdx.x;
    if (bkt >= nb) return;
    uint32_t a = start[bkt], b = start[bkt + 1];
    for (uint32_t i = a + 1; i < b; i++) {
        uint32_t v = order[i];
        uint32_t j = i;
        while (j > a && order[j - 1] > v) { order[j] = order[j - 1]; j--; }
        order[j] = v;
    }
}

template <class T> static hipError_t sgrow(void *&ptr, size_t &cap, size_t need_elems)
{
    size_t need = need_elems * sizeof(T);
    if (need <= cap && ptr) return hipSuccess;
    if (ptr) { hipError_t e = hipFree(ptr); if (e != hipSuccess) return e; ptr = nullptr; }
    size_t n = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&ptr, n);
    cap = (e == hipSuccess) ? n : 0;
    return e;
}

static bool g_lcg_ready[64] = {false};
static std::mutex g_lcg_mu;                   // contexts on several host threads

}  // namespace dna

using namespace dna;

// the affine maps of 2^k LCG steps in constant memory, once per device
static int lcg_const_init(int device)
{
    std::lock_guard<std::mutex> lk(g_lcg_mu);
    if (!g_lcg_ready[device]) {
        uint32_t A[32], Cc[32];
        uint32_t a = 1664525u, cc = 1013904223u;
        for (int k = 0; k < 32; k++) { A[k] = a; Cc[k] = cc; cc = a * cc + cc; a = a * a; }
        DNA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_lcg_a), A, sizeof(A)));
        DNA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_lcg_c), Cc, sizeof(Cc)));
        g_lcg_ready[device] = true;
    }
    return DNA_OK;
}

static ScoreGeom score_geom(const dna_dna_target *t)
{
    ScoreGeom G;
    G.x_half = (t->n_bp_per - 1) * t->rise * 0.5;
    G.rise_inv = 1 / t->rise;
    G.grid_off = -((t->grid_N - 1) * t->spacing_nm) * 0.5;
    G.inv_spacing = 1 / t->spacing_nm;
    G.r_bb = t->r_bb; G.x0 = t->x0; G.rise = t->rise;
    G.grid_N = t->grid_N; G.n_bp_per = t->n_bp_per; G.n_bp = t->n_bp; G.n_fibers = t->n_fibers;
    return G;
}

static bool target_ok(const dna_dna_target *t)
{
    return t->fy && t->fz && t->rbb0 && t->rbb1 && t->n_bp_per != 0 && t->grid_N != 0 &&
           (uint64_t)t->n_fibers * t->n_bp_per == t->n_bp && t->n_fibers == t->grid_N * t->grid_N;
}

// closed-form expected DSB (ssb-dsb.ts:398-403), summed over fibers in order, from the
// per-fiber strand hit counts fk[2f], fk[2f+1]
static double dsb_expected(const dna_dna_target *t, const uint32_t *fk)
{
    double p_pair = 21.0 / (double)t->n_bp_per;
    if (p_pair > 1) p_pair = 1;
    double ex = 0;
    for (uint32_t f = 0; f < t->n_fibers; f++) {
        uint64_t k0 = fk[2 * f], k1 = fk[2 * f + 1];
        if (k0 > 0 && k1 > 0) ex += 1 - std::pow(1 - p_pair, (double)(k0 * k1));
    }
    return ex;
}

static uint32_t lcg_advance(uint32_t s, uint64_t m)
{
    uint32_t a = 1664525u, cc = 1013904223u;
    for (int k = 0; k < 64 && m; k++) {
        if (m & 1ull) s = a * s + cc;
        cc = a * cc + cc; a = a * a; m >>= 1;
    }
    return s;
}

// Scoring with an optional indirect pass first (ind != NULL): the alive-OH list oh
// (host, or the resident list of dna_irt_oh_final when NULL) scored by scoreIndirectSSB
// (ssb-dsb.ts:43-121: no duplicate skipping, p = 0.4) into its own bitmap, the LCG then
// continuing into the direct pass, the two bitmaps OR-combined (combineHits,
// ssb-dsb.ts:413-418) before clustering (app.ts:176-186).
static int score_impl(dna_ctx *c, const dna_dna_target *t, const float *rad, uint64_t rad_n, const float *oh,
                      uint64_t n_oh, uint32_t *lcg_state, int copy_hits, dna_score_out *out, dna_indirect_out *ind)
{
    if (!c || !t || !lcg_state || !out) { set_error("dna_score: NULL argument"); return DNA_ERR_ARG; }
    if (!target_ok(t)) { set_error("dna_score: inconsistent DNA target"); return DNA_ERR_ARG; }
    if (rad_n >= 0xFFFFFFF0ull || n_oh >= 0xFFFFFFF0ull) { set_error("dna_score: too many entries"); return DNA_ERR_CAPACITY; }
    if (ind && !oh && n_oh > c->oh_n) { set_error("dna_score_oh: oh == NULL but fewer OH are resident"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    memset(out, 0, sizeof(*out));
    if (ind) memset(ind, 0, sizeof(*ind));
    int rc = lcg_const_init(c->device); if (rc) return rc;
    const ScoreGeom G = score_geom(t);
    // target arrays
    size_t ntg = 2 * (size_t)t->n_fibers + 4 * (size_t)t->n_bp_per;
    DNA_HIP(sgrow<float>(c->sc.tgt, c->sc.cap_tgt, ntg));
    float *d_fy = (float *)c->sc.tgt, *d_fz = d_fy + t->n_fibers, *d_r0 = d_fz + t->n_fibers, *d_r1 = d_r0 + 2 * (size_t)t->n_bp_per;
    DNA_HIP(hipMemcpyAsync(d_fy, t->fy, t->n_fibers * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(d_fz, t->fz, t->n_fibers * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(d_r0, t->rbb0, 2 * (size_t)t->n_bp_per * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(d_r1, t->rbb1, 2 * (size_t)t->n_bp_per * 4, hipMemcpyHostToDevice, st));
    const float4 *d_rad;
    // host inputs share one upload buffer: the radicals, then the OH list
    const uint64_t n_in = (rad ? rad_n : 0) + ((ind && oh) ? n_oh : 0);
    if (n_in) DNA_HIP(sgrow<float4>(c->sc.rad_in, c->sc.cap_rad_in, n_in + 1));
    if (rad) {
        DNA_HIP(hipMemcpyAsync(c->sc.rad_in, rad, rad_n * 16, hipMemcpyHostToDevice, st));
        d_rad = (const float4 *)c->sc.rad_in;
    } else {
        if (rad_n > c->rad_n) { set_error("dna_score: rad == NULL but fewer radicals are resident"); return DNA_ERR_ARG; }
        d_rad = c->d_rad;
    }
    uint64_t nwords = (2 * (uint64_t)t->n_bp + 31) / 32;
    const uint64_t nw2 = nwords + (nwords & 1);
    // direct bitmap | counters (8 u64) | fiber counts | indirect bitmap | indirect counters (8 u64)
    DNA_HIP(sgrow<uint32_t>(c->sc.hitbits, c->sc.cap_hitbits, 2 * nw2 + 64 + 2 * (size_t)t->n_fibers + 2));
    uint32_t *bits = (uint32_t *)c->sc.hitbits;
    unsigned long long *cnt = (unsigned long long *)(bits + nw2);
    uint32_t *fiber_k = (uint32_t *)(cnt + 8);
    uint32_t *ibits = fiber_k + ((2 * (size_t)t->n_fibers + 2 + 1) & ~(size_t)1);
    unsigned long long *icnt = (unsigned long long *)(ibits + nw2);
    const uint64_t m = rad_n > n_oh ? rad_n : n_oh;
    DNA_HIP(sgrow<uint32_t>(c->sc.flags, c->sc.cap_flags, 3 * (m + 1) + 16));
    uint32_t *flag = (uint32_t *)c->sc.flags, *rank = flag + m + 1;
    DNA_HIP(sgrow<uint64_t>(c->sc.rank, c->sc.cap_rank, m + 1 + m / 1024 + 16));
    uint64_t *hitidx = (uint64_t *)c->sc.rank;
    uint32_t *scan_tmp = (uint32_t *)(hitidx + m + 1);
    DNA_HIP(hipEventRecord(c->ev[5], st));
    DNA_HIP(hipMemsetAsync(bits, 0, nw2 * 4 + 8 * 8, st));
    const uint32_t B = 256;
    uint32_t s_dir = *lcg_state;
    uint64_t ind_reach = 0;
    if (ind) {
        DNA_HIP(hipMemsetAsync(ibits, 0, nw2 * 4 + 8 * 8, st));
        const float4 *d_oh = c->d_oh;
        if (oh && n_oh) {
            d_oh = (const float4 *)c->sc.rad_in + (rad ? rad_n : 0);
            DNA_HIP(hipMemcpyAsync((void *)d_oh, oh, n_oh * 16, hipMemcpyHostToDevice, st));
        }
        if (n_oh) {
            uint32_t gi = (uint32_t)((n_oh + B - 1) / B);
            hipLaunchKernelGGL(k_ssb_classify, dim3(gi), dim3(B), 0, st, d_oh, (uint64_t)0, n_oh, G, d_fy, d_fz, d_r0, d_r1,
                               flag, hitidx, icnt, 0);
            DNA_HIP(hipMemcpyAsync(rank, flag, n_oh * 4, hipMemcpyDeviceToDevice, st));
            DNA_HIP(exclusive_scan_u32(rank, (uint32_t)n_oh, scan_tmp, st));
            hipLaunchKernelGGL(k_ssb_hits, dim3(gi), dim3(B), 0, st, flag, rank, hitidx, n_oh, *lcg_state, 0.4,
                               (uint64_t)t->n_bp, 1, ibits, icnt, (const uint32_t *)nullptr);
        }
        unsigned long long hi[4] = {0, 0, 0, 0};
        DNA_HIP(hipMemcpyAsync(hi, icnt, sizeof(hi), hipMemcpyDeviceToHost, st));
        DNA_HIP(hipStreamSynchronize(st));
        ind->candidates = hi[0]; ind->in_reach = hi[1]; ind->ssb0 = hi[2]; ind->ssb1 = hi[3];
        ind_reach = hi[1];
        s_dir = lcg_advance(*lcg_state, ind_reach);     // the direct pass continues the sequence
    }
    uint32_t gb = (uint32_t)((rad_n + B - 1) / B);
    if (rad_n) {
        hipLaunchKernelGGL(k_ssb_classify, dim3(gb), dim3(B), 0, st, d_rad, (uint64_t)0, rad_n, G, d_fy, d_fz, d_r0, d_r1,
                           flag, hitidx, cnt, 1);
        DNA_HIP(hipMemcpyAsync(rank, flag, rad_n * 4, hipMemcpyDeviceToDevice, st));
        DNA_HIP(exclusive_scan_u32(rank, (uint32_t)rad_n, scan_tmp, st));
        hipLaunchKernelGGL(k_ssb_hits, dim3(gb), dim3(B), 0, st, flag, rank, hitidx, rad_n, s_dir, 0.15,
                           (uint64_t)t->n_bp, 0, bits, cnt, (const uint32_t *)nullptr);
    }
    if (ind) {
        hipLaunchKernelGGL(k_or_bits, dim3((uint32_t)((nwords + B - 1) / B)), dim3(B), 0, st, bits, ibits, nwords);
        DNA_HIP(hipGetLastError());
    }
    uint32_t *ovf = fiber_k + 2 * t->n_fibers;
    DNA_HIP(hipMemsetAsync(ovf, 0, 4, st));
    hipLaunchKernelGGL(k_cluster, dim3(t->n_fibers), dim3(CL_BLOCK), 0, st, bits, G, fiber_k, cnt, ovf);
    DNA_HIP(hipGetLastError());
    unsigned long long h_cnt[8];
    std::vector<uint32_t> fk(2 * (size_t)t->n_fibers + 1);
    DNA_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof(h_cnt), hipMemcpyDeviceToHost, st));
    DNA_HIP(hipMemcpyAsync(fk.data(), fiber_k, (2 * (size_t)t->n_fibers + 1) * 4, hipMemcpyDeviceToHost, st));
    out->hits = nullptr;
    if (copy_hits) {
        uint64_t nh = 2 * (uint64_t)t->n_bp;
        DNA_HIP(sgrow<uint8_t>(c->sc.hits_u8, c->sc.cap_hits_u8, nh));
        if (nh > c->sc.hcap_hits) {
            if (c->sc.h_hits) (void)hipHostFree(c->sc.h_hits);
            DNA_HIP(hipHostMalloc((void **)&c->sc.h_hits, nh, hipHostMallocDefault));
            c->sc.hcap_hits = nh;
        }
        hipLaunchKernelGGL(k_bits_to_u8, dim3((uint32_t)((nh + B - 1) / B)), dim3(B), 0, st, bits, nh, (uint8_t *)c->sc.hits_u8);
        DNA_HIP(hipMemcpyAsync(c->sc.h_hits, c->sc.hits_u8, nh, hipMemcpyDeviceToHost, st));
        out->hits = c->sc.h_hits;
    }
    DNA_HIP(hipEventRecord(c->ev[6], st));
    DNA_HIP(hipStreamSynchronize(st));
    if (fk[2 * t->n_fibers]) { set_error("dna_score: more than 4096 hits on one strand of one fiber"); return DNA_ERR_CAPACITY; }
    out->candidates = h_cnt[0]; out->in_reach = h_cnt[1]; out->ssb_count = h_cnt[2];
    out->dsb = h_cnt[3]; out->ssb0 = h_cnt[4]; out->ssb1 = h_cnt[5];
    out->dsb_expected = dsb_expected(t, fk.data());
    // advance the caller's LCG by the draws consumed (indirect, then direct)
    *lcg_state = lcg_advance(s_dir, h_cnt[1]);
    float ms = 0; (void)hipEventElapsedTime(&ms, c->ev[5], c->ev[6]);
    out->t_ms = ms;
    return DNA_OK;
}

extern "C" int dna_score(dna_ctx *c, const dna_dna_target *t, const float *rad, uint64_t rad_n,
                         uint32_t *lcg_state, int copy_hits, dna_score_out *out)
{
    return score_impl(c, t, rad, rad_n, nullptr, 0, lcg_state, copy_hits, out, nullptr);
}

extern "C" int dna_score_oh(dna_ctx *c, const dna_dna_target *t, const float *rad, uint64_t rad_n, const float *oh,
                            uint64_t n_oh, uint32_t *lcg_state, int copy_hits, dna_score_out *out,
                            dna_indirect_out *ind)
{
    if (!ind) { set_error("dna_score_oh: NULL argument"); return DNA_ERR_ARG; }
    return score_impl(c, t, rad, rad_n, oh, n_oh, lcg_state, copy_hits, out, ind);
}

extern "C" int dna_hash_cells(dna_ctx *c, const float *pos4, uint64_t n, float cell_nm, uint32_t bits,
                              uint32_t *key_out, uint32_t *cell_start_out, uint32_t *order_out)
{
    if (!c || (!pos4 && n) || !key_out || !cell_start_out || !order_out) { set_error("dna_hash_cells: NULL argument"); return DNA_ERR_ARG; }
    if (bits < 1 || bits > 26 || !(cell_nm > 0.0f)) { set_error("dna_hash_cells: bits must be 1..26 and cell_nm > 0"); return DNA_ERR_ARG; }
    if (n >= 0xFFFFFFF0ull) { set_error("dna_hash_cells: too many points"); return DNA_ERR_CAPACITY; }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    uint32_t nb = 1u << bits;
    size_t words = 4 * (size_t)n + 2 * ((size_t)nb + 1) + nb / 1024 + 64;   // pos(4n) key, start, cursor, order...
    DNA_HIP(sgrow<uint32_t>(c->hs.buf, c->hs.cap, words + 2 * (size_t)n));
    float4 *d_pos = (float4 *)c->hs.buf;
    uint32_t *d_key = (uint32_t *)(d_pos + n), *d_start = d_key + n, *d_cur = d_start + nb + 1, *d_ord = d_cur + nb + 1;
    uint32_t *d_tmp = d_ord + n;
    DNA_HIP(hipMemcpyAsync(d_pos, pos4, n * 16, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemsetAsync(d_start, 0, 2 * ((size_t)nb + 1) * 4, st));
    const uint32_t B = 256;
    if (n) hipLaunchKernelGGL(k_hash_key, dim3((uint32_t)((n + B - 1) / B)), dim3(B), 0, st, d_pos, n, cell_nm, nb - 1, d_key, d_start);
    DNA_HIP(exclusive_scan_u32(d_start, nb, d_tmp, st));
    if (n) hipLaunchKernelGGL(k_hash_scatter, dim3((uint32_t)((n + B - 1) / B)), dim3(B), 0, st, d_key, n, d_start, d_cur, d_ord);
    hipLaunchKernelGGL(k_hash_sort_buckets, dim3((nb + B - 1) / B), dim3(B), 0, st, d_start, nb, d_ord);
    DNA_HIP(hipGetLastError());
    DNA_HIP(hipMemcpyAsync(key_out, d_key, n * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipMemcpyAsync(cell_start_out, d_start, ((size_t)nb + 1) * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipMemcpyAsync(order_out, d_ord, n * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    return DNA_OK;
}

// ------------------------------------------------------------ scoring without allocation
// dna_score_noalloc keeps every scratch buffer of the direct scoring in one fixed array per
// GPU: the target copy, the hit bitmap, counters, and per-radical flags, ranks and hit
// indices for one chunk of DNA_SCORE_NOALLOC_CHUNK radicals.  The input is scored chunk by
// chunk: the duplicate walk-back reads the input itself, and the in-reach events of earlier
// chunks (a device-side carry) offset each chunk's LCG draws, so the result equals
// dna_score's.  Host code uses the array only through the address looked up once per
// device (score_scratch_init, from dna_ctx_create), never through its name.
#define NA_TGT_BYTES (((2u * DNA_SCORE_NOALLOC_MAX_FIBERS + 4u * DNA_SCORE_NOALLOC_MAX_BP_PER) * 4u + 255u) & ~255u)
#define NA_BITS_BYTES ((2ull * DNA_SCORE_NOALLOC_MAX_BP + 63) / 64 * 8)
#define NA_MISC_BYTES 16384u            // counters (8 u64) at 0, carry at 64, fiber counts at 128
#define NA_FLAG_BYTES ((((size_t)DNA_SCORE_NOALLOC_CHUNK + 1) * 4 + 255) & ~(size_t)255)
#define NA_TMP_BYTES ((((size_t)DNA_SCORE_NOALLOC_CHUNK / 1024 + 16) * 4 + 255) & ~(size_t)255)
#define NA_HIT_BYTES ((((size_t)DNA_SCORE_NOALLOC_CHUNK + 1) * 8 + 255) & ~(size_t)255)
#define NA_SCRATCH_BYTES (NA_TGT_BYTES + NA_BITS_BYTES + NA_MISC_BYTES + 2 * NA_FLAG_BYTES + NA_TMP_BYTES + NA_HIT_BYTES)
static_assert(128 + 4 * (2 * DNA_SCORE_NOALLOC_MAX_FIBERS + 1) <= NA_MISC_BYTES, "fiber counts fit");

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

namespace dna {
static std::mutex g_na_mu;
static void *g_na_ptr[64];                    // the array's address on each device

// Look up the array's device address on `device` (current), once per device.
int score_scratch_init(int device, void **ptr)
{
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("scoring scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return DNA_ERR_HIP;
        }
        g_na_ptr[device] = p;
    }
    *ptr = g_na_ptr[device];
    return DNA_OK;
}
}  // namespace dna

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_carry_add(uint32_t *carry, const uint32_t *chunk_total) { *carry += *chunk_total; }

extern "C" int dna_score_noalloc(dna_ctx *c, const dna_dna_target *t, const float *rad_dev, uint64_t rad_n,
                                 uint32_t *lcg_state, dna_score_out *out)
{
    if (!c || !t || !lcg_state || !out) { set_error("dna_score_noalloc: NULL argument"); return DNA_ERR_ARG; }
    if (!target_ok(t)) { set_error("dna_score_noalloc: inconsistent DNA target"); return DNA_ERR_ARG; }
    if (t->n_fibers > DNA_SCORE_NOALLOC_MAX_FIBERS || t->n_bp_per > DNA_SCORE_NOALLOC_MAX_BP_PER ||
        t->n_bp > DNA_SCORE_NOALLOC_MAX_BP) {
        set_error("dna_score_noalloc: the target exceeds the fixed scratch (at most 1024 fibers, 16384 bp per "
                  "fiber, 2^22 bp in all)");
        return DNA_ERR_CAPACITY;
    }
    if (rad_n >= 0xFFFFFFF0ull) { set_error("dna_score_noalloc: too many entries"); return DNA_ERR_CAPACITY; }
    if (!rad_dev && rad_n > c->rad_n) {
        set_error("dna_score_noalloc: rad_dev == NULL but fewer radicals are resident"); return DNA_ERR_ARG;
    }
    if (!c->na_scratch) { set_error("dna_score_noalloc: no scoring scratch on this device"); return DNA_ERR_HIP; }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    memset(out, 0, sizeof(*out));
    int rc = lcg_const_init(c->device); if (rc) return rc;
    const ScoreGeom G = score_geom(t);
    // carve the array (at its device address)
    unsigned char *b = (unsigned char *)c->na_scratch;
    float *d_fy = (float *)b, *d_fz = d_fy + t->n_fibers, *d_r0 = d_fz + t->n_fibers, *d_r1 = d_r0 + 2 * (size_t)t->n_bp_per;
    uint32_t *bits = (uint32_t *)(b + NA_TGT_BYTES);
    unsigned char *misc = b + NA_TGT_BYTES + NA_BITS_BYTES;
    unsigned long long *cnt = (unsigned long long *)misc;
    uint32_t *carry = (uint32_t *)(misc + 64), *fiber_k = (uint32_t *)(misc + 128);
    uint32_t *flag = (uint32_t *)(misc + NA_MISC_BYTES), *rank = (uint32_t *)((unsigned char *)flag + NA_FLAG_BYTES);
    uint32_t *scan_tmp = (uint32_t *)((unsigned char *)rank + NA_FLAG_BYTES);
    uint64_t *hitidx = (uint64_t *)((unsigned char *)scan_tmp + NA_TMP_BYTES);
    const float4 *d_rad = rad_dev ? (const float4 *)rad_dev : c->d_rad;
    DNA_HIP(hipEventRecord(c->ev[5], st));
    DNA_HIP(hipMemcpyAsync(d_fy, t->fy, t->n_fibers * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(d_fz, t->fz, t->n_fibers * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(d_r0, t->rbb0, 2 * (size_t)t->n_bp_per * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(d_r1, t->rbb1, 2 * (size_t)t->n_bp_per * 4, hipMemcpyHostToDevice, st));
    // the array keeps the last call's contents: clear the bitmap, counters, carry and fiber
    // counts with a fill kernel
    const uint64_t nwords = (2 * (uint64_t)t->n_bp + 31) / 32;
    hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, st, bits, (size_t)nwords, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(8), dim3(256), 0, st, (uint32_t *)misc, (size_t)NA_MISC_BYTES / 4, 0u);
    DNA_HIP(hipGetLastError());
    const uint32_t B = 256, CH = DNA_SCORE_NOALLOC_CHUNK;
    const uint32_t s0 = *lcg_state;
    for (uint64_t off = 0; off < rad_n; off += CH) {
        const uint32_t nc = (uint32_t)(rad_n - off < CH ? rad_n - off : CH);
        const uint32_t gb = (nc + B - 1) / B;
        hipLaunchKernelGGL(k_ssb_classify, dim3(gb), dim3(B), 0, st, d_rad, off, off + nc, G, d_fy, d_fz, d_r0, d_r1,
                           flag, hitidx, cnt, 1);
        DNA_HIP(hipMemcpyAsync(rank, flag, (size_t)nc * 4, hipMemcpyDeviceToDevice, st));
        DNA_HIP(exclusive_scan_u32(rank, nc, scan_tmp, st));
        hipLaunchKernelGGL(k_ssb_hits, dim3(gb), dim3(B), 0, st, flag, rank, hitidx, (uint64_t)nc, s0, 0.15,
                           (uint64_t)t->n_bp, 0, bits, cnt, (const uint32_t *)carry);
        hipLaunchKernelGGL(k_carry_add, dim3(1), dim3(1), 0, st, carry, rank + nc);
        DNA_HIP(hipGetLastError());
    }
    uint32_t *ovf = fiber_k + 2 * t->n_fibers;
    hipLaunchKernelGGL(k_cluster, dim3(t->n_fibers), dim3(CL_BLOCK), 0, st, bits, G, fiber_k, cnt, ovf);
    DNA_HIP(hipGetLastError());
    unsigned long long h_cnt[8];
    uint32_t fk[2 * DNA_SCORE_NOALLOC_MAX_FIBERS + 1];
    DNA_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof(h_cnt), hipMemcpyDeviceToHost, st));
    DNA_HIP(hipMemcpyAsync(fk, fiber_k, (2 * (size_t)t->n_fibers + 1) * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipEventRecord(c->ev[6], st));
    DNA_HIP(hipStreamSynchronize(st));
    if (fk[2 * t->n_fibers]) { set_error("dna_score_noalloc: more than 4096 hits on one strand of one fiber"); return DNA_ERR_CAPACITY; }
    out->candidates = h_cnt[0]; out->in_reach = h_cnt[1]; out->ssb_count = h_cnt[2];
    out->dsb = h_cnt[3]; out->ssb0 = h_cnt[4]; out->ssb1 = h_cnt[5];
    out->dsb_expected = dsb_expected(t, fk);
    out->hits = nullptr;
    *lcg_state = lcg_advance(s0, h_cnt[1]);
    float ms = 0; (void)hipEventElapsedTime(&ms, c->ev[5], c->ev[6]);
    out->t_ms = ms;
    return DNA_OK;
}
