// transport.hip — Phase A (primaries) and Phase B (secondaries) of the Geant4-DNA
// electron transport with in-kernel prechemistry, for gfx950.
//
// Physics follows src/shaders/primary.wgsl:114-663 and secondary.wgsl:83-450 (helpers
// helpers.wgsl:15-190) expression for expression in FP32, with the elementary functions
// of include/dna_fmath.h, so each history is bit-identical to the CPU restatement.
//
// Execution model (differs from the reference's one-dispatch-per-step wavefront):
//  * both phases are persistent, flattened loops: every lane performs one physics step
//    per iteration and, when its particle ends, immediately takes the next one from a
//    global work counter, so lanes never idle behind a long history;
//  * Phase B applies up to max_sec_steps step calls per secondary (the reference's
//    MAX_SEC_STEPS dispatches, src/gpu/dispatch.ts:137) inside that loop;
//  * the small tables (2.5k floats) are staged in LDS per workgroup; the Born CDF rows
//    (580 KB) stay in L2 and are searched by bisection (rows are non-decreasing, so
//    the result equals the reference's linear scan, helpers.wgsl:101-105);
//  * radicals are appended unordered with a 64-bit key (primary, emitter, sequence) and
//    put in canonical order afterwards by a counting scatter (canonical.hip);
//  * tallies are per-lane registers reduced per wavefront; dose is a u32 fixed-point
//    grid (100 units/eV), summed with device atomics, exact and order-independent.
#include "dna_internal.h"
#include "../../include/dna_fmath.h"

namespace dna {

#define PI_F     3.14159265f
#define TWO_PI_F 6.2831853f
#define NW_F     33.4f

__constant__ float c_SB[5] = {10.79f, 13.39f, 16.05f, 32.30f, 539.0f};
__constant__ float c_EX[5] = {8.22f, 10.00f, 11.24f, 12.61f, 13.77f};
__constant__ float c_DEA_E[14] = {0.0f, 4.2f, 4.3f, 5.8f, 6.0f, 6.2f, 6.4f, 7.0f, 7.5f, 7.7f,
                                  8.4f, 10.2f, 11.2f, 12.3f};
__constant__ float c_DEA_S[14] = {0.0f, 0.0f, 0.0f, 2.5e-5f, 2.8e-4f, 6.0e-4f, 6.7e-4f, 3.7e-4f,
                                  1.09e-4f, 7.0e-5f, 1.4e-4f, 2.0e-5f, 6.0e-5f, 0.0f};

// Small tables staged in LDS (everything except the Born CDF/transfer arrays).
struct SmallTab {
    float XE[DNA_XN], XI[DNA_XN], XC[DNA_XN], XL[DNA_XN];
    float XSF[5][DNA_XN], XEF[5][DNA_XN];
    float XAE[DNA_N_AE], XAC[DNA_N_AE * DNA_N_AC];
    float XWE[DNA_N_WE];
    float XVE[DNA_N_VIB], XVS[DNA_N_VIB], XVMF[DNA_N_VIB * 9], VIB_LEV[9];
    float LOG_XE0, INV_LOG_XE_STEP, LOG_XAE0, INV_LOG_XAE_STEP;
    float DEA_E[14], DEA_S[14];
    uint8_t GUIDE[DNA_GUIDE_PAD];
};

__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float clampf_(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }

__device__ void stage_tables(SmallTab &S, const dna_tables *T, const uint8_t *guide)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    auto cp = [&](float *dst, const float *src, int n) { for (int i = tid; i < n; i += nt) dst[i] = src[i]; };
    cp(S.XE, T->XE, DNA_XN); cp(S.XI, T->XI, DNA_XN); cp(S.XC, T->XC, DNA_XN); cp(S.XL, T->XL, DNA_XN);
    cp(&S.XSF[0][0], &T->XSF[0][0], 5 * DNA_XN); cp(&S.XEF[0][0], &T->XEF[0][0], 5 * DNA_XN);
    cp(S.XAE, T->XAE, DNA_N_AE); cp(S.XAC, T->XAC, DNA_N_AE * DNA_N_AC); cp(S.XWE, T->XWE, DNA_N_WE);
    cp(S.XVE, T->XVE, DNA_N_VIB); cp(S.XVS, T->XVS, DNA_N_VIB); cp(S.XVMF, T->XVMF, DNA_N_VIB * 9);
    cp(S.VIB_LEV, T->VIB_LEV, 9);
    cp(S.DEA_E, c_DEA_E, 14); cp(S.DEA_S, c_DEA_S, 14);
    for (int i = tid; i < DNA_GUIDE_PAD / 4; i += nt) ((uint32_t *)S.GUIDE)[i] = ((const uint32_t *)guide)[i];
    if (tid == 0) {
        S.LOG_XE0 = T->LOG_XE0; S.INV_LOG_XE_STEP = T->INV_LOG_XE_STEP;
        S.LOG_XAE0 = T->LOG_XAE0; S.INV_LOG_XAE_STEP = T->INV_LOG_XAE_STEP;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- lookups
// Index of the log-energy grid (helpers.wgsl:18-20); shared by xs_all, shell and level fractions.
__device__ __forceinline__ void xs_index(const SmallTab &S, float E, uint32_t &i, float &f)
{
    float t = (dna_logf(E) - S.LOG_XE0) * S.INV_LOG_XE_STEP;
    i = (uint32_t)clampf_(floorf(t), 0.0f, (float)(DNA_XN - 2));
    f = t - (float)i;
}

__device__ __forceinline__ void xs_all(const SmallTab &S, float E, float &si, float &sc, float &sl)
{
    if (E <= S.XE[0]) { si = 0.0f; sc = 0.0f; sl = S.XL[0]; return; }
    if (E >= S.XE[DNA_XN - 1]) { si = S.XI[DNA_XN-1]; sc = S.XC[DNA_XN-1]; sl = S.XL[DNA_XN-1]; return; }
    uint32_t i; float f; xs_index(S, E, i, f);
    uint32_t j = i + 1;
    si = fmx(0.0f, S.XI[i] + (S.XI[j] - S.XI[i]) * f);
    sc = fmx(0.0f, S.XC[i] + (S.XC[j] - S.XC[i]) * f);
    sl = fmx(0.0f, S.XL[i] + (S.XL[j] - S.XL[i]) * f);
}

__device__ __forceinline__ void xs_shell_fracs(const SmallTab &S, float E, float sf[5])
{
    if (E <= S.XE[0]) { for (int k = 0; k < 5; k++) sf[k] = S.XSF[k][0]; return; }
    if (E >= S.XE[DNA_XN - 1]) { for (int k = 0; k < 5; k++) sf[k] = S.XSF[k][DNA_XN-1]; return; }
    uint32_t i; float f; xs_index(S, E, i, f);
    for (int k = 0; k < 5; k++) sf[k] = fmx(0.0f, S.XSF[k][i] + (S.XSF[k][i+1] - S.XSF[k][i]) * f);
}

__device__ __forceinline__ void xs_exc_fracs(const SmallTab &S, float E, float ef[5])
{
    if (E <= S.XE[0]) { for (int k = 0; k < 5; k++) ef[k] = S.XEF[k][0]; return; }
    if (E >= S.XE[DNA_XN - 1]) { for (int k = 0; k < 5; k++) ef[k] = S.XEF[k][85]; return; }
    uint32_t i; float t0; xs_index(S, E, i, t0);      // t0 = t - f32(i) before the clamp
    if (i > 85u) i = 85u;
    float f = (i == 85u) ? 0.0f : t0;
    uint32_t j = (i + 1u < 85u) ? i + 1u : 85u;
    for (int k = 0; k < 5; k++) ef[k] = fmx(0.0f, S.XEF[k][i] + (S.XEF[k][j] - S.XEF[k][i]) * f);
}

__device__ __forceinline__ float xs_el_cos(const SmallTab &S, float E, float r)
{
    float rr = clampf_(r, 0.0f, 1.0f);
    float jf = rr * (float)(DNA_N_AC - 1);
    uint32_t jlo = (uint32_t)floorf(jf);
    uint32_t jhi = (jlo + 1u < DNA_N_AC - 1u) ? jlo + 1u : DNA_N_AC - 1u;
    float fj = jf - (float)jlo;
    float tE = (dna_logf(fmx(E, S.XAE[0])) - S.LOG_XAE0) * S.INV_LOG_XAE_STEP;
    uint32_t iE = (uint32_t)clampf_(floorf(tE), 0.0f, (float)(DNA_N_AE - 2));
    float fE = clampf_(tE - (float)iE, 0.0f, 1.0f);
    uint32_t b0 = iE * DNA_N_AC, b1 = (iE + 1u) * DNA_N_AC;
    float c00 = S.XAC[b0 + jlo], c01 = S.XAC[b0 + jhi];
    float c10 = S.XAC[b1 + jlo], c11 = S.XAC[b1 + jhi];
    float cE0 = c00 + (c01 - c00) * fj;
    float cE1 = c10 + (c11 - c10) * fj;
    return clampf_(cE0 + (cE1 - cE0) * fE, -1.0f, 1.0f);
}

// sample_W_sec (helpers.wgsl:88-117): nearest Born energy row, then the last CDF entry
// <= r among the first 249 (bisection on the non-decreasing row), then interpolation.
__device__ __forceinline__ float sample_W_sec(const SmallTab &S, const dna_tables *T, float E,
                                              uint32_t shell, float r)
{
    uint32_t lo = 0, hi = DNA_N_WE - 1;
    while (hi - lo > 1u) { uint32_t mid = (lo + hi) / 2u; if (S.XWE[mid] <= E) lo = mid; else hi = mid; }
    uint32_t iE = (fabsf(S.XWE[hi] - E) < fabsf(S.XWE[lo] - E)) ? hi : lo;
    const float *C = T->XWC[shell] + iE * DNA_N_WR;
    const float *W = T->XWT[shell] + iE * DNA_N_WR;
    // count = #{ j < 249 : C[j] <= r }  (a prefix, the row being non-decreasing), bisected
    // from the guide-table bracket of r (r * 64 is exact, so k = floor(64 r))
    const uint8_t *g = S.GUIDE + (shell * DNA_N_WE + iE) * (DNA_GUIDE_K + 1);
    uint32_t kq = (uint32_t)(r * (float)DNA_GUIDE_K);
    kq = kq < DNA_GUIDE_K - 1u ? kq : DNA_GUIDE_K - 1u;
    uint32_t a = g[kq], b = g[kq + 1];          // invariant: C[j] <= r for j < a; C[j] > r for j >= b
    while (a < b) { uint32_t m = (a + b) >> 1; if (C[m] <= r) a = m + 1; else b = m; }
    uint32_t jlo = a == 0 ? 0u : a - 1u;
    uint32_t jhi = (jlo + 1u < DNA_N_WR - 1u) ? jlo + 1u : DNA_N_WR - 1u;
    float c0 = C[jlo], c1 = C[jhi], e0 = W[jlo], e1 = W[jhi];
    if (c1 <= c0) return e0;
    float f = clampf_((r - c0) / (c1 - c0), 0.0f, 1.0f);
    return e0 + (e1 - e0) * f;
}

// The vibrational grid XVE is strictly increasing (checked when the tables are built), so
// the reference's linear scans for the first i with XVE[i] <= E < XVE[i+1]
// (helpers.wgsl:135-170) find the bin a bisection finds: the largest i with XVE[i] <= E,
// when it is below the last point.  Returns DNA_N_VIB - 1 when there is no such bin.
__device__ __forceinline__ uint32_t vib_bin(const SmallTab &S, float E)
{
    if (!(E >= S.XVE[0])) return DNA_N_VIB - 1u;
    uint32_t a = 0, b = DNA_N_VIB;                  // XVE[a] <= E < XVE[b] (XVE[N] = +inf)
    while (b - a > 1u) { uint32_t m = (a + b) >> 1; if (S.XVE[m] <= E) a = m; else b = m; }
    return a;
}

__device__ __forceinline__ float xs_vib_total(const SmallTab &S, float E)
{
    if (E < 2.0f || E > 100.0f) return 0.0f;
    const uint32_t i = vib_bin(S, E);
    if (i >= DNA_N_VIB - 1u) return 0.0f;
    float f = (E - S.XVE[i]) / (S.XVE[i+1] - S.XVE[i]);
    return fmx(0.0f, S.XVS[i] + (S.XVS[i+1] - S.XVS[i]) * f);
}

__device__ __forceinline__ uint32_t sample_vib_mode(const SmallTab &S, float E, float r)
{
    uint32_t bin = vib_bin(S, E);
    if (bin >= DNA_N_VIB - 1u) bin = 0;
    uint32_t base = bin * 9u;
    float cs[9], cum = 0.0f;
    for (uint32_t k = 0; k < 9; k++) {
        float frac = (S.VIB_LEV[k] > E) ? 0.0f : S.XVMF[base + k];
        cum = cum + frac;
        cs[k] = cum;
    }
    if (cum <= 0.0f) return 0;
    float thr = r * cum;
    for (uint32_t k = 0; k < 9; k++) if (thr < cs[k]) return k;
    return 8;
}

// Melton DEA (helpers.wgsl:172-186).  The 14 energies are strictly increasing, so the
// first i with E_i <= E < E_{i+1} is the number of E_1..E_13 at or below E; the pair
// (E_i, S_i) is then read from LDS.
__device__ __forceinline__ float xs_dea(const SmallTab &S, float E)
{
    if (E < 4.0f || E > 13.0f) return 0.0f;
    int i = 0;
#pragma unroll
    for (int j = 1; j < 14; j++) i += (E >= S.DEA_E[j]) ? 1 : 0;
    if (i >= 13) return 0.0f;
    float f = (E - S.DEA_E[i]) / (S.DEA_E[i+1] - S.DEA_E[i]);
    return fmx(0.0f, S.DEA_S[i] + (S.DEA_S[i+1] - S.DEA_S[i]) * f);
}

// ---------------------------------------------------------------- RNG
struct Rng {
    uint32_t s0, s1, s2, s3;
    __device__ __forceinline__ uint32_t next() {
        uint32_t x = s0 + s3;
        uint32_t r = ((x << 7) | (x >> 25)) + s0;
        uint32_t t = s1 << 9;
        s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t; s3 = (s3 << 11) | (s3 >> 21);
        return r;
    }
    __device__ __forceinline__ float uf() { return (float)(next() >> 1u) / 2147483648.0f; }
};

__device__ __forceinline__ float meesungnoen_sigma(float k)
{
    if (k <= 0.1f) return 0.0f;
    float r = -4.06217193e-08f;
    r = r*k + 3.06848412e-06f; r = r*k - 9.93217814e-05f; r = r*k + 1.80172797e-03f;
    r = r*k - 2.01135480e-02f; r = r*k + 1.42939448e-01f; r = r*k - 6.48348714e-01f;
    r = r*k + 1.85227848e+00f; r = r*k - 3.36450378e+00f; r = r*k + 4.37785068e+00f;
    r = r*k - 4.20557339e+00f; r = r*k + 3.81679083e+00f; r = r*k - 2.34069784e-01f;
    return fmx(r, 0.0f) * 0.6267f;
}

// ---------------------------------------------------------------- per-lane sink
struct Sink {
    float4   *rad;
    uint64_t *key;
    uint32_t *rad_ctr;
    uint32_t  rad_cap;
    uint64_t  keybase;     // pid << 40 | k << 20
    uint32_t  seq;
    uint32_t  cnt[6];      // OH, e-aq, H, H3O+, DNA, H2
    uint32_t *dose;
    float     box, vs;
    int       vc;
    uint32_t  dvi, dacc;   // dose voxel being accumulated and its pending fixed-point sum
};

// Per-wavefront allocation state in LDS (4 wavefronts per workgroup at most): radical
// slots {base, fill, chunk} and work items {next, end, chunk}.  A wavefront reserves radical
// slots in chunks of d.rad_chunk (>= 192, the most one call emits) with one atomic and hands
// them out to its lanes, so emitting lanes almost never wait on a device-wide atomic; slots
// left unused in a chunk keep the invalid key written before the launch and are skipped by
// the canonical scatter.  Work items come in chunks of d.work_chunk (>= 64).  Both sizes
// are launch options (autotuned by the host); results do not depend on them.
// Different lanes lead successive calls, so every access goes through a volatile view:
// no value may be kept in a register from one call to the next.  The view is typed as LDS:
// through a generic volatile pointer the accesses compile to flat instructions, and each of
// those waits for every global load and store of the wavefront still in flight.
__shared__ uint32_t s_wrad_[12], s_wwork_[12];
typedef __attribute__((address_space(3))) volatile uint32_t lds_vu32;
#define s_wrad ((lds_vu32 *)s_wrad_)
#define s_wwork ((lds_vu32 *)s_wwork_)

__device__ __forceinline__ void wave_state_init(uint32_t rad_chunk, uint32_t work_chunk)
{
    if ((threadIdx.x & 63u) == 0) {
        const uint32_t w = threadIdx.x >> 6;
        s_wrad[3 * w] = 0u; s_wrad[3 * w + 1] = rad_chunk; s_wrad[3 * w + 2] = rad_chunk;   // empty: the first use reserves
        s_wwork[3 * w] = 0u; s_wwork[3 * w + 1] = 0u; s_wwork[3 * w + 2] = work_chunk;
    }
}

// Append n (1..3) radicals.  Slots of the lanes emitting together follow the lower-numbered
// emitting lanes' (ranks from ballots per n, so lanes reaching here with different n are
// handled too); the canonical scatter reorders slots anyway.
__device__ __forceinline__ void emit_n(Sink &k, int n, const float4 *e)
{
    const uint64_t m1 = __ballot(n == 1), m2 = __ballot(n == 2), m3 = __ballot(n == 3);
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t all = m1 | m2 | m3;
    const int leader = __ffsll((unsigned long long)all) - 1;
    const uint32_t rank = __popcll(m1 & below) + 2u * __popcll(m2 & below) + 3u * __popcll(m3 & below);
    uint32_t base = 0;
    if ((int)lane == leader) {
        const uint32_t w = threadIdx.x >> 6;
        const uint32_t tot = (uint32_t)(__popcll(m1) + 2 * __popcll(m2) + 3 * __popcll(m3));   // <= 192
        const uint32_t ch = s_wrad[3 * w + 2];
        uint32_t f = s_wrad[3 * w + 1];
        if (f + tot > ch) {
            const uint32_t b = atomicAdd(k.rad_ctr, ch);
            // the next claim would wrap the 32-bit counter: flag it for the host (which reports
            // DNA_ERR_CAPACITY); slots at or beyond rad_cap are never written
            if (b >= 0xFFFFFFFFu - ch) atomicOr(k.rad_ctr + (C_WRAP - C_RAD), 1u);
            s_wrad[3 * w] = b; f = 0u;
        }
        base = s_wrad[3 * w] + f;
        s_wrad[3 * w + 1] = f + tot;
    }
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader) + rank;   // the leader's value
    for (int i = 0; i < n; i++) {
        uint32_t idx = base + (uint32_t)i;
        if (idx < k.rad_cap) { k.rad[idx] = e[i]; k.key[idx] = k.keybase | (uint64_t)(k.seq + i); }
    }
    k.seq += (uint32_t)n;
}

// Next work item for each calling lane from the wavefront's chunk of work items (one
// atomic per chunk).  Items are handed out in increasing order, so a lane that
// receives an item past the end knows that no earlier item of its wavefront is left.
__device__ __forceinline__ uint32_t wave_claim_chunked(uint32_t *ctr)
{
    const uint64_t m = __ballot(1);
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    uint32_t old = 0, avail = 0, fresh = 0;
    if ((int)lane == leader) {
        const uint32_t w = threadIdx.x >> 6, cnt = (uint32_t)__popcll(m);
        const uint32_t nx = s_wwork[3 * w], end = s_wwork[3 * w + 1], ch = s_wwork[3 * w + 2];
        old = nx; avail = end - nx;
        if (avail >= cnt) { s_wwork[3 * w] = nx + cnt; }
        else {
            fresh = atomicAdd(ctr, ch);
            s_wwork[3 * w] = fresh + (cnt - avail);
            s_wwork[3 * w + 1] = fresh + ch;
        }
    }
    old = (uint32_t)__builtin_amdgcn_readlane((int)old, leader);
    avail = (uint32_t)__builtin_amdgcn_readlane((int)avail, leader);
    fresh = (uint32_t)__builtin_amdgcn_readlane((int)fresh, leader);
    return rank < avail ? old + rank : fresh + (rank - avail);
}

// Next work item for each calling lane; lanes calling together share one atomic.
__device__ __forceinline__ uint32_t wave_claim(uint32_t *ctr)
{
    const uint64_t m = __ballot(1);
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    return (uint32_t)__builtin_amdgcn_readlane((int)base, leader) + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// The dose grid takes one atomic per voxel a lane leaves, not one per deposit: deposits
// are truncated one by one, u32(dep * 100) (primary.wgsl:111), and summed per lane while
// the voxel is unchanged; integer sums make the grid identical.
__device__ __forceinline__ void flush_dose(Sink &k)
{
    if (k.dacc) atomicAdd(&k.dose[k.dvi], k.dacc);
    k.dacc = 0u;
    k.dvi = 0xFFFFFFFFu;
}
__device__ __forceinline__ void emit2(Sink &k, float4 a, float4 b) { float4 e[2] = {a, b}; emit_n(k, 2, e); }
__device__ __forceinline__ void emit3(Sink &k, float4 a, float4 b, float4 c) { float4 e[3] = {a, b, c}; emit_n(k, 3, e); }

__device__ __forceinline__ void deposit(Sink &k, float px, float py, float pz, float dep)
{
    if (dep <= 0.0f) return;
    float vs = k.vs, box = k.box;
    int ix = (int)floorf((px + box) / vs);
    int iy = (int)floorf((py + box) / vs);
    int iz = (int)floorf((pz + box) / vs);
    int n = k.vc;
    if (ix < 0 || ix >= n || iy < 0 || iy >= n || iz < 0 || iz >= n) return;
    uint32_t vi = (uint32_t)((iz * n + iy) * n + ix);
    if (vi != k.dvi) { flush_dose(k); k.dvi = vi; }
    k.dacc += (uint32_t)(dep * 100.0f);
}

__device__ uint32_t dna_near(const dna_params &p, float px, float py, float pz)
{
    if (p.dna_enable == 0u) return 0u;
    const float r_damage = 0.29f;
    uint32_t N = p.dna_grid_n;
    float spacing = p.dna_spacing;
    float grid_off = -((float)(N - 1u) * spacing) * 0.5f;
    float inv_s = 1.0f / spacing;
    int fi = (int)rintf((py - grid_off) * inv_s);
    int fj = (int)rintf((pz - grid_off) * inv_s);
    if (fi < 0 || fi >= (int)N || fj < 0 || fj >= (int)N) return 0u;
    float fy = grid_off + (float)fi * spacing;
    float fz = grid_off + (float)fj * spacing;
    float y_rel = py - fy, z_rel = pz - fz;
    float r2 = y_rel * y_rel + z_rel * z_rel;
    float R_reach = p.dna_r_bb + r_damage;
    if (r2 > R_reach * R_reach) return 0u;
    int bp_est = (int)rintf((px - p.dna_x0) / p.dna_rise);
    int b0 = bp_est - 1 > 0 ? bp_est - 1 : 0;
    int b1 = bp_est + 1;
    const float d_phase = 0.5983986258506775f;
    float r_bb = p.dna_r_bb;
    for (int b = b0; b <= b1; b++) {
        float bx = p.dna_x0 + (float)b * p.dna_rise;
        float phi = (float)b * d_phase;
        float sn, cs; dna_sincosf(phi, &sn, &cs);
        float s0y = r_bb * cs, s0z = r_bb * sn;
        float dx = px - bx, dy0 = y_rel - s0y, dz0 = z_rel - s0z;
        if (dx*dx + dy0*dy0 + dz0*dz0 < r_damage * r_damage) return 1u;
        float sn1, cs1; dna_sincosf(phi + PI_F, &sn1, &cs1);
        float s1y = r_bb * cs1, s1z = r_bb * sn1;
        float dy1 = y_rel - s1y, dz1 = z_rel - s1z;
        if (dx*dx + dy1*dy1 + dz1*dz1 < r_damage * r_damage) return 1u;
    }
    return 0u;
}

__device__ __forceinline__ void gauss3(Rng &s, float sig, float &dx, float &dy, float &dz)
{
    float u1 = fmx(s.uf(), 1e-30f); float u2 = s.uf();
    float r = sqrtf(-2.0f * dna_logf(u1)) * sig;
    float sn, cs; dna_sincosf(TWO_PI_F * u2, &sn, &cs);
    dx = r * cs; dy = r * sn;
    float u3 = fmx(s.uf(), 1e-30f); float u4 = s.uf();
    dz = sqrtf(-2.0f * dna_logf(u3)) * sig * dna_cosf(TWO_PI_F * u4);
}

// H2Ovib channels after electron-hole recombination (primary.wgsl:247-276)
__device__ __forceinline__ void h2ovib(Sink &k, Rng &s, float x, float y, float z, float pid)
{
    float r_vd = s.uf();
    if (r_vd < 0.1365f) {
        k.cnt[0] += 2; k.cnt[5] += 1;
        emit3(k, make_float4(x, y, z, 0.0f + pid), make_float4(x, y, z, 0.0f + pid), make_float4(x, y, z, 7.0f + pid));
    } else if (r_vd < 0.494f) {
        k.cnt[0] += 1; k.cnt[2] += 1;
        emit2(k, make_float4(x, y, z, 0.0f + pid), make_float4(x, y, z, 2.0f + pid));
    } else if (r_vd < 0.650f) {
        k.cnt[2] += 2;
        emit2(k, make_float4(x, y, z, 2.0f + pid), make_float4(x, y, z, 2.0f + pid));
    }
}

// H2O+ / e-aq pair with Onsager recombination (primary.wgsl:229-288, 473-536, 560-619;
// secondary.wgsl:192-255)
__device__ void hole_pair(Sink &k, Rng &s, float mx, float my, float mz, float ex, float ey, float ez,
                          float sig_e, float eaq_code, float pid)
{
    if (sig_e > 0.0f) {
        float dx, dy, dz; gauss3(s, sig_e, dx, dy, dz);
        ex += dx; ey += dy; ez += dz;
    }
    float rx = ex - mx, ry = ey - my, rz = ez - mz;
    float r_sep = fmx(sqrtf(rx*rx + ry*ry + rz*rz), 1e-6f);
    float p_recomb = 1.0f - dna_expf(-0.711f / r_sep);
    float r_recomb = s.uf();
    if (r_recomb < p_recomb) {
        h2ovib(k, s, mx, my, mz, pid);
    } else {
        k.cnt[0] += 1; k.cnt[1] += 1; k.cnt[3] += 1;
        emit3(k, make_float4(mx, my, mz, 0.0f + pid), make_float4(ex, ey, ez, eaq_code + pid),
              make_float4(mx, my, mz, 3.0f + pid));
    }
}

__device__ void excitation_products(Sink &k, Rng &s, uint32_t lvl, float px, float py, float pz, float pid)
{
    float r_ch = s.uf();
    if (lvl == 0u) {
        if (r_ch < 0.65f) {
            k.cnt[0] += 1; k.cnt[2] += 1;
            emit2(k, make_float4(px, py, pz, 0.0f + pid), make_float4(px, py, pz, 2.0f + pid));
        }
    } else if (lvl == 1u) {
        if (r_ch < 0.175f) {
        } else if (r_ch < 0.2075f) {
            k.cnt[0] += 2; k.cnt[5] += 1;
            emit3(k, make_float4(px, py, pz, 0.0f + pid), make_float4(px, py, pz, 0.0f + pid),
                  make_float4(px, py, pz, 7.0f + pid));
        } else if (r_ch < 0.7075f) {
            float dx, dy, dz; gauss3(s, 1.1547005f, dx, dy, dz);
            float mx = px + dx, my = py + dy, mz = pz + dz;
            hole_pair(k, s, mx, my, mz, mx, my, mz, meesungnoen_sigma(1.7f), 1.0f, pid);
        } else if (r_ch < 0.961f) {
            k.cnt[0] += 1; k.cnt[2] += 1;
            emit2(k, make_float4(px, py, pz, 0.0f + pid), make_float4(px, py, pz, 2.0f + pid));
        } else {
            k.cnt[2] += 2;
            emit2(k, make_float4(px, py, pz, 2.0f + pid), make_float4(px, py, pz, 2.0f + pid));
        }
    } else {
        if (r_ch < 0.50f) {
            float dx, dy, dz; gauss3(s, 1.1547005f, dx, dy, dz);
            float mx = px + dx, my = py + dy, mz = pz + dz;
            hole_pair(k, s, mx, my, mz, mx, my, mz, meesungnoen_sigma(1.7f), 1.0f, pid);
        }
    }
}

__device__ __forceinline__ void rotate(float &dx, float &dy, float &dz, float cos_t, float sin_t, float cp, float sp)
{
    if (fabsf(dz) > 0.99999f) {
        float sw = (dz > 0.0f) ? 1.0f : -1.0f;
        dx = sin_t * cp; dy = sin_t * sp * sw; dz = cos_t * sw;
    } else {
        float q = sqrtf(1.0f - dz * dz);
        float inv = 1.0f / q;
        float nx = dx * cos_t + sin_t * (dx * dz * cp - dy * sp) * inv;
        float ny = dy * cos_t + sin_t * (dy * dz * cp + dx * sp) * inv;
        float nz = dz * cos_t - q * sin_t * cp;
        float len = sqrtf(nx*nx + ny*ny + nz*nz);
        dx = nx / len; dy = ny / len; dz = nz / len;
    }
}

__device__ __forceinline__ uint32_t pick_shell(const SmallTab &S, Rng &s, float E)
{
    float sf[5]; xs_shell_fracs(S, E, sf);
    float fsum = fmx(sf[0] + sf[1] + sf[2] + sf[3] + sf[4], 1e-12f);
    float r_sh = s.uf() * fsum;
    if (r_sh < sf[0]) return 0;
    if (r_sh < sf[0] + sf[1]) return 1;
    if (r_sh < sf[0] + sf[1] + sf[2]) return 2;
    if (r_sh < sf[0] + sf[1] + sf[2] + sf[3]) return 3;
    return 4;
}

__device__ __forceinline__ uint32_t pick_level(const SmallTab &S, Rng &s, float E)
{
    float ef[5]; xs_exc_fracs(S, E, ef);
    float efsum = fmx(ef[0] + ef[1] + ef[2] + ef[3] + ef[4], 1e-12f);
    float r_ex = s.uf() * efsum;
    if (r_ex < ef[0]) return 0;
    if (r_ex < ef[0] + ef[1]) return 1;
    if (r_ex < ef[0] + ef[1] + ef[2]) return 2;
    if (r_ex < ef[0] + ef[1] + ef[2] + ef[3]) return 3;
    return 4;
}

// wavefront-wide sum of a u32, then one atomic per wavefront
__device__ __forceinline__ void wave_flush(uint32_t *dst, uint32_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// ---------------------------------------------------------------- Phase A
struct PrimState {
    float x, y, z, E, dx, dy, dz, sx, sy, sz;
    float path, prod, mxr;
    uint32_t ni, nx, esc, step, nsec;
    Rng s;
};

__device__ void prim_init(const TransportDev &d, const dna_params &p, uint32_t idx, PrimState &st, Sink &k)
{
    // seeding of src/gpu/buffers.ts:91-104: (i+1)*2654435761 + seed*1013904223 in doubles,
    // ToUint32, then three rounds of the 0x45d9f3b mixer per state word
    double v = (double)(idx + 1u) * 2654435761.0 + (double)p.seed * 1013904223.0;
    uint32_t s = (uint32_t)(uint64_t)fmod(v, 4294967296.0);
    uint32_t w[4];
    for (uint32_t j = 0; j < 4; j++) {
        s ^= s >> 16; s *= 0x45d9f3bu; s ^= s >> 16;
        w[j] = s + j * 0x9E3779B9u;
    }
    st.s.s0 = w[0]; st.s.s1 = w[1]; st.s.s2 = w[2]; st.s.s3 = w[3];
    st.E = d.energies ? d.energies[idx] : p.be;
    st.x = 0.0f; st.y = 0.0f; st.z = 0.0f;
    float sh = p.start_half;
    if (sh > 0.0f) {
        st.x = (st.s.uf() * 2.0f - 1.0f) * sh;
        st.y = (st.s.uf() * 2.0f - 1.0f) * sh;
        st.z = (st.s.uf() * 2.0f - 1.0f) * sh;
    }
    st.sx = st.x; st.sy = st.y; st.sz = st.z;
    float u1 = fmx(st.s.uf(), 1e-10f);
    float u2 = st.s.uf();
    float cos_i = 1.0f - 2.0f * u1;
    float sin_i = sqrtf(fmx(0.0f, 1.0f - cos_i * cos_i));
    float sn, cs; dna_sincosf(TWO_PI_F * u2, &sn, &cs);
    st.dx = sin_i * cs; st.dy = sin_i * sn; st.dz = cos_i;
    if (p.dir_mode == 1u) { st.dx = 0.0f; st.dy = 0.0f; st.dz = 1.0f; }
    st.path = 0.0f; st.prod = 0.0f; st.mxr = 0.0f;
    st.ni = 0; st.nx = 0; st.esc = 0; st.step = 0; st.nsec = 0;
    k.keybase = (uint64_t)idx << 40;
    k.seq = 0;
}

// One iteration of the primary loop (primary.wgsl:150-651).  Returns true when the
// history has ended (loop exit).
__device__ bool prim_step(const TransportDev &d, const dna_params &p, const SmallTab &S, uint32_t idx,
                          PrimState &st, Sink &k, uint32_t *dbg4)
{
    if (st.step >= p.ms) return true;
    float box = p.box, ce = p.ce;
    if (st.E < ce) return true;
    if (fabsf(st.x) >= box || fabsf(st.y) >= box || fabsf(st.z) >= box) { st.esc = 1; return true; }
    st.step++;
    float E = st.E;
    bool productive = E > 12.0f;
    float xi, xc, xl; xs_all(S, E, xi, xc, xl);
    float s_vib = xs_vib_total(S, E), s_dea = xs_dea(S, E);
    float s_tot = xi + xc + xl + s_vib + s_dea;
    if (s_tot <= 0.0f) return true;
    float lambda = 1.0f / (NW_F * s_tot);
    Rng &s = st.s;
    float dist = -dna_logf(fmx(s.uf(), 1e-10f)) * lambda;
    float r_el = s.uf();
    float cos_el = xs_el_cos(S, E, r_el);
    float sin_el = sqrtf(fmx(0.0f, 1.0f - cos_el * cos_el));
    float sp, cp; dna_sincosf(TWO_PI_F * s.uf(), &sp, &cp);
    rotate(st.dx, st.dy, st.dz, cos_el, sin_el, cp, sp);
    st.x += st.dx * dist; st.y += st.dy * dist; st.z += st.dz * dist;
    st.path += dist;
    if (productive) st.prod += dist;
    float rdx = st.x - st.sx, rdy = st.y - st.sy, rdz = st.z - st.sz;
    float r = sqrtf(rdx*rdx + rdy*rdy + rdz*rdz);
    if (r > st.mxr) st.mxr = r;
    if (fabsf(st.x) >= box || fabsf(st.y) >= box || fabsf(st.z) >= box) { st.esc = 1; return true; }
    float px = st.x, py = st.y, pz = st.z;
    float pid = (float)idx * 8.0f;
    float r_type = s.uf() * s_tot;
    if (r_type < xi && E > c_SB[0]) {
        uint32_t shell = pick_shell(S, s, E);
        float bind = c_SB[shell];
        if (E > bind) {
            float r_w = s.uf();
            float Wmax = (E + bind) * 0.5f;
            float Wt = clampf_(sample_W_sec(S, d.tab, E, shell, r_w), bind, Wmax);
            float W_sec = fmx(Wt - bind, 0.0f);
            E -= Wt;
            st.ni++;
            *dbg4 += (uint32_t)(Wt * 10.0f);
            deposit(k, px, py, pz, bind);
            if (W_sec <= ce) deposit(k, px, py, pz, W_sec);
            if (dna_near(p, px, py, pz)) k.cnt[4]++;
            float mdx, mdy, mdz; gauss3(s, 1.1547005f, mdx, mdy, mdz);
            float mx = px + mdx, my = py + mdy, mz = pz + mdz;
            if (W_sec <= ce) {
                hole_pair(k, s, mx, my, mz, px, py, pz, meesungnoen_sigma(W_sec), 5.0f, pid);
            } else {
                float r_track = 0.04f * dna_powf(W_sec, 1.3f);
                float r_tail = meesungnoen_sigma(fmn(W_sec, 7.4f));
                float r_sep = fmx(r_track + r_tail, 1e-3f);
                float p_recomb = 1.0f - dna_expf(-0.711f / r_sep);
                float r_recomb = s.uf();
                if (r_recomb < p_recomb) {
                    h2ovib(k, s, mx, my, mz, pid);
                } else {
                    k.cnt[0] += 1; k.cnt[3] += 1;
                    emit2(k, make_float4(mx, my, mz, 0.0f + pid), make_float4(mx, my, mz, 3.0f + pid));
                }
                // secondary emission, G4DNABornAngle (primary.wgsl:362-413)
                float E_before = E + bind + W_sec;
                float cos_s;
                if (W_sec < 50.0f) cos_s = 2.0f * s.uf() - 1.0f;
                else if (W_sec <= 200.0f) {
                    if (s.uf() <= 0.1f) cos_s = 2.0f * s.uf() - 1.0f;
                    else cos_s = s.uf() * 0.70710678f;
                } else {
                    float sin2 = clampf_((1.0f - W_sec / E_before) / (1.0f + W_sec / 1022000.0f), 0.0f, 1.0f);
                    cos_s = sqrtf(fmx(0.0f, 1.0f - sin2));
                }
                float sin_s = sqrtf(fmx(0.0f, 1.0f - cos_s * cos_s));
                float sps, cps; dna_sincosf(TWO_PI_F * s.uf(), &sps, &cps);
                float sdx = st.dx, sdy = st.dy, sdz = st.dz;
                rotate(sdx, sdy, sdz, cos_s, sin_s, cps, sps);
                float p_inc = sqrtf(E_before), p_sec = sqrtf(W_sec);
                float pfx = p_inc * st.dx - p_sec * sdx;
                float pfy = p_inc * st.dy - p_sec * sdy;
                float pfz = p_inc * st.dz - p_sec * sdz;
                float pfl = sqrtf(pfx*pfx + pfy*pfy + pfz*pfz);
                if (pfl > 1e-10f) { st.dx = pfx / pfl; st.dy = pfy / pfl; st.dz = pfz / pfl; }
                uint32_t r0 = s.next(), r1 = s.next(), r2 = s.next(), r3 = s.next();
                st.nsec++;
                uint32_t q = wave_claim(&d.ctr[C_SEC]);
                if (q < d.sec_cap) {
                    SecRec rec;
                    rec.posE = make_float4(px, py, pz, W_sec);
                    rec.dir = make_float4(sdx, sdy, sdz, 0.0f);
                    rec.rng = make_uint4(r0, r1, r2, r3);
                    rec.pid = idx; rec.k = st.nsec; rec.pad0 = 0; rec.pad1 = 0;
                    d.sec[q] = rec;
                }
            }
        }
    } else if (r_type < xi + xc) {
        uint32_t lvl = pick_level(S, s, E);
        float dep = fmn(c_EX[lvl], E);
        E -= dep;
        deposit(k, px, py, pz, dep);
        st.nx++;
        excitation_products(k, s, lvl, px, py, pz, pid);
        if (dna_near(p, px, py, pz)) k.cnt[4]++;
    } else if (r_type < xi + xc + s_vib) {
        uint32_t mode = sample_vib_mode(S, E, s.uf());
        float dep = fmn(S.VIB_LEV[mode], E);
        E -= dep;
        deposit(k, px, py, pz, dep);
    } else if (r_type < xi + xc + s_vib + s_dea) {
        deposit(k, px, py, pz, E);
        k.cnt[0] += 1; k.cnt[5] += 1;
        emit3(k, make_float4(px, py, pz, 0.0f + pid), make_float4(px, py, pz, 6.0f + pid),
              make_float4(px, py, pz, 7.0f + pid));
        E = 0.0f;
    }
    st.E = E;
    return false;
}

// lanes: the active lanes of each 64-lane wavefront (1..64); the others only help stage
// the tables.  Fewer lanes per wavefront means more wavefronts for a small batch, fewer
// branch paths per wavefront and less waiting behind the longest history in it.
__global__ void __launch_bounds__(256) k_phase_a(TransportDev d, uint32_t lanes)
{
    __shared__ SmallTab S;
    stage_tables(S, d.tab, d.guide);
    wave_state_init(d.rad_chunk, d.work_chunk);
    const dna_params p = d.p;
    Sink k;
    k.rad = d.rad_raw; k.key = d.rad_key; k.rad_ctr = &d.ctr[C_RAD]; k.rad_cap = d.rad_cap;
    k.dose = d.dose; k.box = p.box; k.vs = 2.0f * p.box / (float)p.vc; k.vc = (int)p.vc;
    k.dvi = 0xFFFFFFFFu; k.dacc = 0u;
    for (int i = 0; i < 6; i++) k.cnt[i] = 0;
    uint32_t dbg[5] = {0, 0, 0, 0, 0};
    PrimState st;
    const bool lane_on = (threadIdx.x & 63u) < lanes;
    uint32_t idx = lane_on ? wave_claim(&d.ctr[C_PNEXT]) : 0xFFFFFFFFu;
    bool active = lane_on && idx < p.n;
    if (active) prim_init(d, p, idx, st, k);
    while (active) {
        bool done = prim_step(d, p, S, idx, st, k, &dbg[4]);
        if (done) {
            if (st.esc == 0u && st.E > 0.0f) deposit(k, st.x, st.y, st.z, st.E);
            dna_result r;
            r.path = st.path; r.prod = st.prod; r.fE = st.E; r.ni = st.ni; r.nx = st.nx;
            r.esc = st.esc; r.mx = st.mxr; r._pad = 0;
            d.results[idx] = r;
            d.prim_nsec[idx] = st.nsec;
            d.prim_nrad[idx] = k.seq;
            dbg[0] += st.ni; dbg[1] += st.nx; dbg[2] += st.esc; dbg[3] += 1;
            idx = wave_claim(&d.ctr[C_PNEXT]);
            active = idx < p.n;
            if (active) prim_init(d, p, idx, st, k);
        }
    }
    flush_dose(k);
    for (int i = 0; i < 6; i++) wave_flush(&d.ctr[i], k.cnt[i]);
    for (int i = 0; i < 5; i++) wave_flush(&d.ctr[C_DBG + i], dbg[i]);
}

// ---------------------------------------------------------------- Phase B
struct SecState {
    float x, y, z, E, dx, dy, dz, pid;
    uint32_t calls;
    Rng s;
};

// One dispatch of secondary.wgsl:83-450 for this particle.  Returns true when the
// particle is dead (the reference's dir_alive.w = 0) or out of step calls.
__device__ bool sec_step(const TransportDev &d, const dna_params &p, const SmallTab &S, SecState &st,
                         Sink &k, uint32_t sst[5])
{
    if (st.calls >= p.max_sec_steps) return true;
    st.calls++;
    float box = p.box, ce = p.ce;
    float E = st.E, px = st.x, py = st.y, pz = st.z;
    Rng &s = st.s;
    float pid = st.pid;
    if (E < ce) {
        deposit(k, px, py, pz, E);
        k.cnt[1] += 1;
        float sig = meesungnoen_sigma(E);
        float ex = px, ey = py, ez = pz;
        if (sig > 0.0f) { float dx, dy, dz; gauss3(s, sig, dx, dy, dz); ex += dx; ey += dy; ez += dz; }
        float4 e = make_float4(ex, ey, ez, 5.0f + pid);
        emit_n(k, 1, &e);
        sst[1]++;
        return true;
    }
    if (fabsf(px) >= box || fabsf(py) >= box || fabsf(pz) >= box) { sst[2]++; return true; }
    float xi, xc, xl; xs_all(S, E, xi, xc, xl);
    float s_vib = xs_vib_total(S, E), s_dea = xs_dea(S, E);
    float s_tot = xi + xc + xl + s_vib + s_dea;
    if (s_tot <= 0.0f) return true;
    float lambda = 1.0f / (NW_F * s_tot);
    float dist = -dna_logf(fmx(s.uf(), 1e-10f)) * lambda;
    float r_el = s.uf();
    float cos_el = xs_el_cos(S, E, r_el);
    float sin_el = sqrtf(fmx(0.0f, 1.0f - cos_el * cos_el));
    float sp, cp; dna_sincosf(TWO_PI_F * s.uf(), &sp, &cp);
    rotate(st.dx, st.dy, st.dz, cos_el, sin_el, cp, sp);
    px += st.dx * dist; py += st.dy * dist; pz += st.dz * dist;
    sst[3]++;
    float r_type = s.uf() * s_tot;
    if (r_type < xi && E > c_SB[0]) {
        uint32_t shell = pick_shell(S, s, E);
        float bind = c_SB[shell];
        if (E > bind) {
            float r_w = s.uf();
            float Wmax = (E + bind) * 0.5f;
            float Wt = clampf_(sample_W_sec(S, d.tab, E, shell, r_w), bind, Wmax);
            float W_sec = fmx(Wt - bind, 0.0f);
            E -= Wt;
            sst[4]++;
            deposit(k, px, py, pz, Wt);
            float mdx, mdy, mdz; gauss3(s, 1.1547005f, mdx, mdy, mdz);
            hole_pair(k, s, px + mdx, py + mdy, pz + mdz, px, py, pz, meesungnoen_sigma(W_sec), 5.0f, pid);
            if (dna_near(p, px, py, pz)) k.cnt[4]++;
        }
    } else if (r_type < xi + xc) {
        uint32_t lvl = pick_level(S, s, E);
        float dep = fmn(c_EX[lvl], E);
        E -= dep;
        deposit(k, px, py, pz, dep);
        excitation_products(k, s, lvl, px, py, pz, pid);
        if (dna_near(p, px, py, pz)) k.cnt[4]++;
    } else if (r_type < xi + xc + s_vib) {
        uint32_t mode = sample_vib_mode(S, E, s.uf());
        float dep = fmn(S.VIB_LEV[mode], E);
        E -= dep;
        deposit(k, px, py, pz, dep);
    } else if (r_type < xi + xc + s_vib + s_dea) {
        deposit(k, px, py, pz, E);
        k.cnt[0] += 1; k.cnt[5] += 1;
        emit3(k, make_float4(px, py, pz, 0.0f + pid), make_float4(px, py, pz, 6.0f + pid),
              make_float4(px, py, pz, 7.0f + pid));
        E = 0.0f;
    }
    st.E = E; st.x = px; st.y = py; st.z = pz;
    return false;
}

__device__ __forceinline__ void sec_load(const TransportDev &d, uint32_t q, SecState &st, Sink &k)
{
    SecRec r = d.sec[q];
    st.x = r.posE.x; st.y = r.posE.y; st.z = r.posE.z; st.E = r.posE.w;
    st.dx = r.dir.x; st.dy = r.dir.y; st.dz = r.dir.z;
    st.s.s0 = r.rng.x; st.s.s1 = r.rng.y; st.s.s2 = r.rng.z; st.s.s3 = r.rng.w;
    st.pid = (float)((uint32_t)((float)(r.pid + 1u)) - 1u) * 8.0f;   // secondary.wgsl:88
    st.calls = 0;
    k.keybase = ((uint64_t)r.pid << 40) | ((uint64_t)r.k << 20);
    k.seq = 0;
}

// The next secondary record to transport: records beyond the secondary cap (canonical rank
// sec_base[pid] + k - 1 >= sec_keep) are claimed, given no radicals and skipped.
__device__ __forceinline__ uint32_t claim_kept(const TransportDev &d, uint32_t sec_n)
{
    for (;;) {
        const uint32_t q = wave_claim_chunked(&d.ctr[C_SNEXT]);
        if (q >= sec_n || !d.sec_base) return q;
        const SecRec &r = d.sec[q];
        if (d.sec_base[r.pid] + r.k - 1u < d.sec_keep) return q;
        d.sec_nrad[q] = 0u;
    }
}

__global__ void __launch_bounds__(256) k_phase_b(TransportDev d, uint32_t sec_n)
{
    __shared__ SmallTab S;
    stage_tables(S, d.tab, d.guide);
    wave_state_init(d.rad_chunk, d.work_chunk);
    const dna_params p = d.p;
    Sink k;
    k.rad = d.rad_raw; k.key = d.rad_key; k.rad_ctr = &d.ctr[C_RAD]; k.rad_cap = d.rad_cap;
    k.dose = d.dose; k.box = p.box; k.vs = 2.0f * p.box / (float)p.vc; k.vc = (int)p.vc;
    k.dvi = 0xFFFFFFFFu; k.dacc = 0u;
    for (int i = 0; i < 6; i++) k.cnt[i] = 0;
    uint32_t sst[5] = {0, 0, 0, 0, 0};
    SecState st;
    uint32_t q = claim_kept(d, sec_n);
    bool active = q < sec_n;
    if (active) sec_load(d, q, st, k);
    while (active) {
        bool done = sec_step(d, p, S, st, k, sst);
        if (done) {
            d.sec_nrad[q] = k.seq;
            q = claim_kept(d, sec_n);
            active = q < sec_n;
            if (active) sec_load(d, q, st, k);
        }
    }
    flush_dose(k);
    for (int i = 0; i < 6; i++) wave_flush(&d.ctr[i], k.cnt[i]);
    for (int i = 1; i < 5; i++) wave_flush(&d.ctr[C_SST + i], sst[i]);
}

static int g_num_cu = 0;
int num_cu()
{
    if (!g_num_cu) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            g_num_cu = n;
        else
            g_num_cu = 256;                   // MI355X: 256 CUs
    }
    return g_num_cu;
}

uint32_t phase_a_lanes(uint32_t n, int opt)
{
    if (opt > 0) return (uint32_t)opt;
    // the widest wavefronts that still give every CU one (measured on MI355X at 1024, 4096
    // and 16384 primaries: one wavefront per CU beats both fewer, wider wavefronts and one
    // per SIMD)
    const uint32_t target = (uint32_t)num_cu();
    uint32_t L = 64u;
    while (L > 1u && (n + L - 1u) / L < target) L >>= 1;
    return L;
}

// One 64-lane block per wavefront, `lanes` of them active, one primary per active lane up
// to the device's resident capacity (further primaries are claimed as lanes finish).
static uint32_t phase_a_blocks(uint32_t n, uint32_t lanes)
{
    const uint32_t lanes_max = (uint32_t)num_cu() * 4u * 64u * 4u;
    const uint32_t m = n < lanes_max ? n : lanes_max;
    return (m + lanes - 1u) / lanes;
}

size_t transport_slot_headroom(uint32_t n, uint32_t lanes, uint32_t rad_chunk)
{
    // every wavefront of either phase may leave up to rad_chunk reserved slots unused
    const size_t waves_a = phase_a_blocks(n, lanes);
    const size_t waves_b = (size_t)num_cu() * 4u * 4u;
    return (waves_a + waves_b) * rad_chunk;
}

hipError_t launch_phase_a(const TransportDev &d, uint32_t lanes, hipStream_t st)
{
    if (lanes < 1u || lanes > 64u) return hipErrorInvalidValue;
    const uint32_t blocks = phase_a_blocks(d.p.n, lanes);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_phase_a, dim3(blocks), dim3(64), 0, st, d, lanes);
    return hipGetLastError();
}

// The table lookups of the step evaluated for given energies (and uniforms r), with the
// staged LDS tables the kernels use: which 0 xs_all -> 3 values, 1 shell fractions -> 5,
// 2 excitation-level fractions -> 5, 3 vibrational total -> 1, 4 DEA -> 1,
// 5 elastic cos(theta) at (E, r) -> 1.
__global__ void __launch_bounds__(256) k_xs_probe(const dna_tables *T, const uint8_t *guide, const float *E,
                                                  const float *r, uint32_t n, int which, float *out)
{
    __shared__ SmallTab S;
    stage_tables(S, T, guide);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float e = E[i];
        if (which == 0) { float a, b, c; xs_all(S, e, a, b, c); out[3 * i] = a; out[3 * i + 1] = b; out[3 * i + 2] = c; }
        else if (which == 1) { float f[5]; xs_shell_fracs(S, e, f); for (int k = 0; k < 5; k++) out[5 * i + k] = f[k]; }
        else if (which == 2) { float f[5]; xs_exc_fracs(S, e, f); for (int k = 0; k < 5; k++) out[5 * i + k] = f[k]; }
        else if (which == 3) out[i] = xs_vib_total(S, e);
        else if (which == 4) out[i] = xs_dea(S, e);
        else out[i] = xs_el_cos(S, e, r[i]);
    }
}

hipError_t launch_xs_probe(const dna_tables *T, const uint8_t *guide, const float *E, const float *r, uint32_t n,
                           int which, float *out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks > 1024u) blocks = 1024u;
    hipLaunchKernelGGL(k_xs_probe, dim3(blocks), dim3(256), 0, st, T, guide, E, r, n, which, out);
    return hipGetLastError();
}

hipError_t launch_phase_b(const TransportDev &d, uint32_t sec_n, hipStream_t st)
{
    uint32_t lanes_max = (uint32_t)num_cu() * 4u * 64u * 4u;
    uint32_t lanes = sec_n < lanes_max ? sec_n : lanes_max;
    uint32_t blocks = (lanes + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_phase_b, dim3(blocks), dim3(256), 0, st, d, sec_n);
    return hipGetLastError();
}

}  // namespace dna
