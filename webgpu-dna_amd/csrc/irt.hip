// irt.hip — Phase C, independent-reaction-time chemistry (public/irt-worker.js:370-699)
// for many tracks (primaries) and many instances (worker invocations) in one launch.
//
// Pipeline (all on the device; the host only sizes buffers from small per-primary
// count arrays):
//   decode   w -> (species code, primary id) per radical entry (irt-worker.js:382-394);
//            code 7 entries count initial H2, code 4 entries are skipped
//   group    entries of one (instance, primary) become one track, in input order
//            (the worker's per-pid list, irt-worker.js:391-393)
//   k_irt    one workgroup per track:
//            1. product displacement of the initial species (irt-worker.js:475-537)
//            2. all initial pairs i<j, enumerated circularly so every lane gets the same
//               count: pair (i, i+d mod n) for d = 1..(n-1)/2 (+ n/2 for i < n/2, n even);
//               reaction-time sampling in FP64 (irt-worker.js:541-558, 116-184)
//            3. event loop: the earliest event whose two reactants are alive is found by
//               a workgroup min-reduction over the event list, ordered by (t, a, b); the
//               reaction, its products and checkpoint snapshots are applied by one lane;
//               each product is re-paired with every alive particle by the whole
//               workgroup (irt-worker.js:566-645, 430-465)
//            4. per-instance tallies with 64-bit atomics
// Random numbers come from the counter-based streams of include/dna_irt_math.h, so the
// result is independent of thread scheduling and identical to the CPU restatement.
#include "dna_internal.h"
#include "dna_ctx.h"
#include "../../include/dna_irt_math.h"
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>

namespace dna {

struct IrtConst {
    double sigma[9], sigma_eff[9], rc[9], kdif[9], kact[9], prob[9], kobs[9];
    double D[6], sqrtD[6];
    double R_CUT2, sig_oh, sig_h, sig_eaq;
    float R_CUT_F;                    // R_CUT rounded up to FP32 (a bound, not the test)
    float CELL_R;                     // reach covered by m cells of the pair hash (> R_CUT)
    int rxnMap[36];
    int type[9], nprod[9], prods[9][2];
};
__constant__ IrtConst c_irt;

struct IrtTrack {
    uint32_t inst, pid, start, n;     // bucket range in idx[]
    uint32_t h2;                      // initial H2 markers
    uint32_t cap;                     // particle capacity of this track (products included)
    uint32_t cap_global;              // the worker's CAP = 2*maxN + 512 of the instance
    uint32_t ev_cap;
    uint32_t idx;                     // position in track order (per-track outputs)
    uint64_t part_off, ev_off;
};

struct IrtArgs {
    const float4 *rad;
    const uint32_t *idx;          // bucketed entry indices
    const IrtTrack *tracks;
    float4 *pos;                  // (x, y, z, species) per particle slot
    uint8_t *alive;               // per particle slot: species + 1 while alive, 0 once reacted
    // 6 species lists of capacity cap per track (may hold reacted particles): each entry is
    // the particle's x, y, z and index (lpos; contiguous reads for the re-pairing screen) and
    // its birth time rounded down to FP32, bits in lst (0 for the initial particles; read only
    // for the appended part of a list)
    uint32_t *lst;
    float4 *lpos;
    double *tb;
    double *ev_t;
    uint2 *ev_ab;
    uint4 *pend;                  // stage-2 queue, same per-track offset/capacity as events
    float4 *sorted;               // species-sorted copy of the initial particles
    uint32_t lds_slots;           // float4 slots of dynamic LDS for the sorted copy (pairs kernel)
    uint32_t bm_max;              // event blocks per track at most (1..BM_MAX)
    int drain_at;                 // the initial pairs' drain: 0 in k_irt_loop, 1 in k_irt_pairs,
                                  // 2 in k_irt_drain
    int cell_m;                   // initial-pair hash: 0 = one cell per species, m = cells of
                                  // CELL_R / m along the track's principal axis
    uint32_t *tstate;             // TSTATE u32 per track (launch order): pairs kernel -> loop kernel
    uint32_t *order;              // optional: the track (launch index) of each loop-kernel
                                  // workgroup, heaviest first; null = launch order
    uint32_t al_lds;              // bytes of dynamic LDS for the loop kernel's alive bits (0: global)
    unsigned long long *tally;    // per instance: [ckpt][5], then 9 rxn, n_reacted, n_events
    uint32_t *track_tally;        // optional per-track [ckpt][5] species counts (track order)
    uint32_t *final_ntot;         // optional: particles per track at the end (track order); the
                                  // final alive bytes are then left in alive[]
    double *tdeath;               // optional (with final_ntot): reaction time per particle slot
    uint32_t *overflow;
    double tc[DNA_MAX_CKPT];
    int nc;
    uint64_t seed;
    const uint64_t *seed_ptr;     // when set, the seed is read here (graph replays)
};
#define TALLY_STRIDE (DNA_MAX_CKPT * 5 + 9 + 5 + 8 + 2)   // tallies, rxn, n_re, n_ev, 3 work counters,
                                                       // 8 phase timers, all pairs, re-pair slots scanned

// ------------------------------------------------------------ samplers (FP64)
// sampleIRT (irt-worker.js:116-184) is split in two.  Stage 1 is everything except the
// SamplePDC rejection loop; it returns a reaction time t >= 0, -1 (no reaction) or -2
// (the pair needs SamplePDC).  Stage-2 pairs are queued and the whole workgroup drains
// the queue afterwards, each lane taking a new pair as soon as its trial loop ends, so
// the long, uneven rejection loops (~100-200 trials at 100-500 nm) do not stall the
// other lanes of a wavefront.  Draw indices are those of the one-pass sampler (U at
// draw 6, trial t at draws 7+2t, 8+2t), so the result is unchanged.
#define STAGE_PDC (-2.0)

__device__ double irt_stage1(double r0, int ri, double D, uint64_t ps)
{
    const IrtConst &C = c_irt;
    if (C.type[ri] == 0) {
        double sigma = C.sigma[ri], rc = C.rc[ri];
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
    double rc = C.rc[ri];
    double sig_chk = C.sigma[ri], r0_chk = r0;
    if (rc != 0) { r0_chk = -rc / (1 - dna_exp(rc / r0)); sig_chk = C.sigma_eff[ri]; }
    if (sig_chk > r0_chk) return (C.prob[ri] > dna_u01(ps, 6u)) ? 0.0 : -1.0;
    double Winf = sig_chk / r0_chk * C.kobs[ri] / C.kdif[ri];
    if (!(Winf > dna_u01(ps, 6u))) return -1;
    return STAGE_PDC;
}

// Stage 0: true when sampleIRT can return a time for this pair (contact, or the first
// uniform below the survival probability).  Same expressions as irt_stage1, so the
// decision is identical; most pairs stop here with one draw and no transcendental.
__device__ __forceinline__ bool irt_may_react(double r0, int ri, double D, uint64_t ps)
{
    const IrtConst &C = c_irt;
    if (C.type[ri] == 0) {
        double sigma = C.sigma[ri], rc = C.rc[ri];
        if (sigma <= 0 || D <= 0) return false;
        if (r0 <= sigma) return true;
        double r0e = r0;
        if (rc != 0) r0e = -rc / (1 - dna_exp(rc / r0));
        if (r0e <= sigma) return true;
        double Winf = sigma / r0e;
        double U = dna_u01(ps, 6u);
        return !(U <= 0 || U >= Winf);
    }
    double rc = C.rc[ri];
    double sig_chk = C.sigma[ri], r0_chk = r0;
    if (rc != 0) { r0_chk = -rc / (1 - dna_exp(rc / r0)); sig_chk = C.sigma_eff[ri]; }
    if (sig_chk > r0_chk) return true;
    double Winf = sig_chk / r0_chk * C.kobs[ri] / C.kdif[ri];
    return Winf > dna_u01(ps, 6u);
}

// Constants of one reaction in registers, and irt_may_react on them.
struct RxK { int type; double sigma, rc, sigma_eff, prob, kobs, kdif, D; };

__device__ __forceinline__ RxK rx_constants(int ri, double D)
{
    const IrtConst &C = c_irt;
    RxK k;
    k.type = C.type[ri]; k.sigma = C.sigma[ri]; k.rc = C.rc[ri]; k.sigma_eff = C.sigma_eff[ri];
    k.prob = C.prob[ri]; k.kobs = C.kobs[ri]; k.kdif = C.kdif[ri]; k.D = D;
    return k;
}

__device__ __forceinline__ bool may_react_k(const RxK &k, double r0, uint64_t ps)
{
    if (k.type == 0) {
        if (k.sigma <= 0 || k.D <= 0) return false;
        if (r0 <= k.sigma) return true;
        double r0e = r0;
        if (k.rc != 0) r0e = -k.rc / (1 - dna_exp(k.rc / r0));
        if (r0e <= k.sigma) return true;
        double Winf = k.sigma / r0e;
        double U = dna_u01(ps, 6u);
        return !(U <= 0 || U >= Winf);
    }
    double sig_chk = k.sigma, r0_chk = r0;
    if (k.rc != 0) { r0_chk = -k.rc / (1 - dna_exp(k.rc / r0)); sig_chk = k.sigma_eff; }
    if (sig_chk > r0_chk) return true;
    double Winf = sig_chk / r0_chk * k.kobs / k.kdif;
    return Winf > dna_u01(ps, 6u);
}

// Necessary condition for may_react_k, cheap enough to run on every candidate first.
// With R = max(rc, 0), the Onsager distance satisfies r0e >= r0 - R (rc < 0: 1 - e^-x <= x
// gives r0e >= r0; rc > 0: e^x - 1 <= x e^x gives r0e >= r0 - rc).  So a pair at
// r0 - R > sc (sc = the contact radius sigma or sigma_eff) is not in contact, and it can
// pass U < Winf only if (r0 - R) U < S, where S = sigma (TDC) or sigma_chk kobs/kdif (PDC).
// The test runs in FP32 with relative margins of 1e-5, far above its rounding (< 1e-6),
// on a lower bound of r0; a pair it rejects is one the exact FP64 test rejects too.
struct RxB { float S, sc, R; };

__device__ __forceinline__ RxB rx_bound(const RxK &k)
{
    RxB b;
    double sg = (k.type == 0 || k.rc == 0) ? k.sigma : k.sigma_eff;
    b.S = k.type == 0 ? (float)k.sigma : (float)(sg * k.kobs / k.kdif);
    b.sc = (float)sg;
    b.R = k.rc > 0 ? (float)k.rc : 0.f;
    b.S *= 1.00001f; b.sc *= 1.00001f; b.R *= 1.00001f;
    return b;
}

// rmin: a lower bound of the pair distance; U: a lower bound of the pair's draw 6 (0 = keep:
// a PDC pair with U = 0 passes any Winf > 0)
__device__ __forceinline__ bool cannot_react(const RxB &b, float rmin, float U)
{
    float a = rmin - b.R;
    if (!(a > b.sc)) return false;
    return a * U * 0.99999f > b.S;
}

// Lower bound of draw n, dna_u01(ps, n) = (x >> 11) 2^-53: its top 23 bits, taken as the FP32
// 1.m - 1 with m those bits (exact, and no integer-to-float conversion).
__device__ __forceinline__ float u01_lo(uint64_t ps, uint32_t n)
{
    const uint64_t x = dna_mix64(ps + DNA_GOLDEN * (uint64_t)(n + 1u));
    return __uint_as_float(0x3F800000u | ((uint32_t)(x >> 32) >> 9)) - 1.0f;
}

// Square root with the hardware instruction (1 ulp); only used inside bounds whose margins
// are far above that error.
__device__ __forceinline__ float sqrt_bound(float x) { return __builtin_amdgcn_sqrtf(x); }

// SamplePDC parameters a, b (irt-worker.js:145-160)
__device__ void pdc_ab(double r0, int ri, double D, double &a, double &b)
{
    const IrtConst &C = c_irt;
    double sc = C.sigma[ri], rc = C.rc[ri], kobs = C.kobs[ri], kact = C.kact[ri];
    if (rc == 0) { a = (1.0 / sc) * kact / kobs; b = (r0 - sc) / 2; return; }
    double v = kact / (4 * DNA_PI * sc * sc * dna_exp(-rc / sc));
    double alpha = v + rc * D / (sc * sc * (1 - dna_exp(-rc / sc)));
    double sh, ch; dna_sinhcosh(rc / (2 * sc), &sh, &ch);
    a = 4 * sc * sc * alpha / (D * rc * rc) * sh * sh;
    double sh0, ch0; dna_sinhcosh(rc / (2 * r0), &sh0, &ch0);
    b = rc / 4 * (ch0 / sh0 - ch / sh);
}

// ------------------------------------------------------------ grouping kernels
__device__ __forceinline__ bool decode_w(float wf, uint32_t &pid, int &sp)
{
    double w = floor((double)wf + 0.5);
    if (w >= 0.0 && w < 9.0e18) {             // integer-valued and non-negative: the same
        const uint64_t wi = (uint64_t)w;      // fmod and floor(w / 8) in integer form
        sp = (int)(wi & 7u);
        if ((wi >> 3) >= 4294967295ull) return false;
        pid = (uint32_t)(wi >> 3);
        return true;
    }
    double spd = fmod(w, 8.0);
    if (spd < 0) return false;
    sp = (int)spd;
    double pd = floor(w / 8.0);
    if (pd >= 4294967295.0) return false;
    pid = (uint32_t)pd;
    return true;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__device__ __forceinline__ uint32_t find_inst(const uint64_t *off, uint32_t n_inst, uint64_t i)
{
    uint32_t lo = 0, hi = n_inst;       // off[lo] <= i < off[hi]
    while (hi - lo > 1) { uint32_t m = (lo + hi) >> 1; if (off[m] <= i) lo = m; else hi = m; }
    return lo;
}

// largest primary id; one atomic per block (same-address atomics from every wavefront
// serialise at the L2)
__global__ void __launch_bounds__(256) k_pidmax(const float4 *rad, uint64_t n, uint32_t *pidmax)
{
    __shared__ uint32_t wm[4];
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t pid; int sp;
        if (decode_w(rad[i].w, pid, sp) && pid > m) m = pid;
    }
    for (int o = 32; o > 0; o >>= 1) { uint32_t t = __shfl_xor(m, o, 64); m = t > m ? t : m; }
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / 64; w++) m = wm[w] > m ? wm[w] : m;
        atomicMax(pidmax, m);
    }
}

// kept[i] = 1 for a particle entry; key[i] = inst*(pidmax+1) + pid; counts and markers
// Runs of equal tags in a wavefront: grouped inputs put one primary's entries on
// consecutive lanes, so counting with one atomic per run (not per entry) removes the
// contention of thousands of same-address atomics.
__device__ __forceinline__ uint32_t run_length_if_head(uint64_t tag)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t prev = __shfl_up(tag, 1, 64);
    bool head = lane == 0 || prev != tag;
    unsigned long long H = __ballot(head);
    if (!head) return 0;
    unsigned long long above = lane == 63 ? 0ull : (H & ~((2ull << lane) - 1ull));
    uint32_t next = above ? (uint32_t)(__ffsll((long long)above) - 1) : 64u;
    return next - lane;
}

// kept[i] = 1 for a particle entry; key[i] = inst*(pidmax+1) + pid; counts and markers
__global__ void k_count(const float4 *rad, uint64_t n, const uint64_t *off, uint32_t n_inst, uint32_t K1,
                        uint32_t *kept, uint32_t *key, uint32_t *cnt, uint32_t *h2)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint32_t kk = 0, cat = 0;        // cat: 0 skipped, 1 particle, 2 H2 marker
    if (i < n) {
        uint32_t pid; int sp;
        if (decode_w(rad[i].w, pid, sp) && sp != 4) {
            uint32_t inst = find_inst(off, n_inst, i);
            kk = inst * K1 + pid;
            cat = sp == 7 ? 2u : 1u;
        }
        kept[i] = cat == 1u;
        key[i] = kk;
    }
    uint64_t tag = i < n ? (((uint64_t)kk << 2) | cat) : ~0ull;
    uint32_t len = run_length_if_head(tag);
    if (len && i < n && cat) atomicAdd(cat == 2u ? &h2[kk] : &cnt[kk], len);
}

__global__ void k_first(const uint32_t *kept, const uint32_t *S, const uint32_t *key, uint64_t n,
                        uint32_t *ck, uint32_t *firstj)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    bool on = i < n && kept[i];
    uint32_t k = on ? key[i] : 0;
    uint32_t j = on ? S[i] : 0;
    if (on) ck[j] = k;
    uint64_t tag = on ? (uint64_t)k : (~0ull - (i & 1ull));
    uint32_t len = run_length_if_head(tag);
    if (len && on) atomicMin(&firstj[k], j);
}

__global__ void k_changes(const uint32_t *ck, uint32_t m, uint32_t *changes)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = (j > 0 && j < m && ck[j] != ck[j - 1]) ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(changes, c);
}

// (key, index) pairs of every entry for the stable sort of interleaved input: entries that
// are not particles take the key `past`, after every track
__global__ void k_sort_input(const uint32_t *kept, const uint32_t *key, uint64_t n, uint32_t past, uint32_t *keys,
                             uint32_t *vals)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = kept[i] ? key[i] : past;
    vals[i] = (uint32_t)i;
}

// input grouped by primary: an entry's rank in its track is its rank among the kept
// entries minus that of the track's first one
__global__ void k_bucket(const uint32_t *kept, const uint32_t *S, const uint32_t *key, uint64_t n,
                         const uint32_t *bstart, const uint32_t *firstj, uint32_t *idx)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n || !kept[i]) return;
    uint32_t k = key[i];
    idx[bstart[k] + S[i] - firstj[k]] = (uint32_t)i;
}

// ------------------------------------------------------------ per-track kernels
// Two launches per batch, both in launch order (largest tracks first):
//   k_irt_pairs  one workgroup per track (64, 128 or 256 threads, the size autotuned): the
//                displacement of the initial species, the sorted copy (in LDS, by species
//                and hash cell), the initial pairs of every reaction channel and their FP32
//                screens; the pairs that may react are queued in HBM with their FP64
//                distance, then the whole workgroup samples them (the drain: the rest of
//                sampleIRT, SamplePDC's rejection loops included) into the event list.  A
//                track's drain overlaps the pair enumeration of the other workgroups on its CU.
//   k_irt_loop   one wavefront per track, 16 tracks per CU (128 VGPRs, under 10 KB of LDS):
//                the event loop.  The
//                next event comes from per-block minima of the event list kept in LDS: the
//                global minimum is popped, and when it is stale (a reactant has reacted) it is
//                dropped and only its block is rescanned (lazy deletion).  The events that one
//                reaction's re-pairing pushes are staged in LDS and merged into the minima.
//                A wavefront needs no workgroup barrier between the serial steps of a
//                reaction, and 16 independent event chains share each CU to hide their
//                latency.
#define IRT_BLOCK 256                // largest workgroup of k_irt_pairs
#define IRT_NT (blockDim.x)
#define BM_MAX 128u                  // event blocks per track at most (a block holds 64 2^k events)
#define STAGE_MAX 64u                // events staged per reaction (more: blocks rescanned)
#define AL_LDS_MAX 2048u             // bytes of LDS for the loop kernel's alive bits (16,384 slots)
#define TSTATE 128                   // u32 per track handed from k_irt_pairs to k_irt_loop:
                                     // [0] queued pairs, [1..6] species list lengths,
                                     // [7..9] initial OH, e-aq, H, [10] H2 markers, [11] n,
                                     // [12] windows on, [13..15] axis, [16] umin, [17] cinv,
                                     // [18] cells, [19] cells per bucket, [20] buckets,
                                     // [21] events of the initial pairs drained by k_irt_pairs,
                                     // [32..127] bucket ends, 2 x u16 per word, 16 words per species
#define WB_MAX 32u                   // buckets of the re-pairing windows

__device__ __forceinline__ bool ev_less(double t1, uint64_t ab1, double t2, uint64_t ab2)
{
    return t1 < t2 || (t1 == t2 && ab1 < ab2);
}

// Wavefront minimum of (t, ab) in every lane; returns the lane that holds it.  Event times
// are non-negative, so their bit patterns order like the values: one 64-bit butterfly
// finds the least time, and only lanes sharing it (a tie, e.g. contact reactions at the
// same instant) take a second butterfly on ab.
// Cross-lane steps of the wavefront (all 64 lanes active) with DPP moves and lane reads:
// no LDS round trip per step, unlike shuffles.
// A 64-bit value of the lane given by a DPP control (the lane's own where the pattern has no
// source).
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x)
{
    const int lo = (int)(uint32_t)x, hi = (int)(uint32_t)(x >> 32);
    const uint32_t l = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    const uint32_t h = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return ((uint64_t)h << 32) | l;
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, int lane)
{
    const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    const uint32_t h = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
    return ((uint64_t)h << 32) | l;
}
__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
// the least value over the wavefront, in every lane: minima within quads, half-rows and rows
// of 16 lanes by DPP, then over the four rows
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x)
{
    x = umin64(x, dpp_u64<0xB1>(x));            // quad_perm [1, 0, 3, 2]
    x = umin64(x, dpp_u64<0x4E>(x));            // quad_perm [2, 3, 0, 1]
    x = umin64(x, dpp_u64<0x141>(x));           // row_half_mirror
    x = umin64(x, dpp_u64<0x140>(x));           // row_mirror
    return umin64(umin64(readlane_u64(x, 0), readlane_u64(x, 16)), umin64(readlane_u64(x, 32), readlane_u64(x, 48)));
}
// inclusive prefix sum over the wavefront's lanes: within rows of 16 by shifted DPP adds
// (lanes without a source add 0), then row 0's and rows 0-1's totals into the rows above
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);    // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);    // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);    // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);    // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);   // row_bcast:15 -> rows 1, 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);   // row_bcast:31 -> rows 2, 3
    return x;
}

__device__ __forceinline__ uint32_t wave_min_ev(double &t, uint64_t &ab)
{
    const uint64_t tk = wave_min_u64((uint64_t)__double_as_longlong(t));
    const bool at = (uint64_t)__double_as_longlong(t) == tk;
    const uint64_t m = __ballot(at);
    uint64_t k;
    if (m & (m - 1ull)) k = wave_min_u64(at ? ab : ~0ull);  // a tie on t: least ab among the tied lanes
    else k = readlane_u64(ab, (int)__ffsll((long long)m) - 1);
    const uint32_t src = (uint32_t)__ffsll((long long)__ballot(at && ab == k)) - 1u;
    t = __longlong_as_double((long long)tk);
    ab = k;
    return src;
}

// shared state of k_irt_pairs
#define NX_MAX 128u                       // cells per track of the initial-pair hash
struct PairShared {
    uint32_t npend;                       // queued pairs
    uint32_t gcnt[8], goff[8];            // species groups of the sorted copy
    int      bri[9], nblk;                // channel blocks (reactions with particles)
    unsigned long long c[3];              // initial OH, e-aq, H
    // the hash: nx cells of width >= CELL_R / m along the unit axis ax, starting at umin.
    // The sorted copy is ordered by (species, cell); once it is filled, cs[s * nx + k] is
    // the end of the run of species s in cell k (its start is the previous entry's end).
    uint32_t nx, m;
    float    ax[3], umin, cinv;
    double   red[4][9];                   // per-wavefront partial sums
    float    rlo[4], rhi[4];
    uint32_t cs[6 * NX_MAX + 1];
    uint32_t cu[NX_MAX + 1];              // units of the current channel block before each cell
    // the drain of the queued pairs (pdc_drain): event count, queue and long-list cursors,
    // the long list (in the dynamic LDS of the sorted copy, no longer needed by then)
    uint32_t n_ev, pend_next, long_n, long_next, long_cap;
    uint32_t *long_item;
    int8_t   rmap[36];
    double   Dd[6];
};

// shared state of k_irt_loop (one wavefront)
struct LoopShared {
    static constexpr uint32_t LONG_CAP = 128;   // >= 2 x 64, the small-queue drain lists every pair
    uint32_t n_ev;
    uint32_t n_total;
    double   win_t;
    uint64_t win_ab;
    int      np;
    uint32_t prod_idx[2];
    unsigned long long cnt[5];   // running OH, e-aq, H, H2O2, H2
    unsigned long long rxc[9];   // reaction counts
    unsigned long long n_re;
    uint32_t npend;              // queued stage-2 (SamplePDC) pairs
    uint32_t pend_next;          // drain cursor
    uint32_t long_n, long_next;  // pairs handed from the per-lane to the wavefront schedule
    uint32_t long_item[LONG_CAP];
    // reaction tables by species pair (row * 6 + col) and by species, copied from c_irt
    // so that lanes with different species read them from LDS
    RxB      bnd[36];                     // cannot_react bounds of the pair's reaction
    int8_t   rmap[36];                    // reaction index or -1
    int8_t   nprod[9], prods[9][2];
    double   Dd[6], sqD[6];
    float    Df[6];
    // the products of the current reaction
    uint32_t prod_sp[2];
    float4   prod_pos[2];
    // species lists: length and reacted entries; products are appended, reacted particles
    // stay until a list is compacted (when they outnumber the alive ones)
    uint32_t llen[6], ldead[6], lcomp;
    // re-pairing windows (win): the first n0[s] entries of list s are initial particles in
    // hash-cell order, in buckets of wg cells; wend[s][b] is the end of bucket b
    uint32_t win, wnx, wg, wnb;
    float    wax[3], wumin, wcinv;
    uint32_t n0[6];
    uint16_t wend[6][WB_MAX];
    // re-pairing segments (product p, partner species sj, a list range) flattened
    int      nseg;
    uint32_t seg_u0[25];                  // first flattened index
    uint32_t seg_base[24];                // slot of flattened index u: seg_base + u (mod 2^32)
    uint8_t  seg_ps[24];                  // p << 4 | part << 3 | sj (part 1: appended entries)
    // re-pairings that passed the bound: the partner's list entry (x, y, z) and
    // p << 31 | sj << 28 | index
    uint32_t rq[4][128];
    uint32_t n_init;                      // particles 0 .. n_init - 1 are initial (born at 0)
    float sg0[6];                         // per species: the step bound sig of a partner born at 0
    // event selection: blocks of bs events, the minimum live (t, a << 32 | b) of each block
    // (a minimum may have gone stale since), and the events pushed in the current reaction
    uint32_t bs, nblk, e0;
    double   bm_t[BM_MAX];
    uint64_t bm_ab[BM_MAX];
    uint32_t nt_n;
    uint32_t nt_e[STAGE_MAX];
    double   nt_t[STAGE_MAX];
    uint64_t nt_ab[STAGE_MAX];
    uint32_t snap[DNA_MAX_CKPT * 5];      // species counts at each checkpoint
    // per-track sums kept here, not in registers: clock ticks of the event scan, the
    // reaction, the re-pairing and its drain, loop iterations, the worker's re-pairings,
    // re-pairing list slots scanned
    unsigned long long acc[7];
};

// shared state of k_irt_drain (the initial pairs' drain in a kernel of its own)
struct DrainShared {
    uint32_t n_ev, npend, pend_next, long_n, long_next;
    uint32_t long_item[1024];                   // pairs handed from the per-lane to the wavefront schedule
    int8_t   rmap[36];
    double   Dd[6];
};

extern __shared__ float4 dyn_lds[];
typedef __attribute__((address_space(3))) volatile uint32_t lds_vu32;

// The loop kernel's alive flags, one bit per particle slot of its track (the species of a
// slot is P[k].w).  LDS_BITS: in the dynamic LDS, so that every test and update is an LDS
// access (a flag array reachable through either address space compiles to flat accesses,
// whose waits also wait for every global load in flight); else in global memory, in the
// sorted-copy region, which the pairs kernel no longer needs (launches whose largest track
// exceeds the LDS budget).  Updated by one lane at a time, read by all after the wavefront
// synchronises.
template <bool LDS_BITS>
struct AliveBits {
    uint32_t *g;                                      // the global words (LDS_BITS false)
    __device__ __forceinline__ uint32_t *w() const { return LDS_BITS ? (uint32_t *)dyn_lds : g; }
    __device__ __forceinline__ bool operator[](uint32_t k) const { return (w()[k >> 5] >> (k & 31u)) & 1u; }
    __device__ __forceinline__ void clear(uint32_t k) const { w()[k >> 5] &= ~(1u << (k & 31u)); }
    __device__ __forceinline__ void set(uint32_t k) const { w()[k >> 5] |= 1u << (k & 31u); }
};

// words of the alive bits of a track of cap slots (whole 64-slot chunks)
__host__ __device__ __forceinline__ uint32_t alive_words(uint32_t cap) { return 2u * ((cap + 63u) / 64u); }

// capacity of the long list of the SamplePDC drain
__device__ __forceinline__ uint32_t long_cap(const LoopShared &) { return LoopShared::LONG_CAP; }
__device__ __forceinline__ uint32_t long_cap(const PairShared &sh) { return sh.long_cap; }
__device__ __forceinline__ uint32_t long_cap(const DrainShared &) { return 1024u; }

// An event of the loop kernel: stored in HBM and staged for the block minima.
template <class SH>
__device__ __forceinline__ void push_event(const IrtArgs &A, const IrtTrack &T, SH &sh, double t, uint32_t a,
                                           uint32_t b)
{
    const uint32_t e = atomicAdd(&sh.n_ev, 1u);
    if (e < T.ev_cap) {
        A.ev_t[T.ev_off + e] = t; A.ev_ab[T.ev_off + e] = make_uint2(a, b);
        if constexpr (std::is_same<SH, LoopShared>::value) {
            const uint32_t s = atomicAdd(&sh.nt_n, 1u);
            if (s < STAGE_MAX) { sh.nt_e[s] = e; sh.nt_t[s] = t; sh.nt_ab[s] = ((uint64_t)a << 32) | b; }
        }
    } else atomicOr(A.overflow, 2u);
}

// A queued pair: indices a < b with their species in the top 4 bits (track capacities are
// below 2^28 particles), and r0.
template <class SH>
__device__ __forceinline__ void push_pending(const IrtArgs &A, const IrtTrack &T, SH &sh,
                                             uint32_t a, uint32_t b, uint32_t sa, uint32_t sb, double r0)
{
    uint32_t q = atomicAdd(&sh.npend, 1u);
    if (q < T.ev_cap)
        A.pend[T.ev_off + q] = make_uint4(a | (sa << 28), b | (sb << 28), (uint32_t)__double2loint(r0),
                                          (uint32_t)__double2hiint(r0));
    else atomicOr(A.overflow, 2u);
}

// Stage 2 for queued pairs: the complete sampleIRT (irt-worker.js:116-184) — stage 1
// again (erfcinv for a reacting TDC pair, the contact test), then, for PDC pairs that
// survive the first draw, the SamplePDC rejection loop (irt-worker.js:91-113).  An
// accepted X gives t = X / D; the event is kept if t_base + t < 1 us.  Two schedules:
//  * per-lane: each lane owns one queued pair and takes the next as soon as its own
//    trial loop ends (large queues: the initial pairs of a track);
//  * per-wavefront: the 64 lanes run 64 consecutive trials of one pair at once and the
//    first accepted trial wins, the same X as the sequential loop (small queues: the
//    re-pairing of one reaction's products).
// Called by all threads after a barrier; the caller resets the queue.
struct PendPair { uint32_t a, b; double r0; int ri; double D; uint64_t ps; };

// A queued pair with its r0 (r0 < 0: a re-pairing beyond R_CUT, no event).
template <class SH>
__device__ __forceinline__ PendPair load_pending(const IrtArgs &A, const IrtTrack &T, const SH &sh,
                                                 uint64_t tk, uint32_t item)
{
    uint4 e = A.pend[T.ev_off + item];
    PendPair q;
    q.a = e.x & 0x0FFFFFFFu; q.b = e.y & 0x0FFFFFFFu;
    q.r0 = __hiloint2double((int)e.w, (int)e.z);
    const uint32_t si = e.x >> 28, sj = e.y >> 28;
    q.ri = sh.rmap[si * 6 + sj];
    q.D = sh.Dd[si] + sh.Dd[sj];
    q.ps = dna_pair_state(tk, q.a, q.b);
    return q;
}

// SamplePDC constants of one pair (irt-worker.js:91-94), with the per-trial subexpressions
// that do not depend on the trial computed once: S = p + q M, the branch threshold p / S
// and 2 b / a (the same operations on the same operands as inside the trial, so the values
// are identical)
struct PdcC { double a, b, p, q, M, S, thr, tba; };
__device__ __forceinline__ PdcC pdc_setup(const PendPair &pr)
{
    PdcC c;
    pdc_ab(pr.r0, pr.ri, pr.D, c.a, c.b);
    c.p = 2.0 * sqrt(2.0 * c.b / c.a);
    c.q = 2.0 / sqrt(2.0 * c.b / c.a);
    const double M1 = 1.0 / (c.a * c.a), M2 = 3.0 * c.b / c.a;
    c.M = M1 > M2 ? M1 : M2;
    c.S = c.p + c.q * c.M;
    c.thr = c.p / c.S;
    c.tba = 2.0 * c.b / c.a;
    return c;
}

// One SamplePDC trial (irt-worker.js:95-111), first part: the proposal X and the second
// uniform U, and a decision by bounds in FP32: 0 reject, 1 accept, 2 undecided (pdc_exact
// decides).  The acceptance value is lam = E f with E = exp(-b^2/X),
// f = 1 - a sqrt(pi X) erfcx(z), z = b/sqrt(X) + a sqrt(X).  Most trials are decided by the
// bounds, whose margins are far above FP32 rounding:
//  * always lam <= E (a > 0, erfcx > 0): E is taken with a 1e-4 relative and 1e-37
//    absolute margin;
//  * for z < 6, 1/(z + sqrt(z^2 + 2)) < (sqrt(pi)/2) erfcx(z) <= 1/(z + sqrt(z^2 + 4/pi))
//    (Abramowitz & Stegun 7.1.13) bracket f; z and a sqrt(pi X) are taken with 1e-5
//    relative margins (the bounds decrease in z) and f with 1e-5 absolute ones;
//  * for z >= 6, f is small (1 - a sqrt(pi X) erfcx(z) cancels to 1e-5 .. 1e-14 in most
//    trials there) and is bounded without the cancellation (f_asym): 1e-5 relative and
//    1e-13 absolute margins.
// (The bounds take their FP32 divisions, square roots and the reciprocal of X with the
// hardware instructions, each within a few ulp (< 1e-6 relative): far inside the 1e-5 and
// 1e-4 margins, and the exact path decides every trial they do not.)
__device__ __forceinline__ float fdiv_bound(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }

// f = 1 - a sqrt(pi X) erfcx(z) for large z without its cancellation: with
// S = sqrt(pi) z erfcx(z) and z sqrt(X) = b + a X,
//   f = 1 - S a X / (b + a X) = (b + a X (1 - S)) / (b + a X),
// and from the asymptotic series of erfc (Abramowitz & Stegun 7.1.23; for real z its
// remainder has the sign of, and is at most, the first omitted term)
//   1 - S = r/2 - 3r^2/4 + 15r^3/8 - 105r^4/16 + R,  0 <= R <= 945 r^5/32,  r = 1/z^2.
// Every term is positive (b >= 0), so FP32 keeps f to a few 1e-7 relative even where f is
// 1e-14 (the FP32 evaluation of the bracket [lo, hi] adds at most ~2e-6 relative; the
// callers' margins cover it).  The FP64 routine's f carries the rounding of its own
// cancellation, at most ~1e-15 absolute (its erfcx and a sqrt(pi X) to a few ulp): the
// callers' 1e-13 absolute margin covers that.
__device__ __forceinline__ void f_asym(float aX, float b, float z, float &lo, float &hi)
{
    const float r = __builtin_amdgcn_rcpf(z * z);
    const float t4 = r * (0.5f - r * (0.75f - r * (1.875f - r * 6.5625f)));
    const float rm = 29.53125f * r * r * r * r * r;
    const float den = __builtin_amdgcn_rcpf(aX + b);
    lo = (b + aX * t4) * den;
    hi = (b + aX * (t4 + rm)) * den;
}

__device__ __forceinline__ int pdc_fast(const PdcC &c, uint64_t ps, uint32_t trial, double &X, double &U2)
{
    // draws 7 + 2 trial and 8 + 2 trial of the pair stream: dna_u01(ps, n) is the finaliser
    // of ps + G (n + 1), so the second argument is the first plus G
    const uint64_t x0 = ps + DNA_GOLDEN * (uint64_t)(8u + 2u * trial);
    const double U = (double)(dna_mix64(x0) >> 11) * 1.1102230246251565404e-16;
    double v;
    if (U < c.thr) v = U * c.S / 2;
    else v = 2 / ((1 - U) * c.S / c.M);
    X = v * v;
    U2 = (double)(dna_mix64(x0 + DNA_GOLDEN) >> 11) * 1.1102230246251565404e-16;
    const bool c1 = X <= c.tba, c2 = X >= c.tba;
    const float Xf = (float)X, rX = __builtin_amdgcn_rcpf(Xf);
    // b^2 / X to a few ulp: the exponent's error (< 3e-7 |b^2/X|) moves E by under 3e-5
    // of itself while E is above the 1e-37 absolute margin (|b^2/X| < 87)
    const float E = expf(-((float)(c.b * c.b) * rX));
    const float Ehi = E * 1.0001f + 1e-37f;
    const float fU = (float)U2;
    // the value compared with lam: U (X below 2b/a) or U M / X (above); X = 2b/a exactly
    // (both tests) and NaN go to the exact path
    const float w = c1 ? fU : fU * (float)c.M * rX;
    const float whi = w * 1.0001f, wlo = w * 0.9999f;
    const float bf = (float)c.b;
    const float sq = __builtin_amdgcn_sqrtf(Xf);
    const float z = fdiv_bound(bf, sq) + (float)c.a * sq;
    const float K = (float)c.a * sq * 1.7724539f;                       // a sqrt(pi X)
    const float zl = z * 0.99999f, zh = z * 1.00001f;
    const float Lh = fdiv_bound(1.1283792f, zh + __builtin_amdgcn_sqrtf(zh * zh + 2.f));        // erfcx(z) >  Lh
    const float Ul = fdiv_bound(1.1283792f, zl + __builtin_amdgcn_sqrtf(zl * zl + 1.2732395f)); // erfcx(z) <= Ul
    float fhi = 1.f - K * 0.99999f * Lh * 0.99999f + 1e-5f;
    float flo = 1.f - K * 1.00001f * Ul * 1.00001f - 1e-5f;
    float gl, gh;
    f_asym((float)c.a * Xf, bf, z, gl, gh);
    if (z >= 6.f) { fhi = gh * 1.00001f + 1e-13f; flo = gl * 0.99999f - 1e-13f; }
    const float lhi = Ehi * fmaxf(fhi, 0.f) + 1e-37f;
    const float llo = (E * 0.9999f - 1e-37f) * fmaxf(flo, 0.f);
    const bool one = c1 != c2, rejE = wlo > Ehi, zok = bf >= 0.f;
    const bool rej = one && (rejE || (zok && wlo > lhi));
    const bool acc = one && !rejE && zok && !(wlo > lhi) && whi < llo && llo > 0.f;
    return rej ? 0 : acc ? 1 : 2;
}

// exp(z^2) erfc(z) for z >= 0 in FP32 with a relative bound m on its distance to dna_erfcx
// (include/dna_irt_math.h): below 6 the same Chebyshev series in FP32 (largest error
// 3.9e-7 over 4e6 points against the FP64 routine, from coefficient and operation rounding),
// from 6 on the alternating asymptotic series sqrt(pi) z erfcx(z) = 1 - 1/(2z^2) + 3/(4z^4)
// - 15/(8z^6) + 105/(16z^8), whose remainder is below the first omitted term 945/(32z^10)
// (at most 4.9e-7 from z = 6; the FP64 routine's 30-step continued fraction is far closer to
// the function than that).  m = 1e-5 on top covers the FP32 rounding and expf.
__constant__ float c_cheb32[28] = {
    -1.3026537197817094f, 6.4196979235649026e-1f, 1.9476473204185836e-2f,
    -9.561514786808631e-3f, -9.46595344482036e-4f, 3.66839497852761e-4f,
    4.2523324806907e-5f, -2.0278578112534e-5f, -1.624290004647e-6f,
    1.303655835580e-6f, 1.5626441722e-8f, -8.5238095915e-8f, 6.529054439e-9f,
    5.059343495e-9f, -9.91364156e-10f, -2.27365122e-10f, 9.6467911e-11f,
    2.394038e-12f, -6.886027e-12f, 8.94487e-13f, 3.13092e-13f, -1.12708e-13f,
    3.81e-16f, 7.106e-15f, -1.523e-15f, -9.4e-17f, 1.21e-16f, -2.8e-17f};
__device__ __forceinline__ void erfcx_bounds(float z, float &lo, float &hi)
{
    if (z < 6.f) {
        const float t = 2.f / (2.f + z), ty = 4.f * t - 2.f;
        float d = 0.f, dd = 0.f;
        for (int j = 27; j > 0; j--) { const float tmp = d; d = ty * d - dd + c_cheb32[j]; dd = tmp; }
        const float e = t * expf(0.5f * (c_cheb32[0] + ty * d) - dd);
        lo = e * 0.99999f; hi = e * 1.00001f;
    } else {
        const float r = 1.f / (z * z);
        const float sm = 1.f - r * (0.5f - r * (0.75f - r * (1.875f - r * 6.5625f)));
        const float rem = 29.53125f * r * r * r * r * r;
        const float q = 0.56418958f / z;                   // 1 / (sqrt(pi) z)
        lo = (sm - rem) * q * 0.99999f; hi = (sm + rem) * q * 1.00001f;
    }
}

// Second decision for a trial the first bounds left undecided (pdc_fast returned 2): the
// exact path's own FP64 quantities b^2/X, z, a X and a sqrt(pi X) rounded to FP32 (relative
// error 6e-8: taken with 1e-6 margins; the exponent b^2/X below 87 moves E by under 5.5e-6
// with expf, taken with 5e-5); for z < 6 erfcx from erfcx_bounds at the ends of z's
// interval (erfcx decreases), for z >= 6 f from f_asym with 4e-6 relative and 1e-13
// absolute margins; U (M / X) in FP64.  0 reject, 1 accept, 2 still undecided (pdc_exact).
__device__ __forceinline__ int pdc_refine(const PdcC &c, double X, double U)
{
    const double a = c.a, b = c.b;
    const bool c1 = X <= 2.0 * b / a, c2 = X >= 2.0 * b / a;
    if (c1 == c2 || !(b >= 0.0)) return 2;
    const double sqX = sqrt(X);
    const float E = expf(-(float)(b * b / X));
    const float Ehi = E * 1.00005f + 1e-37f, Elo = E * 0.99995f - 1e-37f;
    const float z = (float)(b / sqX + a * sqX);
    float fhi, flo;
    if (z >= 6.f) {
        float gl, gh;
        f_asym((float)(a * X), (float)b, z, gl, gh);
        fhi = gh * 1.000004f + 1e-13f; flo = gl * 0.999996f - 1e-13f;
    } else {
        const float K = (float)(a * sqrt(DNA_PI * X));
        float lo, hi, lo2, hi2;
        erfcx_bounds(z * 1.000001f, lo, hi2);
        erfcx_bounds(z * 0.999999f, lo2, hi);
        fhi = 1.f - K * 0.999999f * lo + 1e-6f;
        flo = 1.f - K * 1.000001f * hi - 1e-6f;
    }
    const float lhi = Ehi * fmaxf(fhi, 0.f) + 1e-37f;
    const float llo = Elo * fmaxf(flo, 0.f);
    const float w = (float)(c1 ? U : U * c.M / X);
    const float whi = w * 1.000001f, wlo = w * 0.999999f;
    if (wlo > lhi) return 0;
    if (whi < llo && llo > 0.f) return 1;
    return 2;
}

// the exact acceptance test of a trial that the bounds left undecided
__device__ __forceinline__ bool pdc_exact(const PdcC &c, double X, double U)
{
    const double a = c.a, b = c.b, M = c.M;
    double sqX = sqrt(X);
    double lam = dna_exp(-b * b / X) * (1.0 - a * sqrt(DNA_PI * X) * dna_erfcx(b / sqX + a * sqX));
    return (X <= 2.0 * b / a && U <= lam) || (X >= 2.0 * b / a && U * M / X <= lam);
}

__device__ __forceinline__ bool pdc_trial(const PdcC &c, uint64_t ps, uint32_t trial, double &X)
{
    double U;
    int d = pdc_fast(c, ps, trial, X, U);
    if (d == 2) d = pdc_refine(c, X, U);
    return d == 2 ? pdc_exact(c, X, U) : d == 1;
}

#define LANE_TRIALS 64u

// per-lane schedule; a pair still running after LANE_TRIALS trials is handed to the
// wavefront schedule (sh.long_item / sh.long_n) and resumed there at that trial.  Lanes
// take new pairs together, every REFILL trials: a lane whose pair ends waits for the next
// refill, so the first stage and the SamplePDC setup (a global load and FP64 code) run once
// per refill for all idle lanes instead of in almost every trial.
#define REFILL 8u
template <class SH>
__device__ void pdc_drain_lanes(const IrtArgs &A, const IrtTrack &T, uint64_t tk, SH &sh,
                                uint32_t npend, double t_base)
{
    PdcC c; c.a = 1; c.b = 1; c.p = 0; c.q = 0; c.M = 0; c.S = 1; c.thr = 0; c.tba = 0;
    PendPair pp; pp.a = 0; pp.b = 0; pp.D = 1; pp.ps = 0; pp.r0 = 0; pp.ri = 0;
    uint32_t item = 0, trial = 0;
    bool have = false, more = true;
    for (;;) {
        if (!have && more) {
            for (;;) {                                   // pairs with no trials end here
                item = atomicAdd(&sh.pend_next, 1u);
                if (item >= npend) { more = false; break; }
                pp = load_pending(A, T, sh, tk, item);
                const double t1 = pp.r0 < 0 ? -1.0 : irt_stage1(pp.r0, pp.ri, pp.D, pp.ps);
                if (t1 == STAGE_PDC) { c = pdc_setup(pp); trial = 0; have = true; break; }
                if (t1 >= 0) { double ta = t_base + t1; if (ta < DNA_T_END_NS) push_event(A, T, sh, ta, pp.a, pp.b); }
            }
        }
        if (!__ballot(have)) break;
        for (uint32_t k = 0; k < REFILL; k++) {
            if (!have) continue;
            double X = -1;
            const bool acc = pdc_trial(c, pp.ps, trial, X);
            trial++;
            if (acc) {
                if (X > 0) { double ta = t_base + X / pp.D; if (ta < DNA_T_END_NS) push_event(A, T, sh, ta, pp.a, pp.b); }
                have = false;
            } else if (trial == LANE_TRIALS) {
                // to the wavefront schedule; when its list is full the lane goes on itself
                const uint32_t l = atomicAdd(&sh.long_n, 1u);
                if (l < long_cap(sh)) { sh.long_item[l] = item; have = false; }
            } else if (trial == 10000u) have = false;
        }
    }
}

// wavefront schedule over items [from the queue, or the long list when use_long; those
// resume at trial long_start]
template <class SH>
__device__ void pdc_drain_waves(const IrtArgs &A, const IrtTrack &T, uint64_t tk, SH &sh,
                                uint32_t npend, double t_base, bool use_long, uint32_t long_start = LANE_TRIALS)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nlong = sh.long_n < long_cap(sh) ? sh.long_n : long_cap(sh);
    for (;;) {
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(use_long ? &sh.long_next : &sh.pend_next, 1u);
        slot = (uint32_t)__builtin_amdgcn_readlane((int)slot, 0);
        if (slot >= (use_long ? nlong : npend)) break;
        uint32_t item = use_long ? sh.long_item[slot] : slot;
        PendPair pp = load_pending(A, T, sh, tk, item);
        uint32_t start = 0;
        if (!use_long) {
            double t1 = pp.r0 < 0 ? -1.0 : irt_stage1(pp.r0, pp.ri, pp.D, pp.ps);
            if (t1 != STAGE_PDC) {
                if (lane == 0 && t1 >= 0) { double ta = t_base + t1; if (ta < DNA_T_END_NS) push_event(A, T, sh, ta, pp.a, pp.b); }
                continue;
            }
        } else start = long_start;
        const PdcC c = pdc_setup(pp);
        double X = -1;
        for (uint32_t base = start; base < 10000u; base += 64u) {
            uint32_t trial = base + lane;
            double Xl = 0;
            bool acc = false;
            if (trial < 10000u) acc = pdc_trial(c, pp.ps, trial, Xl);
            unsigned long long m = __ballot(acc);
            if (m) {
                const int first = __ffsll((long long)m) - 1;
                X = __longlong_as_double((long long)readlane_u64((uint64_t)__double_as_longlong(Xl), first));
                break;
            }
        }
        if (lane == 0 && X > 0) { double ta = t_base + X / pp.D; if (ta < DNA_T_END_NS) push_event(A, T, sh, ta, pp.a, pp.b); }
    }
}

// small queues (fewer than twice the workgroup's threads, so within the long list): the first
// stage of every pair on the first wavefront, a lane per pair; the pairs that need SamplePDC
// trials are listed and run on the wavefront schedule from trial 0
template <class SH>
__device__ void pdc_drain_small(const IrtArgs &A, const IrtTrack &T, uint64_t tk, SH &sh,
                                uint32_t npend, double t_base)
{
    if (threadIdx.x < 64u) {
        for (uint32_t item = threadIdx.x; item < npend; item += 64u) {
            const PendPair pp = load_pending(A, T, sh, tk, item);
            const double t1 = pp.r0 < 0 ? -1.0 : irt_stage1(pp.r0, pp.ri, pp.D, pp.ps);
            if (t1 == STAGE_PDC) sh.long_item[atomicAdd(&sh.long_n, 1u)] = item;
            else if (t1 >= 0) { double ta = t_base + t1; if (ta < DNA_T_END_NS) push_event(A, T, sh, ta, pp.a, pp.b); }
        }
    }
    __syncthreads();
    if (sh.long_n) pdc_drain_waves(A, T, tk, sh, npend, t_base, true, 0u);
}

// drain the whole queue; all threads call it after a barrier
template <class SH>
__device__ void pdc_drain(const IrtArgs &A, const IrtTrack &T, uint64_t tk, SH &sh, double t_base)
{
    const uint32_t npend = sh.npend < T.ev_cap ? sh.npend : T.ev_cap;
    if (npend >= 2u * IRT_NT) {
        pdc_drain_lanes(A, T, tk, sh, npend, t_base);
        __syncthreads();
        pdc_drain_waves(A, T, tk, sh, npend, t_base, true);
    } else {
        pdc_drain_small(A, T, tk, sh, npend, t_base);
    }
}

// The exact part for one staged re-pairing (product p, partner jj of species sj at the
// list entry's position pj, the partner's own position): the partner's Brownian step, r0,
// the first sampling stage, and the queueing of pairs that can react.  Only a partner born
// in a reaction reads its birth time; the initial particles are born at 0.
__device__ __forceinline__ void repair_exact(const IrtArgs &A, const IrtTrack &T, LoopShared &sh, uint64_t tk,
                                             double t_new, float4 pj, uint32_t e)
{
    const double *tb = A.tb + T.part_off;
    const uint32_t p = e >> 31, sj = (e >> 28) & 7u, jj = e & 0x0FFFFFFFu;
    const float4 pi = sh.prod_pos[p];
    const uint32_t sp = sh.prod_sp[p], pc = sp * 6u + sj;
    const int rj = sh.rmap[pc];
    const uint32_t idx = sh.prod_idx[p];
    const uint32_t a = idx < jj ? idx : jj, b = idx < jj ? jj : idx;
    const uint64_t ps = dna_pair_state(tk, a, b);
    const double dt = t_new - (jj < sh.n_init ? 0.0 : tb[jj]);
    {
        // FP32 screen of the same step: each normal g = sqrt(-2 ln(u + 1e-30)) cos(6.2831853 u')
        // is within 2e-3 + 1e-4 |g| of the FP64 one (hardware log2/cos in revolutions: the
        // input rounding moves -2 ln u by < 2e-7, which the square root turns into < 4.5e-4 near
        // 0; relative errors are ~1e-6), so r0 >= r0_f - M below, and a pair that the bound
        // rejects at r0_f - M is rejected by the exact test too
        float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z, sg = 0.f, gsum = 0.f;
        if (dt > 0) {
            sg = sqrt_bound((float)(2.0 * sh.Dd[sj] * dt));
            float g[3];
            for (int k = 0; k < 3; k++) {
                const float ua = (float)(dna_u01(ps, 2u * k) + 1e-30), ub = (float)dna_u01(ps, 2u * k + 1u);
                const float L = fmaxf(0.f, -1.3862944f * __builtin_amdgcn_logf(ua));      // -2 ln ua
                g[k] = sqrt_bound(L) * __builtin_amdgcn_cosf(ub);                        // cos(2 pi ub)
                gsum += fabsf(g[k]);
            }
            dx -= g[0] * sg; dy -= g[1] * sg; dz -= g[2] * sg;
        }
        const float r0f = sqrt_bound(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        const float M = sg * (6e-3f + 2e-4f * gsum) + 1e-5f * (r0f + sg * gsum) + 1e-6f;
        const float r0lo = r0f - M;
        if (r0lo > c_irt.R_CUT_F || cannot_react(sh.bnd[pc], r0lo, u01_lo(ps, 6u))) return;
    }
    double xj = pj.x, yj = pj.y, zj = pj.z;
    if (dt > 0) {
        double sig = sqrt(2.0 * sh.Dd[sj] * dt);
        double g0 = sqrt(-2 * dna_log(dna_u01(ps, 0) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 1));
        double g1 = sqrt(-2 * dna_log(dna_u01(ps, 2) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 3));
        double g2 = sqrt(-2 * dna_log(dna_u01(ps, 4) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 5));
        xj += g0 * sig; yj += g1 * sig; zj += g2 * sig;
    }
    double ddx = (double)pi.x - xj, ddy = (double)pi.y - yj, ddz = (double)pi.z - zj;
    double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
    if (r2 > c_irt.R_CUT2) return;
    const double r0 = sqrt(r2);
    if (!may_react_k(rx_constants(rj, sh.Dd[sp] + sh.Dd[sj]), r0, ps)) return;
    push_pending(A, T, sh, a, b, a == idx ? sp : sj, a == idx ? sj : sp, r0);
}

// Re-pair the products of the reaction at t_new with every alive particle that has a
// reaction channel with them (irt-worker.js:430-465).  The (product, partner species, list
// range) segments of one reaction (build_segments) are flattened into one index range over
// the species lists; list entries that have reacted are skipped.
// Product 0 is not paired with product 1: it does not exist yet at that point of the
// worker's loop (irt-worker.js:618-644).  Each pair first meets the necessary condition
// of cannot_react on a lower bound of its distance after the partner's Brownian step;
// the survivors are staged per wavefront and finished 64 at a time (the FP64 step,
// r0 = |product - partner| (irt-worker.js:443-456), the first sampling stage), so that
// part runs on full wavefronts; pairs that can react are queued for the drain.
template <class AL>
__device__ void pair_products(const IrtArgs &A, const IrtTrack &T, LoopShared &sh, const AL &al,
                              uint64_t tk, double t_new)
{
    const float *lbr = (const float *)A.lst + 6 * T.part_off;
    const float4 *lps = A.lpos + 6 * T.part_off;
    // one wavefront (the loop kernel's workgroup): the segment count, the flattened total and
    // the chunk index are scalars, so the chunk loop is a scalar loop
    const int ns = __builtin_amdgcn_readfirstlane(sh.nseg), np = __builtin_amdgcn_readfirstlane(sh.np);
    const uint32_t tot = __builtin_amdgcn_readfirstlane(sh.seg_u0[ns]);
    const uint32_t pidx1 = __builtin_amdgcn_readfirstlane(sh.prod_idx[1]);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    // written and read by different lanes between wave barriers: volatile, and typed as LDS
    // so that the accesses stay LDS instructions (a generic volatile pointer compiles to
    // flat accesses, which wait for every global load in flight)
    lds_vu32 *qx = (lds_vu32 *)sh.rq[0], *qy = (lds_vu32 *)sh.rq[1], *qz = (lds_vu32 *)sh.rq[2];
    lds_vu32 *qe = (lds_vu32 *)sh.rq[3];
    uint32_t qn = 0;                                   // wavefront-uniform
    // The list entries of the next 64 flattened indices are loaded before the current 64
    // are screened, so their latency overlaps the screen.  A lane's index only grows, so its
    // segment is found by moving forward from the previous one.
    int sg = 0;
    uint32_t ps_n = 0;
    float bt_n = 0.f;
    float4 pj_n = make_float4(0.f, 0.f, 0.f, 0.f);
    auto fetch = [&](uint32_t u) {
        while (sg + 1 < ns && u >= sh.seg_u0[sg + 1]) sg++;
        ps_n = sh.seg_ps[sg];
        const uint32_t slot = sh.seg_base[sg] + u;
        pj_n = lps[slot];                              // x, y, z, the particle index
        bt_n = (ps_n & 8u) ? lbr[slot] : 0.f;          // birth time rounded down (appended part)
    };
    if (lane < tot) fetch(lane);
    for (uint32_t u0 = 0; u0 < tot; u0 += 64u) {
        const uint32_t u = u0 + lane;
        const uint32_t psg = ps_n;
        const float4 pj = pj_n;
        const float bt = bt_n;
        if (u + 64u < tot) fetch(u + 64u);
        const uint32_t jj = __float_as_uint(pj.w);
        bool keep = false;
        uint32_t ent = 0;
        if (u < tot) {
            const uint32_t p = psg >> 4, sj = psg & 7u;
            const uint32_t idx = sh.prod_idx[p];
            if (al[jj] && jj != idx && !(p == 0 && np == 2 && jj == pidx1)) {
                const uint32_t pc = sh.prod_sp[p] * 6u + sj;
                const float4 pi = sh.prod_pos[p];
                const uint32_t a = idx < jj ? idx : jj, b = idx < jj ? jj : idx;
                const uint64_t ps = dna_pair_state(tk, a, b);
                // t_new - (birth time rounded down) >= the exact step time, so the bound's
                // step is at least the true one: the screen stays conservative.  A partner of
                // the initial part is born at 0: its step bound is the reaction's, per species
                double dt = t_new;
                float sgm = sh.sg0[sj];
                if (psg & 8u) {
                    dt = t_new - (double)bt;
                    sgm = dt > 0 ? sqrt_bound((float)(2.0 * (double)sh.Df[sj] * dt)) : 0.f;
                }
                float fx = pi.x - pj.x, fy = pi.y - pj.y, fz = pi.z - pj.z;
                const float r = sqrt_bound(fmaf(fz, fz, fmaf(fy, fy, fx * fx))) * 0.99999f;
                // first without any draw of the step: each randn() is at most
                // sqrt(-2 ln(1e-30)) = 11.754 in magnitude, so the step is at most 20.36 sig; a
                // partner beyond R_CUT by that much cannot pair, and the draw of the first
                // sampling stage alone may already rule the pair out
                const float rloose = r - sgm * 20.4f;
                ent = (p << 31) | (sj << 28) | jj;
                if (!(rloose > c_irt.R_CUT_F)) {
                    const float U = u01_lo(ps, 6u);
                    if (!cannot_react(sh.bnd[pc], rloose, U)) {
                        // then with the step's own draws: each randn() is at most
                        // sqrt(-2 ln(u + 1e-30)), so the step is at most
                        // sig * sqrt(-2 (ln u0 + ln u2 + ln u4)) (FP32, inflated by 1e-4;
                        // u01_lo <= the draw, so -ln of it is an upper bound; the three logs
                        // are taken as one of the FP32 product, within 2.4e-7 of their sum;
                        // u01_lo = 0 gives an infinite step and keeps the pair)
                        float rmin = r;
                        if (dt > 0) {
                            const float m2 = -2.f * logf(u01_lo(ps, 0) * u01_lo(ps, 2) * u01_lo(ps, 4));
                            rmin -= sgm * sqrt_bound(m2 * 1.0001f + 1e-4f) * 1.0001f;
                        }
                        keep = !(rmin > c_irt.R_CUT_F) && !cannot_react(sh.bnd[pc], rmin, U);
                    }
                }
            }
        }
        const uint64_t m = __ballot(keep);
        if (keep) {
            const uint32_t k = qn + (uint32_t)__popcll(m & below);
            qx[k] = __float_as_uint(pj.x); qy[k] = __float_as_uint(pj.y); qz[k] = __float_as_uint(pj.z); qe[k] = ent;
        }
        qn += (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (qn >= 64u) {
            repair_exact(A, T, sh, tk, t_new,
                         make_float4(__uint_as_float(qx[lane]), __uint_as_float(qy[lane]), __uint_as_float(qz[lane]), 0.f),
                         qe[lane]);
            __builtin_amdgcn_wave_barrier();
            if (lane < qn - 64u) {
                qx[lane] = qx[64u + lane]; qy[lane] = qy[64u + lane]; qz[lane] = qz[64u + lane]; qe[lane] = qe[64u + lane];
            }
            qn -= 64u;
            __builtin_amdgcn_wave_barrier();
        }
    }
    // the remainder (< 64)
    if (lane < qn)
        repair_exact(A, T, sh, tk, t_new,
                     make_float4(__uint_as_float(qx[lane]), __uint_as_float(qy[lane]), __uint_as_float(qz[lane]), 0.f),
                     qe[lane]);
    __syncthreads();
}

// Drop reacted entries from the species lists flagged in sh.lcomp (the loop kernel's one
// wavefront).  Stable: the entries keep their order, so the initial part of a list stays in
// cell order, and the ends of its buckets are moved with it.  In place: a chunk of 64
// entries is read (the loads complete before the ballot that the stores depend on) before
// any of its slots is rewritten, and an entry only ever moves to a lower position.
template <class AL>
__device__ void compact_lists(const IrtArgs &A, const IrtTrack &T, LoopShared &sh, const AL &al)
{
    uint32_t *lst = A.lst + 6 * T.part_off;
    float4 *lps = A.lpos + 6 * T.part_off;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t mask = sh.lcomp;
    for (int s = 0; s < 6; s++) {
        if (!(mask & (1u << s))) continue;
        uint32_t *L = lst + (size_t)s * T.cap;
        float4 *LP = lps + (size_t)s * T.cap;
        // the ranges: the buckets of the initial part in order (they end at n0), then the
        // appended products up to the list's length
        const uint32_t len = sh.llen[s], nb = sh.wnb;
        uint32_t dst = 0, i0 = 0, nkeep = 0;
        for (uint32_t r = 0; r <= nb; r++) {
            const uint32_t i1 = r < nb ? (uint32_t)sh.wend[s][r] : len;
            for (uint32_t c0 = i0; c0 < i1; c0 += 64u) {
                const uint32_t i = c0 + lane;
                const bool in = i < i1;
                const uint32_t bt = in ? L[i] : 0u;                 // birth time bits
                const float4 q = in ? LP[i] : make_float4(0.f, 0.f, 0.f, 0.f);
                const bool live = in && al[__float_as_uint(q.w)];
                const uint64_t m = __ballot(live);
                if (live) { const uint32_t d = dst + (uint32_t)__popcll(m & below); L[d] = bt; LP[d] = q; }
                dst += (uint32_t)__popcll(m);
            }
            __builtin_amdgcn_wave_barrier();
            if (r < nb) { if (lane == 0) sh.wend[s][r] = (uint16_t)dst; nkeep = dst; }
            i0 = i1;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) { sh.llen[s] = dst; sh.ldead[s] = 0; sh.n0[s] = nkeep; }
        __builtin_amdgcn_wave_barrier();
    }
}

// projection onto a hash axis (explicit FMAs: the same value wherever it is computed)
__device__ __forceinline__ float axis_u(float a0, float a1, float a2, const float4 &p)
{
    return fmaf(a2, p.z, fmaf(a1, p.y, a0 * p.x));
}

// the cell of axis coordinate u: the same expression as the pairs kernel's keys, monotone
// in u
__device__ __forceinline__ uint32_t hash_cell(float u, float umin, float cinv, uint32_t nx)
{
    const float c = (u - umin) * cinv;
    return c > 0.f ? (c < (float)(nx - 1) ? (uint32_t)c : nx - 1) : 0u;   // NaN: cell 0
}

// The re-pairing segments of the current reaction (the loop kernel's one wavefront): lane
// (p * 6 + sj) * 2 + part covers product p and partner species sj, part 0 the initial part
// of list sj, part 1 its appended products.  With windows, part 0 is cut to the buckets
// within reach of the product: an initial particle (born at t = 0) whose axis coordinate
// differs from the product's by more than reach = (R_CUT + 20.4 sig) (1.001) + 1 nm, with
// sig = sqrt(2 D_sj t_new) the bound of its Brownian step, is farther than R_CUT + 20.4 sig
// (the axis is a unit vector; the margins cover the FP32 coordinates), which is exactly the
// distance at which pair_products' first screen drops it without a draw: leaving it out
// changes nothing.  Also counts the re-pairings of the worker (every alive partner with a
// channel, public/irt-worker.js:430-465) and the list slots to scan.
__device__ __forceinline__ void build_segments(LoopShared &sh, double t_new, uint32_t cap)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t pc = lane >> 1, part = lane & 1u, p = pc / 6u, sj = pc - 6u * (pc / 6u);
    const int np = sh.np;
    if (lane < 6u) sh.sg0[lane] = t_new > 0 ? sqrt_bound((float)(2.0 * (double)sh.Df[lane] * t_new)) : 0.f;
    uint32_t lo = 0, len = 0, nalive = 0;
    if (lane < 24u && (int)p < np) {
        const uint32_t sp = sh.prod_sp[p];
        const uint32_t ll = sh.llen[sj], n0 = sh.n0[sj];
        if (sh.rmap[sp * 6 + sj] >= 0 && ll > 0) {
            if (part) {
                lo = n0; len = ll - n0;
                nalive = ll - sh.ldead[sj] - (sp == sj ? 1u : 0u) - ((p == 0 && np == 2 && sh.prod_sp[1] == sj) ? 1u : 0u);
            } else if (n0) {
                const float u = axis_u(sh.wax[0], sh.wax[1], sh.wax[2], sh.prod_pos[p]);
                const float sg = t_new > 0 ? sqrt_bound((float)(2.0 * (double)sh.Df[sj] * t_new)) * 1.00001f : 0.f;
                const float reach = (c_irt.R_CUT_F + 20.4f * sg) * 1.001f + 1.0f;
                const uint32_t blo = hash_cell(u - reach, sh.wumin, sh.wcinv, sh.wnx) / sh.wg;
                const uint32_t bhi = hash_cell(u + reach, sh.wumin, sh.wcinv, sh.wnx) / sh.wg;
                lo = blo ? sh.wend[sj][blo - 1] : 0u;
                len = sh.wend[sj][bhi] - lo;
            }
        }
    }
    const uint32_t x = wave_scan_u32(len);
    const uint64_t m = __ballot(len > 0);
    if (len > 0) {
        const uint32_t k = (uint32_t)__popcll(m & below);
        // list slot sj * cap + lo + (u - u0); below 2^32, as 6 cap is for particle indices
        // of 28 bits, so the 32-bit sum with the base taken modulo 2^32 is exact
        sh.seg_u0[k] = x - len; sh.seg_base[k] = sj * cap + lo - (x - len);
        sh.seg_ps[k] = (uint8_t)((p << 4) | (part << 3) | sj);
    }
    const uint32_t na = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_u32(nalive), 63);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    if (lane == 0) {
        const int ns = __popcll(m);
        sh.nseg = ns; sh.seg_u0[ns] = total;
        sh.acc[5] += na; sh.acc[6] += total;
    }
}

// exclusive scan of a[0..len) in LDS by wavefront 0, the total into a[len]; the caller
// synchronizes the workgroup before and after
__device__ __forceinline__ void wave0_scan(uint32_t *a, uint32_t len)
{
    if (threadIdx.x >= 64) return;
    const uint32_t lane = threadIdx.x;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < len; base += 64u) {
        const uint32_t i = base + lane;
        const uint32_t v = i < len ? a[i] : 0u;
        uint32_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (i < len) a[i] = carry + x - v;
        carry += __shfl(x, 63, 64);
    }
    if (lane == 0) a[len] = carry;
}

// the (species s, cell k) run of the sorted copy
__device__ __forceinline__ uint32_t run_end(const PairShared &sh, uint32_t s, uint32_t k) { return sh.cs[s * sh.nx + k]; }
__device__ __forceinline__ uint32_t run_start(const PairShared &sh, uint32_t s, uint32_t k)
{
    const uint32_t q = s * sh.nx + k;
    return q ? sh.cs[q - 1] : 0u;
}

// The pairs of one cell k of a channel block (species sa x sb): the rows are the sa
// particles of cell k, each row in units of 64 partners.
//  * sa != sb: the partners are the sb particles of cells k-m .. k+m;
//  * sa == sb: first the pairs inside cell k, enumerated circularly (row i with
//    i + d mod n for d = 1..(n-1)/2, and d = n/2 for i < n/2 when n is even: every
//    unordered pair once, with equal rows), then the particles of cells k+1 .. k+m.
// Every pair of particles whose cells are at most m apart is visited exactly once.
struct CellWin { uint32_t r0, nrow, cprA, cpr, wlo, wlen; };
__device__ __forceinline__ CellWin cell_window(const PairShared &sh, uint32_t sa, uint32_t sb, uint32_t k)
{
    CellWin w;
    const uint32_t nx = sh.nx, m = sh.m;
    const uint32_t khi = k + m < nx ? k + m : nx - 1;
    w.r0 = run_start(sh, sa, k);
    w.nrow = run_end(sh, sa, k) - w.r0;
    if (sa == sb) {
        w.cprA = w.nrow >= 2 ? ((w.nrow - 1) / 2 + 1 + 63) / 64 : 0u;
        w.wlo = run_end(sh, sa, k);
        w.wlen = run_end(sh, sa, khi) - w.wlo;
    } else {
        w.cprA = 0;
        w.wlo = run_start(sh, sb, k > m ? k - m : 0u);
        w.wlen = run_end(sh, sb, khi) - w.wlo;
    }
    w.cpr = w.cprA + (w.wlen + 63) / 64;
    return w;
}

// Initial pairs of the channel blocks (see k_irt_pairs step 2), with the sorted copy in
// LDS (SG_LDS: DS loads) or in global memory.  A pair is dropped without FP64 work when an
// FP32 bound decides it: the FP32 squared distance is within 3.1e-7 (relative) of the FP64
// one (the differences' roundings, doubled by the squares, and the three roundings of the
// product and the two FMAs, all of non-negative terms), so r2 above R_CUT^2 (1 + 1e-5) or below
// R_CUT^2 (1 - 1e-5) in FP32 settles the cutoff test, and the bound of cannot_react takes
// sqrt(r2) (1 - 1e-5) and the draw's 23-bit lower bound.  The exact FP64 distance is
// computed near the cutoff and for the pairs that are queued.  The cells leave out only
// pairs farther apart than R_CUT (see the hash in k_irt_pairs), so the queued pairs are
// those of the full enumeration.
template <bool SG_LDS>
__device__ __forceinline__ void pair_blocks(const IrtArgs &A, const IrtTrack &T, PairShared &sh, uint64_t tk,
                                            unsigned long long &wcand, unsigned long long &wsamp)
{
    const float4 *Sg = SG_LDS ? (const float4 *)dyn_lds : A.sorted + T.part_off;
    // block loop outermost: the reaction and its constants are wavefront-uniform
    // (readfirstlane -> scalar registers), so the per-pair path has no table loads.  The
    // wavefront index is a scalar too: the unit loop, its cell search and the unit's row and
    // chunk then run on the scalar unit, not as per-lane copies on the VALU.
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nb = __builtin_amdgcn_readfirstlane(sh.nblk);
    uint32_t ncand = 0, nsamp = 0;                    // this lane's, per channel block
    const uint32_t nx = __builtin_amdgcn_readfirstlane(sh.nx);
    const int BA[9] = {0, 0, 0, 1, 1, 1, 2, 1, 3}, BB[9] = {0, 1, 2, 1, 2, 3, 2, 4, 5};
    const double RC2 = c_irt.R_CUT2;
    const float RC2_hi = (float)(RC2 * 1.00001), RC2_lo = (float)(RC2 * 0.99999);
    for (int q = 0; q < nb; q++) {
        const int ri = __builtin_amdgcn_readfirstlane(sh.bri[q]);
        const uint32_t sa = BA[ri], sb = BB[ri];
        // units of each cell, then their running sum
        for (uint32_t k = tid; k < nx; k += IRT_NT) {
            const CellWin w = cell_window(sh, sa, sb, k);
            sh.cu[k] = w.nrow * w.cpr;
        }
        __syncthreads();
        wave0_scan(sh.cu, nx);
        __syncthreads();
        const uint32_t units = __builtin_amdgcn_readfirstlane(sh.cu[nx]);
        const RxB kb = rx_bound(rx_constants(ri, c_irt.D[sa] + c_irt.D[sb]));
        uint32_t k = 0, ku0 = 0, ku1 = sh.cu[1];
        CellWin w = cell_window(sh, sa, sb, 0);
        for (uint32_t ul = wave; ul < units; ul += IRT_NT / 64u) {
            if (ul >= ku1) {                          // the next cell with units
                do { k++; ku0 = ku1; ku1 = sh.cu[k + 1]; } while (ul >= ku1);
                w = cell_window(sh, sa, sb, k);
            }
            const uint32_t loc = ul - ku0;
            const uint32_t i = loc / w.cpr, c = loc - i * w.cpr;
            const float4 pi = Sg[w.r0 + i];
            uint32_t j;
            bool valid;
            if (c < w.cprA) {
                const uint32_t n = w.nrow, h = (n - 1) / 2;
                const uint32_t dmax = h + (((n & 1u) == 0 && i < n / 2) ? 1u : 0u);
                const uint32_t d = c * 64u + lane + 1u;
                valid = d <= dmax;
                uint32_t jl = i + d; if (jl >= n) jl -= n;
                j = w.r0 + jl;
            } else {
                const uint32_t jj = (c - w.cprA) * 64u + lane;
                valid = jj < w.wlen;
                j = w.wlo + jj;
            }
            if (!valid) continue;
            ncand++;
            const float4 pj = Sg[j];
            const float fx = pi.x - pj.x, fy = pi.y - pj.y, fz = pi.z - pj.z;
            const float r2f = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
            if (r2f > RC2_hi) continue;
            if (!(r2f < RC2_lo)) {
                double ddx = (double)pi.x - (double)pj.x, ddy = (double)pi.y - (double)pj.y, ddz = (double)pi.z - (double)pj.z;
                if (ddx * ddx + ddy * ddy + ddz * ddz > RC2) continue;
            }
            nsamp++;
            const uint32_t oi = (uint32_t)__float_as_int(pi.w), oj = (uint32_t)__float_as_int(pj.w);
            const uint32_t a = oi < oj ? oi : oj, b = oi < oj ? oj : oi;
            const uint64_t ps = dna_pair_state(tk, a, b);
            if (cannot_react(kb, sqrt_bound(r2f) * 0.99999f, u01_lo(ps, 6u))) continue;
            double ddx = (double)pi.x - (double)pj.x, ddy = (double)pi.y - (double)pj.y, ddz = (double)pi.z - (double)pj.z;
            push_pending(A, T, sh, a, b, oi < oj ? sa : sb, oi < oj ? sb : sa, sqrt(ddx * ddx + ddy * ddy + ddz * ddz));
        }
        __syncthreads();                              // cu is rewritten by the next block
        wcand += ncand; wsamp += nsamp; ncand = 0; nsamp = 0;
    }
}

template <bool DEATH>
__global__ void __launch_bounds__(IRT_BLOCK, 4) k_irt_pairs(IrtArgs A)
{
    const IrtConst &C = c_irt;
    __shared__ PairShared sh;
    const IrtTrack T = A.tracks[blockIdx.x];          // launch order: largest tracks first
    const uint32_t tid = threadIdx.x;
    const uint32_t n = T.n;
    const uint32_t *ord = A.idx + T.start;            // the track's entries in input order
    float4 *P = A.pos + T.part_off;
    uint8_t *al = A.alive + T.part_off;
    double *tb = A.tb + T.part_off;
    const uint64_t tk = dna_track_key(A.seed_ptr ? *A.seed_ptr : A.seed, T.inst, T.pid);
    if (tid == 0) { sh.npend = 0; sh.c[0] = 0; sh.c[1] = 0; sh.c[2] = 0; }
    if (tid < 8) sh.gcnt[tid] = 0;
    __syncthreads();
    // phase timers (100 MHz wall clock, thread 0): load, pairs
    const unsigned long long tp0 = wall_clock64();
    // 1. load + displacement (irt-worker.js:475-537)
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    for (uint32_t k = tid; k < n; k += IRT_NT) {
        float4 e = A.rad[ord[k]];
        double x = e.x, y = e.y, z = e.z;
        int s = (int)fmod(floor((double)e.w + 0.5), 8.0);
        uint64_t ps = dna_pair_state(tk, k, k);
        uint32_t dn = 0;
        double sig = 0.0;
        int spc; uint8_t alv = 1;
        if (s == 5) spc = 1;
        else if (s == 0 || s == 3) { spc = s; if (dna_u01(ps, dn++) < 0.5) sig = C.sig_oh; }
        else if (s == 1) { spc = 1; sig = C.sig_eaq; }
        else if (s == 2) { spc = 2; sig = C.sig_h; }
        else if (s == 6) spc = 5;
        else { spc = s; alv = 0; }
        if (sig != 0.0) {
            double v;
            v = dna_u01(ps, dn++); for (int q = 1; q < 6; q++) v += dna_u01(ps, dn++);
            x += sig * ((v - 3) * 1.4142135623730951);
            v = dna_u01(ps, dn++); for (int q = 1; q < 6; q++) v += dna_u01(ps, dn++);
            y += sig * ((v - 3) * 1.4142135623730951);
            v = dna_u01(ps, dn++); for (int q = 1; q < 6; q++) v += dna_u01(ps, dn++);
            z += sig * ((v - 3) * 1.4142135623730951);
        }
        P[k] = make_float4((float)x, (float)y, (float)z, (float)spc);
        al[k] = alv ? (uint8_t)(spc + 1) : (uint8_t)0;
        if (DEATH) A.tdeath[T.part_off + k] = __longlong_as_double(0x7FF0000000000000ll);   // +inf
        tb[k] = 0.0;
        if (alv) { if (spc == 0) c0++; else if (spc == 1) c1++; else if (spc == 2) c2++; }
        if (alv && spc >= 0 && spc < 6) atomicAdd(&sh.gcnt[spc], 1u);
    }
    for (int o = 32; o > 0; o >>= 1) { c0 += __shfl_xor(c0, o, 64); c1 += __shfl_xor(c1, o, 64); c2 += __shfl_xor(c2, o, 64); }
    if ((tid & 63) == 0) { atomicAdd(&sh.c[0], c0); atomicAdd(&sh.c[1], c1); atomicAdd(&sh.c[2], c2); }
    // 2. the sorted copy: the alive initial particles ordered by (species, cell), original
    // index in .w.  Without the hash (cell_m = 0) a species is one cell.  With it, the cells
    // are slabs of width 1 / cinv >= CELL_R / m across the track's principal axis (the top
    // eigenvector of the positions' covariance, by power iteration): two particles within
    // R_CUT have axis coordinates within R_CUT (the axis is a unit vector), so their cells
    // are at most m apart, and the cell windows of pair_blocks keep every pair that the
    // full enumeration would sample (CELL_R covers the FP32 error of the coordinate and of
    // the cell index, and clamping the index to the nx cells only brings cells closer).
    // The coordinate, then the key, is kept in tb[] between the passes, so that every pass
    // sees the same key.
    const uint32_t wave = tid >> 6, lane = tid & 63u, nw = IRT_NT / 64u;
    const bool hashed = A.cell_m > 0;
    __syncthreads();
    if (hashed) {
        double m1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t k = tid; k < n; k += IRT_NT) {
            if (!al[k]) continue;
            const float4 q = P[k];
            const double x = q.x, y = q.y, z = q.z;
            m1[0] += x; m1[1] += y; m1[2] += z;
            m1[3] += x * x; m1[4] += x * y; m1[5] += x * z; m1[6] += y * y; m1[7] += y * z; m1[8] += z * z;
        }
        for (int v = 0; v < 9; v++) {
            double r = m1[v];
            for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
            if (lane == 0) sh.red[wave][v] = r;
        }
        __syncthreads();
        if (tid == 0) {
            double S[9];
#pragma unroll
            for (int v = 0; v < 9; v++) { S[v] = 0; for (uint32_t w = 0; w < nw; w++) S[v] += sh.red[w][v]; }
            double N = 0;
            for (int g = 0; g < 6; g++) N += sh.gcnt[g];
            double a0 = 1.0, a1 = 0.0, a2 = 0.0;
            if (N >= 2) {
                const double mx = S[0] / N, my = S[1] / N, mz = S[2] / N;
                const double cxx = S[3] / N - mx * mx, cxy = S[4] / N - mx * my, cxz = S[5] / N - mx * mz;
                const double cyy = S[6] / N - my * my, cyz = S[7] / N - my * mz, czz = S[8] / N - mz * mz;
                // start on the axis of the largest variance
                if (cyy > cxx && cyy >= czz) { a0 = 0.0; a1 = 1.0; }
                else if (czz > cxx && czz > cyy) { a0 = 0.0; a2 = 1.0; }
                for (int it = 0; it < 32; it++) {
                    const double w0 = cxx * a0 + cxy * a1 + cxz * a2;
                    const double w1 = cxy * a0 + cyy * a1 + cyz * a2;
                    const double w2 = cxz * a0 + cyz * a1 + czz * a2;
                    const double nrm = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
                    if (!(nrm > 0) || !(nrm < 1e300)) break;
                    a0 = w0 / nrm; a1 = w1 / nrm; a2 = w2 / nrm;
                }
            }
            sh.ax[0] = (float)a0; sh.ax[1] = (float)a1; sh.ax[2] = (float)a2;
        }
        __syncthreads();
        float lo = 3e38f, hi = -3e38f;
        for (uint32_t k = tid; k < n; k += IRT_NT) {
            if (!al[k]) continue;
            const float u = axis_u(sh.ax[0], sh.ax[1], sh.ax[2], P[k]);
            tb[k] = (double)u;
            lo = fminf(lo, u); hi = fmaxf(hi, u);
        }
        for (int o = 32; o > 0; o >>= 1) { lo = fminf(lo, __shfl_xor(lo, o, 64)); hi = fmaxf(hi, __shfl_xor(hi, o, 64)); }
        if (lane == 0) { sh.rlo[wave] = lo; sh.rhi[wave] = hi; }
        __syncthreads();
        if (tid == 0) {
            for (uint32_t w = 1; w < nw; w++) { lo = fminf(lo, sh.rlo[w]); hi = fmaxf(hi, sh.rhi[w]); }
            const uint32_t m = (uint32_t)A.cell_m;
            float cinv = (float)m / C.CELL_R;
            const float ext = (hi > lo && hi - lo < 1e30f) ? hi - lo : 0.f;   // non-finite: one cell
            // at most NX_MAX cells: wider ones for long tracks (m cells still span CELL_R)
            if (ext * cinv > (float)(NX_MAX - 1)) cinv = (float)(NX_MAX - 1) / ext;
            sh.nx = (uint32_t)(ext * cinv) + 1u;
            if (sh.nx > NX_MAX) sh.nx = NX_MAX;
            sh.m = m; sh.umin = lo; sh.cinv = cinv;
        }
    } else if (tid == 0) { sh.nx = 1; sh.m = 0; }
    __syncthreads();
    const uint32_t nx = sh.nx;
    for (uint32_t k = tid; k <= 6 * nx; k += IRT_NT) sh.cs[k] = 0;
    __syncthreads();
    for (uint32_t k = tid; k < n; k += IRT_NT) {
        const uint8_t av = al[k];
        if (!av) continue;
        uint32_t key = (uint32_t)(av - 1) * nx;
        if (hashed) {
            key += hash_cell((float)tb[k], sh.umin, sh.cinv, nx);
            tb[k] = (double)key;
        }
        atomicAdd(&sh.cs[key], 1u);
    }
    __syncthreads();
    wave0_scan(sh.cs, 6 * nx);
    __syncthreads();
    if (tid == 0) {
        for (int g = 0; g < 6; g++) sh.goff[g] = sh.cs[g * nx];
        // channel blocks (irt-worker.js:216-226): the reactions whose species are present
        const int BA[9] = {0, 0, 0, 1, 1, 1, 2, 1, 3}, BB[9] = {0, 1, 2, 1, 2, 3, 2, 4, 5};
        int nb = 0;
        for (int r = 0; r < 9; r++) {
            const uint32_t nA = sh.gcnt[BA[r]], nB = sh.gcnt[BB[r]];
            if (BA[r] == BB[r] ? nA >= 2 : (nA && nB)) sh.bri[nb++] = r;
        }
        sh.nblk = nb;
    }
    __syncthreads();
    unsigned long long wcand = 0, wsamp = 0;
    float4 *Sg = (n <= A.lds_slots) ? dyn_lds : A.sorted + T.part_off;
    for (uint32_t k = tid; k < n; k += IRT_NT) {
        const uint8_t av = al[k];
        if (!av) continue;
        uint32_t key = (uint32_t)(av - 1);
        if (hashed) { key = (uint32_t)tb[k]; tb[k] = 0.0; }
        const float4 q = P[k];
        const uint32_t slot = atomicAdd(&sh.cs[key], 1u);
        Sg[slot] = make_float4(q.x, q.y, q.z, __int_as_float((int)k));
    }
    __syncthreads();
    const unsigned long long tp1 = wall_clock64();
    // 3. initial pairs (irt-worker.js:541-558) of the species blocks that have a reaction
    // channel, in units of one row particle x 64 partners, so every lane of a wavefront
    // evaluates the same reaction with the same constants.  Pairs are identified by their
    // original indices, so neither the enumeration order nor the cells change any draw or
    // the event order.
    if (n <= A.lds_slots) pair_blocks<true>(A, T, sh, tk, wcand, wsamp);
    else pair_blocks<false>(A, T, sh, tk, wcand, wsamp);
    // species lists of the alive initial particles = the species groups of Sg
    for (uint32_t g = tid; g < n; g += IRT_NT) {
        int spc = 0;
        while (spc < 5 && g >= sh.goff[spc + 1]) spc++;
        if (g < sh.goff[spc] + sh.gcnt[spc]) {
            const float4 q = Sg[g];
            const size_t slot = 6 * T.part_off + (size_t)spc * T.cap + (g - sh.goff[spc]);
            A.lst[slot] = 0u;                          // born at 0
            A.lpos[slot] = q;                          // x, y, z, the particle index
        }
    }
    __syncthreads();
    const unsigned long long tp2 = wall_clock64();
    // 4. the queued pairs: the rest of sampleIRT (irt-worker.js:116-184) in FP64, SamplePDC's
    // rejection loops included, into the track's event list, by the whole workgroup; a track's
    // drain overlaps the pair enumeration of the other workgroups on its CU.  The long list
    // takes the dynamic LDS of the sorted copy.
    if (tid == 0) {
        sh.n_ev = 0; sh.pend_next = 0; sh.long_n = 0; sh.long_next = 0;
        sh.long_cap = A.lds_slots * 4u < 1024u ? A.lds_slots * 4u : 1024u;
        sh.long_item = (uint32_t *)dyn_lds;
    }
    if (tid < 36) sh.rmap[tid] = (int8_t)C.rxnMap[tid];
    if (tid < 6) sh.Dd[tid] = C.D[tid];
    __syncthreads();
    if (A.drain_at == 1) {
        pdc_drain(A, T, tk, sh, 0.0);
        __syncthreads();
    }
    const unsigned long long tp3 = wall_clock64();
    for (int o = 32; o > 0; o >>= 1) { wcand += __shfl_xor(wcand, o, 64); wsamp += __shfl_xor(wsamp, o, 64); }
    if ((tid & 63) == 0) {
        unsigned long long *g = A.tally + (size_t)T.inst * TALLY_STRIDE + DNA_MAX_CKPT * 5 + 11;
        if (wcand) atomicAdd(&g[0], wcand);
        if (wsamp) atomicAdd(&g[1], wsamp);
    }
    // hand-over to the loop kernel (same launch index).  With the hash, each species list
    // starts with its initial particles in cell order: the loop kernel's re-pairing reads
    // only the buckets (runs of consecutive cells) within reach of a product.
    uint32_t *st = A.tstate + (size_t)TSTATE * blockIdx.x;
    bool win = hashed && nx > 1;
    for (int g = 0; g < 6; g++) win = win && sh.gcnt[g] < 65536u;
    const uint32_t wg = (nx + WB_MAX - 1) / WB_MAX, wnb = (nx + wg - 1) / wg;
    if (win)
        for (uint32_t k = tid; k < 6 * WB_MAX / 2; k += IRT_NT) {
            const uint32_t s = k / (WB_MAX / 2), b = 2 * (k % (WB_MAX / 2));
            uint32_t e0 = 0, e1 = 0;
            if (b < wnb) e0 = run_end(sh, s, (b + 1) * wg < nx ? (b + 1) * wg - 1 : nx - 1) - sh.goff[s];
            if (b + 1 < wnb) e1 = run_end(sh, s, (b + 2) * wg < nx ? (b + 2) * wg - 1 : nx - 1) - sh.goff[s];
            st[32 + k] = e0 | (e1 << 16);
        }
    if (tid == 0) {
        st[0] = sh.npend;
        for (int s = 0; s < 6; s++) st[1 + s] = sh.gcnt[s];
        st[7] = (uint32_t)sh.c[0]; st[8] = (uint32_t)sh.c[1]; st[9] = (uint32_t)sh.c[2];
        st[10] = T.h2; st[11] = n;
        st[12] = win ? 1u : 0u;
        st[21] = sh.n_ev;
        if (win) {
            st[13] = __float_as_uint(sh.ax[0]); st[14] = __float_as_uint(sh.ax[1]); st[15] = __float_as_uint(sh.ax[2]);
            st[16] = __float_as_uint(sh.umin); st[17] = __float_as_uint(sh.cinv);
            st[18] = nx; st[19] = wg; st[20] = wnb;
        }
        unsigned long long *g = A.tally + (size_t)T.inst * TALLY_STRIDE + DNA_MAX_CKPT * 5 + 14;
        atomicAdd(&g[0], tp1 - tp0); atomicAdd(&g[1], tp2 - tp1); atomicAdd(&g[6], tp3 - tp2);
        // the worker's own enumeration: every pair of alive initial particles with a channel
        const int BA[9] = {0, 0, 0, 1, 1, 1, 2, 1, 3}, BB[9] = {0, 1, 2, 1, 2, 3, 2, 4, 5};
        unsigned long long all = 0;
        for (int r = 0; r < 9; r++) {
            const unsigned long long nA = sh.gcnt[BA[r]], nB = sh.gcnt[BB[r]];
            all += BA[r] == BB[r] ? nA * (nA - (nA ? 1ull : 0ull)) / 2 : nA * nB;
        }
        if (all) atomicAdd(&g[8], all);
    }
}

// The minimum of the live events of block j into the block minima (all lanes of the
// wavefront; lane 0 writes it).  Events of reacted particles are skipped.
template <class AL>
__device__ __forceinline__ void block_min(const IrtArgs &A, const IrtTrack &T, LoopShared &sh, const AL &al,
                                          uint32_t j)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t bs = sh.bs, ne = sh.n_ev < T.ev_cap ? sh.n_ev : T.ev_cap;
    const uint32_t end = (j + 1) * bs < ne ? (j + 1) * bs : ne;
    double bt = 1e300;
    uint64_t bab = ~0ull;
    for (uint32_t e = j * bs + lane; e < end; e += 64u) {
        const uint2 ab = A.ev_ab[T.ev_off + e];
        const double t = A.ev_t[T.ev_off + e];
        if (!al[ab.x] || !al[ab.y]) continue;
        const uint64_t k = ((uint64_t)ab.x << 32) | ab.y;
        if (ev_less(t, k, bt, bab)) { bt = t; bab = k; }
    }
    wave_min_ev(bt, bab);
    if (lane == 0) { sh.bm_t[j] = bt; sh.bm_ab[j] = bab; }
}

// Double the event blocks until bm_max of them cover ne events: block j of the doubled size
// holds the old blocks 2j and 2j+1, so its minimum is the lesser of theirs (the loop
// kernel's one wavefront; at most BM_MAX / 2 <= 64 new blocks, one per lane).
__device__ __forceinline__ void grow_blocks(LoopShared &sh, uint32_t ne, uint32_t bm_max)
{
    const uint32_t lane = threadIdx.x & 63u;
    static_assert(BM_MAX <= 128u, "one new block per lane");
    while ((uint64_t)sh.bs * bm_max < ne) {
        const uint32_t nb = sh.nblk, nb2 = (nb + 1u) / 2u;
        double t = 1e300;
        uint64_t k = ~0ull;
        if (lane < nb2) {
            t = sh.bm_t[2 * lane]; k = sh.bm_ab[2 * lane];
            if (2 * lane + 1 < nb && ev_less(sh.bm_t[2 * lane + 1], sh.bm_ab[2 * lane + 1], t, k)) {
                t = sh.bm_t[2 * lane + 1]; k = sh.bm_ab[2 * lane + 1];
            }
        }
        __syncthreads();
        for (uint32_t j = lane; j < BM_MAX; j += 64u) {
            sh.bm_t[j] = j < nb2 ? t : 1e300;
            sh.bm_ab[j] = j < nb2 ? k : ~0ull;
        }
        __syncthreads();
        if (lane == 0) { sh.bs *= 2u; sh.nblk = nb2; }
        __syncthreads();
    }
}

// The initial pairs' drain as a kernel of its own (DNA_OPT_IRT_DRAIN = 2): one
// DRAIN_NT-thread workgroup per track, light registers and 4 KB of LDS, between the pairs
// and the loop kernels; the event count goes to the loop kernel through tstate[21].
#define DRAIN_NT 128
__global__ void __launch_bounds__(DRAIN_NT) k_irt_drain(IrtArgs A)
{
    __shared__ DrainShared sh;
    const IrtTrack T = A.tracks[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    uint32_t *st = A.tstate + (size_t)TSTATE * blockIdx.x;
    const uint64_t tk = dna_track_key(A.seed_ptr ? *A.seed_ptr : A.seed, T.inst, T.pid);
    const unsigned long long tp0 = wall_clock64();
    if (tid == 0) { sh.n_ev = 0; sh.npend = st[0]; sh.pend_next = 0; sh.long_n = 0; sh.long_next = 0; }
    if (tid < 36) sh.rmap[tid] = (int8_t)c_irt.rxnMap[tid];
    if (tid < 6) sh.Dd[tid] = c_irt.D[tid];
    __syncthreads();
    pdc_drain(A, T, tk, sh, 0.0);
    __syncthreads();
    if (tid == 0) {
        st[21] = sh.n_ev;
        unsigned long long *pf = A.tally + (size_t)T.inst * TALLY_STRIDE + DNA_MAX_CKPT * 5 + 14;
        atomicAdd(&pf[6], wall_clock64() - tp0);
    }
}

// The loop kernel's workgroup order (DNA_OPT_IRT_ORDER): within each run of ORDER_CHUNK
// tracks (about one resident wave of loop workgroups: 16 per CU), the tracks by decreasing
// queued pairs + initial particles, ties by launch index.  A SIMD's arbiter issues for its
// oldest wavefront first, so a workgroup's dispatch rank sets how much of its SIMD it gets
// while the others still run: the heaviest chains take the first ranks.  Bitonic sort of
// (~weight, index) in LDS, one workgroup per chunk; the order changes only when each
// track's wavefront runs, not what it computes.
#define ORDER_CHUNK 4096u
#define ORDER_NT 1024
__global__ void __launch_bounds__(ORDER_NT) k_irt_order(const uint32_t *tstate, uint32_t ntr, uint32_t *order)
{
    __shared__ unsigned long long key[ORDER_CHUNK];
    const uint32_t base = blockIdx.x * ORDER_CHUNK;
    const uint32_t len = ntr - base < ORDER_CHUNK ? ntr - base : ORDER_CHUNK;
    for (uint32_t i = threadIdx.x; i < ORDER_CHUNK; i += ORDER_NT) {
        unsigned long long k = ~0ull;
        if (i < len) {
            const uint32_t *st = tstate + (size_t)TSTATE * (base + i);
            const uint32_t w = st[0] + st[11];
            k = ((unsigned long long)(~w) << 32) | (base + i);
        }
        key[i] = k;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= ORDER_CHUNK; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < ORDER_CHUNK; i += ORDER_NT) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long a = key[i], b = key[l];
                    if ((a > b) == ((i & k) == 0)) { key[i] = b; key[l] = a; }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < len; i += ORDER_NT) order[base + i] = (uint32_t)key[i];
}

template <bool DEATH, bool LDS_BITS>
__global__ void __launch_bounds__(64, 4) k_irt_loop(IrtArgs A)
{
    const IrtConst &C = c_irt;
    __shared__ LoopShared sh;
    const uint32_t li = A.order ? A.order[blockIdx.x] : blockIdx.x;     // the track's launch index
    const IrtTrack T = A.tracks[li];
    const uint32_t lane = threadIdx.x;
    float4 *P = A.pos + T.part_off;
    uint8_t *const al_global = A.alive + T.part_off;
    double *tb = A.tb + T.part_off;
    const uint64_t tk = dna_track_key(A.seed_ptr ? *A.seed_ptr : A.seed, T.inst, T.pid);
    const uint32_t *st = A.tstate + (size_t)TSTATE * li;
    AliveBits<LDS_BITS> al;
    al.g = (uint32_t *)(A.sorted + T.part_off);
    const unsigned long long tp0 = wall_clock64();
    if (lane == 0) {
        sh.npend = 0; sh.pend_next = 0; sh.long_n = 0; sh.long_next = 0;
        sh.n_ev = st[21]; sh.n_total = st[11]; sh.n_init = st[11]; sh.n_re = 0; sh.nt_n = 0; sh.lcomp = 0;
        for (int s = 0; s < 6; s++) { sh.llen[s] = st[1 + s]; sh.ldead[s] = 0; }
        sh.win = st[12];
        for (int s = 0; s < 6; s++) sh.n0[s] = sh.win ? st[1 + s] : 0u;
        sh.wnb = 0; sh.wnx = 1; sh.wg = 1; sh.wumin = 0.f; sh.wcinv = 0.f;
        sh.wax[0] = 1.f; sh.wax[1] = 0.f; sh.wax[2] = 0.f;
        if (sh.win) {
            sh.wax[0] = __uint_as_float(st[13]); sh.wax[1] = __uint_as_float(st[14]); sh.wax[2] = __uint_as_float(st[15]);
            sh.wumin = __uint_as_float(st[16]); sh.wcinv = __uint_as_float(st[17]);
            sh.wnx = st[18]; sh.wg = st[19]; sh.wnb = st[20];
        }
        sh.cnt[0] = st[7]; sh.cnt[1] = st[8]; sh.cnt[2] = st[9]; sh.cnt[3] = 0; sh.cnt[4] = st[10];
        for (int r = 0; r < 9; r++) sh.rxc[r] = 0;
    }
    if (lane < 36) {
        const int r = C.rxnMap[lane];
        sh.rmap[lane] = (int8_t)r;
        RxB bz; bz.S = 0.f; bz.sc = 0.f; bz.R = 0.f;
        sh.bnd[lane] = r >= 0 ? rx_bound(rx_constants(r, C.D[lane / 6] + C.D[lane % 6])) : bz;
    }
    if (lane < 9) { sh.nprod[lane] = (int8_t)C.nprod[lane]; sh.prods[lane][0] = (int8_t)C.prods[lane][0]; sh.prods[lane][1] = (int8_t)C.prods[lane][1]; }
    if (lane < 6) { sh.Dd[lane] = C.D[lane]; sh.sqD[lane] = C.sqrtD[lane]; sh.Df[lane] = (float)C.D[lane]; }
    if (st[12])
        for (uint32_t k = lane; k < 6 * WB_MAX / 2; k += 64u) {
            const uint32_t wv = st[32 + k], s6 = k / (WB_MAX / 2), b = 2 * (k % (WB_MAX / 2));
            sh.wend[s6][b] = (uint16_t)(wv & 0xFFFFu); sh.wend[s6][b + 1] = (uint16_t)(wv >> 16);
        }
    for (uint32_t i = lane; i < DNA_MAX_CKPT * 5; i += 64u) sh.snap[i] = 0u;
    for (uint32_t j = lane; j < BM_MAX; j += 64u) { sh.bm_t[j] = 1e300; sh.bm_ab[j] = ~0ull; }
    __syncthreads();
    // the alive bits from the pairs kernel's alive bytes
    {
        const uint32_t nt0 = sh.n_total, nw = alive_words(T.cap);
        uint32_t *aw = al.w();
        for (uint32_t c0 = 0; 2u * (c0 / 64u) < nw; c0 += 64u) {
            const uint32_t k = c0 + lane;
            const uint64_t m = __ballot(k < nt0 && al_global[k] != 0);
            if (lane == 0) { aw[c0 / 32u] = (uint32_t)m; aw[c0 / 32u + 1u] = (uint32_t)(m >> 32); }
        }
        __syncthreads();
    }
    // 1. the queued initial pairs (unless k_irt_pairs drained them): the rest of sampleIRT
    // (irt-worker.js:116-184) in FP64, then every event into the block minima
    if (A.drain_at == 0) {
        if (lane == 0) sh.npend = st[0];
        __syncthreads();
        pdc_drain(A, T, tk, sh, 0.0);
        __syncthreads();
        if (lane == 0) { sh.npend = 0; sh.pend_next = 0; sh.long_n = 0; sh.long_next = 0; sh.nt_n = 0; }
        __syncthreads();
    }
    const unsigned long long n_init_ev = sh.n_ev;
    if (lane == 0) {
        // blocks of bs events, bs the least 64 * 2^k that covers the events so far with
        // bm_max blocks; it doubles when the events outgrow them (grow_blocks)
        const uint32_t ne = sh.n_ev < T.ev_cap ? sh.n_ev : T.ev_cap;
        uint32_t bs = 64u;
        while ((uint64_t)bs * A.bm_max < ne) bs *= 2u;
        sh.bs = bs;
        sh.nblk = (ne + bs - 1) / bs;
    }
    __syncthreads();
    for (uint32_t j = 0; j < sh.nblk; j++) block_min(A, T, sh, al, j);
    __syncthreads();
    const unsigned long long tp1 = wall_clock64();
    if (lane == 0) for (int k = 0; k < 7; k++) sh.acc[k] = 0;
    int cp = 0;
    // 2. event loop (irt-worker.js:566-645)
    for (;;) {
        const unsigned long long q0 = wall_clock64();
        if (lane == 0) sh.acc[4]++;
        // the next event: the least block minimum whose reactants are both alive.  Each
        // reaction refreshes the minima it makes stale (below), so the first candidate is
        // live; a stale one would be replaced by its block's rescan
        double t0;
        uint64_t ab0;
        uint32_t bjw = 0;
        for (;;) {
            double bt = 1e300;
            uint64_t bab = ~0ull;
            uint32_t bj = 0;
            for (uint32_t j = lane; j < sh.nblk; j += 64u) {
                const double t = sh.bm_t[j];
                const uint64_t k = sh.bm_ab[j];
                if (ev_less(t, k, bt, bab)) { bt = t; bab = k; bj = j; }
            }
            bj = (uint32_t)__builtin_amdgcn_readlane((int)bj, (int)wave_min_ev(bt, bab));
            t0 = bt; ab0 = bab; bjw = bj;
            if (!(bt < 1e300)) break;
            if (al[(uint32_t)(bab >> 32)] && al[(uint32_t)bab]) break;
            block_min(A, T, sh, al, bj);
            __syncthreads();
        }
        if (!(t0 < 1e300)) break;
        // The reaction takes the minimum of block bjw: that block's first 64 events are
        // loaded now, while lane 0 runs the reaction, and reduced right after it
        const uint32_t bs_w = sh.bs, ne_w = sh.n_ev < T.ev_cap ? sh.n_ev : T.ev_cap;
        const uint32_t end_w = (bjw + 1u) * bs_w < ne_w ? (bjw + 1u) * bs_w : ne_w;
        const uint32_t ew = bjw * bs_w + lane;
        uint2 pab = make_uint2(0u, 0u);
        double pt = 1e300;
        if (ew < end_w) { pab = A.ev_ab[T.ev_off + ew]; pt = A.ev_t[T.ev_off + ew]; }
        const unsigned long long q1 = wall_clock64();
        if (lane == 0) sh.acc[0] += q1 - q0;
        if (lane == 0) {
            sh.win_t = t0; sh.win_ab = ab0; sh.np = 0;
            sh.npend = 0; sh.pend_next = 0; sh.long_n = 0; sh.long_next = 0;
            sh.nt_n = 0; sh.e0 = sh.n_ev;
            while (cp < A.nc && t0 >= A.tc[cp]) { for (int s = 0; s < 5; s++) sh.snap[cp * 5 + s] = (uint32_t)sh.cnt[s]; cp++; }
            uint32_t a = (uint32_t)(ab0 >> 32), b = (uint32_t)ab0;
            const float4 pa = P[a], pb = P[b];
            int si = (int)pa.w, sj = (int)pb.w;
            int ri = sh.rmap[(si < sj ? si : sj) * 6 + (si < sj ? sj : si)];
            al.clear(a); al.clear(b);
            if (DEATH) { A.tdeath[T.part_off + a] = t0; A.tdeath[T.part_off + b] = t0; }
            sh.ldead[si]++; sh.ldead[sj]++;
            sh.n_re++; sh.rxc[ri]++;
            if (si == 0) sh.cnt[0]--; else if (si == 1) sh.cnt[1]--; else if (si == 2) sh.cnt[2]--; else if (si == 4) sh.cnt[3]--;
            if (sj == 0) sh.cnt[0]--; else if (sj == 1) sh.cnt[1]--; else if (sj == 2) sh.cnt[2]--; else if (sj == 4) sh.cnt[3]--;
            if (ri == 3 || ri == 4 || ri == 6) sh.cnt[4]++;
            int np = sh.nprod[ri];
            if (np > 0) {
                // one product sits at the diffusion-weighted midpoint; two keep the reactants'
                // positions (no FP64 divisions then)
                double mx = 0, my = 0, mz = 0;
                if (np == 1) {
                    const double sqDi = sh.sqD[si], sqDj = sh.sqD[sj];
                    const double wsum = sqDi + sqDj;
                    mx = (sqDj * pa.x + sqDi * pb.x) / wsum;
                    my = (sqDj * pa.y + sqDi * pb.y) / wsum;
                    mz = (sqDj * pa.z + sqDi * pb.z) / wsum;
                }
                uint32_t nt = sh.n_total;
                int made = 0;
                for (int pp = 0; pp < np; pp++) {
                    uint32_t id = nt;
                    if (id >= T.cap_global) break;
                    if (id >= T.cap) { atomicOr(A.overflow, 1u); break; }
                    int ps = sh.prods[ri][pp];
                    float4 q;
                    if (np == 1) q = make_float4((float)mx, (float)my, (float)mz, (float)ps);
                    else if (pp == 0) q = make_float4(pa.x, pa.y, pa.z, (float)ps);
                    else q = make_float4(pb.x, pb.y, pb.z, (float)ps);
                    P[id] = q; al.set(id); tb[id] = t0;
                    if (DEATH) A.tdeath[T.part_off + id] = __longlong_as_double(0x7FF0000000000000ll);
                    const size_t slot = 6 * T.part_off + (size_t)ps * T.cap + sh.llen[ps]++;
                    A.lst[slot] = __float_as_uint(__double2float_rd(t0));
                    A.lpos[slot] = make_float4(q.x, q.y, q.z, __uint_as_float(id));
                    if (ps == 0) sh.cnt[0]++; else if (ps == 1) sh.cnt[1]++; else if (ps == 2) sh.cnt[2]++; else if (ps == 4) sh.cnt[3]++;
                    sh.prod_idx[made] = id; sh.prod_sp[made] = (uint32_t)ps; sh.prod_pos[made] = q;
                    made++;
                    nt++;
                }
                sh.np = made;
                sh.n_total = nt;
            }
            // the partner lists of the re-pairing that are mostly reacted are compacted first
            uint32_t lc = 0;
            for (int pp = 0; pp < sh.np; pp++) {
                const uint32_t sp = sh.prod_sp[pp];
                for (int sj2 = 0; sj2 < 6; sj2++)
                    if (sh.rmap[sp * 6 + sj2] >= 0 && 2u * sh.ldead[sj2] > sh.llen[sj2] && sh.llen[sj2] >= 64u)
                        lc |= 1u << sj2;
            }
            sh.lcomp = lc;
        }
        __syncthreads();
        // the block minima this reaction made stale: block bjw (its minimum was the reacting
        // pair), and any other block whose minimum has either reactant, so the next
        // selection finds live minima
        {
            double bt = 1e300;
            uint64_t bab = ~0ull;
            if (ew < end_w && al[pab.x] && al[pab.y]) { bt = pt; bab = ((uint64_t)pab.x << 32) | pab.y; }
            for (uint32_t e = ew + 64u; e < end_w; e += 64u) {
                const uint2 q = A.ev_ab[T.ev_off + e];
                const double t = A.ev_t[T.ev_off + e];
                if (!al[q.x] || !al[q.y]) continue;
                const uint64_t k = ((uint64_t)q.x << 32) | q.y;
                if (ev_less(t, k, bt, bab)) { bt = t; bab = k; }
            }
            wave_min_ev(bt, bab);
            if (lane == 0) { sh.bm_t[bjw] = bt; sh.bm_ab[bjw] = bab; }
            const uint32_t ra = (uint32_t)(ab0 >> 32), rb = (uint32_t)ab0;
            for (uint32_t j0 = 0; j0 < sh.nblk; j0 += 64u) {
                const uint32_t j = j0 + lane;
                bool stale = false;
                if (j < sh.nblk && j != bjw) {
                    const uint64_t k = sh.bm_ab[j];
                    const uint32_t ka = (uint32_t)(k >> 32), kb = (uint32_t)k;
                    stale = ka == ra || ka == rb || kb == ra || kb == rb;
                }
                uint64_t m = __ballot(stale);
                while (m) {
                    const uint32_t jj = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    block_min(A, T, sh, al, j0 + jj);
                    m &= m - 1ull;
                }
            }
            __syncthreads();
        }
        const unsigned long long q2 = wall_clock64();
        if (lane == 0) sh.acc[1] += q2 - q1;
        if (sh.np) {
            if (sh.lcomp) {
                compact_lists(A, T, sh, al);
                if (lane == 0) sh.lcomp = 0;
            }
            build_segments(sh, sh.win_t, T.cap);
            __syncthreads();
            pair_products(A, T, sh, al, tk, sh.win_t);
        }
        __syncthreads();
        const unsigned long long q3 = wall_clock64();
        if (sh.npend) {
            pdc_drain(A, T, tk, sh, sh.win_t);
            __syncthreads();
        }
        // merge this reaction's events into the block minima: staged (at most STAGE_MAX
        // events, so at most two blocks), else by rescanning the blocks they went to
        {
            const uint32_t ne = sh.n_ev < T.ev_cap ? sh.n_ev : T.ev_cap, e0 = sh.e0;
            if ((uint64_t)sh.bs * A.bm_max < ne) grow_blocks(sh, ne, A.bm_max);
            const uint32_t bs = sh.bs;
            if (ne > e0) {
                const uint32_t j0 = e0 / bs, j1 = (ne - 1) / bs;
                if (sh.nt_n > STAGE_MAX) {
                    for (uint32_t j = j0; j <= j1; j++) block_min(A, T, sh, al, j);
                } else {
                    const uint32_t ns = sh.nt_n;
                    for (uint32_t j = j0; j <= j1; j++) {
                        double t = 1e300;
                        uint64_t k = ~0ull;
                        if (lane < ns && sh.nt_e[lane] / bs == j) { t = sh.nt_t[lane]; k = sh.nt_ab[lane]; }
                        wave_min_ev(t, k);
                        if (lane == 0 && ev_less(t, k, sh.bm_t[j], sh.bm_ab[j])) { sh.bm_t[j] = t; sh.bm_ab[j] = k; }
                    }
                }
                if (lane == 0 && j1 + 1 > sh.nblk) sh.nblk = j1 + 1;
                __syncthreads();
            }
        }
        const unsigned long long q4 = wall_clock64();
        if (lane == 0) { sh.acc[2] += q3 - q2; sh.acc[3] += q4 - q3; }
    }
    const unsigned long long tp2 = wall_clock64();
    if (A.final_ntot) {                       // the final state for dna_irt_oh_final
        const uint32_t nt = sh.n_total;
        for (uint32_t k = lane; k < nt; k += 64u) al_global[k] = al[k] ? (uint8_t)((uint32_t)P[k].w + 1u) : (uint8_t)0;
        if (lane == 0) A.final_ntot[T.idx] = nt;
    }
    if (lane == 0) {
        while (cp < A.nc) { for (int s = 0; s < 5; s++) sh.snap[cp * 5 + s] = (uint32_t)sh.cnt[s]; cp++; }
    }
    __syncthreads();
    unsigned long long *g = A.tally + (size_t)T.inst * TALLY_STRIDE;
    for (int i = (int)lane; i < A.nc * 5; i += 64) {
        if (sh.snap[i]) atomicAdd(&g[i], (unsigned long long)sh.snap[i]);
        if (A.track_tally) A.track_tally[(size_t)T.idx * A.nc * 5 + i] = sh.snap[i];
    }
    if (lane == 0) {
        for (int r = 0; r < 9; r++) if (sh.rxc[r]) atomicAdd(&g[DNA_MAX_CKPT * 5 + r], sh.rxc[r]);
        atomicAdd(&g[DNA_MAX_CKPT * 5 + 9], sh.n_re);
        atomicAdd(&g[DNA_MAX_CKPT * 5 + 10], n_init_ev);
        if (sh.acc[5]) atomicAdd(&g[DNA_MAX_CKPT * 5 + 13], sh.acc[5]);
        if (sh.acc[6]) atomicAdd(&g[DNA_MAX_CKPT * 5 + 23], sh.acc[6]);
        unsigned long long *pf = g + DNA_MAX_CKPT * 5 + 14;
        atomicAdd(&pf[6], tp1 - tp0); atomicAdd(&pf[2], sh.acc[0]); atomicAdd(&pf[3], sh.acc[1]);
        atomicAdd(&pf[4], sh.acc[2]); atomicAdd(&pf[5], sh.acc[4]); atomicAdd(&pf[7], sh.acc[3]);
        (void)tp2;
    }
}

// ------------------------------------------------------------ final alive OH
// One workgroup per track (launch order), after a k_irt launch that kept its final state.
__global__ void __launch_bounds__(256) k_oh_count(const IrtTrack *tracks, const uint8_t *alive, const uint32_t *ntot,
                                                  uint32_t *cnt)
{
    const IrtTrack T = tracks[blockIdx.x];
    const uint8_t *al = alive + T.part_off;
    const uint32_t nt = ntot[T.idx];
    uint32_t c = 0;
    for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) c += al[k] == 1u ? 1u : 0u;   // species 0 + 1
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ uint32_t w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) { uint32_t t = 0; for (uint32_t i = 0; i < (blockDim.x >> 6); i++) t += w[i]; cnt[T.idx] = t; }
}

// The alive OH of a track in particle order from off[T.idx], each moved by one Brownian
// step over DNA_T_END_NS - t_birth with draws 32-37 of its own stream (tk, k, k).
__global__ void __launch_bounds__(256) k_oh_write(const IrtTrack *tracks, const float4 *pos, const uint8_t *alive,
                                                  const double *tb, const uint32_t *ntot, const uint32_t *off,
                                                  uint64_t seed, float4 *out)
{
    const IrtConst &C = c_irt;
    const IrtTrack T = tracks[blockIdx.x];
    const uint8_t *al = alive + T.part_off;
    const uint32_t nt = ntot[T.idx];
    const uint64_t tk = dna_track_key(seed, T.inst, T.pid);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __shared__ uint32_t wc[4];
    uint32_t base = off[T.idx];
    for (uint32_t c0 = 0; c0 < nt; c0 += blockDim.x) {
        const uint32_t k = c0 + threadIdx.x;
        const bool is_oh = k < nt && al[k] == 1u;
        const uint64_t m = __ballot(is_oh);
        if (lane == 0) wc[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, tot = 0;
        for (uint32_t i = 0; i < nw; i++) { if (i < w) before += wc[i]; tot += wc[i]; }
        if (is_oh) {
            const float4 q = pos[T.part_off + k];
            double x = q.x, y = q.y, z = q.z;
            const double dt = DNA_T_END_NS - tb[T.part_off + k];
            if (dt > 0) {
                const uint64_t ps = dna_pair_state(tk, k, k);
                const double sig = sqrt(2.0 * C.D[0] * dt);
                const double g0 = sqrt(-2 * dna_log(dna_u01(ps, 32) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 33));
                const double g1 = sqrt(-2 * dna_log(dna_u01(ps, 34) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 35));
                const double g2 = sqrt(-2 * dna_log(dna_u01(ps, 36) + 1e-30)) * dna_cos(6.2831853 * dna_u01(ps, 37));
                x += g0 * sig; y += g1 * sig; z += g2 * sig;
            }
            const uint32_t rank = before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            out[base + rank] = make_float4((float)x, (float)y, (float)z, 0.0f);
        }
        base += tot;
        __syncthreads();
    }
}

// Every particle slot of a track at the end, from off[T.idx]: (x, y, z, pid * 8 + rad_buf
// code, t_birth, t_death) as doubles; species index 5 (OH-) is rad_buf code 6.
__global__ void __launch_bounds__(256) k_particles_write(const IrtTrack *tracks, const float4 *pos, const double *tb,
                                                         const double *td, const uint32_t *ntot, const uint32_t *off,
                                                         double *out)
{
    const IrtTrack T = tracks[blockIdx.x];
    const uint32_t nt = ntot[T.idx];
    double *o = out + 6 * (size_t)off[T.idx];
    for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) {
        const float4 q = pos[T.part_off + k];
        const int s = (int)q.w;
        const int code = s == 5 ? 6 : s;
        o[6 * k] = q.x; o[6 * k + 1] = q.y; o[6 * k + 2] = q.z;
        o[6 * k + 3] = (double)((uint64_t)T.pid * 8u + (uint64_t)code);
        o[6 * k + 4] = tb[T.part_off + k]; o[6 * k + 5] = td[T.part_off + k];
    }
}

// ------------------------------------------------------------ host side
static bool g_const_ready[64] = {false};

static void build_const(IrtConst &C)
{
    static const double IRT_D[6] = {2.2, 4.9, 7.0, 9.46, 2.3, 5.3};
    static const double VDW_R[6] = {0.22, 0.50, 0.19, 0.25, 0.21, 0.33};
    static const double CHARGE[6] = {0, -1, 0, 1, 0, -1};
    static const int RA[9] = {0, 0, 0, 1, 1, 1, 2, 1, 3}, RB[9] = {0, 1, 2, 1, 2, 3, 2, 4, 5};
    static const double RK[9] = {0.55e10, 2.95e10, 1.55e10, 0.636e10, 2.50e10, 2.11e10, 0.503e10, 1.10e10, 1.13e11};
    static const int TY[9] = {1, 1, 1, 0, 0, 1, 0, 1, 0};
    static const int NP[9] = {1, 1, 0, 2, 1, 1, 0, 2, 0};
    static const int PR[9][2] = {{4, 0}, {5, 0}, {0, 0}, {5, 5}, {5, 0}, {2, 0}, {0, 0}, {5, 0}, {0, 0}};
    const double ONSAGER = 1.44 / (0.02527 * 80.1), Rs = 0.29, K_CONV = 1e24 / (6.022e23 * 1e9);
    for (int s = 0; s < 6; s++) { C.D[s] = IRT_D[s]; C.sqrtD[s] = sqrt(IRT_D[s]); }
    for (int r = 0; r < 9; r++) {
        int a = RA[r], b = RB[r];
        double D_sum = IRT_D[a] + IRT_D[b];
        bool self = a == b;
        double rc = CHARGE[a] * CHARGE[b] * ONSAGER;
        double k_nm = RK[r] * K_CONV;
        C.rc[r] = rc; C.kobs[r] = k_nm; C.type[r] = TY[r]; C.nprod[r] = NP[r];
        C.prods[r][0] = PR[r][0]; C.prods[r][1] = PR[r][1];
        if (TY[r] == 0) {
            double sig = k_nm / (4 * DNA_PI * D_sum);
            if (self) sig *= 2;
            C.sigma[r] = sig; C.sigma_eff[r] = sig; C.kdif[r] = 0; C.kact[r] = 0; C.prob[r] = 1;
        } else {
            double sc = VDW_R[a] + VDW_R[b];
            C.sigma[r] = sc;
            double se, kd;
            if (rc == 0) { se = sc; kd = 4 * DNA_PI * D_sum * sc; }
            else { se = -rc / (1 - dna_exp(rc / sc)); kd = 4 * DNA_PI * D_sum * se; }
            if (self) kd /= 2;
            C.sigma_eff[r] = se; C.kdif[r] = kd;
            double ka = kd * k_nm / (kd - k_nm);
            C.kact[r] = ka;
            C.prob[r] = Rs / (Rs + (kd / ka) * (se + Rs));
        }
    }
    for (int i = 0; i < 36; i++) C.rxnMap[i] = -1;
    for (int r = 0; r < 9; r++) { C.rxnMap[RA[r] * 6 + RB[r]] = r; C.rxnMap[RB[r] * 6 + RA[r]] = r; }
    double R_CUT = 1.45 + 2 * sqrt(8 * 9.46 * 1000);
    C.R_CUT2 = R_CUT * R_CUT;
    C.R_CUT_F = (float)R_CUT * 1.00001f;
    // 0.1% and 1 nm above R_CUT: far above the FP32 error of a projection onto the axis
    // (a few ulp of coordinates below 10^5 nm) and of the cell index
    C.CELL_R = (float)(R_CUT * 1.001 + 1.0);
    static const double g[13] = {-4.06217193e-08, 3.06848412e-06, -9.93217814e-05, 1.80172797e-03,
        -2.01135480e-02, 1.42939448e-01, -6.48348714e-01, 1.85227848e+00, -3.36450378e+00,
        4.37785068e+00, -4.20557339e+00, 3.81679083e+00, -2.34069784e-01};
    double rm = 0, pw[13];
    pw[0] = 1.0; for (int i = 1; i < 13; i++) pw[i] = pw[i - 1] * 1.7;
    for (int i = 12; i >= 0; i--) rm += g[12 - i] * pw[i];
    C.sig_eaq = rm * 0.62665706865775006;
    C.sig_oh = 0.8 / sqrt(3.0);
    C.sig_h = (17.0 / 18.0) * 2.4 / sqrt(3.0);
}

template <class T> static hipError_t dgrow(void *&ptr, size_t &cap, size_t need_elems)
{
    size_t need = need_elems * sizeof(T);
    if (need <= cap && ptr) return hipSuccess;
    if (ptr) { hipError_t e = hipFree(ptr); if (e != hipSuccess) return e; ptr = nullptr; }
    size_t n = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&ptr, n);
    cap = (e == hipSuccess) ? n : 0;
    return e;
}

}  // namespace dna

using namespace dna;

// Scratch of the grouping stage, carved out of one allocation.
struct GroupScratch {
    uint32_t *kept, *key, *S, *ck, *cnt, *h2, *firstj, *bstart, *idx, *misc;
};


static std::mutex g_const_mu;                 // contexts on several host threads

static int ensure_const(dna_ctx *c)
{
    std::lock_guard<std::mutex> lk(g_const_mu);
    if (!g_const_ready[c->device]) {
        IrtConst C; build_const(C);
        DNA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_irt), &C, sizeof(C)));
        g_const_ready[c->device] = true;
    }
    return DNA_OK;
}

// The track list of one set of inputs.  Entries of one (instance, primary) form a track;
// the grouping scratch lives in the context until the next IRT call.
struct IrtGroup {
    uint32_t n_inst = 0;
    uint64_t rad_n = 0, K = 0, K1 = 0;
    uint32_t run = 0, maxn_all = 0, gb = 0;
    bool contiguous = true;
    std::vector<IrtTrack> tr;
    std::vector<IrtTrack> launch;    // tr by decreasing particle count: the launch order
    std::vector<uint32_t> bstart, maxN;
    GroupScratch g;
};

// Group on the device and build the track list on the host.  sel (optional, instance 0
// only): keep only the tracks of primaries pid < n_sel with sel[pid] != 0; every other
// track is grouped as usual (capacities and RNG keys are those of the full run, so a
// selected track's history is the one the full run gives it).
static int irt_group(dna_ctx *c, const float4 *d_rad, uint64_t rad_n, const std::vector<uint64_t> &off,
                     const uint8_t *sel, uint32_t n_sel, IrtGroup &G)
{
    hipStream_t st = c->stream;
    const uint32_t n_inst = (uint32_t)off.size() - 1;
    G.n_inst = n_inst; G.rad_n = rad_n;
    if (rad_n >= 0xFFFFFFF0ull) { set_error("dna_run_irt: too many radical entries"); return DNA_ERR_CAPACITY; }
    size_t misc_words = 64 + 2 * (n_inst + 1);
    const size_t scan_words = (size_t)rad_n / 1024 + 16;
    // stage 1 scratch: kept, key, S (rad_n + 1 each), ck (rad_n), misc, the scan's partial
    // sums.  All of it in irt.part: irt.scratch holds the final state that
    // DNA_OPT_IRT_FINAL keeps, which grouping for a captured plan must not touch.
    size_t need1 = 4 * ((size_t)rad_n + 1) + misc_words + scan_words;
    DNA_HIP(dgrow<uint32_t>(c->irt.part, c->irt.cap_part, need1));
    uint32_t *base = (uint32_t *)c->irt.part;
    GroupScratch &g = G.g;
    g.kept = base; g.key = g.kept + rad_n + 1; g.S = g.key + rad_n + 1; g.ck = g.S + rad_n + 1;
    g.misc = g.ck + rad_n + 1;
    uint32_t *scan_tmp = g.misc + misc_words;
    uint64_t *d_off = (uint64_t *)(((uintptr_t)(g.misc + 16) + 7) & ~(uintptr_t)7);
    DNA_HIP(hipMemcpyAsync(d_off, off.data(), off.size() * 8, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemsetAsync(g.misc, 0, 16 * 4, st));
    const uint32_t B = 256;
    G.gb = (uint32_t)((rad_n + B - 1) / B);
    hipLaunchKernelGGL(k_pidmax, dim3(G.gb < 1024 ? G.gb : 1024), dim3(B), 0, st, d_rad, rad_n, g.misc);
    uint32_t h_misc[16];
    DNA_HIP(hipMemcpyAsync(h_misc, g.misc, 16 * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    const uint64_t K1 = (uint64_t)h_misc[0] + 1, K = K1 * n_inst;
    G.K1 = K1; G.K = K;
    if (K > (1ull << 28)) { set_error("dna_run_irt: primary-id range too large"); return DNA_ERR_CAPACITY; }
    // stage 2 scratch: per-key arrays cnt, h2, firstj, bstart
    DNA_HIP(dgrow<uint32_t>(c->irt.tracks, c->irt.cap_tracks, 4 * (K + 1)));
    uint32_t *kb = (uint32_t *)c->irt.tracks;
    g.cnt = kb; g.h2 = kb + (K + 1); g.firstj = kb + 2 * (K + 1); g.bstart = kb + 3 * (K + 1);
    DNA_HIP(hipMemsetAsync(g.cnt, 0, 2 * (K + 1) * 4, st));
    DNA_HIP(hipMemsetAsync(g.firstj, 0xFF, (K + 1) * 4, st));
    hipLaunchKernelGGL(k_count, dim3(G.gb), dim3(B), 0, st, d_rad, rad_n, d_off, n_inst, (uint32_t)K1,
                       g.kept, g.key, g.cnt, g.h2);
    DNA_HIP(hipGetLastError());
    // exclusive scan of kept flags -> S (the scan helper scans in place)
    DNA_HIP(hipMemcpyAsync(g.S, g.kept, rad_n * 4, hipMemcpyDeviceToDevice, st));
    DNA_HIP(exclusive_scan_u32(g.S, (uint32_t)rad_n, scan_tmp, st));
    hipLaunchKernelGGL(k_first, dim3(G.gb), dim3(B), 0, st, g.kept, g.S, g.key, rad_n, g.ck, g.firstj);
    uint32_t m_kept = 0;
    DNA_HIP(hipMemcpyAsync(&m_kept, g.S + rad_n, 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    DNA_HIP(hipMemsetAsync(g.misc + 1, 0, 4, st));
    if (m_kept > 1)
        hipLaunchKernelGGL(k_changes, dim3((m_kept + B - 1) / B), dim3(B), 0, st, g.ck, m_kept, g.misc + 1);
    std::vector<uint32_t> cnt(K), h2(K);
    DNA_HIP(hipMemcpyAsync(cnt.data(), g.cnt, K * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipMemcpyAsync(h2.data(), g.h2, K * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipMemcpyAsync(h_misc, g.misc, 16 * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    G.maxN.assign(n_inst, 0);
    G.tr.clear();
    G.bstart.assign(K + 1, 0);
    uint32_t run = 0;
    for (uint64_t k = 0; k < K; k++) {
        G.bstart[k] = run;
        if (cnt[k]) {
            IrtTrack t; memset(&t, 0, sizeof(t));
            t.inst = (uint32_t)(k / K1); t.pid = (uint32_t)(k % K1); t.start = run; t.n = cnt[k]; t.h2 = h2[k];
            if (cnt[k] > G.maxN[t.inst]) G.maxN[t.inst] = cnt[k];
            run += cnt[k];
            if (sel && !(t.inst == 0 && t.pid < n_sel && sel[t.pid])) continue;
            G.tr.push_back(t);
        }
    }
    G.bstart[K] = run;
    G.run = run;
    const uint32_t ntr = (uint32_t)G.tr.size();
    G.contiguous = ntr == 0 || h_misc[1] == ntr - 1;
    G.maxn_all = 0;
    for (auto &t : G.tr) G.maxn_all = t.n > G.maxn_all ? t.n : G.maxn_all;
    if (G.maxn_all >= (1u << 27) - 512u) {       // particle indices carry the species in bits 28-31
        set_error("dna_run_irt: a primary has 2^27 or more entries");
        return DNA_ERR_CAPACITY;
    }
    return DNA_OK;
}

// Per-track capacities of one attempt: particle slots (after an overflow the worker's
// CAP = 2 maxN + 512 of the instance) and event slots ev_mult n + 1024.
static void irt_size(IrtGroup &G, int attempt, int ev_mult, uint64_t &part, uint64_t &evs)
{
    part = 0; evs = 0;
    for (auto &t : G.tr) {
        uint32_t capg = 2 * G.maxN[t.inst] + 512;
        uint32_t capt = 2 * t.n + 512;
        if (attempt > 0) capt = capg;
        t.cap_global = capg; t.cap = capt < capg ? capt : capg;
        t.part_off = part; part += t.cap;
        t.ev_cap = (uint32_t)std::min<uint64_t>((uint64_t)ev_mult * t.n + 1024, 0xFFFFFFF0ull);
        t.ev_off = evs; evs += t.ev_cap;
    }
    // launch the largest tracks first, so that the last workgroups to start are short ones
    for (size_t i = 0; i < G.tr.size(); i++) G.tr[i].idx = (uint32_t)i;
    G.launch = G.tr;
    std::stable_sort(G.launch.begin(), G.launch.end(), [](const IrtTrack &x, const IrtTrack &y) { return x.n > y.n; });
}

// Device buffers of one launch and their byte sizes.
struct IrtBufs {
    char *scratch = nullptr, *events = nullptr;
    unsigned long long *tally = nullptr;   // n_inst x TALLY_STRIDE, then the overflow word
    IrtTrack *tr = nullptr;
    uint32_t *bst = nullptr, *idx = nullptr, *tt = nullptr;
    uint32_t *ovf = nullptr;
    size_t tally_bytes = 0;
};

// per particle slot: position 16, sorted copy 16, birth time 8, species lists 24, alive 1,
// death time 8, then (256-aligned) the list positions 96
static size_t lpos_offset(uint64_t part) { return ((size_t)part * (16 + 16 + 8 + 4 * 6 + 1 + 8) + 8 + 255) & ~(size_t)255; }
static size_t bytes_scratch(uint64_t part, uint64_t rad_n)
{
    return std::max<size_t>(lpos_offset(part) + (size_t)part * 6 * 16 + 256, rad_n / 1024 * 4 + 64);
}
// events: times, pairs and the SamplePDC queue (32 bytes per event slot), then the per-track
// hand-over words and the loop kernel's workgroup order
static size_t bytes_events(uint64_t evs, uint32_t ntr)
{
    return evs * 32 + (size_t)ntr * (TSTATE + 1) * 4 + 256;
}
static size_t bytes_block(const IrtGroup &G, size_t tally_bytes, size_t tt_bytes)
{
    return tally_bytes + 256 + G.tr.size() * sizeof(IrtTrack) + (G.K + 1) * 4 + (size_t)G.run * 4 + tt_bytes + 1536;
}

// carve tally | tracks | bucket starts | bucket indices | per-track tallies out of one block
static void carve_block(const IrtGroup &G, char *base, size_t tally_bytes, bool want_tt, IrtBufs &b)
{
    b.tally_bytes = tally_bytes;
    b.tally = (unsigned long long *)base;
    b.ovf = (uint32_t *)(base + tally_bytes);
    b.tr = (IrtTrack *)(base + ((tally_bytes + 8 + 255) & ~(size_t)255));
    b.bst = (uint32_t *)((char *)b.tr + ((G.tr.size() * sizeof(IrtTrack) + 255) & ~(size_t)255));
    b.idx = b.bst + ((G.K + 1 + 63) & ~(uint64_t)63);
    b.tt = want_tt ? b.idx + (((uint64_t)G.run + 63) & ~(uint64_t)63) : nullptr;
}

// upload the track list and put the entries' indices in b.idx, grouped by track (key
// order) and in input order within a track: for input grouped by primary, k_bucket
// places each entry at its rank among the kept entries; for interleaved primaries, a
// stable radix sort of (key, index) over every entry (the entries that are not particles
// take the key K, past every track) gives the same order for tracks of any size
static int irt_prepare(dna_ctx *c, const IrtGroup &G, const IrtBufs &b)
{
    hipStream_t st = c->stream;
    const uint32_t ntr = (uint32_t)G.tr.size();
    DNA_HIP(hipMemsetAsync(b.tally, 0, b.tally_bytes + 8, st));
    DNA_HIP(hipMemcpyAsync(b.tr, G.launch.data(), ntr * sizeof(IrtTrack), hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(b.bst, G.bstart.data(), (G.K + 1) * 4, hipMemcpyHostToDevice, st));
    const GroupScratch &g = G.g;
    if (G.contiguous) {
        hipLaunchKernelGGL(k_bucket, dim3(G.gb), dim3(256), 0, st, g.kept, g.S, g.key, G.rad_n, b.bst, g.firstj,
                           b.idx);
        DNA_HIP(hipGetLastError());
        return DNA_OK;
    }
    const uint32_t n = (uint32_t)G.rad_n;
    unsigned bits = 1;
    while (bits < 32 && (1ull << bits) <= G.K) bits++;
    size_t tmp_bytes = 0;
    DNA_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                      (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0u, bits, st));
    const size_t arr = ((size_t)n * 4 + 255) & ~(size_t)255;
    DNA_HIP(dgrow<char>(c->irt.sort, c->irt.cap_sort, 4 * arr + tmp_bytes + 256));
    char *base = (char *)c->irt.sort;
    uint32_t *kin = (uint32_t *)base, *vin = (uint32_t *)(base + arr);
    uint32_t *kout = (uint32_t *)(base + 2 * arr), *vout = (uint32_t *)(base + 3 * arr);
    hipLaunchKernelGGL(k_sort_input, dim3(G.gb), dim3(256), 0, st, g.kept, g.key, G.rad_n, (uint32_t)G.K, kin, vin);
    DNA_HIP(hipGetLastError());
    DNA_HIP(rocprim::radix_sort_pairs(base + 4 * arr, tmp_bytes, (const uint32_t *)kin, kout, (const uint32_t *)vin,
                                      vout, n, 0u, bits, st));
    DNA_HIP(hipMemcpyAsync(b.idx, vout, (size_t)G.run * 4, hipMemcpyDeviceToDevice, st));
    return DNA_OK;
}

static IrtArgs irt_args(const IrtGroup &G, const float4 *d_rad, const IrtBufs &b, uint64_t part, uint64_t evs,
                        const double *t_ckpt, int nc, uint64_t seed)
{
    IrtArgs A;
    memset(&A, 0, sizeof(A));
    char *sp = b.scratch;
    A.rad = d_rad; A.idx = b.idx; A.tracks = b.tr;
    A.pos = (float4 *)sp; A.sorted = (float4 *)(sp + part * 16); A.tb = (double *)(sp + part * 32);
    A.lst = (uint32_t *)(sp + part * 40); A.alive = (uint8_t *)(sp + part * 64);
    A.lpos = (float4 *)(sp + lpos_offset(part));
    A.tdeath = nullptr;                         // set with final_ntot, in the space behind alive[]
    A.ev_t = (double *)b.events; A.ev_ab = (uint2 *)(b.events + evs * 8);
    A.pend = (uint4 *)(b.events + evs * 16);
    A.tally = b.tally; A.track_tally = b.tt; A.overflow = b.ovf;
    for (int k = 0; k < nc; k++) A.tc[k] = t_ckpt[k];
    A.nc = nc; A.seed = seed; A.seed_ptr = nullptr;
    A.bm_max = BM_MAX;
    // the sorted copy lives in LDS for tracks of up to 3072 particles (48 KB), in global
    // memory above; the drain's long list reuses that LDS afterwards (at least 128 slots:
    // 512 list entries, twice the largest workgroup, as the small-queue drain needs)
    A.lds_slots = std::max<uint32_t>(std::min<uint32_t>(G.maxn_all, 3072u), 128u);
    // the loop kernel keeps its track's alive bits in LDS when the largest track's fit in
    // AL_LDS_MAX bytes (16 one-wavefront workgroups per CU with LoopShared), else (al_lds 0)
    // in global memory
    uint32_t capmax = 0;
    for (auto &t : G.tr) capmax = t.cap > capmax ? t.cap : capmax;
    const uint32_t alb = 4u * alive_words(capmax);
    A.al_lds = alb <= AL_LDS_MAX ? ((alb + 15u) & ~15u) : 0u;
    A.tstate = (uint32_t *)(b.events + evs * 32);
    A.order = A.tstate + (size_t)G.tr.size() * TSTATE;
    return A;
}

static void irt_launch_pairs(const IrtArgs &A, uint32_t grid, int block, hipStream_t st);

// one IRT launch: the pairs kernel, the drain kernel when the drain runs on its own, the
// loop kernel's workgroup order, the event loop
static hipError_t irt_launch(const IrtGroup &G, const IrtArgs &A, uint32_t grid, int block, hipStream_t st,
                             hipEvent_t before_loop = nullptr)
{
    irt_launch_pairs(A, grid, block, st);
    if (A.drain_at == 2) hipLaunchKernelGGL(k_irt_drain, dim3(grid), dim3(DRAIN_NT), 0, st, A);
    if (A.order) hipLaunchKernelGGL(k_irt_order, dim3((grid + ORDER_CHUNK - 1) / ORDER_CHUNK), dim3(ORDER_NT), 0,
                                    st, (const uint32_t *)A.tstate, grid, A.order);
    if (before_loop) {
        const hipError_t e = hipEventRecord(before_loop, st);
        if (e != hipSuccess) return e;
    }
    // the death-time stores are compiled only into the variants that keep the final state;
    // alive bits in LDS when the largest track's fit in al_lds bytes, else in global memory
    const bool lb = A.al_lds > 0;
    if (A.tdeath) {
        if (lb) hipLaunchKernelGGL((k_irt_loop<true, true>), dim3(grid), dim3(64), (size_t)A.al_lds, st, A);
        else hipLaunchKernelGGL((k_irt_loop<true, false>), dim3(grid), dim3(64), 0, st, A);
    } else {
        if (lb) hipLaunchKernelGGL((k_irt_loop<false, true>), dim3(grid), dim3(64), (size_t)A.al_lds, st, A);
        else hipLaunchKernelGGL((k_irt_loop<false, false>), dim3(grid), dim3(64), 0, st, A);
    }
    return hipSuccess;
}

static void irt_launch_pairs(const IrtArgs &A, uint32_t grid, int block, hipStream_t st)
{
    const size_t shmem = (size_t)A.lds_slots * 16;
    if (A.tdeath) hipLaunchKernelGGL((k_irt_pairs<true>), dim3(grid), dim3(block), shmem, st, A);
    else hipLaunchKernelGGL((k_irt_pairs<false>), dim3(grid), dim3(block), shmem, st, A);
}

// Workgroup size and initial-pair cells (both only shape k_irt_pairs): the options, or
// autotuned per track-size class (log2 of the mean particles per track): the first launch
// of a class times k_irt_pairs for every candidate pair on up to 3072 of its tracks, twice,
// and keeps the fastest best-of-two, so one disturbed trial (another context's work on
// the GPU, a cold cache) does not settle the class for the process.  Results do not depend
// on either; the trial tallies are cleared afterwards.  Small batches (under 1024 tracks)
// take a rule instead: 128 threads below 512 particles per track, else 256, and the cells
// of R_CUT / 2 from 4096 particles per track (tracks longer than a few R_CUT).
static int irt_pick_launch(dna_ctx *c, const IrtGroup &G, IrtArgs &A, const IrtBufs &b, int &block)
{
    hipStream_t st = c->stream;
    const uint32_t ntr = (uint32_t)G.tr.size();
    const uint64_t mean = G.run / ntr;
    auto cell_arg = [](int v) { return v < 0 ? 0 : v; };
    int cell = c->irt_cell;
    block = c->irt_block;
    if (block && cell) { A.cell_m = cell_arg(cell); return DNA_OK; }
    int cls = 0;
    for (uint64_t m = mean; m > 1; m >>= 1) cls++;
    auto it = c->irt_tuned.find(cls);
    if (it != c->irt_tuned.end()) {
        block = it->second.first; A.cell_m = cell_arg(it->second.second); return DNA_OK;
    }
    if (ntr < 1024) {
        if (!block) block = mean < 512 ? 128 : 256;
        if (!cell) cell = mean >= 4096 ? 2 : -1;
        A.cell_m = cell_arg(cell);
        return DNA_OK;
    }
    const uint32_t Kt = std::min<uint32_t>(ntr, 3072u);
    std::vector<int> bs, cs;
    if (block) bs.push_back(block); else bs = {256, 128, 64};
    if (cell) cs.push_back(cell); else cs = {-1, 1, 2, 4};
    std::vector<float> t(bs.size() * cs.size(), 1e30f);
    for (int round = 0; round < 2; round++)
        for (size_t i = 0; i < bs.size(); i++)
            for (size_t j = 0; j < cs.size(); j++) {
                A.cell_m = cell_arg(cs[j]);
                DNA_HIP(hipEventRecord(c->ev[0], st));
                irt_launch_pairs(A, Kt, bs[i], st);
                DNA_HIP(hipGetLastError());
                DNA_HIP(hipEventRecord(c->ev[1], st));
                DNA_HIP(hipEventSynchronize(c->ev[1]));
                float ms = 0;
                DNA_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                float &tk = t[i * cs.size() + j];
                tk = ms < tk ? ms : tk;
            }
    float best = 1e30f;
    for (size_t i = 0; i < bs.size(); i++)
        for (size_t j = 0; j < cs.size(); j++)
            if (t[i * cs.size() + j] < best) { best = t[i * cs.size() + j]; block = bs[i]; cell = cs[j]; }
    c->irt_tuned[cls] = std::make_pair(block, cell);
    A.cell_m = cell_arg(cell);
    DNA_HIP(hipMemsetAsync(b.tally, 0, b.tally_bytes + 8, st));
    return DNA_OK;
}

// host tallies -> dna_chem_out of every instance
static void irt_outputs(const IrtGroup &G, const unsigned long long *ht, const uint32_t *n_therm, const float *E,
                        int nc, float ms, float kms, dna_chem_out *outs)
{
    for (uint32_t i = 0; i < G.n_inst; i++) {
        const unsigned long long *g2 = &ht[(size_t)i * TALLY_STRIDE];
        dna_chem_out &o = outs[i];
        double per100 = (double)n_therm[i] * (double)E[i] / 100.0;
        for (int k = 0; k < nc; k++) for (int s = 0; s < 5; s++) {
            o.tally[k][s] = g2[k * 5 + s];
            o.G[k][s] = per100 > 0 ? (double)o.tally[k][s] / per100 : 0.0;
        }
        for (int r = 0; r < 9; r++) o.rxn_counts[r] = g2[DNA_MAX_CKPT * 5 + r];
        o.n_reacted = g2[DNA_MAX_CKPT * 5 + 9];
        o.n_events = g2[DNA_MAX_CKPT * 5 + 10];
        o.n_pair_cand = g2[DNA_MAX_CKPT * 5 + 11];
        o.n_pair_sampled = g2[DNA_MAX_CKPT * 5 + 12];
        o.n_repair = g2[DNA_MAX_CKPT * 5 + 13];
        for (int q = 0; q < 8; q++) o.prof[q] = g2[DNA_MAX_CKPT * 5 + 14 + q];
        o.n_pair_all = g2[DNA_MAX_CKPT * 5 + 22];
        o.n_repair_scanned = g2[DNA_MAX_CKPT * 5 + 23];
        // static FP64 cost model of the worker's algorithm (counted on the restated worker
        // source), whatever the cells leave out: distance and cut-off test 9 per pair of the
        // all-pairs scan, sqrt + first sampling stage 6, Brownian re-pair step
        // (3 x [log 20 + cos 25 + sqrt + 4] + sqrt(2 D dt) 3 + distance 9 + sampling 6) 167
        o.algo_flops = 9.0 * (double)o.n_pair_all + 6.0 * (double)o.n_pair_sampled + 167.0 * (double)o.n_repair;
        o.t_ms = ms;
        o.t_kernel_ms = kms;
    }
}

static void init_outs(uint32_t n_inst, const std::vector<uint64_t> &off, const double *t_ckpt, int nc, dna_chem_out *outs)
{
    for (uint32_t i = 0; i < n_inst; i++) {
        memset(&outs[i], 0, sizeof(dna_chem_out));
        outs[i].n_ckpt = nc;
        for (int k = 0; k < nc; k++) outs[i].t_ckpt[k] = t_ckpt[k];
        outs[i].chem_n = off[i + 1] - off[i];
    }
}

// One IRT grouped and prepared on the context's stream, its kernels not yet launched: what
// irt_stage leaves for irt_finish.  It refers to the context's IRT buffers and to the
// radical entries it was staged from.
struct IrtStaged {
    IrtGroup G;
    std::vector<uint64_t> off;
    std::vector<uint32_t> n_therm;
    std::vector<float> E;
    std::vector<double> tck;
    int nc = 0;
    const float4 *d_rad = nullptr;
    uint64_t rad_n = 0;
    bool want_tt = false;
    bool empty = false;                 // no entries or no tracks: nothing to launch
    bool resident = false;              // staged from the context's resident radicals
    uint64_t rad_epoch = 0;             // the context's radical epoch when staged
    int attempt = 0, ev_mult = 8;
    uint64_t part = 0, evs = 0;
    IrtBufs b;
    size_t tally_bytes = 0, tt_bytes = 0;
    int block = 256, cell_m = 0;
    bool launched = false;              // the kernels are enqueued (irt_launch_staged)
    uint64_t seed = 0;
    IrtArgs A;
    std::vector<unsigned long long> ht;  // read-back: tallies, overflow word
    std::vector<uint32_t> htt;           // read-back: per-track tallies
};

namespace dna {
void irt_staged_drop(dna_ctx *c)
{
    IrtStaged *S = (IrtStaged *)c->irt_staged;
    // a launched IRT's read-back copies still target S->ht / S->htt: let them land first
    if (S && S->launched) (void)hipStreamSynchronize(c->stream);   // teardown path: nothing to report to
    delete S;
    c->irt_staged = nullptr;
}
}  // namespace dna

// the kernel arguments of a staged IRT (the launch-choice trials take seed 0: their output
// is discarded)
static IrtArgs staged_args(dna_ctx *c, const IrtStaged &S, uint64_t seed)
{
    IrtArgs A = irt_args(S.G, S.d_rad, S.b, S.part, S.evs, S.tck.data(), S.nc, seed);
    A.bm_max = (uint32_t)c->irt_evblocks;
    A.drain_at = c->irt_drain;
    if (!c->irt_order) A.order = nullptr;
    if (c->irt_final) {
        A.final_ntot = c->d_ntot;
        A.tdeath = (double *)((char *)c->irt.scratch + ((S.part * 65 + 7) & ~(uint64_t)7));
    }
    return A;
}

// capacities, buffers, the track list on the device and the launch choice of S.attempt
static int irt_stage_attempt(dna_ctx *c, IrtStaged &S)
{
    const uint32_t ntr = (uint32_t)S.G.tr.size();
    irt_size(S.G, S.attempt, S.ev_mult, S.part, S.evs);
    DNA_HIP(dgrow<char>(c->irt.scratch, c->irt.cap_scratch, bytes_scratch(S.part, S.rad_n)));
    DNA_HIP(dgrow<char>(c->irt.events, c->irt.cap_events, bytes_events(S.evs, ntr)));
    S.tally_bytes = (size_t)S.G.n_inst * TALLY_STRIDE * 8;
    S.tt_bytes = S.want_tt ? (size_t)ntr * S.nc * 5 * 4 : 0;
    DNA_HIP(dgrow<char>(c->irt.tally, c->irt.cap_tally, bytes_block(S.G, S.tally_bytes, S.tt_bytes)));
    S.b = IrtBufs();
    S.b.scratch = (char *)c->irt.scratch; S.b.events = (char *)c->irt.events;
    carve_block(S.G, (char *)c->irt.tally, S.tally_bytes, S.want_tt, S.b);
    int rc = irt_prepare(c, S.G, S.b); if (rc) return rc;
    IrtArgs A = staged_args(c, S, 0);
    rc = irt_pick_launch(c, S.G, A, S.b, S.block); if (rc) return rc;
    S.cell_m = A.cell_m;
    return DNA_OK;
}

// Everything of an IRT before its kernels: grouping of the entries into tracks (host round
// trips for the counts), capacities, the track list and bucket indices on the device, the
// launch choice.  It replaces any earlier final state: the buffers are reused.
static int irt_stage(dna_ctx *c, const float4 *d_rad, uint64_t rad_n, const std::vector<uint64_t> &off,
                     const uint32_t *n_therm, const float *E, const double *t_ckpt, int nc, bool want_tt,
                     IrtStaged &S, const uint8_t *sel = nullptr, uint32_t n_sel = 0)
{
    hipStream_t st = c->stream;
    const uint32_t n_inst = (uint32_t)off.size() - 1;
    int rc = ensure_const(c); if (rc) return rc;
    c->final_ready = false;                   // this call replaces any earlier final state
    S.off = off; S.n_therm.assign(n_therm, n_therm + n_inst); S.E.assign(E, E + n_inst);
    S.tck.assign(t_ckpt, t_ckpt + nc); S.nc = nc;
    S.d_rad = d_rad; S.rad_n = rad_n; S.want_tt = want_tt;
    S.G.n_inst = n_inst;
    DNA_HIP(hipEventRecord(c->ev[5], st));
    if (rad_n == 0) { S.empty = true; return DNA_OK; }
    rc = irt_group(c, d_rad, rad_n, off, sel, n_sel, S.G); if (rc) return rc;
    if (S.G.tr.empty()) { S.empty = true; return DNA_OK; }
    if (c->irt_final) DNA_HIP(dgrow<uint32_t>(*(void **)&c->d_ntot, c->cap_ntot, (uint32_t)S.G.tr.size()));
    return irt_stage_attempt(c, S);
}

// The IRT kernels of a staged IRT and the read-back of its tallies, enqueued on the
// context's stream; c->ev_loop is recorded before the event-loop kernel.
static int irt_launch_staged(dna_ctx *c, IrtStaged &S, uint64_t seed)
{
    hipStream_t st = c->stream;
    S.seed = seed;
    S.launched = true;
    if (S.empty) return DNA_OK;
    const uint32_t ntr = (uint32_t)S.G.tr.size();
    S.A = staged_args(c, S, seed);
    S.A.cell_m = S.cell_m;
    c->irt_last_block = S.block; c->irt_last_cell = S.A.cell_m ? S.A.cell_m : -1;
    DNA_HIP(hipEventRecord(c->ev[7], st));
    DNA_HIP(irt_launch(S.G, S.A, ntr, S.block, st, c->ev_loop));
    DNA_HIP(hipEventRecord(c->ev[4], st));
    DNA_HIP(hipGetLastError());
    // the tallies and the overflow word behind them in one copy
    S.ht.assign((size_t)S.G.n_inst * TALLY_STRIDE + 1, 0ull);
    DNA_HIP(hipMemcpyAsync(S.ht.data(), S.b.tally, S.tally_bytes + 8, hipMemcpyDeviceToHost, st));
    S.htt.assign(S.want_tt ? (size_t)ntr * S.nc * 5 : 0, 0u);
    if (S.want_tt) DNA_HIP(hipMemcpyAsync(S.htt.data(), S.b.tt, S.tt_bytes, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipEventRecord(c->ev[6], st));
    return DNA_OK;
}

// Wait for a launched IRT and fill the outputs (a launch whose capacities overflowed is
// staged again with larger ones and repeated).  track_tally (optional): per-primary species
// counts [n_sel][nc][5] of instance 0.
static int irt_collect(dna_ctx *c, IrtStaged &S, dna_chem_out *outs, uint32_t *track_tally = nullptr,
                       uint32_t n_sel = 0)
{
    hipStream_t st = c->stream;
    const uint32_t n_inst = S.G.n_inst, nc = (uint32_t)S.nc;
    init_outs(n_inst, S.off, S.tck.data(), S.nc, outs);
    c->final_ready = false;
    for (uint32_t i = 0; i < n_inst; i++) outs[i].n_tracks = 0;
    for (auto &t : S.G.tr) outs[t.inst].n_tracks++;
    if (S.empty) {
        if (c->irt_final) { c->final_ready = true; c->final_ntr = 0; c->final_seed = S.seed; }
        return DNA_OK;
    }
    const uint32_t ntr = (uint32_t)S.G.tr.size();
    for (;;) {
        DNA_HIP(hipStreamSynchronize(st));
        const uint32_t ovf = (uint32_t)S.ht[(size_t)n_inst * TALLY_STRIDE];
        if (ovf) {
            if (ovf & 2u) S.ev_mult *= 4;
            if (++S.attempt >= 4) break;
            int rc = irt_stage_attempt(c, S); if (rc) return rc;
            rc = irt_launch_staged(c, S, S.seed); if (rc) return rc;
            continue;
        }
        if (track_tally) {
            memset(track_tally, 0, (size_t)n_sel * nc * 5 * 4);
            for (uint32_t j = 0; j < ntr; j++)
                if (S.G.tr[j].inst == 0 && S.G.tr[j].pid < n_sel)
                    memcpy(track_tally + (size_t)S.G.tr[j].pid * nc * 5, &S.htt[(size_t)j * nc * 5], (size_t)nc * 5 * 4);
        }
        float ms = 0, kms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[5], c->ev[6]);     // timings only
        (void)hipEventElapsedTime(&kms, c->ev[7], c->ev[4]);
        irt_outputs(S.G, S.ht.data(), S.n_therm.data(), S.E.data(), S.nc, ms, kms, outs);
        if (c->irt_final) {
            c->final_ready = true; c->final_ntr = ntr; c->final_seed = S.seed; c->final_tracks = S.b.tr;
            c->final_pos = S.A.pos; c->final_alive = S.A.alive; c->final_tb = S.A.tb; c->final_td = S.A.tdeath;
        }
        return DNA_OK;
    }
    set_error("dna_run_irt: event/particle capacity kept overflowing");
    return DNA_ERR_CAPACITY;
}

static int irt_finish(dna_ctx *c, IrtStaged &S, uint64_t seed, dna_chem_out *outs,
                      uint32_t *track_tally = nullptr, uint32_t n_sel = 0)
{
    int rc = irt_launch_staged(c, S, seed); if (rc) return rc;
    return irt_collect(c, S, outs, track_tally, n_sel);
}

// track_tally (optional): per-primary species counts [n_sel][nc][5] of instance 0.
static int run_irt_device(dna_ctx *c, const float4 *d_rad, uint64_t rad_n, const std::vector<uint64_t> &off,
                          const uint32_t *n_therm, const float *E, const double *t_ckpt, int nc, uint64_t seed,
                          dna_chem_out *outs, const uint8_t *sel = nullptr, uint32_t n_sel = 0,
                          uint32_t *track_tally = nullptr)
{
    irt_staged_drop(c);                       // its buffers are about to be reused
    IrtStaged S;
    int rc = irt_stage(c, d_rad, rad_n, off, n_therm, E, t_ckpt, nc, track_tally != nullptr, S, sel, n_sel);
    if (rc) return rc;
    return irt_finish(c, S, seed, outs, track_tally, n_sel);
}

static int stage_rad(dna_ctx *c, const float *rad, uint64_t rad_n, const float4 **d_rad)
{
    if (rad) {
        DNA_HIP(dgrow<float4>(c->irt.rad_in, c->irt.cap_rad_in, rad_n + 1));
        DNA_HIP(hipMemcpyAsync(c->irt.rad_in, rad, rad_n * 16, hipMemcpyHostToDevice, c->stream));
        *d_rad = (const float4 *)c->irt.rad_in;
    } else {
        if (rad_n > c->rad_n) { set_error("dna_run_irt: rad == NULL but fewer radicals are resident"); return DNA_ERR_ARG; }
        *d_rad = c->d_rad;
    }
    return DNA_OK;
}

extern "C" int dna_run_irt(dna_ctx *c, const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                           const double *t_ckpt, int n_ckpt, uint64_t seed, dna_chem_out *out)
{
    if (!c || !out || !t_ckpt || n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) { set_error("dna_run_irt: bad arguments"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(c->device));
    // the host buffer (if any) is staged into irt.rad_in; the bucket/track block lives in irt.tally
    const float4 *d_rad = nullptr;
    int rc = stage_rad(c, rad, rad_n, &d_rad);
    if (rc) return rc;
    std::vector<uint64_t> off = {0, rad_n};
    return run_irt_device(c, d_rad, rad_n, off, &n_therm, &E_eV, t_ckpt, n_ckpt, seed, out);
}

// dna_run_irt in two calls, so that a caller can group and prepare one batch while the GPU
// still runs another batch's IRT kernels
extern "C" int dna_irt_stage(dna_ctx *c, const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                             const double *t_ckpt, int n_ckpt)
{
    if (!c || !t_ckpt || n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) { set_error("dna_irt_stage: bad arguments"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(c->device));
    irt_staged_drop(c);
    const float4 *d_rad = nullptr;
    int rc = stage_rad(c, rad, rad_n, &d_rad);
    if (rc) return rc;
    IrtStaged *S = new IrtStaged();
    S->resident = rad == nullptr;
    S->rad_epoch = c->rad_epoch;
    std::vector<uint64_t> off = {0, rad_n};
    rc = irt_stage(c, d_rad, rad_n, off, &n_therm, &E_eV, t_ckpt, n_ckpt, false, *S);
    if (rc) { delete S; return rc; }
    c->irt_staged = S;
    return DNA_OK;
}

// the staged IRT of c, checked for dna_irt_launch / dna_irt_finish (fn names the caller)
static IrtStaged *staged_for_launch(dna_ctx *c, const char *fn)
{
    IrtStaged *S = (IrtStaged *)c->irt_staged;
    if (!S || S->launched) {
        set_error(std::string(fn) + ": no IRT staged on this context (dna_irt_stage, and no IRT call since)");
        return nullptr;
    }
    if (S->resident && S->rad_epoch != c->rad_epoch) {
        irt_staged_drop(c);
        set_error(std::string(fn) + ": the resident radicals changed since dna_irt_stage");
        return nullptr;
    }
    return S;
}

extern "C" int dna_irt_finish(dna_ctx *c, uint64_t seed, dna_chem_out *out)
{
    if (!c || !out) { set_error("dna_irt_finish: bad arguments"); return DNA_ERR_ARG; }
    IrtStaged *S = staged_for_launch(c, "dna_irt_finish");
    if (!S) return DNA_ERR_ARG;
    DNA_HIP(hipSetDevice(c->device));
    const int rc = irt_finish(c, *S, seed, out);
    irt_staged_drop(c);
    return rc;
}

extern "C" int dna_irt_launch(dna_ctx *c, uint64_t seed)
{
    if (!c) { set_error("dna_irt_launch: bad arguments"); return DNA_ERR_ARG; }
    IrtStaged *S = staged_for_launch(c, "dna_irt_launch");
    if (!S) return DNA_ERR_ARG;
    DNA_HIP(hipSetDevice(c->device));
    const int rc = irt_launch_staged(c, *S, seed);
    if (rc) irt_staged_drop(c);
    return rc;
}

extern "C" int dna_irt_wait(dna_ctx *c, int what, dna_chem_out *out)
{
    if (!c || (what != DNA_IRT_WAIT_LOOP && what != DNA_IRT_WAIT_ALL) || (what == DNA_IRT_WAIT_ALL && !out)) {
        set_error("dna_irt_wait: bad arguments"); return DNA_ERR_ARG;
    }
    IrtStaged *S = (IrtStaged *)c->irt_staged;
    if (!S || !S->launched) { set_error("dna_irt_wait: no IRT launched on this context (dna_irt_launch)"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(c->device));
    if (what == DNA_IRT_WAIT_LOOP) {
        if (!S->empty) DNA_HIP(hipEventSynchronize(c->ev_loop));
        return DNA_OK;
    }
    const int rc = irt_collect(c, *S, out);
    irt_staged_drop(c);
    return rc;
}

static int check_batch_args(const char *fn, const uint64_t *inst_off, uint32_t n_inst, std::vector<uint64_t> &off)
{
    off.assign(inst_off, inst_off + n_inst + 1);
    for (uint32_t i = 0; i < n_inst; i++)
        if (off[i + 1] < off[i]) { set_error(std::string(fn) + ": offsets must be non-decreasing"); return DNA_ERR_ARG; }
    for (auto &o : off) o -= inst_off[0];
    return DNA_OK;
}

extern "C" int dna_run_irt_batch(dna_ctx *c, const float *rad, const uint64_t *inst_off, uint32_t n_inst,
                                 const uint32_t *n_therm, const float *E_eV, const double *t_ckpt, int n_ckpt,
                                 uint64_t seed, dna_chem_out *out)
{
    if (!c || !out || !rad || !inst_off || !n_therm || !E_eV || !t_ckpt || n_inst == 0 || n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) {
        set_error("dna_run_irt_batch: bad arguments"); return DNA_ERR_ARG;
    }
    std::vector<uint64_t> off;
    int rc = check_batch_args("dna_run_irt_batch", inst_off, n_inst, off); if (rc) return rc;
    DNA_HIP(hipSetDevice(c->device));
    uint64_t rad_n = off[n_inst];
    const float4 *d_rad = nullptr;
    rc = stage_rad(c, rad + 4 * inst_off[0], rad_n, &d_rad);
    if (rc) return rc;
    return run_irt_device(c, d_rad, rad_n, off, n_therm, E_eV, t_ckpt, n_ckpt, seed, out);
}

extern "C" int dna_run_irt_tracks(dna_ctx *c, const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                                  const uint8_t *select, uint32_t n_prim, const double *t_ckpt, int n_ckpt,
                                  uint64_t seed, uint32_t *track_tally, dna_chem_out *out)
{
    if (!c || !out || !t_ckpt || n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT || ((select || track_tally) && n_prim == 0)) {
        set_error("dna_run_irt_tracks: bad arguments"); return DNA_ERR_ARG;
    }
    DNA_HIP(hipSetDevice(c->device));
    const float4 *d_rad = nullptr;
    int rc = stage_rad(c, rad, rad_n, &d_rad);
    if (rc) return rc;
    if (track_tally && rad_n == 0) memset(track_tally, 0, (size_t)n_prim * n_ckpt * 5 * 4);
    std::vector<uint64_t> off = {0, rad_n};
    std::vector<uint8_t> all;
    const uint8_t *sel = select;
    if (!sel && track_tally) { all.assign(n_prim, 1); sel = all.data(); }
    return run_irt_device(c, d_rad, rad_n, off, &n_therm, &E_eV, t_ckpt, n_ckpt, seed, out, sel,
                          sel ? n_prim : 0, track_tally);
}

// ------------------------------------------------------------ captured batches
// A plan holds its own copy of the radicals, the track list and every buffer of the launch,
// and one instantiated hipGraph: clear the tallies (a fill kernel: no memset nodes), copy
// the seed in, the IRT kernel, copy tallies and the overflow word out.  A replay with a new
// seed runs the same instances with fresh random streams.
struct dna_irt_plan {
    dna_ctx *c = nullptr;
    IrtGroup G;
    std::vector<uint64_t> off;
    std::vector<uint32_t> n_therm;
    std::vector<float> E;
    std::vector<double> tck;
    int nc = 0;
    float4 *rad = nullptr;
    char *buf = nullptr;
    unsigned long long *h_tally = nullptr;   // pinned: tallies then the overflow word
    uint64_t *h_seed = nullptr;              // pinned
    uint64_t *d_seed = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int block = 256, cell = -1, ev_mult = 8, attempt = 0;
    size_t tally_bytes = 0;
};

static void plan_release_graph(dna_irt_plan *P)
{
    if (P->exec) { (void)hipGraphExecDestroy(P->exec); P->exec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
    if (P->buf) { (void)hipFree(P->buf); P->buf = nullptr; }
}

// (re)build buffers and graph for the current ev_mult / attempt
static int plan_build(dna_irt_plan *P)
{
    dna_ctx *c = P->c;
    hipStream_t st = c->stream;
    plan_release_graph(P);
    irt_staged_drop(c);                       // the grouping scratch is reused
    int rc = irt_group(c, P->rad, P->G.rad_n, P->off, nullptr, 0, P->G); if (rc) return rc;
    IrtGroup &G = P->G;
    const uint32_t ntr = (uint32_t)G.tr.size();
    uint64_t part = 0, evs = 0;
    irt_size(G, P->attempt, P->ev_mult, part, evs);
    P->tally_bytes = (size_t)G.n_inst * TALLY_STRIDE * 8;
    const size_t s_bytes = (bytes_scratch(part, G.rad_n) + 255) & ~(size_t)255;
    const size_t e_bytes = (bytes_events(evs, ntr) + 255) & ~(size_t)255;
    const size_t b_bytes = bytes_block(G, P->tally_bytes, 0);
    DNA_HIP(hipMalloc((void **)&P->buf, s_bytes + e_bytes + b_bytes + 256));
    IrtBufs b;
    b.scratch = P->buf; b.events = P->buf + s_bytes;
    carve_block(G, P->buf + s_bytes + e_bytes, P->tally_bytes, false, b);
    P->d_seed = (uint64_t *)(P->buf + s_bytes + e_bytes + b_bytes);
    rc = irt_prepare(c, G, b); if (rc) return rc;
    IrtArgs A = irt_args(G, P->rad, b, part, evs, P->tck.data(), P->nc, 0);
    A.bm_max = (uint32_t)c->irt_evblocks;
    A.drain_at = c->irt_drain;
    if (!c->irt_order) A.order = nullptr;
    if (ntr) { rc = irt_pick_launch(c, G, A, b, P->block); if (rc) return rc; }
    P->cell = A.cell_m ? A.cell_m : -1;
    A.seed_ptr = P->d_seed;
    DNA_HIP(hipStreamSynchronize(st));
    // capture
    DNA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const size_t words = P->tally_bytes / 8 + 1;
    // the capture is always ended; a failure while recording is reported after it
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, b.tally, words);
    hipError_t ce = hipGetLastError();
    if (ce == hipSuccess) ce = hipMemcpyAsync(P->d_seed, P->h_seed, 8, hipMemcpyHostToDevice, st);
    if (ce == hipSuccess && ntr) { ce = irt_launch(G, A, ntr, P->block, st); if (ce == hipSuccess) ce = hipGetLastError(); }
    if (ce == hipSuccess) ce = hipMemcpyAsync(P->h_tally, b.tally, P->tally_bytes + 8, hipMemcpyDeviceToHost, st);
    hipError_t e = hipStreamEndCapture(st, &P->graph);
    if (e == hipSuccess) e = ce;
    if (e != hipSuccess) {
        if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
        set_error(std::string("dna_irt_plan: capture failed: ") + hipGetErrorString(e));
        return DNA_ERR_HIP;
    }
    DNA_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    return DNA_OK;
}

extern "C" void dna_irt_plan_destroy(dna_irt_plan *P)
{
    if (!P) return;
    (void)hipStreamSynchronize(P->c->stream);
    plan_release_graph(P);
    if (P->rad) (void)hipFree(P->rad);
    if (P->h_tally) (void)hipHostFree(P->h_tally);
    if (P->h_seed) (void)hipHostFree(P->h_seed);
    dna_ctx *c = P->c;
    delete P;
    ctx_release(c);
}

extern "C" int dna_irt_plan_create(dna_ctx *c, const float *rad, const uint64_t *inst_off, uint32_t n_inst,
                                   const uint32_t *n_therm, const float *E_eV, const double *t_ckpt, int n_ckpt,
                                   dna_irt_plan **out)
{
    if (!c || !out || !inst_off || !n_therm || !E_eV || !t_ckpt || n_inst == 0 || n_ckpt < 1 || n_ckpt > DNA_MAX_CKPT) {
        set_error("dna_irt_plan_create: bad arguments"); return DNA_ERR_ARG;
    }
    std::vector<uint64_t> off;
    int rc = check_batch_args("dna_irt_plan_create", inst_off, n_inst, off); if (rc) return rc;
    const uint64_t rad_n = off[n_inst];
    if (rad_n == 0) { set_error("dna_irt_plan_create: no radical entries"); return DNA_ERR_ARG; }
    if (!rad && inst_off[0] + rad_n > c->rad_n) {
        set_error("dna_irt_plan_create: rad == NULL but fewer radicals are resident"); return DNA_ERR_ARG;
    }
    DNA_HIP(hipSetDevice(c->device));
    rc = ensure_const(c); if (rc) return rc;
    dna_irt_plan *P = new dna_irt_plan();
    ctx_retain(c);                                // given back by dna_irt_plan_destroy
    P->c = c; P->off = off; P->nc = n_ckpt;
    P->n_therm.assign(n_therm, n_therm + n_inst);
    P->E.assign(E_eV, E_eV + n_inst);
    P->tck.assign(t_ckpt, t_ckpt + n_ckpt);
    P->G.rad_n = rad_n;
    hipError_t e = hipMalloc((void **)&P->rad, rad_n * 16 + 16);
    if (e == hipSuccess) {
        if (rad) e = hipMemcpy(P->rad, rad + 4 * inst_off[0], rad_n * 16, hipMemcpyHostToDevice);
        else e = hipMemcpy(P->rad, c->d_rad + inst_off[0], rad_n * 16, hipMemcpyDeviceToDevice);
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&P->h_tally, (size_t)n_inst * TALLY_STRIDE * 8 + 64, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&P->h_seed, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        set_error(std::string("dna_irt_plan_create: ") + hipGetErrorString(e));
        dna_irt_plan_destroy(P);
        return DNA_ERR_HIP;
    }
    rc = plan_build(P);
    if (rc) { dna_irt_plan_destroy(P); return rc; }
    *out = P;
    return DNA_OK;
}

extern "C" int dna_irt_plan_run(dna_irt_plan *P, uint64_t seed, dna_chem_out *out)
{
    if (!P || !out) { set_error("dna_irt_plan_run: bad arguments"); return DNA_ERR_ARG; }
    dna_ctx *c = P->c;
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    for (int tries = 0; tries < 4; tries++) {
        DNA_HIP(hipStreamSynchronize(st));
        *P->h_seed = seed;
        DNA_HIP(hipEventRecord(c->ev[5], st));
        DNA_HIP(hipGraphLaunch(P->exec, st));
        DNA_HIP(hipEventRecord(c->ev[6], st));
        DNA_HIP(hipStreamSynchronize(st));
        const uint32_t ovf = (uint32_t)P->h_tally[P->tally_bytes / 8];
        if (ovf) {
            // capacities exceeded: rebuild with the worker's CAP and more event slots
            if (ovf & 2u) P->ev_mult *= 4;
            P->attempt = 1;
            int rc = plan_build(P); if (rc) return rc;
            continue;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[5], c->ev[6]);
        init_outs(P->G.n_inst, P->off, P->tck.data(), P->nc, out);
        for (uint32_t i = 0; i < P->G.n_inst; i++) out[i].n_tracks = 0;
        for (auto &t : P->G.tr) out[t.inst].n_tracks++;
        irt_outputs(P->G, P->h_tally, P->n_therm.data(), P->E.data(), P->nc, ms, ms, out);
        c->irt_last_block = P->block; c->irt_last_cell = P->cell;
        return DNA_OK;
    }
    set_error("dna_irt_plan_run: event/particle capacity kept overflowing");
    return DNA_ERR_CAPACITY;
}

extern "C" int dna_irt_oh_final(dna_ctx *c, float *out_pos4, uint64_t cap, uint64_t *n_out)
{
    if (!c || !n_out) { set_error("dna_irt_oh_final: NULL argument"); return DNA_ERR_ARG; }
    if (!c->final_ready) {
        set_error("dna_irt_oh_final: no IRT run with DNA_OPT_IRT_FINAL = 1 since the last IRT call");
        return DNA_ERR_ARG;
    }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const uint32_t ntr = c->final_ntr;
    if (ntr == 0) { c->oh_n = 0; *n_out = 0; return DNA_OK; }
    // counts per track, exclusive scan (total at [ntr]), then the ordered write
    DNA_HIP(dgrow<uint32_t>(c->irt.part, c->irt.cap_part, 2 * (size_t)ntr + 2 + ((size_t)ntr + 1023) / 1024 + 8));
    uint32_t *cnt = (uint32_t *)c->irt.part, *tmp = cnt + ntr + 2;
    const IrtTrack *tr = (const IrtTrack *)c->final_tracks;
    hipLaunchKernelGGL(k_oh_count, dim3(ntr), dim3(256), 0, st, tr, c->final_alive, c->d_ntot, cnt);
    DNA_HIP(hipGetLastError());
    DNA_HIP(exclusive_scan_u32(cnt, ntr, tmp, st));
    uint32_t total = 0;
    DNA_HIP(hipMemcpyAsync(&total, cnt + ntr, 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    DNA_HIP(dgrow<float4>(*(void **)&c->d_oh, c->cap_oh, (size_t)total + 1));
    hipLaunchKernelGGL(k_oh_write, dim3(ntr), dim3(256), 0, st, tr, c->final_pos, c->final_alive, c->final_tb,
                       c->d_ntot, cnt, c->final_seed, c->d_oh);
    DNA_HIP(hipGetLastError());
    c->oh_n = total;
    *n_out = total;
    if (out_pos4) {
        if (cap < total) {
            DNA_HIP(hipStreamSynchronize(st));
            set_error("dna_irt_oh_final: output capacity below the OH count (see *n_out)");
            return DNA_ERR_CAPACITY;
        }
        DNA_HIP(hipMemcpyAsync(out_pos4, c->d_oh, (size_t)total * 16, hipMemcpyDeviceToHost, st));
    }
    DNA_HIP(hipStreamSynchronize(st));
    return DNA_OK;
}

// The reaction table every IRT run uses (the worker's rxnSigma / rxnSigmaEff / rxnRc /
// rxnKdif / rxnKact / rxnProb, public/irt-worker.js:233-260): out[9][6].  Host only.
extern "C" int dna_irt_reactions(double *out)
{
    if (!out) { set_error("dna_irt_reactions: NULL argument"); return DNA_ERR_ARG; }
    IrtConst C;
    build_const(C);
    for (int r = 0; r < 9; r++) {
        double *o = out + 6 * r;
        o[0] = C.sigma[r]; o[1] = C.sigma_eff[r]; o[2] = C.rc[r]; o[3] = C.kdif[r]; o[4] = C.kact[r]; o[5] = C.prob[r];
    }
    return DNA_OK;
}

extern "C" int dna_irt_particles(dna_ctx *c, double *out6, uint64_t cap, uint64_t *n_out)
{
    if (!c || !n_out) { set_error("dna_irt_particles: NULL argument"); return DNA_ERR_ARG; }
    if (!c->final_ready) {
        set_error("dna_irt_particles: no IRT run with DNA_OPT_IRT_FINAL = 1 since the last IRT call");
        return DNA_ERR_ARG;
    }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const uint32_t ntr = c->final_ntr;
    if (ntr == 0) { *n_out = 0; return DNA_OK; }
    DNA_HIP(dgrow<uint32_t>(c->irt.part, c->irt.cap_part, 2 * (size_t)ntr + 2 + ((size_t)ntr + 1023) / 1024 + 8));
    uint32_t *off = (uint32_t *)c->irt.part, *tmp = off + ntr + 2;
    DNA_HIP(hipMemcpyAsync(off, c->d_ntot, (size_t)ntr * 4, hipMemcpyDeviceToDevice, st));
    DNA_HIP(exclusive_scan_u32(off, ntr, tmp, st));
    uint32_t total = 0;
    DNA_HIP(hipMemcpyAsync(&total, off + ntr, 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    *n_out = total;
    if (!out6) return DNA_OK;
    if (cap < total) { set_error("dna_irt_particles: output capacity below the particle count (see *n_out)"); return DNA_ERR_CAPACITY; }
    double *d = nullptr;
    DNA_HIP(hipMalloc(&d, (size_t)total * 48 + 48));
    hipLaunchKernelGGL(k_particles_write, dim3(ntr), dim3(256), 0, st, (const IrtTrack *)c->final_tracks, c->final_pos,
                       c->final_tb, c->final_td, c->d_ntot, off, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out6, d, (size_t)total * 48, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    DNA_HIP(e);
    return DNA_OK;
}
