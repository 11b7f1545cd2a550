// dna_internal.h — declarations shared by the HIP translation units of libdna_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include "../../include/dna_amd.h"
#include "../../include/dna_tables.h"

#define DNA_HIP(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) { \
    dna::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); return DNA_ERR_HIP; } } while (0)

namespace dna {

void set_error(const std::string &msg);

// Shares of a context (see dna_ctx::refs): surrogates and plans take one at creation and
// give it back when destroyed; the last release frees the context.
void ctx_retain(dna_ctx *c);
void ctx_release(dna_ctx *c);
void irt_staged_drop(dna_ctx *c);          // frees a staged IRT (irt.hip)

// Device address of dna_score_noalloc's fixed array on `device` (the current device),
// looked up once per device.
int score_scratch_init(int device, void **ptr);

// Secondary-electron record of the device work queue (64 bytes).
struct SecRec {
    float4 posE;      // x, y, z, E
    float4 dir;       // dx, dy, dz, unused
    uint4  rng;       // xoshiro128++ state
    uint32_t pid;     // parent primary
    uint32_t k;       // 1-based emission index within the parent
    uint32_t pad0, pad1;
};

// Device buffers and counters used by one transport run.
// Guide table of the Born CDF rows: for shell s, energy row e and k = 0..64,
// GUIDE[s][e][k] = the number of the first 249 entries of the row below k/64 (64 for k = 64
// is replaced by 249): a bisection for r in [k/64, (k+1)/64) can start from
// [GUIDE[k], GUIDE[k+1]] and ends on the same index as over the whole row.
#define DNA_GUIDE_K 64
#define DNA_GUIDE_BYTES (5 * DNA_N_WE * (DNA_GUIDE_K + 1))
#define DNA_GUIDE_PAD ((DNA_GUIDE_BYTES + 3) & ~3)
void build_born_guide(const dna_tables &t, uint8_t *guide /* DNA_GUIDE_PAD bytes */);

struct TransportDev {
    const dna_tables *tab;      // device copy of the tables
    const uint8_t *guide;       // Born CDF guide table (DNA_GUIDE_PAD bytes)
    dna_params p;
    const float *energies;      // [n] or null
    dna_result *results;        // [n]
    uint32_t *dose;             // [vc^3]
    float4   *rad_raw;          // [rad_cap] unordered emissions
    uint64_t *rad_key;          // [rad_cap] pid<<40 | k<<20 | seq
    uint32_t rad_cap;
    SecRec   *sec;              // [sec_cap]
    uint32_t sec_cap;
    uint32_t *prim_nsec;        // [n] secondaries emitted by each primary
    uint32_t *prim_nrad;        // [n] radical entries emitted by each primary
    uint32_t *sec_nrad;         // [sec_cap] radical entries emitted by each secondary
    uint32_t *ctr;              // [32] global counters, layout below
    // secondary cap: with sec_base (exclusive scan of prim_nsec) set, Phase B transports only
    // the secondaries whose canonical rank sec_base[pid] + k - 1 is below sec_keep
    const uint32_t *sec_base;
    uint32_t sec_keep;
    uint32_t rad_chunk;         // radical slots a wavefront reserves at once (>= 192)
    uint32_t work_chunk;        // secondaries a wavefront claims at once (>= 64)
};
// ctr layout
enum { C_OH = 0, C_EAQ = 1, C_H = 2, C_H3O = 3, C_DNA = 4, C_H2 = 5, C_SEC = 6, C_RAD = 7,
       C_SST = 8 /* 8..15 sec_stats */, C_DBG = 16 /* 16..23 dbg */, C_PNEXT = 24, C_SNEXT = 25,
       C_WRAP = 26 /* set when the radical slot counter would pass 2^32 */, C_NCTR = 32 };

int num_cu();                                  // compute units of the current device
uint32_t phase_a_lanes(uint32_t n, int opt);
size_t transport_slot_headroom(uint32_t n, uint32_t lanes, uint32_t rad_chunk);
hipError_t launch_phase_a(const TransportDev &d, uint32_t lanes, hipStream_t st);
hipError_t launch_phase_b(const TransportDev &d, uint32_t sec_n, hipStream_t st);
hipError_t launch_xs_probe(const dna_tables *T, const uint8_t *guide, const float *E, const float *r, uint32_t n,
                           int which, float *out, hipStream_t st);

// canonical ordering: returns the device canonical buffer through out pointers
struct CanonScratch {
    uint32_t *base_p;     // [n+1]
    uint32_t *cnt;        // [P+1]
    uint32_t *block;      // scan partials
    size_t    block_cap;
};
hipError_t launch_canonicalize(const TransportDev &d, uint32_t n, uint32_t sec_n, uint32_t rad_n,
                               CanonScratch &cs, float4 *rad_out, hipStream_t st);

// exclusive scan of u32 (n elements) in place; out[n] = total. tmp holds >= ceil(n/1024)+1 words.
hipError_t exclusive_scan_u32(uint32_t *data, uint32_t n, uint32_t *tmp, hipStream_t st);

}  // namespace dna
