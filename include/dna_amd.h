/*
 * dna_amd.h — C ABI of the MI355X Geant4-DNA chain (transport + prechemistry,
 * IRT chemistry, SSB/DSB scoring, spatial-hash binning, chemistry surrogate).
 *
 * Plain C types only: pointers, sizes, PODs.  Every entry point returns an int
 * status (DNA_OK = 0); on failure dna_last_error() returns a message for the calling
 * thread.  A context owns device buffers and host (pinned) output buffers; output
 * pointers handed back in *_out structs stay valid until the next call on the same
 * context, and the caller copies what it keeps (the reference hands out
 * `.slice(0)` copies of mapped ranges, src/gpu/dispatch.ts:122,157,175,237).
 * Calls on one context are serialised by the caller; use one context per thread.
 *
 * Reference interfaces replaced (allaunthefox/webgpu-dna):
 *   dna_ctx_create/destroy   initGPU + allocateBuffers + createPipelines
 *                            (src/gpu/device.ts:11, src/gpu/buffers.ts:48,
 *                             src/gpu/pipelines.ts:40; ensurePipelines src/app.ts:65)
 *   dna_run_transport        Phase A + Phase B of runAtEnergy (src/gpu/dispatch.ts:85-192),
 *                            kernels src/shaders/primary.wgsl:114 and secondary.wgsl:83
 *   dna_run_irt              runChemistryWorker (src/chemistry/worker.ts:32) ->
 *                            public/irt-worker.js:370-699
 *   dna_run_irt_batch        many independent worker invocations in one launch
 *   dna_score                scoreDirectSSB_events + combineHits + clusterDSB
 *                            (src/scoring/ssb-dsb.ts:135,413,348) with the LCG of
 *                            makeSsbRng (ssb-dsb.ts:421) carried in/out
 *   dna_hash_cells           spatial-hash cell assignment of src/shaders/chemistry.wgsl:66-80,
 *                            110-122 (cell = floor(p / cell), Teschner key mod 2^bits)
 *   dna_surrogate_*          no reference counterpart (learned IRT surrogate)
 */
#ifndef DNA_AMD_H
#define DNA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNA_OK              0
#define DNA_ERR_ARG         1   /* invalid argument                          */
#define DNA_ERR_NO_DEVICE   2   /* no usable GPU (reference: initGPU -> null) */
#define DNA_ERR_HIP         3   /* HIP runtime error                         */
#define DNA_ERR_TABLES      4   /* cross-section table file missing/corrupt  */
#define DNA_ERR_CAPACITY    5   /* a size exceeds a documented maximum        */

/* ------------------------------------------------------------------ types */

typedef struct dna_ctx dna_ctx;   /* opaque context */

/* Transport parameters: the uniform struct P of primary.wgsl:1-7 as packed by
 * writePrimaryParams (src/gpu/dispatch.ts:22-41), plus three fields the reference
 * fixes elsewhere (seed: dispatch.ts:98 passes floor(E_eV); max_sec_steps:
 * MAX_SEC_STEPS=2000, constants.ts:66; dir_mode replaces P._pad3). */
typedef struct dna_params {
    uint32_t n;            /* number of primaries                                  */
    float    box;          /* world half-width (nm)                                */
    float    ce;           /* tracking cut / solvation threshold (eV)              */
    uint32_t ms;           /* max primary steps (reference 65536)                  */
    float    be;           /* beam energy (eV), used when energies == NULL         */
    uint32_t max_sec;      /* secondary cap, 0 = none: only the first max_sec
                              secondaries in canonical (primary, emission) order are
                              transported, the rest count in sec_dropped (the
                              reference's MAX_SEC, dispatch.ts:124-126)              */
    uint32_t vc;           /* dose voxels per axis (reference 128)                 */
    uint32_t max_rad;      /* radical cap, 0 = none: the first max_rad entries of the
                              canonical radical buffer are kept, the rest count in
                              rad_dropped (MAX_RAD, dispatch.ts:190-192).  Device
                              buffers are sized by the library either way.          */
    float    start_half;   /* start positions uniform in +-start_half (0 = origin) */
    uint32_t dna_enable;   /* kernel-level DNA proximity counter (counters[4])     */
    uint32_t dna_grid_n;   /* fibers per grid side                                 */
    uint32_t dir_mode;     /* 0 isotropic start direction (reference), 1 = +z      */
    float    dna_rise;     /* nm per bp                                            */
    float    dna_spacing;  /* fiber spacing (nm)                                   */
    float    dna_x0;       /* x of bp 0                                            */
    float    dna_r_bb;     /* backbone radius (nm)                                 */
    uint32_t seed;         /* RNG seed (reference: floor(E_eV))                    */
    uint32_t max_sec_steps;/* secondary step cap (reference 2000)                  */
} dna_params;

/* Per-primary result, struct R of primary.wgsl:9 (32 bytes). */
typedef struct dna_result {
    float    path;   /* total path length (nm)                 */
    float    prod;   /* path while E > 12 eV (nm)              */
    float    fE;     /* final energy (eV)                      */
    uint32_t ni;     /* ionisations                            */
    uint32_t nx;     /* electronic excitations                 */
    uint32_t esc;    /* 1 if the primary left the world        */
    float    mx;     /* max distance from the start point (nm) */
    uint32_t _pad;
} dna_result;

/* Copy flags for dna_run_transport: which large arrays to copy to host. */
#define DNA_COPY_RESULTS  1u
#define DNA_COPY_DOSE     2u
#define DNA_COPY_RAD      4u
#define DNA_COPY_ALL      7u

typedef struct dna_transport_out {
    const dna_result *results;   /* [n]            (DNA_COPY_RESULTS)                    */
    const uint32_t   *dose;      /* [vc^3], 100 units per eV (DNA_COPY_DOSE)             */
    const float      *rad_buf;   /* [4*rad_n_stored] (x,y,z, pid*8+code) (DNA_COPY_RAD)  */
    uint32_t counters[8];        /* slot map primary.wgsl:24-35                          */
    uint32_t sec_stats[8];       /* [1] cutoff, [2] bounds, [3] steps, [4] tertiary ions */
    uint32_t dbg[8];             /* [0] sum ni, [1] sum nx, [2] escaped, [3] primaries,
                                    [4] sum(u32(W_transfer*10)) over primary ionisations */
    uint64_t dose_sum;           /* sum of dose[] (exact)                                */
    uint64_t rad_n_raw, rad_n_stored, rad_dropped;  /* emitted, kept, beyond max_rad        */
    uint64_t sec_n_raw, sec_n, sec_dropped;         /* emitted, transported, beyond max_sec */
    double   t_prim_ms, t_sec_ms, t_total_ms;   /* device time of the two phases        */
} dna_transport_out;

/* IRT chemistry output for one instance (irt-worker.js:670-698). */
#define DNA_MAX_CKPT 32
#define DNA_N_RXN 9
typedef struct dna_chem_out {
    int      n_ckpt;
    double   t_ckpt[DNA_MAX_CKPT];
    /* species tallies per checkpoint, exact integers: 0 OH, 1 e-aq, 2 H, 3 H2O2, 4 H2 */
    uint64_t tally[DNA_MAX_CKPT][5];
    double   G[DNA_MAX_CKPT][5];       /* tally / (n_therm * E / 100)        */
    uint64_t rxn_counts[DNA_N_RXN];
    uint64_t n_reacted;
    uint64_t chem_n;                   /* rad entries handed in              */
    uint64_t n_tracks;                 /* primaries with at least one radical */
    uint64_t n_events;                 /* candidate reaction events sampled  */
    double   t_ms;                     /* device time of the whole call      */
    double   t_kernel_ms;              /* device time of the per-track kernel */
    /* work: pairs with a reaction channel whose distance was evaluated, pairs within the
     * cut-off (reaction time sampled), product re-pairings with a Brownian step */
    uint64_t n_pair_cand, n_pair_sampled, n_repair;
    double   algo_flops;               /* FP64 operations by the static cost model */
    uint64_t prof[8];                  /* per-phase 100 MHz clock sums over tracks: [0] load
                                          and the initial-pair cells, [1] initial pairs,
                                          [2] event scan, [3] reaction, [4] re-pairing scan,
                                          [5] event-loop iterations, [6] initial sampling
                                          drain, [7] re-pairing drain */
    /* the worker's all-pairs count (pairs of alive initial particles with a reaction
     * channel; n_pair_cand is the part of it that the cells left to visit), and the species
     * list slots the re-pairing scanned (n_repair counts the worker's re-pairings: every
     * alive partner with a channel).  algo_flops counts the worker's work:
     * 9 n_pair_all + 6 n_pair_sampled + 167 n_repair. */
    uint64_t n_pair_all, n_repair_scanned;
} dna_chem_out;

/* DNA fiber target, DNATarget of src/physics/types.ts:23-37 / buildDNATarget
 * (src/physics/dna-geometry.ts:16-68).  Arrays are f32 as in the reference. */
typedef struct dna_dna_target {
    double rise, r_bb, spacing_nm, x0, L_nm;
    uint32_t n_bp_per, n_fibers, n_bp, grid_N;
    const float *fy, *fz;      /* [n_fibers]     */
    const float *rbb0, *rbb1;  /* [2 * n_bp_per] */
} dna_dna_target;

typedef struct dna_score_out {
    uint64_t ssb_count, candidates, in_reach;   /* DirectSSBResult (types.ts:131-136)   */
    uint64_t dsb, ssb0, ssb1;                    /* DSBClusterResult (types.ts:139-144)  */
    double   dsb_expected;
    const uint8_t *hits;                         /* [2*n_bp] combined hits, if requested */
    double   t_ms;
} dna_score_out;

/* scoreIndirectSSB result (src/physics/types.ts:122-128) without the hit array */
typedef struct dna_indirect_out {
    uint64_t ssb0, ssb1, candidates, in_reach;
} dna_indirect_out;

/* ------------------------------------------------------------ entry points */

/* Create a context on HIP device `device`, loading the cross-section table file
 * (NULL = the cross_sections.bin shipped next to the library).  dna_ctx_destroy releases
 * the caller's handle; surrogates and captured IRT plans made on the context keep it alive
 * until they are destroyed too, so the handles may be destroyed in any order.  A context
 * must not be used after dna_ctx_destroy. */
int  dna_ctx_create(int device, const char *table_path, struct dna_ctx **out);
void dna_ctx_destroy(struct dna_ctx *ctx);
const char *dna_last_error(void);
const char *dna_version(void);

/* Phase A + B: transport n primaries (energies[i] per primary, or params->be for all
 * when energies == NULL) with in-kernel prechemistry.  The radical buffer is left on
 * the device in canonical order (by primary id, then by emitting particle — the
 * primary first, then its secondaries in emission order — then emission order) for
 * the chained dna_run_irt / dna_score calls that pass rad == NULL. */
int dna_run_transport(struct dna_ctx *ctx, const dna_params *p, const float *energies,
                      uint32_t copy_flags, dna_transport_out *out);

/* Phase C: IRT to the last checkpoint (reference 1000 ns) for every primary found in
 * rad (host pointer, 4*rad_n floats) or, when rad == NULL, in the device-resident
 * radical buffer of the last dna_run_transport.  seed keys the counter-based RNG. */
int dna_run_irt(struct dna_ctx *ctx, const float *rad, uint64_t rad_n, uint32_t n_therm,
                float E_eV, const double *t_ckpt, int n_ckpt, uint64_t seed,
                dna_chem_out *out);

/* dna_run_irt in two calls.  No counterpart in the reference (its worker runs one batch at
 * a time): it is for callers that pipeline batches, grouping and preparing the next batch's
 * IRT while the GPU still runs the current one's kernels.  dna_irt_stage does everything of
 * dna_run_irt before its kernels (grouping of the entries into tracks, with host round trips
 * for the counts; capacities; the track list on the device; the launch choice) on the
 * context's stream; dna_irt_finish launches the kernels and fills *out as dna_run_irt does,
 * with the same results for the same seed (t_ms counts from dna_irt_stage).  One IRT can be
 * staged per context; any other IRT call on the context, and destroying it, drop it.  With
 * rad == NULL, a dna_run_transport between the two calls makes dna_irt_finish fail
 * (DNA_ERR_ARG). */
int dna_irt_stage(struct dna_ctx *ctx, const float *rad, uint64_t rad_n, uint32_t n_therm,
                  float E_eV, const double *t_ckpt, int n_ckpt);
int dna_irt_finish(struct dna_ctx *ctx, uint64_t seed, dna_chem_out *out);

/* dna_irt_finish in two calls (no counterpart in the reference either): dna_irt_launch
 * enqueues the staged IRT's kernels on the context's stream and returns; dna_irt_wait with
 * DNA_IRT_WAIT_LOOP blocks until the kernels before the event loop (the pair enumeration
 * and its drain) are done, which is when a pipelining caller can launch another batch's IRT
 * to fill the GPU as this one's event-loop workgroups finish; with DNA_IRT_WAIT_ALL it
 * blocks until the IRT is done and fills *out as dna_irt_finish does. */
#define DNA_IRT_WAIT_LOOP 0
#define DNA_IRT_WAIT_ALL  1
int dna_irt_launch(struct dna_ctx *ctx, uint64_t seed);
int dna_irt_wait(struct dna_ctx *ctx, int what, dna_chem_out *out);

/* Many independent IRT instances in one launch: instance i owns radical entries
 * [inst_off[i], inst_off[i+1]) of rad (host pointer, n_inst+1 offsets). */
int dna_run_irt_batch(struct dna_ctx *ctx, const float *rad, const uint64_t *inst_off,
                      uint32_t n_inst, const uint32_t *n_therm, const float *E_eV,
                      const double *t_ckpt, int n_ckpt, uint64_t seed,
                      dna_chem_out *out /* [n_inst] */);

/* Direct-event SSB scoring + DSB clustering on tgt.  rad == NULL uses the
 * device-resident radical buffer.  *lcg_state is the makeSsbRng state before the call
 * and is advanced by the number of draws consumed.  copy_hits != 0 copies the
 * combined per-strand hit array to out->hits. */
int dna_score(struct dna_ctx *ctx, const dna_dna_target *tgt, const float *rad,
              uint64_t rad_n, uint32_t *lcg_state, int copy_hits, dna_score_out *out);

/* dna_score with the indirect-SSB pass first, as scoreDamageAt10keV does when chemistry
 * positions exist (src/app.ts:176-186): scoreIndirectSSB (src/scoring/ssb-dsb.ts:43-121)
 * on the n_oh alive-OH positions oh (host float4, or NULL for the list left on the device
 * by dna_irt_oh_final), the LCG continuing into the direct pass, and combineHits
 * (ssb-dsb.ts:413-418) before clusterDSB.  ind receives IndirectSSBResult; out the direct
 * counts and the clustering of the combined hits. */
int dna_score_oh(struct dna_ctx *ctx, const dna_dna_target *target, const float *rad, uint64_t rad_n,
                 const float *oh, uint64_t n_oh, uint32_t *lcg_state, int copy_hits,
                 dna_score_out *out, dna_indirect_out *ind);

/* Allocation-free scoring.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  dna_score on device-resident input, with every scratch buffer in one
 * fixed array per GPU (about 17 MB, held by each GPU the library runs on until the process
 * ends); the call allocates nothing.  rad_dev: a device pointer to rad_n (x, y, z, w) float4
 * entries, or NULL for the radicals resident from the last dna_run_transport.  The combined
 * hit array is not returned (out->hits = NULL); counts, dsb_expected and the LCG state are
 * dna_score's.  Limits: targets of at most DNA_SCORE_NOALLOC_MAX_FIBERS fibers,
 * DNA_SCORE_NOALLOC_MAX_BP_PER bp per fiber and DNA_SCORE_NOALLOC_MAX_BP bp in all
 * (DNA_ERR_CAPACITY beyond); any rad_n below 2^32 - 16, scored in chunks of
 * DNA_SCORE_NOALLOC_CHUNK.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two contexts or threads) race, and the caller must serialise them. */
#define DNA_SCORE_NOALLOC_CHUNK      (1u << 20)
#define DNA_SCORE_NOALLOC_MAX_FIBERS 1024u
#define DNA_SCORE_NOALLOC_MAX_BP_PER 16384u
#define DNA_SCORE_NOALLOC_MAX_BP     (1u << 22)
int dna_score_noalloc(struct dna_ctx *ctx, const dna_dna_target *target, const float *rad_dev, uint64_t rad_n,
                      uint32_t *lcg_state, dna_score_out *out);

/* Spatial-hash binning: for n points (x,y,z,w float4, host pointer) compute
 * cell = floor(p / cell_nm) (f32 division) and key = Teschner hash mod 2^hash_bits;
 * then counting-sort the points by key: cell_start[2^hash_bits + 1] (CSR offsets) and
 * order[n] (point indices, stable within a bucket).  Outputs are host pointers. */
int dna_hash_cells(struct dna_ctx *ctx, const float *pos4, uint64_t n, float cell_nm,
                   uint32_t hash_bits, uint32_t *key_out, uint32_t *cell_start_out,
                   uint32_t *order_out);

/* The step's table lookups on the GPU for given energies (host arrays): which 0 =
 * xs_all (helpers.wgsl:15-27) -> out[3n] (sigma_ion, sigma_exc, sigma_el in nm^2),
 * 1 = xs_shell_fracs (:29-43) -> out[5n], 2 = xs_exc_fracs (:45-64) -> out[5n],
 * 3 = xs_vib_total (:135-156) -> out[n], 4 = xs_dea (:172-186) -> out[n],
 * 5 = xs_el_cos (:66-82) at (E[i], r[i]) -> out[n].  The same code and LDS tables the
 * transport kernels use; it replaces the in-shader evaluations that the reference's
 * cross-section experiments probe (experiments/level-1, level-2). */
int dna_xs_probe(struct dna_ctx *ctx, int which, const float *E, const float *r, uint64_t n, float *out);

/* Launch options (no reference counterpart: WebGPU fixes its workgroup sizes in the
 * shaders).  DNA_OPT_IRT_BLOCK: IRT workgroup size, 0 = autotune (default: the first
 * launch of each track-size class times 256/128/64-thread workgroups twice on up to 3072
 * of its tracks and keeps the fastest best-of-two), or 64 / 128 / 256.  DNA_OPT_IRT_BLOCK_LAST (get
 * only): the size the last IRT launch used.  Results do not depend on these options. */
#define DNA_OPT_IRT_BLOCK       1
#define DNA_OPT_IRT_BLOCK_LAST  2
/* DNA_OPT_IRT_FINAL: 1 = dna_run_irt keeps each track's final state on the device for
 * dna_irt_oh_final and dna_irt_particles, and records each particle's death time
 * (default 0). */
#define DNA_OPT_IRT_FINAL       3
/* Transport launch knobs, 0 = autotune (default): batches under 1024 primaries take the
 * rule below; the first batch of each (log2 primaries, log2 mean energy) class times the
 * candidates twice each and keeps the fastest (Phase A lanes on Phase A's time, then the
 * radical chunk on both phases', then the work chunk on Phase B's).  Results do not depend
 * on them.
 * DNA_OPT_PA_LANES: active lanes per 64-lane wavefront of the primary-transport kernel, or
 * 1, 2, 4, ..., 64 (rule and middle candidate: the widest power of two that still gives
 * every CU one wavefront, 16 for 4096 primaries on 256 CUs; candidates half and double).
 * DNA_OPT_RAD_CHUNK: radical slots a wavefront reserves with one atomic, or a power of two
 * from 256 to 4096 (rule 256; candidates 256, 512, 1024).
 * DNA_OPT_WORK_CHUNK: secondaries a Phase B wavefront claims with one atomic, or a power
 * of two from 64 to 1024 (rule 64; candidates 64, 128, 256).
 * DNA_OPT_*_LAST (get only): the values the last dna_run_transport used. */
#define DNA_OPT_PA_LANES        4
#define DNA_OPT_RAD_CHUNK       7
#define DNA_OPT_WORK_CHUNK      8
#define DNA_OPT_PA_LANES_LAST   9
#define DNA_OPT_RAD_CHUNK_LAST  10
#define DNA_OPT_WORK_CHUNK_LAST 11
/* DNA_OPT_MLP_FUSED: 1 (default) = the surrogate's forward pass (dna_surrogate_forward,
 * _grad, _train, _predict) in one fused kernel launch when its shapes allow it (padded
 * input, hidden width and 2 n_out multiples of 16 and at most 256); 0 = one GEMM launch per
 * layer.  Both give the same values. */
#define DNA_OPT_MLP_FUSED       12
/* DNA_OPT_IRT_DRAIN: where the IRT samples its initial pairs' queue (the rest of sampleIRT,
 * SamplePDC's rejection loops included): 0 in the one-wavefront event-loop kernel before
 * its loop, 1 (default) in the pairs kernel's workgroup after the enumeration, 2 in a kernel
 * of its own between the two.  Results do not depend on it. */
#define DNA_OPT_IRT_DRAIN       13
/* DNA_OPT_IRT_ORDER: 1 (default) runs the IRT's event-loop workgroups heaviest track first
 * (queued initial pairs + initial particles, within runs of 4096 tracks), 0 in launch
 * order.  Results do not depend on it. */
#define DNA_OPT_IRT_ORDER       14
/* DNA_OPT_IRT_EVBLOCKS: at most this many event blocks per track (1..128, default 128) in
 * the IRT's event loop, which keeps each block's least event and selects among the blocks:
 * a block holds 64 events, doubled as often as the track's events need.  Fewer blocks make
 * a cheaper selection and a dearer refresh.  Results do not depend on it. */
#define DNA_OPT_IRT_EVBLOCKS    15
/* DNA_OPT_IRT_CELL: cells of the initial-pair search of the IRT (the spatial hash that
 * stands in for the worker's all-pairs scan, public/irt-worker.js:541-558; the cell
 * semantics of src/shaders/chemistry.wgsl:66-80 with one axis): 0 = autotune together with
 * DNA_OPT_IRT_BLOCK (every pair of candidates timed twice, fastest best-of-two kept per
 * track-size class), -1 = none (all pairs of each reaction channel), or m = 1, 2, 4:
 * slabs of width R_CUT / m (R_CUT = 1.45 + 2 sqrt(8 x 9.46 x 1000) nm, plus 0.1% + 1 nm)
 * across the track's principal axis, each particle paired with the particles of the
 * m slabs on either side.  DNA_OPT_IRT_CELL_LAST (get only): the cells the last IRT launch
 * used.  The cells leave out only pairs beyond R_CUT: results do not depend on them. */
#define DNA_OPT_IRT_CELL        5
#define DNA_OPT_IRT_CELL_LAST   6
int dna_set_option(struct dna_ctx *ctx, int option, int value);
int dna_get_option(struct dna_ctx *ctx, int option, int *value);

/* The OH alive at 1 us after the last dna_run_irt made with DNA_OPT_IRT_FINAL = 1, as
 * (x, y, z, 0) float4 in track order (primaries ascending) and particle order within a
 * track, each moved by one Brownian step over 1 us - t_birth (sigma = sqrt(2 D_OH dt) per
 * axis, Box-Muller on draws 32-37 of the particle's own counter-based stream).  The worker
 * has no positions at 1 us (chem_pos_final = null, src/chemistry/worker.ts:69-70); this is
 * the input this library gives to the indirect-SSB scoring (dna_score_oh).  The list stays
 * on the device (oh == NULL in dna_score_oh); out_pos4 (host, may be NULL) receives it when
 * cap >= *n_out, else DNA_ERR_CAPACITY with *n_out set. */
int dna_irt_oh_final(struct dna_ctx *ctx, float *out_pos4, uint64_t cap, uint64_t *n_out);

/* Every particle slot of every track at the end of the last dna_run_irt made with
 * DNA_OPT_IRT_FINAL = 1, in track order: out6[n][6] doubles (x, y, z, pid * 8 + rad_buf
 * code (OH- = 6), t_birth, t_death = +inf if it never reacted).  A particle is alive at a
 * checkpoint t iff t_birth < t <= t_death, which reproduces the checkpoint tallies.  This
 * feeds the reference's 4D snapshot export (src/splat/export.ts:32-112, which takes
 * ChemSnapshot {label, t_ns, pos, alive}, src/physics/types.ts:57-63) from IRT runs.
 * out6 == NULL returns the count only; cap < *n_out gives DNA_ERR_CAPACITY. */
int dna_irt_particles(struct dna_ctx *ctx, double *out6, uint64_t cap, uint64_t *n_out);

/* The reaction table of the IRT (public/irt-worker.js:233-260, the worker's rxnSigma,
 * rxnSigmaEff, rxnRc, rxnKdif, rxnKact, rxnProb): out[9][6] = (sigma nm, sigma_eff nm,
 * rc nm, k_dif and k_act nm^3/ns, contact probability) per reaction in the order of
 * dna_chem_out.rxn_counts.  The worker reports sigma and rc with each reaction count
 * (rxn_info, irt-worker.js:690-694).  Needs no device. */
int dna_irt_reactions(double *out);

/* Captured batches: the instances of dna_run_irt_batch (rad == NULL: entries
 * [inst_off[0], inst_off[n_inst]) of the radicals resident from the last transport),
 * planned once (grouping, buffers, workgroup size) and captured as one hipGraph;
 * dna_irt_plan_run replays it for a seed and fills out[n_inst], identical to
 * dna_run_irt_batch with the same seed.  A plan owns its buffers and a copy of the
 * radicals; a replay that exceeds the planned capacities is re-planned transparently.
 * (The reference runs one Worker per call, src/chemistry/worker.ts:32.) */
typedef struct dna_irt_plan dna_irt_plan;
int  dna_irt_plan_create(struct dna_ctx *ctx, const float *rad, const uint64_t *inst_off,
                         uint32_t n_inst, const uint32_t *n_therm, const float *E_eV,
                         const double *t_ckpt, int n_ckpt, dna_irt_plan **out);
int  dna_irt_plan_run(dna_irt_plan *plan, uint64_t seed, dna_chem_out *out /* [n_inst] */);
void dna_irt_plan_destroy(dna_irt_plan *plan);

/* Phase C on a subset of tracks with per-track output (the reference runs every track;
 * this is the split the surrogate fallback needs).  Same semantics as dna_run_irt, plus:
 * select[pid] != 0 (pid < n_prim) picks the primaries whose tracks run (NULL = all); a
 * selected track's history is the one it has in the full run.  track_tally (host,
 * [n_prim][n_ckpt][5], may be NULL) receives per-primary species counts; with it, only
 * primaries pid < n_prim run.  out aggregates the tracks that ran. */
int dna_run_irt_tracks(struct dna_ctx *ctx, const float *rad, uint64_t rad_n, uint32_t n_therm,
                       float E_eV, const uint8_t *select, uint32_t n_prim, const double *t_ckpt,
                       int n_ckpt, uint64_t seed, uint32_t *track_tally, dna_chem_out *out);

/* ------------------------------------------------- learned chemistry surrogate
 * No reference counterpart.  An MLP n_in -> n_hidden (x n_layers, ReLU) -> 2*n_out in
 * BF16 on the matrix cores with FP32 accumulation; outputs are a mean and a log-variance
 * per target.  Parameters are one flat FP32 vector: for each layer l the weights
 * [din_l][dout_l] (row-major; din_0 = n_in rounded up to a multiple of 32, padded rows
 * zero), then the biases of every layer.  Inputs and targets are normalised on the
 * device with dna_surrogate_set_norm (identity by default). */
typedef struct dna_surrogate dna_surrogate;

int  dna_surrogate_create(struct dna_ctx *ctx, int n_in, int n_hidden, int n_layers, int n_out,
                          uint64_t seed, dna_surrogate **out);
void dna_surrogate_destroy(dna_surrogate *s);
int  dna_surrogate_shape(const dna_surrogate *s, int *n_in, int *n_hidden, int *n_layers,
                         int *n_out, uint64_t *n_params);
int  dna_surrogate_get_params(dna_surrogate *s, float *params);
int  dna_surrogate_set_params(dna_surrogate *s, const float *params);
int  dna_surrogate_set_norm(dna_surrogate *s, const float *x_mean, const float *x_std,
                            const float *y_mean, const float *y_std);
/* raw outputs Y [B][2*n_out] = (normalised mean, log-variance) for X [B][n_in] */
int  dna_surrogate_forward(dna_surrogate *s, const float *X, int B, float *Y);
/* mean Gaussian NLL and its gradient w.r.t. every parameter (flat layout) on one batch */
int  dna_surrogate_grad(dna_surrogate *s, const float *X, const float *T, int B,
                        double *loss, float *grad);
/* Adam, shuffled mini-batches; *mean_loss = mean NLL over the last epoch */
int  dna_surrogate_train(dna_surrogate *s, const float *X, const float *T, int N, int batch,
                         int epochs, float lr, float weight_decay, uint64_t seed,
                         double *mean_loss);
/* Per-track features [n][70]: 64-bin histogram of distance from the track centroid (log
 * bins 0.5 nm .. 20 um, fractions), initial OH, e-aq, H, H3O+, H2 per 100 eV, log10(E/keV).
 * rad == NULL: track p = primary p of the last dna_run_transport (n must equal its n). */
int  dna_track_features(dna_surrogate *s, const float *rad, const uint64_t *inst_off,
                        uint32_t n, const float *E, float *F);
/* Surrogate chemistry for the n primaries of the last dna_run_transport: mean and sigma
 * per target ([n][n_out], may be NULL) and fallback[p] = 1 where a sigma exceeds
 * sigma_max; those tracks go to dna_run_irt_tracks.  t_ms (may be NULL): device ms of
 * features, network, whole call. */
int  dna_surrogate_predict(dna_surrogate *s, const float *E, uint32_t n, float sigma_max,
                           float *mu, float *sigma, uint8_t *fallback, uint32_t *n_fallback,
                           float *t_ms);

#ifdef __cplusplus
}
#endif
#endif /* DNA_AMD_H */
