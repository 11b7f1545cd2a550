/*
 * oracle.h — CPU restatement of the reference Geant4-DNA chain.  TEST
 * INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it as the checker (and as the timed CPU baseline); the product library
 * never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to the
 * allaunthefox/webgpu-dna checkout).  Numerics: transport in FP32 with the elementary
 * functions of include/dna_fmath.h, IRT in FP64, scoring in FP64 with the reference's
 * 32-bit LCG — the same arithmetic the HIP kernels use, so results compare exactly.
 *
 * How each part is pinned (golden vectors, closed forms, transcriptions) is in the
 * docstrings of the tests/ that check it.
 */
#ifndef DNA_ORACLE_H
#define DNA_ORACLE_H

#include <stdint.h>
#include "../include/dna_amd.h"
#include "../include/dna_tables.h"

#ifdef __cplusplus
extern "C" {
#endif

int  oracle_load_tables(const char *path);           /* 0 on success */
const dna_tables *oracle_tables(void);

/* elementary-function probe for accuracy tests: fn 0 logf, 1 expf, 2 sinf, 3 cosf,
 * 4 powf(x,1.3), 10 log, 11 exp, 12 cos, 13 sinh, 14 cosh, 15 erfcx, 16 erfcinv */
void oracle_fmath32(int fn, const float *x, float *y, int64_t n);
void oracle_fmath64(int fn, const double *x, double *y, int64_t n);

/* table lookups of helpers.wgsl, vectorised for golden tests.  which: 0 xs_all (3 out),
 * 1 xs_shell_fracs (5), 2 xs_exc_fracs (5), 3 xs_vib_total (1), 4 xs_dea (1) */
void oracle_xs(int which, const float *E, float *out, int64_t n);
float oracle_xs_el_cos(float E, float r);
float oracle_sample_W(float E, uint32_t shell, float r);
uint32_t oracle_sample_vib_mode(float E, float r);
float oracle_meesungnoen_sigma(float k);

/* RNG: xoshiro128++ (helpers.wgsl:188-190) and the host seeding (buffers.ts:91-104) */
void oracle_seed_primary(uint32_t np, uint32_t seed, uint32_t *state /* 4*np */);
void oracle_rng_draws(const uint32_t *state4, uint32_t n, uint32_t *raw, float *rf);

/* Transport (Phase A + B).  Outputs are allocated by the oracle and freed with
 * oracle_free(); rad_buf is in canonical order. */
typedef struct oracle_transport_out {
    dna_result *results;       /* [n] */
    uint32_t   *dose;          /* [vc^3] */
    float      *rad_buf;       /* [4*rad_n] */
    uint64_t    rad_n;
    uint64_t    sec_n;
    uint32_t    counters[8], sec_stats[8], dbg[8];
    uint64_t    dose_sum;
    uint64_t    rad_n_raw;     /* emitted; rad_n = kept (first max_rad in canonical order) */
    uint64_t    sec_n_raw;     /* emitted; sec_n = transported (first max_sec)            */
} oracle_transport_out;
int  oracle_transport(const dna_params *p, const float *energies, int n_threads,
                      oracle_transport_out *out);
/* transport of a subset of primaries [first, first+count) (CPU-baseline sampling) */
int  oracle_transport_range(const dna_params *p, const float *energies, uint32_t first,
                            uint32_t count, int n_threads, oracle_transport_out *out);
void oracle_free(void *ptr);

/* IRT (irt-worker.js).  Same contract as dna_run_irt. */
int oracle_irt(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
               const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads,
               dna_chem_out *out);
int oracle_irt_batch(const float *rad, const uint64_t *off, uint32_t n_inst, const uint32_t *n_therm,
                     const float *E, const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads,
                     dna_chem_out *out);
/* oracle_irt plus the OH alive at 1 us (positions after one Brownian step to 1 us) of every
 * track in track order; *oh_out is malloc'ed, release with oracle_free */
int oracle_irt_oh(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                  const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads, dna_chem_out *out,
                  float **oh_out, uint64_t *oh_n);
/* oracle_irt plus every particle slot of every track at the end (track order): [n][6]
 * doubles (x, y, z, pid * 8 + code, t_birth, t_death = inf if it never reacted) */
int oracle_irt_particles(const float *rad, uint64_t rad_n, uint32_t n_therm, float E_eV,
                         const double *t_ckpt, int n_ckpt, uint64_t seed, int n_threads, dna_chem_out *out,
                         double **pt_out, uint64_t *pt_n);
/* reaction parameter table (irt-worker.js:240-279): 9 x {sigma, sigma_eff, rc, kdif, kact, prob} */
void oracle_irt_params(double *out54);
/* one IRT reaction-time draw for reaction ri at distance r0 with the RNG stream of
 * (seed, track, a, b): returns t (ns) or -1. */
double oracle_irt_sample(double r0, int ri, uint64_t seed, uint32_t track, uint32_t a, uint32_t b);

/* Scoring (ssb-dsb.ts) */
int oracle_score(const dna_dna_target *tgt, const float *rad, uint64_t rad_n,
                 uint32_t *lcg_state, uint8_t *hits_out /* [2*n_bp] or NULL */,
                 dna_score_out *out);

void oracle_cluster(const dna_dna_target *tgt, const uint8_t *hits, dna_score_out *out);
/* indirect SSB from alive-OH positions first, then direct, combined hits, clustering */
int oracle_score_oh(const dna_dna_target *tgt, const float *rad, uint64_t rad_n, const float *oh, uint64_t n_oh,
                    uint32_t *lcg_state, uint8_t *hits_out, dna_score_out *out, dna_indirect_out *ind);
/* Spatial hash (chemistry.wgsl:66-80,110-122) */
void oracle_hash_cells(const float *pos4, uint64_t n, float cell_nm, uint32_t hash_bits,
                       uint32_t *key_out, uint32_t *cell_start_out, uint32_t *order_out);

#ifdef __cplusplus
}
#endif
#endif
