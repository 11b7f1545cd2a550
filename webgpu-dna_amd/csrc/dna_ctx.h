// dna_ctx.h — the opaque context behind the C ABI.
#pragma once
#include <atomic>
#include <map>
#include <array>
#include "dna_internal.h"

struct IrtBuffers {
    void   *scratch = nullptr;  size_t cap_scratch = 0;   // per-track particle arrays
    void   *events = nullptr;   size_t cap_events = 0;    // per-track candidate events
    void   *tracks = nullptr;   size_t cap_tracks = 0;    // track descriptors
    void   *part = nullptr;     size_t cap_part = 0;      // grouping scratch
    void   *rad_in = nullptr;   size_t cap_rad_in = 0;    // host-supplied radicals
    void   *tally = nullptr;    size_t cap_tally = 0;     // per-instance tallies
    void   *sort = nullptr;     size_t cap_sort = 0;      // stable key sort of interleaved input
};

struct ScoreBuffers {
    void *hitbits = nullptr;  size_t cap_hitbits = 0;
    void *flags = nullptr;    size_t cap_flags = 0;
    void *rank = nullptr;     size_t cap_rank = 0;
    void *hits_u8 = nullptr;  size_t cap_hits_u8 = 0;
    void *rad_in = nullptr;   size_t cap_rad_in = 0;
    void *tgt = nullptr;      size_t cap_tgt = 0;
    uint8_t *h_hits = nullptr; size_t hcap_hits = 0;
};

struct HashBuffers {
    void *buf = nullptr; size_t cap = 0;
};

struct dna_ctx {
    // Owners of the context: the handle returned by dna_ctx_create, plus one per surrogate and
    // per captured IRT plan made on it (they use its stream and buffers).  dna_ctx_destroy
    // drops the creator's share; the memory goes with the last share, so destroying the
    // context before its surrogates and plans leaves them working until they are destroyed.
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    dna_tables h_tab;
    dna_tables *d_tab = nullptr;
    uint8_t *d_guide = nullptr;          // Born CDF guide table
    uint32_t *d_ctr = nullptr, *h_ctr = nullptr;
    unsigned long long *d_u64 = nullptr;
    unsigned long long h_u64_tmp = 0;
    hipEvent_t ev[8] = {};
    // transport
    dna_result *d_results = nullptr;  size_t cap_n_results = 0;
    uint32_t *d_prim_nsec = nullptr;   size_t cap_prim_nsec = 0;
    uint32_t *d_prim_nrad = nullptr;   size_t cap_prim_nrad = 0;
    float    *d_energies = nullptr;    size_t cap_energies = 0;
    uint32_t *d_base_p = nullptr;      size_t cap_base_p = 0;
    uint32_t *d_dose = nullptr;        size_t cap_dose = 0;
    float4   *d_rad_raw = nullptr;     size_t cap_rad_raw = 0;
    uint64_t *d_rad_key = nullptr;     size_t cap_rad_key = 0;
    float4   *d_rad = nullptr;         size_t cap_rad = 0;      // canonical radicals
    dna::SecRec *d_sec = nullptr;      size_t cap_sec = 0;
    uint32_t *d_sec_nrad = nullptr;    size_t cap_sec_nrad = 0;
    uint32_t *d_cnt = nullptr;         size_t cap_cnt = 0;
    uint32_t *d_block = nullptr;       size_t cap_block = 0;
    uint32_t *d_sec_base = nullptr;    size_t cap_sec_base = 0;  // canonical secondary ranks (max_sec)
    uint64_t rad_n = 0;                // canonical radicals resident on the device
    uint32_t last_n = 0;
    // host outputs (pinned)
    dna_result *h_results = nullptr;   size_t hcap_results = 0;
    uint32_t *h_dose = nullptr;        size_t hcap_dose = 0;
    float *h_rad = nullptr;            size_t hcap_rad = 0;
    IrtBuffers irt;
    ScoreBuffers sc;
    HashBuffers hs;
    // launch tuning: IRT workgroup size and initial-pair cells (0 = autotune per track-size
    // class; cells: -1 = none, m = cells of R_CUT / m)
    int irt_block = 0;
    int irt_cell = 0;
    int irt_drain = 1;                 // DNA_OPT_IRT_DRAIN: where the initial pairs are drained
    int irt_order = 1;                 // DNA_OPT_IRT_ORDER: loop workgroups heaviest track first
    int irt_evblocks = 128;            // DNA_OPT_IRT_EVBLOCKS: event blocks per track at most
    int irt_last_block = 0;            // workgroup size of the last IRT launch
    int irt_last_cell = 0;             // cells of the last IRT launch (-1 = none)
    std::map<int, std::pair<int, int>> irt_tuned;   // log2(mean particles per track) -> (size, cells)
    // final IRT state (DNA_OPT_IRT_FINAL) and the alive-OH list made from it
    int irt_final = 0;
    // transport knobs (0 = autotune per (log2 primaries, log2 mean energy) class) and the
    // values of the last run
    int mlp_fused = 1;                 // DNA_OPT_MLP_FUSED: surrogate forward in one launch
    int pa_lanes = 0, rad_chunk = 0, work_chunk = 0;
    int last_pa_lanes = 0, last_rad_chunk = 0, last_work_chunk = 0;
    std::map<std::pair<int, int>, std::array<uint32_t, 3>> tr_tuned;
    bool final_ready = false;
    uint32_t final_ntr = 0;
    uint64_t final_seed = 0;
    const void *final_tracks = nullptr;          // device IrtTrack[final_ntr], launch order
    const float4 *final_pos = nullptr;
    const uint8_t *final_alive = nullptr;
    const double *final_tb = nullptr;
    const double *final_td = nullptr;
    uint32_t *d_ntot = nullptr;        size_t cap_ntot = 0;    // particles per track at the end
    float4 *d_oh = nullptr;            size_t cap_oh = 0;      // alive OH at 1 us
    uint64_t oh_n = 0;
    void *na_scratch = nullptr;        // device address of the fixed scoring array (dna_score_noalloc)
    // an IRT staged by dna_irt_stage for dna_irt_finish (dropped by any other IRT call), and
    // the count of transport runs that rewrote the resident radicals
    void *irt_staged = nullptr;
    uint64_t rad_epoch = 0;
    hipEvent_t ev_loop = nullptr;      // recorded before a launched IRT's event-loop kernel
};
