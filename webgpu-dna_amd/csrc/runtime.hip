// runtime.hip — context, device memory and the transport entry point of the C ABI
// (include/dna_amd.h).  Replaces initGPU/allocateBuffers/createPipelines and the Phase A/B
// part of runAtEnergy (src/gpu/device.ts:11, buffers.ts:48, pipelines.ts:40,
// dispatch.ts:85-192).
#include "dna_internal.h"
#include "dna_ctx.h"
#include <dlfcn.h>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>

namespace dna {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace dna

using namespace dna;

extern "C" const char *dna_last_error(void) { return g_err.c_str(); }
extern "C" const char *dna_version(void) { return "dna_amd 0.1 (gfx950)"; }

static std::string default_table_path()
{
    Dl_info info;
    if (dladdr((void *)&default_table_path, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.rfind('/');
        std::string dir = (s == std::string::npos) ? "." : p.substr(0, s);
        return dir + "/../data/cross_sections.bin";
    }
    return "cross_sections.bin";
}

template <class T> static hipError_t grow(T *&ptr, size_t &cap, size_t need)
{
    if (need <= cap && ptr) return hipSuccess;
    if (ptr) { hipError_t e = hipFree(ptr); if (e != hipSuccess) return e; ptr = nullptr; }
    size_t n = need < 16 ? 16 : need;
    hipError_t e = hipMalloc((void **)&ptr, n * sizeof(T));
    if (e == hipSuccess) cap = n; else cap = 0;
    return e;
}

template <class T> static hipError_t grow_host(T *&ptr, size_t &cap, size_t need)
{
    if (need <= cap && ptr) return hipSuccess;
    if (ptr) { (void)hipHostFree(ptr); ptr = nullptr; }
    size_t n = need < 16 ? 16 : need;
    hipError_t e = hipHostMalloc((void **)&ptr, n * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) cap = n; else cap = 0;
    return e;
}

extern "C" int dna_ctx_create(int device, const char *table_path, dna_ctx **out)
{
    if (!out) { set_error("dna_ctx_create: out is NULL"); return DNA_ERR_ARG; }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { set_error("no HIP device"); return DNA_ERR_NO_DEVICE; }
    if (device < 0 || device >= ndev || device >= 64) { set_error("device index out of range"); return DNA_ERR_ARG; }
    std::string path = table_path ? std::string(table_path) : default_table_path();
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { set_error("cannot open table file " + path); return DNA_ERR_TABLES; }
    std::vector<char> buf(sizeof(dna_tables) + 1);
    size_t got = fread(buf.data(), 1, buf.size(), f);
    fclose(f);
    if (got != sizeof(dna_tables) || sizeof(dna_tables) != DNA_TABLES_BYTES) {
        set_error("table file " + path + " has the wrong size"); return DNA_ERR_TABLES;
    }
    DNA_HIP(hipSetDevice(device));
    dna_ctx *c = new dna_ctx();
    c->device = device;
    memcpy(&c->h_tab, buf.data(), sizeof(dna_tables));
    DNA_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    DNA_HIP(hipMalloc((void **)&c->d_tab, sizeof(dna_tables)));
    DNA_HIP(hipMemcpy(c->d_tab, &c->h_tab, sizeof(dna_tables), hipMemcpyHostToDevice));
    {
        std::vector<uint8_t> guide(DNA_GUIDE_PAD, 0);
        dna::build_born_guide(c->h_tab, guide.data());
        DNA_HIP(hipMalloc((void **)&c->d_guide, DNA_GUIDE_PAD));
        DNA_HIP(hipMemcpy(c->d_guide, guide.data(), DNA_GUIDE_PAD, hipMemcpyHostToDevice));
    }
    DNA_HIP(hipMalloc((void **)&c->d_ctr, C_NCTR * sizeof(uint32_t)));
    DNA_HIP(hipHostMalloc((void **)&c->h_ctr, C_NCTR * sizeof(uint32_t), hipHostMallocDefault));
    DNA_HIP(hipMalloc((void **)&c->d_u64, 64 * sizeof(unsigned long long)));
    for (int i = 0; i < 8; i++) DNA_HIP(hipEventCreate(&c->ev[i]));
    DNA_HIP(hipEventCreateWithFlags(&c->ev_loop, hipEventDisableTiming));
    {
        int rc = score_scratch_init(device, &c->na_scratch);
        if (rc) { ctx_release(c); return rc; }
    }
    *out = c;
    return DNA_OK;
}

namespace dna {
void build_born_guide(const dna_tables &t, uint8_t *guide)
{
    for (int s = 0; s < 5; s++)
        for (int e = 0; e < (int)DNA_N_WE; e++) {
            const float *C = t.XWC[s] + (size_t)e * DNA_N_WR;
            uint8_t *g = guide + ((size_t)s * DNA_N_WE + e) * (DNA_GUIDE_K + 1);
            for (int k = 0; k < DNA_GUIDE_K; k++) {
                const float thr = (float)k / (float)DNA_GUIDE_K;
                uint32_t n = 0;
                while (n < DNA_N_WR - 1 && C[n] < thr) n++;
                g[k] = (uint8_t)n;
            }
            g[DNA_GUIDE_K] = (uint8_t)(DNA_N_WR - 1);
        }
}
}  // namespace dna

extern "C" int dna_xs_probe(dna_ctx *c, int which, const float *E, const float *r, uint64_t n, float *out)
{
    if (!c || !E || !out || which < 0 || which > 5 || (which == 5 && !r)) {
        set_error("dna_xs_probe: bad argument"); return DNA_ERR_ARG;
    }
    if (n >= (1ull << 31)) { set_error("dna_xs_probe: at most 2^31-1 energies"); return DNA_ERR_CAPACITY; }
    if (n == 0) return DNA_OK;
    static const int width[6] = {3, 5, 5, 1, 1, 1};
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    float *dE = nullptr, *dr = nullptr, *dout = nullptr;
    const size_t wout = (size_t)width[which];
    hipError_t e = hipMalloc(&dE, n * 4);
    if (e == hipSuccess && which == 5) e = hipMalloc(&dr, n * 4);
    if (e == hipSuccess) e = hipMalloc(&dout, n * wout * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(dE, E, n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && dr) e = hipMemcpyAsync(dr, r, n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_xs_probe(c->d_tab, c->d_guide, dE, dr, (uint32_t)n, which, dout, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * wout * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dE); (void)hipFree(dr); (void)hipFree(dout);
    DNA_HIP(e);
    return DNA_OK;
}

extern "C" int dna_set_option(dna_ctx *c, int option, int value)
{
    if (!c) { set_error("dna_set_option: NULL context"); return DNA_ERR_ARG; }
    switch (option) {
    case DNA_OPT_IRT_BLOCK:
        if (value != 0 && value != 64 && value != 128 && value != 256) {
            set_error("dna_set_option: DNA_OPT_IRT_BLOCK must be 0 (autotune), 64, 128 or 256");
            return DNA_ERR_ARG;
        }
        c->irt_block = value;
        c->irt_tuned.clear();
        return DNA_OK;
    case DNA_OPT_IRT_CELL:
        if (value != 0 && value != -1 && value != 1 && value != 2 && value != 4) {
            set_error("dna_set_option: DNA_OPT_IRT_CELL must be 0 (autotune), -1 (no cells), 1, 2 or 4");
            return DNA_ERR_ARG;
        }
        c->irt_cell = value;
        c->irt_tuned.clear();
        return DNA_OK;
    case DNA_OPT_IRT_FINAL:
        if (value != 0 && value != 1) { set_error("dna_set_option: DNA_OPT_IRT_FINAL must be 0 or 1"); return DNA_ERR_ARG; }
        c->irt_final = value;
        return DNA_OK;
    case DNA_OPT_PA_LANES:
        if (value < 0 || value > 64 || (value & (value - 1))) {
            set_error("dna_set_option: DNA_OPT_PA_LANES must be 0 (autotune) or a power of two up to 64");
            return DNA_ERR_ARG;
        }
        c->pa_lanes = value;
        c->tr_tuned.clear();
        return DNA_OK;
    case DNA_OPT_IRT_DRAIN:
        if (value < 0 || value > 2) { set_error("dna_set_option: DNA_OPT_IRT_DRAIN must be 0, 1 or 2"); return DNA_ERR_ARG; }
        c->irt_drain = value;
        c->irt_tuned.clear();
        return DNA_OK;
    case DNA_OPT_IRT_ORDER:
        if (value != 0 && value != 1) { set_error("dna_set_option: DNA_OPT_IRT_ORDER must be 0 or 1"); return DNA_ERR_ARG; }
        c->irt_order = value;
        return DNA_OK;
    case DNA_OPT_IRT_EVBLOCKS:
        if (value < 1 || value > 128) { set_error("dna_set_option: DNA_OPT_IRT_EVBLOCKS must be 1..128"); return DNA_ERR_ARG; }
        c->irt_evblocks = value;
        return DNA_OK;
    case DNA_OPT_MLP_FUSED:
        if (value != 0 && value != 1) { set_error("dna_set_option: DNA_OPT_MLP_FUSED must be 0 or 1"); return DNA_ERR_ARG; }
        c->mlp_fused = value;
        return DNA_OK;
    case DNA_OPT_RAD_CHUNK:
        if (value != 0 && (value < 256 || value > 4096 || (value & (value - 1)))) {
            set_error("dna_set_option: DNA_OPT_RAD_CHUNK must be 0 (autotune) or a power of two from 256 to 4096");
            return DNA_ERR_ARG;
        }
        c->rad_chunk = value;
        c->tr_tuned.clear();
        return DNA_OK;
    case DNA_OPT_WORK_CHUNK:
        if (value != 0 && (value < 64 || value > 1024 || (value & (value - 1)))) {
            set_error("dna_set_option: DNA_OPT_WORK_CHUNK must be 0 (autotune) or a power of two from 64 to 1024");
            return DNA_ERR_ARG;
        }
        c->work_chunk = value;
        c->tr_tuned.clear();
        return DNA_OK;
    default:
        set_error("dna_set_option: unknown option");
        return DNA_ERR_ARG;
    }
}

extern "C" int dna_get_option(dna_ctx *c, int option, int *value)
{
    if (!c || !value) { set_error("dna_get_option: NULL argument"); return DNA_ERR_ARG; }
    switch (option) {
    case DNA_OPT_IRT_BLOCK: *value = c->irt_block; return DNA_OK;
    case DNA_OPT_IRT_BLOCK_LAST: *value = c->irt_last_block; return DNA_OK;
    case DNA_OPT_IRT_CELL: *value = c->irt_cell; return DNA_OK;
    case DNA_OPT_IRT_CELL_LAST: *value = c->irt_last_cell; return DNA_OK;
    case DNA_OPT_IRT_FINAL: *value = c->irt_final; return DNA_OK;
    case DNA_OPT_PA_LANES: *value = c->pa_lanes; return DNA_OK;
    case DNA_OPT_RAD_CHUNK: *value = c->rad_chunk; return DNA_OK;
    case DNA_OPT_MLP_FUSED: *value = c->mlp_fused; return DNA_OK;
    case DNA_OPT_IRT_DRAIN: *value = c->irt_drain; return DNA_OK;
    case DNA_OPT_IRT_ORDER: *value = c->irt_order; return DNA_OK;
    case DNA_OPT_IRT_EVBLOCKS: *value = c->irt_evblocks; return DNA_OK;
    case DNA_OPT_WORK_CHUNK: *value = c->work_chunk; return DNA_OK;
    case DNA_OPT_PA_LANES_LAST: *value = c->last_pa_lanes; return DNA_OK;
    case DNA_OPT_RAD_CHUNK_LAST: *value = c->last_rad_chunk; return DNA_OK;
    case DNA_OPT_WORK_CHUNK_LAST: *value = c->last_work_chunk; return DNA_OK;
    default:
        set_error("dna_get_option: unknown option");
        return DNA_ERR_ARG;
    }
}

static void ctx_free(dna_ctx *c)
{
    (void)hipSetDevice(c->device);           // teardown: nothing to report to
    (void)hipStreamSynchronize(c->stream);
    irt_staged_drop(c);
    void *dev[] = {c->d_tab, c->d_guide, c->d_ctr, c->d_u64, c->d_results, c->d_dose, c->d_rad_raw, c->d_rad_key,
                   c->d_rad, c->d_sec, c->d_prim_nsec, c->d_prim_nrad, c->d_sec_nrad, c->d_energies,
                   c->d_base_p, c->d_cnt, c->d_block, c->d_sec_base, c->irt.scratch, c->irt.events, c->irt.tracks,
                   c->irt.part, c->irt.rad_in, c->irt.tally, c->irt.sort, c->sc.hitbits, c->sc.flags, c->sc.rank,
                   c->sc.hits_u8, c->sc.rad_in, c->sc.tgt, c->hs.buf, c->d_ntot, c->d_oh};
    for (void *p : dev) if (p) (void)hipFree(p);
    void *host[] = {c->h_ctr, c->h_results, c->h_dose, c->h_rad, c->sc.h_hits};
    for (void *p : host) if (p) (void)hipHostFree(p);
    for (int i = 0; i < 8; i++) (void)hipEventDestroy(c->ev[i]);
    if (c->ev_loop) (void)hipEventDestroy(c->ev_loop);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

namespace dna {
void ctx_retain(dna_ctx *c) { c->refs.fetch_add(1, std::memory_order_relaxed); }
void ctx_release(dna_ctx *c)
{
    if (c->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) ctx_free(c);
}
}  // namespace dna

// Drops the creator's share: the context is freed now, or when the last surrogate or plan
// made on it is destroyed.
extern "C" void dna_ctx_destroy(dna_ctx *c)
{
    if (c) ctx_release(c);
}

__global__ void k_dose_sum(const uint32_t *dose, size_t n, unsigned long long *out)
{
    unsigned long long s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += dose[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

static int ensure_transport(dna_ctx *c, uint32_t n, size_t rad_cap, size_t sec_cap, uint32_t vc)
{
    size_t v3 = (size_t)vc * vc * vc;
    DNA_HIP(grow(c->d_results, c->cap_n_results, n));
    DNA_HIP(grow(c->d_prim_nsec, c->cap_prim_nsec, n));
    DNA_HIP(grow(c->d_prim_nrad, c->cap_prim_nrad, n));
    DNA_HIP(grow(c->d_energies, c->cap_energies, n));
    DNA_HIP(grow(c->d_base_p, c->cap_base_p, (size_t)n + 1));
    DNA_HIP(grow(c->d_dose, c->cap_dose, v3));
    DNA_HIP(grow(c->d_rad_raw, c->cap_rad_raw, rad_cap));
    DNA_HIP(grow(c->d_rad_key, c->cap_rad_key, rad_cap));
    DNA_HIP(grow(c->d_rad, c->cap_rad, rad_cap));
    DNA_HIP(grow(c->d_sec, c->cap_sec, sec_cap));
    DNA_HIP(grow(c->d_sec_nrad, c->cap_sec_nrad, sec_cap));
    DNA_HIP(grow(c->d_cnt, c->cap_cnt, (size_t)n + sec_cap + 1));
    size_t nb = ((size_t)n + sec_cap + 1 + 1023) / 1024 + 2;
    DNA_HIP(grow(c->d_block, c->cap_block, nb));
    return DNA_OK;
}

// One pass of both transport phases with the given launch knobs, for timing only (the
// outputs are not canonicalized; the real run repeats the work): phase times in ms.
static int transport_trial(dna_ctx *c, TransportDev d, uint32_t lanes, float &ta, float &tb)
{
    hipStream_t st = c->stream;
    const size_t v3 = (size_t)d.p.vc * d.p.vc * d.p.vc;
    DNA_HIP(hipMemsetAsync(c->d_ctr, 0, C_NCTR * sizeof(uint32_t), st));
    DNA_HIP(hipMemsetAsync(c->d_dose, 0, v3 * sizeof(uint32_t), st));
    DNA_HIP(hipMemsetAsync(c->d_rad_key, 0xFF, (size_t)d.rad_cap * sizeof(uint64_t), st));
    DNA_HIP(hipEventRecord(c->ev[0], st));
    DNA_HIP(launch_phase_a(d, lanes, st));
    DNA_HIP(hipEventRecord(c->ev[1], st));
    DNA_HIP(hipMemcpyAsync(c->h_ctr, c->d_ctr, C_NCTR * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    const uint32_t sec_n = c->h_ctr[C_SEC] < d.sec_cap ? c->h_ctr[C_SEC] : d.sec_cap;
    DNA_HIP(hipEventRecord(c->ev[2], st));
    DNA_HIP(launch_phase_b(d, sec_n, st));
    DNA_HIP(hipEventRecord(c->ev[3], st));
    DNA_HIP(hipStreamSynchronize(st));
    DNA_HIP(hipEventElapsedTime(&ta, c->ev[0], c->ev[1]));
    DNA_HIP(hipEventElapsedTime(&tb, c->ev[2], c->ev[3]));
    return DNA_OK;
}

// Transport launch knobs: the options, or autotuned per batch class (log2 of the primaries,
// log2 of the mean energy in eV).  Batches under 1024 primaries take the rule (Phase A lanes
// from phase_a_lanes, 256-slot radical chunks, 64-item work chunks).  Otherwise the first
// batch of a class is run with candidates, each timed twice and kept at its best: the
// Phase A lanes (half, the rule's and double) on Phase A's time, then the radical chunk
// (256, 512, 1024) on both phases', then the work chunk (64, 128, 256) on Phase B's.
// Results do not depend on any of them.
struct TrKnobs { uint32_t lanes, rad_chunk, work_chunk; };
static int transport_knobs(dna_ctx *c, const dna_params &p, double sumE, const float *energies, TrKnobs &k)
{
    const uint32_t rule = phase_a_lanes(p.n, 0);
    k.lanes = c->pa_lanes ? (uint32_t)c->pa_lanes : rule;
    k.rad_chunk = c->rad_chunk ? (uint32_t)c->rad_chunk : 256u;
    k.work_chunk = c->work_chunk ? (uint32_t)c->work_chunk : 64u;
    if (c->pa_lanes && c->rad_chunk && c->work_chunk) return DNA_OK;
    if (p.n < 1024) return DNA_OK;
    int cn = 0, ce = 0;
    for (uint32_t m = p.n; m > 1; m >>= 1) cn++;
    for (double e = sumE / p.n; e >= 2.0; e *= 0.5) ce++;
    const std::pair<int, int> key(cn, ce);
    auto it = c->tr_tuned.find(key);
    if (it != c->tr_tuned.end()) {
        if (!c->pa_lanes) k.lanes = it->second[0];
        if (!c->rad_chunk) k.rad_chunk = it->second[1];
        if (!c->work_chunk) k.work_chunk = it->second[2];
        return DNA_OK;
    }
    std::vector<uint32_t> L, R, W;
    if (c->pa_lanes) L = {k.lanes};
    else { if (rule > 1) L.push_back(rule / 2); L.push_back(rule); if (rule < 64) L.push_back(rule * 2); }
    if (c->rad_chunk) R = {k.rad_chunk}; else R = {256u, 512u, 1024u};
    if (c->work_chunk) W = {k.work_chunk}; else W = {64u, 128u, 256u};
    // buffers for the most headroom any candidate needs (fewest lanes, largest chunk)
    const size_t headroom = transport_slot_headroom(p.n, L.front(), R.back());
    const size_t rad_cap = (size_t)(0.2 * sumE) + 65536 + headroom;
    const size_t sec_cap = (size_t)(0.025 * sumE) + 4096;
    if (rad_cap > (size_t)0xFFFFFFF0u || sec_cap > (size_t)0xFFFFFFF0u) return DNA_OK;   // the run reports it
    int rc = ensure_transport(c, p.n, rad_cap, sec_cap, p.vc);
    if (rc) return rc;
    hipStream_t st = c->stream;
    if (energies) DNA_HIP(hipMemcpyAsync(c->d_energies, energies, p.n * sizeof(float), hipMemcpyHostToDevice, st));
    TransportDev d;
    d.tab = c->d_tab; d.guide = c->d_guide; d.p = p; d.energies = energies ? c->d_energies : nullptr;
    d.results = c->d_results; d.dose = c->d_dose; d.rad_raw = c->d_rad_raw; d.rad_key = c->d_rad_key;
    d.rad_cap = (uint32_t)rad_cap; d.sec = c->d_sec; d.sec_cap = (uint32_t)sec_cap;
    d.prim_nsec = c->d_prim_nsec; d.prim_nrad = c->d_prim_nrad; d.sec_nrad = c->d_sec_nrad; d.ctr = c->d_ctr;
    d.sec_base = nullptr; d.sec_keep = 0xFFFFFFFFu;
    d.rad_chunk = k.rad_chunk; d.work_chunk = k.work_chunk;
    auto best_of = [&](const std::vector<uint32_t> &cand, int which, uint32_t &out) -> int {
        float best = 1e30f;
        std::vector<float> t(cand.size(), 1e30f);
        for (int round = 0; round < 2; round++)
            for (size_t i = 0; i < cand.size(); i++) {
                uint32_t lanes = k.lanes;
                if (which == 0) lanes = cand[i];
                else if (which == 1) d.rad_chunk = cand[i];
                else d.work_chunk = cand[i];
                float ta = 0, tb = 0;
                int r = transport_trial(c, d, lanes, ta, tb); if (r) return r;
                const float v = which == 0 ? ta : which == 1 ? ta + tb : tb;
                t[i] = v < t[i] ? v : t[i];
            }
        for (size_t i = 0; i < cand.size(); i++) if (t[i] < best) { best = t[i]; out = cand[i]; }
        return DNA_OK;
    };
    if (L.size() > 1) { rc = best_of(L, 0, k.lanes); if (rc) return rc; }
    if (R.size() > 1) { rc = best_of(R, 1, k.rad_chunk); if (rc) return rc; }
    d.rad_chunk = k.rad_chunk;
    if (W.size() > 1) { rc = best_of(W, 2, k.work_chunk); if (rc) return rc; }
    c->tr_tuned[key] = {k.lanes, k.rad_chunk, k.work_chunk};
    return DNA_OK;
}

extern "C" int dna_run_transport(dna_ctx *c, const dna_params *p_in, const float *energies,
                                 uint32_t copy_flags, dna_transport_out *out)
{
    if (!c || !p_in || !out) { set_error("dna_run_transport: NULL argument"); return DNA_ERR_ARG; }
    dna_params p = *p_in;
    if (p.n == 0 || p.vc == 0 || !(p.box > 0.0f) || p.ms == 0) { set_error("dna_run_transport: bad params"); return DNA_ERR_ARG; }
    if (p.n >= (1u << 24)) { set_error("dna_run_transport: at most 2^24-1 primaries"); return DNA_ERR_CAPACITY; }
    if ((size_t)p.vc * p.vc * p.vc > (1ull << 31)) { set_error("dna_run_transport: vc too large"); return DNA_ERR_CAPACITY; }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    double sumE = 0.0;
    if (energies) for (uint32_t i = 0; i < p.n; i++) sumE += energies[i] > 0 ? energies[i] : 0.0;
    else sumE = (double)p.n * (p.be > 0 ? p.be : 0.0);
    // internal capacities: reference yields are ~0.13-0.16 radical entries and ~0.0145
    // secondaries per eV; grow and repeat the (deterministic) run if a count overflows.  The
    // caller's max_rad / max_sec are caps on what is kept, applied in canonical order after
    // the run (the reference's MAX_RAD / MAX_SEC with *_dropped counts, dispatch.ts:124-126,
    // 190-192), not buffer sizes.
    TrKnobs kn;
    int krc = transport_knobs(c, p, sumE, energies, kn);
    if (krc) return krc;
    c->last_pa_lanes = (int)kn.lanes; c->last_rad_chunk = (int)kn.rad_chunk; c->last_work_chunk = (int)kn.work_chunk;
    const uint32_t pa_lanes = kn.lanes;
    const size_t headroom = transport_slot_headroom(p.n, pa_lanes, kn.rad_chunk);   // chunk tails left unused
    size_t rad_cap = (size_t)(0.2 * sumE) + 65536 + headroom;
    size_t sec_cap = (size_t)(0.025 * sumE) + 4096;
    // slot counters are 32-bit: refuse a batch that may need 2^32 or more slots
    const size_t HARD = (size_t)0xFFFFFFF0u;
    if (0.2 * sumE + 65536.0 + (double)headroom > (double)HARD || 0.025 * sumE > (double)HARD) {
        set_error("dna_run_transport: the batch may emit 2^32 or more radical slots; split it");
        return DNA_ERR_CAPACITY;
    }
    for (int attempt = 0; attempt < 4; attempt++) {
        if (rad_cap > HARD || sec_cap > HARD) {
            set_error("dna_run_transport: the batch needs 2^32 or more radical or secondary slots; split it");
            return DNA_ERR_CAPACITY;
        }
        int rc = ensure_transport(c, p.n, rad_cap, sec_cap, p.vc);
        if (rc) return rc;
        if (energies) DNA_HIP(hipMemcpyAsync(c->d_energies, energies, p.n * sizeof(float), hipMemcpyHostToDevice, st));
        size_t v3 = (size_t)p.vc * p.vc * p.vc;
        DNA_HIP(hipMemsetAsync(c->d_ctr, 0, C_NCTR * sizeof(uint32_t), st));
        DNA_HIP(hipMemsetAsync(c->d_dose, 0, v3 * sizeof(uint32_t), st));
        DNA_HIP(hipMemsetAsync(c->d_u64, 0, 64 * sizeof(unsigned long long), st));
        DNA_HIP(hipMemsetAsync(c->d_rad_key, 0xFF, rad_cap * sizeof(uint64_t), st));   // unused slots: invalid key
        TransportDev d;
        d.tab = c->d_tab; d.guide = c->d_guide; d.p = p; d.energies = energies ? c->d_energies : nullptr;
        d.results = c->d_results; d.dose = c->d_dose; d.rad_raw = c->d_rad_raw; d.rad_key = c->d_rad_key;
        d.rad_cap = (uint32_t)rad_cap; d.sec = c->d_sec; d.sec_cap = (uint32_t)sec_cap;
        d.prim_nsec = c->d_prim_nsec; d.prim_nrad = c->d_prim_nrad; d.sec_nrad = c->d_sec_nrad; d.ctr = c->d_ctr;
        d.sec_base = nullptr; d.sec_keep = 0xFFFFFFFFu;
        d.rad_chunk = kn.rad_chunk; d.work_chunk = kn.work_chunk;
        DNA_HIP(hipEventRecord(c->ev[0], st));
        DNA_HIP(launch_phase_a(d, pa_lanes, st));
        DNA_HIP(hipEventRecord(c->ev[1], st));
        DNA_HIP(hipMemcpyAsync(c->h_ctr, c->d_ctr, C_NCTR * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        DNA_HIP(hipStreamSynchronize(st));
        uint32_t sec_raw = c->h_ctr[C_SEC];
        if (sec_raw > sec_cap) { sec_cap = (size_t)sec_raw * 5 / 4 + 4096; continue; }
        uint32_t sec_n = sec_raw;                   // secondary records written by Phase A
        uint32_t sec_kept = sec_raw;                // of which Phase B transports
        if (p.max_sec && sec_raw > p.max_sec) {
            // the first max_sec secondaries in canonical (primary, emission) order
            DNA_HIP(grow(c->d_sec_base, c->cap_sec_base, (size_t)p.n + 1));
            DNA_HIP(hipMemcpyAsync(c->d_sec_base, c->d_prim_nsec, (size_t)p.n * 4, hipMemcpyDeviceToDevice, st));
            DNA_HIP(exclusive_scan_u32(c->d_sec_base, p.n, c->d_block, st));
            d.sec_base = c->d_sec_base; d.sec_keep = p.max_sec;
            sec_kept = p.max_sec;
        }
        DNA_HIP(hipEventRecord(c->ev[2], st));
        DNA_HIP(launch_phase_b(d, sec_n, st));
        DNA_HIP(hipEventRecord(c->ev[3], st));
        DNA_HIP(hipMemcpyAsync(c->h_ctr, c->d_ctr, C_NCTR * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        DNA_HIP(hipStreamSynchronize(st));
        if (c->h_ctr[C_WRAP]) {
            set_error("dna_run_transport: the radical slot counter passed 2^32; split the batch");
            return DNA_ERR_CAPACITY;
        }
        uint32_t rad_raw = c->h_ctr[C_RAD];
        if (rad_raw > rad_cap) { rad_cap = (size_t)rad_raw * 5 / 4 + 65536 + headroom; continue; }
        CanonScratch cs; cs.base_p = c->d_base_p; cs.cnt = c->d_cnt; cs.block = c->d_block; cs.block_cap = c->cap_block;
        DNA_HIP(launch_canonicalize(d, p.n, sec_n, rad_raw, cs, c->d_rad, st));
        hipLaunchKernelGGL(k_dose_sum, dim3(1024), dim3(256), 0, st, c->d_dose, v3, c->d_u64);
        DNA_HIP(hipGetLastError());
        DNA_HIP(hipEventRecord(c->ev[4], st));
        // radicals emitted = the total of the per-emitter counts (slots include unused ones);
        // a caller's max_rad keeps the first max_rad of them in canonical order
        uint32_t rad_total = 0;
        DNA_HIP(hipMemcpyAsync(&rad_total, c->d_cnt + (size_t)p.n + sec_n, 4, hipMemcpyDeviceToHost, st));
        DNA_HIP(hipStreamSynchronize(st));
        const uint32_t rad_n = (p.max_rad && rad_total > p.max_rad) ? p.max_rad : rad_total;
        // copies
        unsigned long long dsum = 0;
        DNA_HIP(hipMemcpyAsync(&c->h_u64_tmp, c->d_u64, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        out->results = nullptr; out->dose = nullptr; out->rad_buf = nullptr;
        if (copy_flags & DNA_COPY_RESULTS) {
            DNA_HIP(grow_host(c->h_results, c->hcap_results, p.n));
            DNA_HIP(hipMemcpyAsync(c->h_results, c->d_results, p.n * sizeof(dna_result), hipMemcpyDeviceToHost, st));
            out->results = c->h_results;
        }
        if (copy_flags & DNA_COPY_DOSE) {
            DNA_HIP(grow_host(c->h_dose, c->hcap_dose, v3));
            DNA_HIP(hipMemcpyAsync(c->h_dose, c->d_dose, v3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            out->dose = c->h_dose;
        }
        if ((copy_flags & DNA_COPY_RAD) && rad_n) {
            DNA_HIP(grow_host(c->h_rad, c->hcap_rad, (size_t)rad_n * 4));
            DNA_HIP(hipMemcpyAsync(c->h_rad, c->d_rad, (size_t)rad_n * 16, hipMemcpyDeviceToHost, st));
            out->rad_buf = c->h_rad;
        }
        DNA_HIP(hipStreamSynchronize(st));
        dsum = c->h_u64_tmp;
        for (int i = 0; i < 8; i++) { out->counters[i] = c->h_ctr[i]; out->sec_stats[i] = c->h_ctr[C_SST + i]; out->dbg[i] = c->h_ctr[C_DBG + i]; }
        out->counters[C_RAD] = rad_total;          // emitted, as the reference's counters[7]
        out->dose_sum = dsum;
        out->rad_n_raw = rad_total; out->rad_n_stored = rad_n; out->rad_dropped = rad_total - rad_n;
        out->sec_n_raw = sec_raw; out->sec_n = sec_kept; out->sec_dropped = sec_raw - sec_kept;
        float ta = 0, tb = 0, tt = 0;
        (void)hipEventElapsedTime(&ta, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&tb, c->ev[2], c->ev[3]);
        (void)hipEventElapsedTime(&tt, c->ev[0], c->ev[4]);
        out->t_prim_ms = ta; out->t_sec_ms = tb; out->t_total_ms = tt;
        c->rad_n = rad_n;
        c->rad_epoch++;
        c->last_n = p.n;
        return DNA_OK;
    }
    set_error("dna_run_transport: capacity estimate kept overflowing");
    return DNA_ERR_CAPACITY;
}
