// canonical.hip — puts the transport's radical emissions into canonical order and
// provides the device-wide exclusive scan used for it.
//
// Canonical order: by primary id, then by emitter (the primary itself, then its
// secondaries in the order the primary emitted them), then by emission sequence.  Each
// raw entry carries key = pid << 40 | k << 20 | seq, with k = 0 for the primary and the
// 1-based emission index for a secondary.  The destination of an entry is
//     off[ base_p[pid] + k ] + seq
// where base_p is the exclusive scan of (1 + secondaries) over primaries (the dense
// emitter index) and off the exclusive scan of per-emitter entry counts.  One read and
// one write of 16 bytes per entry plus the key: HBM-bound, no sorting.
#include "dna_internal.h"

namespace dna {

// ---- exclusive scan: 1024 elements per block, three passes
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// block-wide exclusive scan of 4 values per thread (256 threads x 4 = 1024)
__device__ void block_scan4(uint32_t v[4], uint32_t &total, uint32_t *smem /* 4 words */)
{
    uint32_t s = v[0] + v[1] + v[2] + v[3];
    uint32_t incl = wave_incl_scan(s);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 63) smem[w] = incl;
    __syncthreads();
    uint32_t wbase = 0;
    for (int i = 0; i < w; i++) wbase += smem[i];
    total = smem[0] + smem[1] + smem[2] + smem[3];
    uint32_t run = wbase + incl - s;
    for (int i = 0; i < 4; i++) { uint32_t x = v[i]; v[i] = run; run += x; }
}

__global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t *data, uint32_t n, uint32_t *partial)
{
    __shared__ uint32_t sm[4];
    uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t v[4];
    for (int i = 0; i < 4; i++) v[i] = (base + i < n) ? data[base + i] : 0u;
    uint32_t tot;
    block_scan4(v, tot, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// single block scans the partials (any count) in chunks of 1024
__global__ void __launch_bounds__(256) k_scan_partials(uint32_t *partial, uint32_t nb)
{
    __shared__ uint32_t sm[4];
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nb; c += 1024u) {
        uint32_t base = c + threadIdx.x * 4u;
        uint32_t v[4];
        for (int i = 0; i < 4; i++) v[i] = (base + i < nb) ? partial[base + i] : 0u;
        uint32_t tot;
        block_scan4(v, tot, sm);
        for (int i = 0; i < 4; i++) if (base + i < nb) partial[base + i] = v[i] + carry;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

__global__ void __launch_bounds__(256) k_scan_apply(uint32_t *data, uint32_t n, const uint32_t *partial, uint32_t nb)
{
    __shared__ uint32_t sm[4];
    uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t v[4];
    for (int i = 0; i < 4; i++) v[i] = (base + i < n) ? data[base + i] : 0u;
    uint32_t tot;
    block_scan4(v, tot, sm);
    uint32_t off = partial[blockIdx.x];
    for (int i = 0; i < 4; i++) if (base + i < n) data[base + i] = v[i] + off;
    if (blockIdx.x == 0 && threadIdx.x == 0) data[n] = partial[nb];
}

hipError_t exclusive_scan_u32(uint32_t *data, uint32_t n, uint32_t *tmp, hipStream_t st)
{
    uint32_t nb = (n + 1023u) / 1024u;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, st, data, n, tmp);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(256), 0, st, tmp, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, data, n, tmp, nb);
    return hipGetLastError();
}

// ---- canonical scatter
__global__ void k_prim_counts(const uint32_t *prim_nsec, uint32_t n, uint32_t *base_p)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) base_p[i] = 1u + prim_nsec[i];
}

__global__ void k_emitter_counts(const uint32_t *base_p, const uint32_t *prim_nrad, uint32_t n,
                                 const SecRec *sec, const uint32_t *sec_nrad, uint32_t sec_n,
                                 uint32_t *cnt)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[base_p[i]] = prim_nrad[i];
    if (i < sec_n) { uint32_t pid = sec[i].pid, k = sec[i].k; cnt[base_p[pid] + k] = sec_nrad[i]; }
}

__global__ void k_scatter(const float4 *raw, const uint64_t *key, uint32_t rad_n, const uint32_t *base_p,
                          const uint32_t *off, float4 *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rad_n) return;
    uint64_t kk = key[i];
    if (kk == ~0ull) return;               // a slot reserved but not used
    uint32_t pid = (uint32_t)(kk >> 40), k = (uint32_t)(kk >> 20) & 0xFFFFFu, seq = (uint32_t)kk & 0xFFFFFu;
    out[off[base_p[pid] + k] + seq] = raw[i];
}

hipError_t launch_canonicalize(const TransportDev &d, uint32_t n, uint32_t sec_n, uint32_t rad_n,
                               CanonScratch &cs, float4 *rad_out, hipStream_t st)
{
    const uint32_t B = 256;
    hipLaunchKernelGGL(k_prim_counts, dim3((n + B - 1) / B), dim3(B), 0, st, d.prim_nsec, n, cs.base_p);
    hipError_t e = exclusive_scan_u32(cs.base_p, n, cs.block, st);
    if (e != hipSuccess) return e;
    uint32_t P = n + sec_n;
    uint32_t m = n > sec_n ? n : sec_n;
    hipLaunchKernelGGL(k_emitter_counts, dim3((m + B - 1) / B), dim3(B), 0, st, cs.base_p, d.prim_nrad, n,
                       d.sec, d.sec_nrad, sec_n, cs.cnt);
    e = exclusive_scan_u32(cs.cnt, P, cs.block, st);
    if (e != hipSuccess) return e;
    if (rad_n)
        hipLaunchKernelGGL(k_scatter, dim3((rad_n + B - 1) / B), dim3(B), 0, st, d.rad_raw, d.rad_key, rad_n,
                           cs.base_p, cs.cnt, rad_out);
    return hipGetLastError();
}

}  // namespace dna
