// mlp.hip — learned IRT surrogate: an MLP (n_in -> H -> ... -> H -> 2*n_out, ReLU) in
// BF16 on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), trained
// by the kernels below on this library's own per-track IRT output.  No counterpart in the
// reference: it predicts a track's G(t) (mean and log-variance per output) from features
// of its initial radical distribution; tracks whose predicted sigma is too large go to IRT.
//
// Kernels:
//   k_mlp_fused   the whole forward pass for 64 rows per workgroup: input normalisation,
//                 every layer from an LDS activation tile and L2-resident transposed BF16
//                 weights, and the prediction epilogue (one launch for any batch)
//   k_gemm        C[MxN] = epilogue(A[MxK] B[KxN]); A and B are BF16 with arbitrary strides,
//                 so the backward products dX = dY W^T and dW = X^T dY use the kernel (and
//                 the forward when the fused kernel is switched off).  64x64 block tile, K
//                 chunks of 32 staged in LDS as [row][k] (16-byte vector loads whenever a
//                 dimension is contiguous), so each MFMA operand fragment is one 16-byte LDS
//                 read; 4 wavefronts, each a 32x32 sub-tile of 2x2 MFMA 16x16 tiles.
//                 Epilogues: bias, ReLU, ReLU' mask, FP32 and/or BF16 output, and the column
//                 sums of B (the bias gradient of dW = X^T dY).
//   k_nll         Gaussian NLL 0.5 (s + (t - mu)^2 e^-s) per output and its gradient, with
//                 the mini-batch gather and normalisation of the targets
//   k_adam        Adam on FP32 master weights, refreshing both BF16 copies
//   k_features    per-track input features from a radical list
//   k_predict     de-normalised mean / sigma and the per-track IRT fallback flag (the
//                 per-layer path)
// A training step is four launches plus the backward GEMMs (one per layer for dW and the
// bias gradient, one per hidden layer for dX): fused forward, NLL, backward, Adam.
#include "dna_internal.h"
#include "dna_ctx.h"
#include <vector>
#include <cmath>
#include <cstring>

namespace dna {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define GT 64      // block tile (rows and columns)
#define GK 32      // K chunk
#define LDK 40     // LDS row pitch in bf16: 80 B, keeps every 8-element fragment 16-B aligned

enum { ACT_NONE = 0, ACT_RELU = 1 };
// operand staging modes: 0 scalar, 1 K contiguous, 2 rows (M or N) contiguous
enum { LD_SCALAR = 0, LD_KCONT = 1, LD_RCONT = 2 };

struct GemmArgs {
    const bf16 *A; long sam, sak;
    const bf16 *B; long sbk, sbn;
    float *C; long scm;               // optional FP32 output (row stride scm, unit column stride)
    bf16 *Cb; long sbm;               // optional BF16 output
    const float *bias;                // [N] or null
    const bf16 *mask; long smm;       // ReLU' mask source (post-activation of the forward)
    float *colsum;                    // optional: column sums of B over K (the bias gradient of a
                                      // dW = X^T dY product), written by the blocks of row 0
    int M, N, K;
    int act;
    int amode, bmode;
};

// Stage a 64 x 32 tile (rows r0.., k k0..) of a strided BF16 matrix into S[row][k].
__device__ __forceinline__ void stage_tile(bf16 (*S)[LDK], const bf16 *X, long sr, long sk, int R, int K,
                                           int r0, int k0, int mode, int tid)
{
    const bf16 z = (bf16)0.0f;
    if (mode == LD_KCONT) {
        int r = tid >> 2, kv = (tid & 3) * 8;
        int gr = r0 + r, gk = k0 + kv;
        bf16x8 v;
        if (gr < R && gk < K) v = *(const bf16x8 *)(X + gr * sr + gk);
        else for (int j = 0; j < 8; j++) v[j] = z;
        *(bf16x8 *)&S[r][kv] = v;
    } else if (mode == LD_RCONT) {
        int kk = tid >> 3, rv = (tid & 7) * 8;
        int gr = r0 + rv, gk = k0 + kk;
        bf16x8 v;
        if (gr < R && gk < K) v = *(const bf16x8 *)(X + gk * sk + gr);
        else for (int j = 0; j < 8; j++) v[j] = z;
        for (int j = 0; j < 8; j++) S[rv + j][kk] = v[j];
    } else {
        for (int e = tid; e < GT * GK; e += 256) {
            int r = e / GK, kk = e % GK;
            int gr = r0 + r, gk = k0 + kk;
            S[r][kk] = (gr < R && gk < K) ? X[gr * sr + gk * sk] : z;
        }
    }
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) bf16 As[GT][LDK];
    __shared__ __attribute__((aligned(16))) bf16 Bs[GT][LDK];   // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    // MFMA 16x16x32 operands: lane l supplies row l%16 of A (column l%16 of B) at the
    // k-slice (l/16)*8 .. +7; A and B use the same slice, so the sum runs over all 32 k
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const bool csum = g.colsum && blockIdx.y == 0;
    float cs = 0.f;
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        stage_tile(As, g.A, g.sam, g.sak, g.M, g.K, m0, k0, g.amode, tid);
        stage_tile(Bs, g.B, g.sbn, g.sbk, g.N, g.K, n0, k0, g.bmode, tid);
        __syncthreads();
        if (csum && tid < GT)
            for (int kk = 0; kk < GK; kk++) cs += (float)Bs[tid][kk];
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; i++) a[i] = *(const bf16x8 *)&As[wm + i * 16 + fr][fk];
        for (int j = 0; j < 2; j++) b[j] = *(const bf16x8 *)&Bs[wn + j * 16 + fr][fk];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    if (csum && tid < GT && n0 + tid < g.N) g.colsum[n0 + tid] = cs;
    // accumulator: lane l holds column l%16, rows (l/16)*4 + 0..3 of each 16x16 tile
    const int oc = lane & 15, orow = (lane >> 4) * 4;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++) {
                int m = m0 + wm + i * 16 + orow + q, n = n0 + wn + j * 16 + oc;
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][q];
                if (g.bias) v += g.bias[n];
                if (g.act == ACT_RELU) v = v > 0.f ? v : 0.f;
                if (g.mask && !((float)g.mask[m * g.smm + n] > 0.f)) v = 0.f;
                if (g.C) g.C[m * g.scm + n] = v;
                if (g.Cb) g.Cb[m * g.sbm + n] = (bf16)v;
            }
}

#define S_CLAMP 12.0f   // log-variance range used by the loss (gradient zero outside)

// Gaussian NLL over a batch: Y = [mu (no) | s (no)] per row; the target of row b is
// (T[src(b)][o] - mean[o]) * inv_std[o], src(b) = perm[b] when given (mini-batch gather).
__global__ void k_nll(const float *Y, const float *T, const int *perm, const float *tmean, const float *tinv, int B,
                      int no, float inv_count, bf16 *dY, double *loss)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0;
    if (i < B * no) {
        int b = i / no, o = i % no;
        const long r = perm ? perm[b] : b;
        const float tn = (T[r * no + o] - tmean[o]) * tinv[o];
        float mu = Y[(long)b * 2 * no + o], s = Y[(long)b * 2 * no + no + o];
        bool in = s > -S_CLAMP && s < S_CLAMP;
        float sc = fminf(fmaxf(s, -S_CLAMP), S_CLAMP);
        float d = mu - tn;
        float e = expf(-sc);
        l = 0.5 * ((double)sc + (double)d * d * e);
        dY[(long)b * 2 * no + o] = (bf16)(d * e * inv_count);
        dY[(long)b * 2 * no + no + o] = (bf16)(in ? 0.5f * (1.f - d * d * e) * inv_count : 0.f);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0 && l != 0) atomicAdd(loss, l);
}

// layer shapes for the kernels that map a flat parameter index to (layer, row, column):
// weights of layer l at woff[l], din[l] x dout[l] row-major; biases after all weights
struct LayerMap { long woff[9]; int din[9], dout[9]; int nl; long nw; };

// BF16 copies of FP32 parameter i: wb (the flat layout) and, for weights, wt (each layer
// transposed to [dout][din], the fused forward's operand layout)
__device__ __forceinline__ void store_bf16(const LayerMap &L, long i, float x, bf16 *wb, bf16 *wt)
{
    const bf16 v = (bf16)x;
    wb[i] = v;
    if (i >= L.nw) return;
    int l = 0;
    while (l + 1 < L.nl && i >= L.woff[l + 1]) l++;
    const long j = i - L.woff[l];
    const int r = (int)(j / L.dout[l]), c = (int)(j - (long)r * L.dout[l]);
    wt[L.woff[l] + (long)c * L.din[l] + r] = v;
}

__global__ void k_adam(float *w, const float *gr, float *m1, float *m2, bf16 *wb, bf16 *wt, LayerMap L, long n,
                       float lr, float b1, float b2, float eps, float c1, float c2, float wd)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = gr[i] + wd * w[i];
    float a = b1 * m1[i] + (1 - b1) * g;
    float v = b2 * m2[i] + (1 - b2) * g * g;
    m1[i] = a; m2[i] = v;
    float x = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[i] = x;
    store_bf16(L, i, x, wb, wt);
}

__global__ void k_f32_to_bf16(const float *x, bf16 *wb, bf16 *wt, LayerMap L, long n)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) store_bf16(L, i, x[i], wb, wt);
}

// xb[b][k] = bf16((x[src(b)][k] - mean[k]) * inv_std[k]), zero in the padding columns;
// rows come from perm[] when given (mini-batch gather), else in order
__global__ void k_pack_inputs(const float *x, const int *perm, int B, int n_in, int n_pad, const float *mean,
                              const float *inv_std, bf16 *xb)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)B * n_pad) return;
    int b = (int)(i / n_pad), k = (int)(i % n_pad);
    long r = perm ? perm[b] : b;
    xb[i] = k < n_in ? (bf16)((x[r * n_in + k] - mean[k]) * inv_std[k]) : (bf16)0.0f;
}

// ---------------------------------------------------------------- fused forward
// The whole network for 64 rows per workgroup (4 wavefronts) in one launch: the rows'
// FP32 inputs are normalised into a BF16 tile in LDS, then each layer reads its input tile
// from LDS (MFMA A fragments, one 16-byte read each) and its weights from the transposed
// BF16 copy wt ([dout][din] per layer, one 16-byte global read per B fragment; the weights,
// 0.4 MB for 70-256-256-256-160, stay in L2), keeps the layer's output in registers
// (wavefront w owns the 16-column tiles w, w + 4, ...), and, after every wavefront has read
// the tile, writes the ReLU'd BF16 output over it.  The MFMA sequence per output tile is
// k_gemm's (16x16x32 in increasing k, bias, ReLU, BF16 rounding), so every value equals the
// per-layer path.  The output layer's epilogue writes the raw outputs and/or the
// prediction: mean, sigma and the per-row fallback flag (k_predict's arithmetic).
// Shapes: padded input, hidden width and 2 n_out multiples of 16 (32 for the input), at most
// FT_W (FT_OUT for 2 n_out).
#define FT_M 64
#define FT_W 256
#define FT_LD (FT_W + 8)     // LDS row pitch in bf16 (528 B: 16-byte aligned fragments)
#define FT_OUT 192           // 2 n_out at most: the FP32 output tile stays under 64 KB of LDS
#define FT_YLD (FT_OUT + 4)  // LDS row pitch in floats of the output tile

struct FusedArgs {
    const float *X; const int *perm; int B;       // FP32 inputs [*][n_in], rows perm[b] or b
    int n_in, n_in_pad, n_hid, n_layers, n_out;
    const float *xm, *xinv;                       // input normalisation
    const bf16 *wt; const float *w;               // transposed BF16 weights; FP32 parameters (biases)
    bf16 *xb;                                     // optional: the BF16 input tile [B][n_in_pad]
    bf16 *acts;                                   // optional: hidden activations [l][B][n_hid]
    float *Y;                                     // optional: raw outputs [B][2 n_out]
    const float *ym, *ys; float sigma_max;        // optional prediction (mu set): de-normalisation,
    const uint64_t *off;                          // fallback gate, track offsets (empty tracks
    float *mu, *sig; uint8_t *fb; uint32_t *n_fb; // never fall back)
};

__global__ void __launch_bounds__(256) k_mlp_fused(FusedArgs a)
{
    // one LDS tile: the BF16 activations [FT_M][FT_LD], and after the last layer's MFMAs its
    // FP32 outputs [FT_M][FT_YLD]
    extern __shared__ __attribute__((aligned(16))) unsigned char ft_lds[];
    bf16 (*H)[FT_LD] = (bf16 (*)[FT_LD])ft_lds;
    __shared__ uint32_t rowbad[FT_M];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * FT_M;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const int oc = lane & 15, orow = (lane >> 4) * 4;
    const int rows = a.B - r0 < FT_M ? a.B - r0 : FT_M;
    if (tid < FT_M) rowbad[tid] = 0u;
    for (int e = tid; e < FT_M * a.n_in_pad; e += 256) {
        const int r = e / a.n_in_pad, k = e - r * a.n_in_pad;
        float v = 0.f;
        if (r < rows && k < a.n_in) {
            const long src = a.perm ? a.perm[r0 + r] : r0 + r;
            v = (a.X[src * a.n_in + k] - a.xm[k]) * a.xinv[k];
        }
        const bf16 b = (bf16)v;
        H[r][k] = b;
        if (a.xb && r < rows) a.xb[(long)(r0 + r) * a.n_in_pad + k] = b;
    }
    __syncthreads();
    int K = a.n_in_pad;
    const long hh = (long)a.n_hid * a.n_hid, w0 = (long)a.n_in_pad * a.n_hid;
    const int n2 = 2 * a.n_out;
    const long nw = w0 + (long)(a.n_layers - 1) * hh + (long)a.n_hid * n2;
#pragma unroll 1
    for (int l = 0; l <= a.n_layers; l++) {
        const bool last = l == a.n_layers;
        const int N = last ? n2 : a.n_hid, nct = N / 16;
        // the flat parameter layout: weights of every layer, then every bias
        const bf16 *W = a.wt + (l == 0 ? 0 : w0 + (long)(l - 1) * hh);
        const float *bias = a.w + nw + (long)l * a.n_hid;
        f32x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int k0 = 0; k0 < K; k0 += 32) {
            bf16x8 af[4];
#pragma unroll
            for (int rt = 0; rt < 4; rt++) af[rt] = *(const bf16x8 *)&H[rt * 16 + fr][k0 + fk];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int ct = wave + 4 * j;
                if (ct < nct) {
                    const bf16x8 bv = *(const bf16x8 *)(W + (long)(ct * 16 + fr) * K + k0 + fk);
#pragma unroll
                    for (int rt = 0; rt < 4; rt++)
                        acc[rt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[rt], bv, acc[rt][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();                          // the input tile has been read by every wavefront
        // bias (and ReLU + BF16 for a hidden layer) into the LDS tile: lane l holds column
        // l % 16, rows (l / 16) * 4 + q of each 16 x 16 tile
        float *Yf = (float *)ft_lds;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int ct = wave + 4 * j;
            if (ct >= nct) continue;
            const int n = ct * 16 + oc;
            const float bn = bias[n];
#pragma unroll
            for (int rt = 0; rt < 4; rt++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int m = rt * 16 + orow + q;
                    const float v = acc[rt][j][q] + bn;
                    if (last) Yf[m * FT_YLD + n] = v;
                    else H[m][n] = (bf16)(v > 0.f ? v : 0.f);
                }
        }
        __syncthreads();
        if (!last && a.acts) {                    // the activations kept for the backward pass
            const int cpr = a.n_hid / 8;
            for (int e = tid; e < rows * cpr; e += 256) {
                const int r = e / cpr, c = (e - r * cpr) * 8;
                *(bf16x8 *)(a.acts + ((long)l * a.B + r0 + r) * a.n_hid + c) = *(const bf16x8 *)&H[r][c];
            }
        }
        K = N;
    }
    // the output layer from LDS: raw outputs and/or the prediction
    const float *Yf = (const float *)ft_lds;
    for (int e = tid; e < rows * n2; e += 256) {
        const int r = e / n2, n = e - r * n2, gr = r0 + r;
        const float v = Yf[r * FT_YLD + n];
        if (a.Y) a.Y[(long)gr * n2 + n] = v;
        if (a.mu) {
            const int no = a.n_out;
            if (n < no) a.mu[(long)gr * no + n] = a.ym[n] + a.ys[n] * v;
            else {
                const int o = n - no;
                const float sc = fminf(fmaxf(v, -S_CLAMP), S_CLAMP);
                const float sg = a.ys[o] * expf(0.5f * sc);
                a.sig[(long)gr * no + o] = sg;
                if (!(sg <= a.sigma_max)) rowbad[r] = 1u;
            }
        }
    }
    __syncthreads();
    if (a.mu && tid < FT_M) {
        bool f = false;
        if (tid < rows) {
            const int gr = r0 + tid;
            f = rowbad[tid] && a.off[gr + 1] != a.off[gr];
            a.fb[gr] = f ? 1 : 0;
        }
        const uint64_t m = __ballot(f);
        if (tid == 0 && m) atomicAdd(a.n_fb, (uint32_t)__popcll(m));
    }
}

static size_t fused_lds_bytes()
{
    const size_t h = (size_t)FT_M * FT_LD * 2, y = (size_t)FT_M * FT_YLD * 4;
    return h > y ? h : y;
}

// Per-track features (one workgroup per track, entries [off[t], off[t+1]) of rad):
//   [0, 64)  fraction of the track's initial particles in each of 64 log-spaced bins of
//            distance from the track centroid, 0.5 nm .. 20 um
//   64..68   initial OH, e-aq (codes 1 and 5), H, H3O+ and H2 (code-7 markers) per 100 eV
//   69       log10(E / 1 keV)
// Radical codes follow the transport encoding w = pid*8 + code; code 4 is not a particle.
#define NF_HIST 64
#define NF_TOTAL 70
__global__ void __launch_bounds__(256) k_features(const float4 *rad, const uint64_t *off, const float *E, float *F,
                                                  int nfeat)
{
    __shared__ float sx[4], sy[4], sz[4];
    __shared__ unsigned int hist[NF_HIST], spc[5], nn;
    const int t = blockIdx.x, tid = threadIdx.x;
    const uint64_t a = off[t], b = off[t + 1];
    float x = 0, y = 0, z = 0;
    unsigned int cnt = 0;
    if (tid < NF_HIST) hist[tid] = 0;
    if (tid < 5) spc[tid] = 0;
    if (tid == 0) nn = 0;
    __syncthreads();
    for (uint64_t i = a + tid; i < b; i += 256) {
        float4 e = rad[i];
        int code = (int)((uint32_t)(e.w + 0.5f) & 7u);
        if (code == 4 || code == 7) continue;
        x += e.x; y += e.y; z += e.z; cnt++;
    }
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o, 64); y += __shfl_xor(y, o, 64); z += __shfl_xor(z, o, 64); cnt += __shfl_xor(cnt, o, 64);
    }
    if ((tid & 63) == 0) { sx[tid >> 6] = x; sy[tid >> 6] = y; sz[tid >> 6] = z; atomicAdd(&nn, cnt); }
    __syncthreads();
    const float inv = nn ? 1.0f / (float)nn : 0.f;
    const float cx = (sx[0] + sx[1] + sx[2] + sx[3]) * inv, cy = (sy[0] + sy[1] + sy[2] + sy[3]) * inv,
                cz = (sz[0] + sz[1] + sz[2] + sz[3]) * inv;
    const float lo = logf(0.5f), scale = (float)NF_HIST / (logf(20000.f) - logf(0.5f));
    for (uint64_t i = a + tid; i < b; i += 256) {
        float4 e = rad[i];
        int code = (int)((uint32_t)(e.w + 0.5f) & 7u);
        if (code == 4) continue;
        if (code == 7) { atomicAdd(&spc[4], 1u); continue; }
        int s = code == 0 ? 0 : (code == 1 || code == 5) ? 1 : code == 2 ? 2 : code == 3 ? 3 : -1;
        if (s >= 0) atomicAdd(&spc[s], 1u);
        float dx = e.x - cx, dy = e.y - cy, dz = e.z - cz;
        float r = sqrtf(dx * dx + dy * dy + dz * dz);
        int bin = (int)((logf(fmaxf(r, 0.5f)) - lo) * scale);
        bin = bin < 0 ? 0 : bin >= NF_HIST ? NF_HIST - 1 : bin;
        atomicAdd(&hist[bin], 1u);
    }
    __syncthreads();
    float *f = F + (size_t)t * nfeat;
    const float per100 = E[t] > 0.f ? 100.f / E[t] : 0.f;
    if (tid < NF_HIST) f[tid] = hist[tid] * inv;
    if (tid < 5) f[NF_HIST + tid] = spc[tid] * per100;
    if (tid == 0) f[NF_HIST + 5] = E[t] > 0.f ? log10f(E[t] / 1000.f) : 0.f;
}

// radical offsets of primary p in the canonical buffer: cnt[base_p[p]] (p = 0..n)
__global__ void k_pid_off(const uint32_t *base_p, const uint32_t *cnt, uint32_t n, uint64_t *off)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= n) off[p] = cnt[base_p[p]];
}

// mu = ym + ys * mu_n, sigma = ys * exp(s / 2); fallback when any sigma > sigma_max, unless
// the track has no particles.  One wavefront per track.
__global__ void k_predict(const float *Y, int B, int no, const float *ym, const float *ys, const uint64_t *off,
                          float sigma_max, float *mu, float *sig, uint8_t *fb, uint32_t *n_fb)
{
    int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (b >= B) return;
    bool bad = false;
    for (int o = lane; o < no; o += 64) {
        float m = ym[o] + ys[o] * Y[(long)b * 2 * no + o];
        float s = fminf(fmaxf(Y[(long)b * 2 * no + no + o], -S_CLAMP), S_CLAMP);
        float sg = ys[o] * expf(0.5f * s);
        mu[(long)b * no + o] = m;
        sig[(long)b * no + o] = sg;
        bad |= !(sg <= sigma_max);
    }
    bool any = __any(bad);
    if (lane == 0) {
        bool empty = off[b + 1] == off[b];
        uint8_t f = (any && !empty) ? 1 : 0;
        fb[b] = f;
        if (f) atomicAdd(n_fb, 1u);
    }
}

}  // namespace dna

using namespace dna;

// ------------------------------------------------------------------ host side
struct dna_surrogate {
    dna_ctx *ctx = nullptr;
    int n_in = 0, n_in_pad = 0, n_hid = 0, n_layers = 0, n_out = 0;   // the net emits 2*n_out
    std::vector<int> din, dout;                                        // per layer
    std::vector<size_t> woff, boff;                                    // offsets into the flat parameters
    size_t n_params = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // FP32 master, gradient, Adam moments
    bf16 *wb = nullptr;                                               // BF16 weights used by the GEMMs
    bf16 *wt = nullptr;                                               // BF16 weights, each layer transposed
    LayerMap lmap;
    float *norm = nullptr;   // x_mean[n_in], x_inv_std[n_in], y_mean[n_out], y_std[n_out], y_inv_std[n_out]
    long step = 0;
    int cap = 0;             // batch capacity of the buffers below
    bf16 *xb = nullptr, *acts = nullptr, *dh = nullptr, *dh2 = nullptr;
    float *y = nullptr, *xin = nullptr, *tin = nullptr;
    double *loss = nullptr;
    // prediction buffers
    size_t pcap = 0;
    uint64_t *d_off = nullptr;
    float *d_E = nullptr, *d_F = nullptr, *d_mu = nullptr, *d_sig = nullptr;
    uint8_t *d_fb = nullptr;
    uint32_t *d_nfb = nullptr;
};

static int ld_mode(long s_row, long s_k, int R, int K, const void *p)
{
    bool al = ((uintptr_t)p & 15) == 0;
    if (s_k == 1 && s_row % 8 == 0 && K % 8 == 0 && al) return LD_KCONT;
    if (s_row == 1 && s_k % 8 == 0 && R % 8 == 0 && al) return LD_RCONT;
    return LD_SCALAR;
}

static int launch_gemm(hipStream_t st, GemmArgs g)
{
    g.amode = ld_mode(g.sam, g.sak, g.M, g.K, g.A);
    g.bmode = ld_mode(g.sbn, g.sbk, g.N, g.K, g.B);
    dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, g);
    DNA_HIP(hipGetLastError());
    return DNA_OK;
}

template <class T> static hipError_t sg_alloc(T *&p, size_t n)
{
    if (p) { (void)hipFree(p); p = nullptr; }
    return hipMalloc((void **)&p, n * sizeof(T) + 16);
}

static int ensure_batch(dna_surrogate *s, int B)
{
    if (B <= s->cap) return DNA_OK;
    int L = s->n_layers;
    int wmax = s->n_hid > 2 * s->n_out ? s->n_hid : 2 * s->n_out;
    DNA_HIP(sg_alloc(s->xb, (size_t)B * s->n_in_pad));
    DNA_HIP(sg_alloc(s->acts, (size_t)B * s->n_hid * L));
    DNA_HIP(sg_alloc(s->dh, (size_t)B * wmax));
    DNA_HIP(sg_alloc(s->dh2, (size_t)B * wmax));
    DNA_HIP(sg_alloc(s->y, (size_t)B * 2 * s->n_out));
    DNA_HIP(sg_alloc(s->xin, (size_t)B * s->n_in));
    DNA_HIP(sg_alloc(s->tin, (size_t)B * s->n_out));
    s->cap = B;
    return DNA_OK;
}

static float *x_mean(dna_surrogate *s) { return s->norm; }
static float *x_inv(dna_surrogate *s) { return s->norm + s->n_in; }
static float *y_mean(dna_surrogate *s) { return s->norm + 2 * s->n_in; }
static float *y_std(dna_surrogate *s) { return s->norm + 2 * s->n_in + s->n_out; }
static float *y_inv(dna_surrogate *s) { return s->norm + 2 * s->n_in + 2 * s->n_out; }

static int pack_inputs(dna_surrogate *s, const float *x_dev, const int *perm, int B)
{
    long n = (long)B * s->n_in_pad;
    hipLaunchKernelGGL(k_pack_inputs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->ctx->stream, x_dev, perm, B,
                       s->n_in, s->n_in_pad, x_mean(s), x_inv(s), s->xb);
    DNA_HIP(hipGetLastError());
    return DNA_OK;
}

static bool fused_ok(const dna_surrogate *s)
{
    return s->ctx->mlp_fused && s->n_in_pad <= FT_W && s->n_hid <= FT_W && s->n_hid % 16 == 0 &&
           2 * s->n_out <= FT_OUT && (2 * s->n_out) % 16 == 0;
}

// The fused forward over B rows of FP32 inputs x_dev (row perm[b] when perm is given):
// writes the BF16 input tile and the hidden activations when training (want_acts), the raw
// outputs to Y when given, and the prediction when p_mu is given.
static int forward_fused(dna_surrogate *s, const float *x_dev, const int *perm, int B, bool want_acts, float *Y,
                         float sigma_max = 0.f, const uint64_t *off = nullptr, float *p_mu = nullptr,
                         float *p_sig = nullptr, uint8_t *fb = nullptr, uint32_t *n_fb = nullptr)
{
    FusedArgs a; memset(&a, 0, sizeof(a));
    a.X = x_dev; a.perm = perm; a.B = B;
    a.n_in = s->n_in; a.n_in_pad = s->n_in_pad; a.n_hid = s->n_hid; a.n_layers = s->n_layers; a.n_out = s->n_out;
    a.xm = x_mean(s); a.xinv = x_inv(s);
    a.wt = s->wt; a.w = s->w;
    if (want_acts) { a.xb = s->xb; a.acts = s->acts; }
    a.Y = Y;
    if (p_mu) {
        a.ym = y_mean(s); a.ys = y_std(s); a.sigma_max = sigma_max; a.off = off;
        a.mu = p_mu; a.sig = p_sig; a.fb = fb; a.n_fb = n_fb;
    }
    hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)((B + FT_M - 1) / FT_M)), dim3(256), fused_lds_bytes(),
                       s->ctx->stream, a);
    DNA_HIP(hipGetLastError());
    return DNA_OK;
}

// forward on s->xb (B x n_in_pad) -> s->y (FP32, B x 2 n_out); hidden BF16 activations kept in s->acts
static int forward(dna_surrogate *s, int B)
{
    hipStream_t st = s->ctx->stream;
    const bf16 *in = s->xb;
    long ld_in = s->n_in_pad;
    for (int l = 0; l <= s->n_layers; l++) {
        GemmArgs g; memset(&g, 0, sizeof(g));
        g.A = in; g.sam = ld_in; g.sak = 1;
        g.B = s->wb + s->woff[l]; g.sbk = s->dout[l]; g.sbn = 1;
        g.M = B; g.N = s->dout[l]; g.K = s->din[l];
        g.bias = s->w + s->boff[l];
        if (l < s->n_layers) {
            bf16 *out = s->acts + (size_t)l * B * s->n_hid;
            g.act = ACT_RELU; g.Cb = out; g.sbm = s->n_hid;
            int rc = launch_gemm(st, g); if (rc) return rc;
            in = out; ld_in = s->n_hid;
        } else {
            g.C = s->y; g.scm = 2 * s->n_out;
            int rc = launch_gemm(st, g); if (rc) return rc;
        }
    }
    return DNA_OK;
}

// backward from dY (BF16 in s->dh, B x 2 n_out) -> parameter gradients in s->g (overwritten)
static int backward(dna_surrogate *s, int B)
{
    hipStream_t st = s->ctx->stream;
    bf16 *dcur = s->dh, *dnext = s->dh2;
    long ld_d = 2 * s->n_out;
    for (int l = s->n_layers; l >= 0; l--) {
        const bf16 *in = l == 0 ? s->xb : s->acts + (size_t)(l - 1) * B * s->n_hid;
        long ld_in = l == 0 ? s->n_in_pad : s->n_hid;
        // dW[din x dout] = in^T dcur (sum over the batch)
        GemmArgs gw; memset(&gw, 0, sizeof(gw));
        gw.A = in; gw.sam = 1; gw.sak = ld_in;
        gw.B = dcur; gw.sbk = ld_d; gw.sbn = 1;
        gw.M = s->din[l]; gw.N = s->dout[l]; gw.K = B;
        gw.C = s->g + s->woff[l]; gw.scm = s->dout[l];
        gw.colsum = s->g + s->boff[l];            // the bias gradient: column sums of dcur
        int rc = launch_gemm(st, gw); if (rc) return rc;
        if (l == 0) break;
        // gradient at the previous layer's pre-activation: (dcur W^T) * ReLU'(in)
        GemmArgs gd; memset(&gd, 0, sizeof(gd));
        gd.A = dcur; gd.sam = ld_d; gd.sak = 1;
        gd.B = s->wb + s->woff[l]; gd.sbk = 1; gd.sbn = s->dout[l];
        gd.M = B; gd.N = s->din[l]; gd.K = s->dout[l];
        gd.Cb = dnext; gd.sbm = s->n_hid;
        gd.mask = in; gd.smm = ld_in;
        rc = launch_gemm(st, gd); if (rc) return rc;
        bf16 *tmp = dcur; dcur = dnext; dnext = tmp;
        ld_d = s->n_hid;
    }
    return DNA_OK;
}

// loss and gradients of rows perm[b] (or b) of the raw targets t_dev
static int loss_and_backward(dna_surrogate *s, const float *t_dev, const int *perm, int B)
{
    hipStream_t st = s->ctx->stream;
    int no = s->n_out;
    hipLaunchKernelGGL(k_nll, dim3((B * no + 255) / 256), dim3(256), 0, st, s->y, t_dev, perm, y_mean(s), y_inv(s), B,
                       no, 1.0f / (float)(B * no), s->dh, s->loss);
    DNA_HIP(hipGetLastError());
    return backward(s, B);
}

// forward of a training batch: the fused kernel, or packing and one GEMM per layer
static int train_forward(dna_surrogate *s, const float *x_dev, const int *perm, int B)
{
    if (fused_ok(s)) return forward_fused(s, x_dev, perm, B, true, s->y);
    int rc = pack_inputs(s, x_dev, perm, B); if (rc) return rc;
    return forward(s, B);
}

static uint64_t sm64(uint64_t &z)
{
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull; x ^= x >> 27; x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

extern "C" int dna_surrogate_create(dna_ctx *ctx, int n_in, int n_hidden, int n_layers, int n_out, uint64_t seed,
                                    dna_surrogate **out)
{
    if (!ctx || !out || n_in < 1 || n_in > 4096 || n_hidden < 16 || n_hidden > 4096 || n_hidden % 16 || n_layers < 1 ||
        n_layers > 8 || n_out < 1 || n_out > 2048) {
        set_error("dna_surrogate_create: bad arguments"); return DNA_ERR_ARG;
    }
    DNA_HIP(hipSetDevice(ctx->device));
    dna_surrogate *s = new dna_surrogate();
    ctx_retain(ctx);                              // given back by dna_surrogate_destroy
    s->ctx = ctx; s->n_in = n_in; s->n_in_pad = (n_in + 31) / 32 * 32; s->n_hid = n_hidden;
    s->n_layers = n_layers; s->n_out = n_out;
    size_t off = 0;
    for (int l = 0; l <= n_layers; l++) {
        int di = l == 0 ? s->n_in_pad : n_hidden, dd = l == n_layers ? 2 * n_out : n_hidden;
        s->din.push_back(di); s->dout.push_back(dd);
        s->woff.push_back(off); off += (size_t)di * dd;
    }
    for (int l = 0; l <= n_layers; l++) { s->boff.push_back(off); off += s->dout[l]; }
    s->n_params = off;
    memset(&s->lmap, 0, sizeof(s->lmap));
    s->lmap.nl = n_layers + 1; s->lmap.nw = (long)s->boff[0];
    for (int l = 0; l <= n_layers; l++) { s->lmap.woff[l] = (long)s->woff[l]; s->lmap.din[l] = s->din[l]; s->lmap.dout[l] = s->dout[l]; }
    // He-normal weights (Box-Muller over a SplitMix64 stream), zero biases, zero rows for the
    // padded inputs; the output layer starts 10x smaller
    std::vector<float> h(off, 0.f);
    uint64_t z = seed;
    for (int l = 0; l <= n_layers; l++) {
        int fan_in = l == 0 ? n_in : n_hidden;
        double sd = std::sqrt(2.0 / fan_in) * (l == n_layers ? 0.1 : 1.0);
        for (int i = 0; i < s->din[l]; i++)
            for (int j = 0; j < s->dout[l]; j++) {
                double u1 = ((sm64(z) >> 11) + 1) * 0x1.0p-53, u2 = (sm64(z) >> 11) * 0x1.0p-53;
                double gn = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                h[s->woff[l] + (size_t)i * s->dout[l] + j] = (l == 0 && i >= n_in) ? 0.f : (float)(gn * sd);
            }
    }
    std::vector<float> nrm(2 * n_in + 3 * n_out);
    for (int i = 0; i < n_in; i++) { nrm[i] = 0.f; nrm[n_in + i] = 1.f; }
    for (int o = 0; o < n_out; o++) { nrm[2 * n_in + o] = 0.f; nrm[2 * n_in + n_out + o] = 1.f; nrm[2 * n_in + 2 * n_out + o] = 1.f; }
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = sg_alloc(s->w, off);
    if (e == hipSuccess) e = sg_alloc(s->g, off);
    if (e == hipSuccess) e = sg_alloc(s->m1, off);
    if (e == hipSuccess) e = sg_alloc(s->m2, off);
    if (e == hipSuccess) e = sg_alloc(s->wb, off);
    if (e == hipSuccess) e = sg_alloc(s->wt, off);
    if (e == hipSuccess) e = sg_alloc(s->norm, nrm.size());
    if (e == hipSuccess) e = sg_alloc(s->loss, 1);
    if (e == hipSuccess) e = sg_alloc(s->d_nfb, 1);
    if (e == hipSuccess) e = hipMemcpy(s->w, h.data(), off * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->norm, nrm.data(), nrm.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->m1, 0, off * 4);
    if (e == hipSuccess) e = hipMemset(s->m2, 0, off * 4);
    if (e == hipSuccess) e = hipMemset(s->g, 0, off * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((off + 255) / 256)), dim3(256), 0, ctx->stream, s->w, s->wb,
                           s->wt, s->lmap, (long)off);
        e = hipStreamSynchronize(ctx->stream);
    }
    if (e != hipSuccess) {
        set_error(std::string("dna_surrogate_create: ") + hipGetErrorString(e));
        dna_surrogate_destroy(s);
        return DNA_ERR_HIP;
    }
    *out = s;
    return DNA_OK;
}

extern "C" void dna_surrogate_destroy(dna_surrogate *s)
{
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    void *ptrs[] = {s->w, s->g, s->m1, s->m2, s->wb, s->wt, s->norm, s->loss, s->xb, s->acts, s->dh, s->dh2, s->y,
                    s->xin, s->tin, s->d_off, s->d_E, s->d_F, s->d_mu, s->d_sig, s->d_fb, s->d_nfb};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    dna_ctx *c = s->ctx;
    delete s;
    ctx_release(c);
}

extern "C" int dna_surrogate_shape(const dna_surrogate *s, int *n_in, int *n_hidden, int *n_layers, int *n_out,
                                   uint64_t *n_params)
{
    if (!s) { set_error("dna_surrogate_shape: NULL"); return DNA_ERR_ARG; }
    if (n_in) *n_in = s->n_in;
    if (n_hidden) *n_hidden = s->n_hid;
    if (n_layers) *n_layers = s->n_layers;
    if (n_out) *n_out = s->n_out;
    if (n_params) *n_params = s->n_params;
    return DNA_OK;
}

extern "C" int dna_surrogate_get_params(dna_surrogate *s, float *out)
{
    if (!s || !out) { set_error("dna_surrogate_get_params: NULL"); return DNA_ERR_ARG; }
    DNA_HIP(hipStreamSynchronize(s->ctx->stream));
    DNA_HIP(hipMemcpy(out, s->w, s->n_params * 4, hipMemcpyDeviceToHost));
    return DNA_OK;
}

extern "C" int dna_surrogate_set_params(dna_surrogate *s, const float *in)
{
    if (!s || !in) { set_error("dna_surrogate_set_params: NULL"); return DNA_ERR_ARG; }
    DNA_HIP(hipStreamSynchronize(s->ctx->stream));
    DNA_HIP(hipMemcpy(s->w, in, s->n_params * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((s->n_params + 255) / 256)), dim3(256), 0, s->ctx->stream, s->w,
                       s->wb, s->wt, s->lmap, (long)s->n_params);
    DNA_HIP(hipStreamSynchronize(s->ctx->stream));
    return DNA_OK;
}

extern "C" int dna_surrogate_set_norm(dna_surrogate *s, const float *xm, const float *xs, const float *ym,
                                      const float *ys)
{
    if (!s || !xm || !xs || !ym || !ys) { set_error("dna_surrogate_set_norm: NULL"); return DNA_ERR_ARG; }
    int ni = s->n_in, no = s->n_out;
    std::vector<float> nrm(2 * ni + 3 * no);
    for (int i = 0; i < ni; i++) {
        if (!(xs[i] > 0.f)) { set_error("dna_surrogate_set_norm: x_std must be > 0"); return DNA_ERR_ARG; }
        nrm[i] = xm[i]; nrm[ni + i] = 1.f / xs[i];
    }
    for (int o = 0; o < no; o++) {
        if (!(ys[o] > 0.f)) { set_error("dna_surrogate_set_norm: y_std must be > 0"); return DNA_ERR_ARG; }
        nrm[2 * ni + o] = ym[o]; nrm[2 * ni + no + o] = ys[o]; nrm[2 * ni + 2 * no + o] = 1.f / ys[o];
    }
    DNA_HIP(hipStreamSynchronize(s->ctx->stream));
    DNA_HIP(hipMemcpy(s->norm, nrm.data(), nrm.size() * 4, hipMemcpyHostToDevice));
    return DNA_OK;
}

// Raw network outputs Y [B x 2 n_out] (normalised mean, log-variance) for X [B x n_in].
extern "C" int dna_surrogate_forward(dna_surrogate *s, const float *X, int B, float *Y)
{
    if (!s || !X || !Y || B < 1) { set_error("dna_surrogate_forward: bad arguments"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(s->ctx->device));
    int rc = ensure_batch(s, B); if (rc) return rc;
    hipStream_t st = s->ctx->stream;
    DNA_HIP(hipMemcpyAsync(s->xin, X, (size_t)B * s->n_in * 4, hipMemcpyHostToDevice, st));
    if (fused_ok(s)) { rc = forward_fused(s, s->xin, nullptr, B, false, s->y); if (rc) return rc; }
    else {
        rc = pack_inputs(s, s->xin, nullptr, B); if (rc) return rc;
        rc = forward(s, B); if (rc) return rc;
    }
    DNA_HIP(hipMemcpyAsync(Y, s->y, (size_t)B * 2 * s->n_out * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    return DNA_OK;
}

// Mean loss and the full parameter gradient (flat layout of get_params) on one batch.
extern "C" int dna_surrogate_grad(dna_surrogate *s, const float *X, const float *T, int B, double *loss, float *grad)
{
    if (!s || !X || !T || B < 1) { set_error("dna_surrogate_grad: bad arguments"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(s->ctx->device));
    int rc = ensure_batch(s, B); if (rc) return rc;
    hipStream_t st = s->ctx->stream;
    DNA_HIP(hipMemcpyAsync(s->xin, X, (size_t)B * s->n_in * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(s->tin, T, (size_t)B * s->n_out * 4, hipMemcpyHostToDevice, st));
    rc = train_forward(s, s->xin, nullptr, B); if (rc) return rc;
    DNA_HIP(hipMemsetAsync(s->loss, 0, sizeof(double), st));
    rc = loss_and_backward(s, s->tin, nullptr, B); if (rc) return rc;
    double h_loss = 0;
    DNA_HIP(hipMemcpyAsync(&h_loss, s->loss, sizeof(double), hipMemcpyDeviceToHost, st));
    if (grad) DNA_HIP(hipMemcpyAsync(grad, s->g, s->n_params * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipStreamSynchronize(st));
    if (loss) *loss = h_loss / (double)B / (double)s->n_out;
    return DNA_OK;
}

// Adam over `epochs` shuffled passes of N samples (X [N x n_in], T [N x n_out], host),
// mini-batches of `batch` (a final partial batch is dropped).  Inputs and targets are
// normalised with the statistics of dna_surrogate_set_norm.  *mean_loss: mean NLL of the
// last epoch.
extern "C" int dna_surrogate_train(dna_surrogate *s, const float *X, const float *T, int N, int batch, int epochs,
                                   float lr, float weight_decay, uint64_t seed, double *mean_loss)
{
    if (!s || !X || !T || N < 1 || batch < 1 || batch > N || epochs < 1 || !(lr > 0.f)) {
        set_error("dna_surrogate_train: bad arguments"); return DNA_ERR_ARG;
    }
    DNA_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    int rc = ensure_batch(s, batch); if (rc) return rc;
    float *dX = nullptr, *dT = nullptr;
    int *dP = nullptr;
    DNA_HIP(hipMalloc((void **)&dX, (size_t)N * s->n_in * 4));
    DNA_HIP(hipMalloc((void **)&dT, (size_t)N * s->n_out * 4));
    DNA_HIP(hipMalloc((void **)&dP, (size_t)N * 4));
    DNA_HIP(hipMemcpyAsync(dX, X, (size_t)N * s->n_in * 4, hipMemcpyHostToDevice, st));
    DNA_HIP(hipMemcpyAsync(dT, T, (size_t)N * s->n_out * 4, hipMemcpyHostToDevice, st));
    std::vector<int> perm(N);
    for (int i = 0; i < N; i++) perm[i] = i;
    uint64_t z = seed ^ 0x5DEECE66Dull;
    const int nb = N / batch;
    double last = 0;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = N - 1; i > 0; i--) { int j = (int)(sm64(z) % (uint64_t)(i + 1)); std::swap(perm[i], perm[j]); }
        DNA_HIP(hipMemcpyAsync(dP, perm.data(), (size_t)N * 4, hipMemcpyHostToDevice, st));
        DNA_HIP(hipMemsetAsync(s->loss, 0, sizeof(double), st));
        for (int bi = 0; bi < nb; bi++) {
            const int *pb = dP + (size_t)bi * batch;
            rc = train_forward(s, dX, pb, batch); if (rc) break;
            rc = loss_and_backward(s, dT, pb, batch); if (rc) break;
            s->step++;
            float c1 = 1.f - (float)std::pow((double)b1, (double)s->step);
            float c2 = 1.f - (float)std::pow((double)b2, (double)s->step);
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((s->n_params + 255) / 256)), dim3(256), 0, st, s->w, s->g, s->m1,
                               s->m2, s->wb, s->wt, s->lmap, (long)s->n_params, lr, b1, b2, 1e-8f, c1, c2, weight_decay);
        }
        if (rc) break;
        // the host permutation is reshuffled next epoch, so this epoch's upload must have
        // been consumed: read the loss back, which waits for the stream
        double l = 0;
        hipError_t e = hipMemcpyAsync(&l, s->loss, sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) { set_error(std::string("dna_surrogate_train: ") + hipGetErrorString(e)); rc = DNA_ERR_HIP; break; }
        last = l / ((double)nb * batch * s->n_out);
    }
    (void)hipStreamSynchronize(st);          // errors were checked at each epoch's read-back
    (void)hipFree(dX); (void)hipFree(dT); (void)hipFree(dP);
    if (rc) return rc;
    if (mean_loss) *mean_loss = last;
    return DNA_OK;
}

static int ensure_pred(dna_surrogate *s, size_t n)
{
    if (n <= s->pcap) return DNA_OK;
    DNA_HIP(sg_alloc(s->d_off, n + 1));
    DNA_HIP(sg_alloc(s->d_E, n));
    DNA_HIP(sg_alloc(s->d_F, n * NF_TOTAL));
    DNA_HIP(sg_alloc(s->d_mu, n * s->n_out));
    DNA_HIP(sg_alloc(s->d_sig, n * s->n_out));
    DNA_HIP(sg_alloc(s->d_fb, n));
    s->pcap = n;
    return DNA_OK;
}

// features of n tracks into s->d_F from radicals d_rad with device offsets s->d_off
static int features_dev(dna_surrogate *s, const float4 *d_rad, uint32_t n)
{
    hipLaunchKernelGGL(k_features, dim3(n), dim3(256), 0, s->ctx->stream, d_rad, s->d_off, s->d_E, s->d_F, NF_TOTAL);
    DNA_HIP(hipGetLastError());
    return DNA_OK;
}

// Per-track features [n][70].  rad == NULL: the n primaries resident from the last transport
// (track p = primary p, inst_off ignored); else host radicals with n+1 instance offsets.
extern "C" int dna_track_features(dna_surrogate *s, const float *rad, const uint64_t *inst_off, uint32_t n,
                                  const float *E, float *F)
{
    if (!s || !E || !F || n == 0 || (rad && !inst_off)) { set_error("dna_track_features: bad arguments"); return DNA_ERR_ARG; }
    dna_ctx *c = s->ctx;
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    int rc = ensure_pred(s, n); if (rc) return rc;
    const float4 *d_rad = nullptr;
    float4 *tmp = nullptr;
    if (!rad) {
        if (n != c->last_n || !c->d_rad) { set_error("dna_track_features: n must equal the primaries of the last transport"); return DNA_ERR_ARG; }
        hipLaunchKernelGGL(k_pid_off, dim3((n + 256) / 256), dim3(256), 0, st, c->d_base_p, c->d_cnt, n, s->d_off);
        d_rad = c->d_rad;
    } else {
        std::vector<uint64_t> o(inst_off, inst_off + n + 1);
        for (uint32_t i = 0; i < n; i++) if (o[i + 1] < o[i]) { set_error("dna_track_features: offsets must be non-decreasing"); return DNA_ERR_ARG; }
        uint64_t m = inst_off[n] - inst_off[0];
        for (auto &v : o) v -= inst_off[0];
        DNA_HIP(hipMalloc((void **)&tmp, m * 16 + 16));
        DNA_HIP(hipMemcpyAsync(tmp, rad + 4 * inst_off[0], m * 16, hipMemcpyHostToDevice, st));
        DNA_HIP(hipMemcpyAsync(s->d_off, o.data(), (n + 1) * 8, hipMemcpyHostToDevice, st));
        d_rad = tmp;
    }
    DNA_HIP(hipMemcpyAsync(s->d_E, E, (size_t)n * 4, hipMemcpyHostToDevice, st));
    rc = features_dev(s, d_rad, n);
    if (rc == DNA_OK) {
        hipError_t e = hipMemcpyAsync(F, s->d_F, (size_t)n * NF_TOTAL * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) { set_error(std::string("dna_track_features: ") + hipGetErrorString(e)); rc = DNA_ERR_HIP; }
    }
    if (tmp) { (void)hipStreamSynchronize(st); (void)hipFree(tmp); }
    return rc;
}

// Surrogate chemistry for the n primaries resident from the last transport: features,
// network, de-normalised mean G and sigma per output ([n][n_out], host, may be NULL) and
// fallback[p] = 1 where some sigma exceeds sigma_max (those tracks are for IRT).
// Requires n_in == 70.  t_ms (optional, 3 values): features, network, whole call.
extern "C" int dna_surrogate_predict(dna_surrogate *s, const float *E, uint32_t n, float sigma_max, float *mu,
                                     float *sigma, uint8_t *fallback, uint32_t *n_fallback, float *t_ms)
{
    if (!s || !E || !fallback || n == 0 || s->n_in != NF_TOTAL) { set_error("dna_surrogate_predict: bad arguments"); return DNA_ERR_ARG; }
    dna_ctx *c = s->ctx;
    if (n != c->last_n || !c->d_rad) { set_error("dna_surrogate_predict: n must equal the primaries of the last transport"); return DNA_ERR_ARG; }
    DNA_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    int rc = ensure_pred(s, n); if (rc) return rc;
    const int chunk = 65536;
    rc = ensure_batch(s, (int)std::min<uint32_t>(n, chunk)); if (rc) return rc;
    DNA_HIP(hipEventRecord(c->ev[0], st));
    DNA_HIP(hipMemcpyAsync(s->d_E, E, (size_t)n * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pid_off, dim3((n + 256) / 256), dim3(256), 0, st, c->d_base_p, c->d_cnt, n, s->d_off);
    rc = features_dev(s, c->d_rad, n); if (rc) return rc;
    DNA_HIP(hipEventRecord(c->ev[1], st));
    DNA_HIP(hipMemsetAsync(s->d_nfb, 0, 4, st));
    if (fused_ok(s) && n <= 0x7FFFFFFFu) {
        // one launch for every track: normalisation, the layers and the prediction epilogue
        rc = forward_fused(s, s->d_F, nullptr, (int)n, false, nullptr, sigma_max, s->d_off, s->d_mu, s->d_sig, s->d_fb,
                           s->d_nfb);
        if (rc) return rc;
    } else
        for (uint32_t b0 = 0; b0 < n; b0 += chunk) {
            int B = (int)std::min<uint32_t>(chunk, n - b0);
            rc = pack_inputs(s, s->d_F + (size_t)b0 * NF_TOTAL, nullptr, B); if (rc) return rc;
            rc = forward(s, B); if (rc) return rc;
            hipLaunchKernelGGL(k_predict, dim3((B + 3) / 4), dim3(256), 0, st, s->y, B, s->n_out, y_mean(s), y_std(s),
                               s->d_off + b0, sigma_max, s->d_mu + (size_t)b0 * s->n_out, s->d_sig + (size_t)b0 * s->n_out,
                               s->d_fb + b0, s->d_nfb);
            DNA_HIP(hipGetLastError());
        }
    DNA_HIP(hipEventRecord(c->ev[2], st));
    DNA_HIP(hipMemcpyAsync(fallback, s->d_fb, n, hipMemcpyDeviceToHost, st));
    uint32_t nf = 0;
    DNA_HIP(hipMemcpyAsync(&nf, s->d_nfb, 4, hipMemcpyDeviceToHost, st));
    if (mu) DNA_HIP(hipMemcpyAsync(mu, s->d_mu, (size_t)n * s->n_out * 4, hipMemcpyDeviceToHost, st));
    if (sigma) DNA_HIP(hipMemcpyAsync(sigma, s->d_sig, (size_t)n * s->n_out * 4, hipMemcpyDeviceToHost, st));
    DNA_HIP(hipEventRecord(c->ev[3], st));
    DNA_HIP(hipStreamSynchronize(st));
    if (n_fallback) *n_fallback = nf;
    if (t_ms) {
        float a = 0, b = 0, w = 0;
        (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&w, c->ev[0], c->ev[3]);
        t_ms[0] = a; t_ms[1] = b; t_ms[2] = w;
    }
    return DNA_OK;
}
