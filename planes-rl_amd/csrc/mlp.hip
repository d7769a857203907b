// mlp.hip -- surrogate fitness model: MLP 35 -> 256 -> 256 -> 256 -> 4 (ReLU hidden layers)
// trained on the (genome, fitness) pairs the flight simulation produces.
//
// Arithmetic: BF16 operands on the matrix cores (v_mfma_f32_32x32x16_bf16, gfx950), FP32
// accumulation, FP32 master weights and Adam state; activations stored as BF16 between
// layers.  Heads: 0 mean distance, 1 CVaR of the worst quarter, 2 worst distance (all in
// units of 10 m, regression, masked for unprintable designs), 3 oversize logit (BCE).
//
// Two GEMM kernels, both with K-contiguous operands loaded 16 bytes per lane:
//   gemm_nt     C = A * Bt^T for the forward pass (A = activations, Bt = the layer's
//               transposed weight copy; epilogue bias + ReLU -> BF16, or bias -> FP32) and
//               the input gradients (A = output gradient, Bt = the weights; epilogue the ReLU
//               mask of the layer input -> BF16); training also writes each BF16 result
//               transposed, the layout the weight gradients read;
//   gemm_wgrad  dW = Xin^T dY over the batch from those transposes (features x batch), eight
//               waves per 32 x 32 tile on batch slices summed in wave order through LDS, the
//               bias gradient (column sums of dY) from an all-ones MFMA operand alongside:
//               training is deterministic.
// A training step is 14 launches (embedding, 4 forward, loss, 4 weight + bias gradients,
// 3 input gradients, Adam); prediction runs 65536 rows per pass on buffers of its own, its
// first layer embedding the genomes on the way into LDS (gemm_lds).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/planes_hip.h"

#include <algorithm>

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 35, K0 = 64, HID = 256, NOUT = 4, NOP = 32;  // padded input / output widths (x32)
constexpr int NL = 4;
const int LK[NL] = {K0, HID, HID, HID};
const int LN[NL] = {HID, HID, HID, NOP};

__device__ __forceinline__ __bf16 to_bf(float x) { return (__bf16)x; }  // round to nearest even
__device__ __forceinline__ float from_bf(__bf16 x) { return (float)x; }

enum Epi { EPI_BIAS_RELU_BF16 = 0, EPI_BIAS_F32 = 1, EPI_MASK_BF16 = 2 };

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

// Epilogue of a wave's (32 WM) x (32 WN) accumulator tiles at (m0, n0): bias + ReLU -> BF16,
// bias -> FP32, or the ReLU mask -> BF16; BF16 results also transposed into Ct when given,
// four rows at a time as one 8-byte store.
template <int EPI, int WM, int WN>
__device__ __forceinline__ void store_tiles(const f16v (&acc)[WM][WN], int m0, int n0, int l,
                                            const float* __restrict__ bias, const __bf16* __restrict__ mask, int ldm,
                                            void* __restrict__ C, int ldc, __bf16* __restrict__ Ct, int ldct) {
    const int lr = l & 31;
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int mg = m0 + 32 * i + 8 * g + 4 * (l >> 5), n = n0 + 32 * j + lr;
                __bf16 o[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int m = mg + q;
                    float v = acc[i][j][4 * g + q];
                    if (EPI == EPI_BIAS_RELU_BF16) {
                        o[q] = to_bf(fmaxf(v + bias[n], 0.0f));
                        reinterpret_cast<__bf16*>(C)[(size_t)m * ldc + n] = o[q];
                    } else if (EPI == EPI_BIAS_F32) {
                        reinterpret_cast<float*>(C)[(size_t)m * ldc + n] = v + bias[n];
                    } else if (EPI == EPI_MASK_BF16) {
                        const float keep = from_bf(mask[(size_t)m * ldm + n]) > 0.0f ? 1.0f : 0.0f;
                        o[q] = to_bf(v * keep);
                        reinterpret_cast<__bf16*>(C)[(size_t)m * ldc + n] = o[q];
                    }
                }
                if ((EPI == EPI_BIAS_RELU_BF16 || EPI == EPI_MASK_BF16) && Ct) {
                    bf4 t = {o[0], o[1], o[2], o[3]};
                    *reinterpret_cast<bf4*>(Ct + (size_t)n * ldct + mg) = t;
                }
            }
}

// C[M x N] = A[M x K] * B[K x N] with both operands K-contiguous: A row-major (lda), and B
// given as its transpose Bt[N x K] (row-major, ldb): B(k, n) = Bt[n * ldb + k].  The forward
// pass (B = the layer's transposed weight copy) and the input gradient (dY * W^T: Bt = W
// itself) have this form.  A workgroup of WGM x WGN waves computes a (32 WM WGM) x
// (32 WN WGN) tile; each wave a (32 WM) x (32 WN) tile of v_mfma_f32_32x32x16_bf16 products,
// its operands one 16-byte load per lane and MFMA tile (lane l: row l % 32, k = k0 + 8 (l / 32)
// .. + 7).  M, N multiples of the workgroup tile, K a multiple of 16.  Accumulator element r
// of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32 of its 32 x 32 tile.
// BF16 epilogues optionally also write the transpose, Ct[n * ldct + m] (the operand layout
// the weight gradients read), four rows at a time as one 8-byte store.
template <int EPI, int WGM, int WGN, int WM, int WN>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_nt(const __bf16* __restrict__ A, int lda,
                                                          const __bf16* __restrict__ Bt, int ldb, int K,
                                                          const float* __restrict__ bias,
                                                          const __bf16* __restrict__ mask, int ldm, void* __restrict__ C,
                                                          int ldc, __bf16* __restrict__ Ct, int ldct) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * (32 * WM * WGM) + (w / WGN) * (32 * WM);
    const int n0 = blockIdx.y * (32 * WN * WGN) + (w % WGN) * (32 * WN);
    const int lr = l & 31, kq = 8 * (l >> 5);
    f16v acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
    const __bf16* pa = A + (size_t)(m0 + lr) * lda + kq;
    const __bf16* pb = Bt + (size_t)(n0 + lr) * ldb + kq;
    // operands of the next D k-steps in flight (a ring of registers, refilled D steps ahead):
    // the loop issues loads D steps before their MFMAs instead of waiting on each
    constexpr int D = (WM + WN) <= 2 ? 8 : 4;
    const int ns = K / 16;
    bf8 ra[D][WM], rb[D][WN];
#pragma unroll
    for (int d = 0; d < D; d++)
        if (d < ns) {
#pragma unroll
            for (int i = 0; i < WM; i++) ra[d][i] = *reinterpret_cast<const bf8*>(pa + (size_t)32 * i * lda + 16 * d);
#pragma unroll
            for (int j = 0; j < WN; j++) rb[d][j] = *reinterpret_cast<const bf8*>(pb + (size_t)32 * j * ldb + 16 * d);
        }
    for (int s0 = 0; s0 < ns; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            const int st = s0 + d;
            if (st < ns) {
                bf8 a[WM], b[WN];
#pragma unroll
                for (int i = 0; i < WM; i++) a[i] = ra[d][i];
#pragma unroll
                for (int j = 0; j < WN; j++) b[j] = rb[d][j];
                if (st + D < ns) {
                    const int k1 = 16 * (st + D);
#pragma unroll
                    for (int i = 0; i < WM; i++) ra[d][i] = *reinterpret_cast<const bf8*>(pa + (size_t)32 * i * lda + k1);
#pragma unroll
                    for (int j = 0; j < WN; j++) rb[d][j] = *reinterpret_cast<const bf8*>(pb + (size_t)32 * j * ldb + k1);
                }
#pragma unroll
                for (int i = 0; i < WM; i++)
#pragma unroll
                    for (int j = 0; j < WN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
    }
    store_tiles<EPI, WM, WN>(acc, m0, n0, l, bias, mask, ldm, C, ldc, Ct, ldct);
}

// gemm_nt with the operands staged through LDS: a workgroup of WGM x WGN waves computes a
// (32 WM WGM) x (32 WN WGN) tile; each K chunk of 64 of its A rows and Bt rows is loaded with
// 16-byte vector loads covering whole 128-byte rows (each operand element read from global
// memory once per workgroup instead of once per wave), written to one of two LDS buffers (rows
// padded by 16 bytes), and the waves' MFMA operands are 16-byte LDS reads.  The next chunk's
// global loads are in flight while the current one is multiplied; one barrier per chunk.
// M, N multiples of the workgroup tile, K a multiple of 64.
template <int EPI, int WGM, int WGN, int WM, int WN>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_lds(const __bf16* __restrict__ A, int lda,
                                                           const __bf16* __restrict__ Bt, int ldb, int K,
                                                           const float* __restrict__ bias,
                                                           const __bf16* __restrict__ mask, int ldm,
                                                           void* __restrict__ C, int ldc, __bf16* __restrict__ Ct,
                                                           int ldct) {
    constexpr int NT = 64 * WGM * WGN;
    constexpr int BM = 32 * WM * WGM, BN = 32 * WN * WGN, BK = 64, PITCH = BK + 8;
    constexpr int VA = BM * 8 / NT, VB = BN * 8 / NT;  // staging vectors per thread and chunk
    static_assert(VA >= 1 && VB >= 1 && VA * NT == BM * 8 && VB * NT == BN * 8, "staging tiles");
    __shared__ __attribute__((aligned(16))) __bf16 sA[2][BM * PITCH];
    __shared__ __attribute__((aligned(16))) __bf16 sB[2][BN * PITCH];
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int bm0 = blockIdx.x * BM, bn0 = blockIdx.y * BN;
    const int wm0 = (w / WGN) * (32 * WM), wn0 = (w % WGN) * (32 * WN);
    const int lr = l & 31, kq = 8 * (l >> 5);
    // staging: thread t moves vectors t + NT v (8 per 128-byte row: row = idx / 8, column 8 (idx % 8))
    bf8 na[VA], nb[VB];
    auto load = [&](int k0) {
#pragma unroll
        for (int v = 0; v < VA; v++) {
            const int idx = t + NT * v, r = idx >> 3, c = 8 * (idx & 7);
            na[v] = *reinterpret_cast<const bf8*>(A + (size_t)(bm0 + r) * lda + k0 + c);
        }
#pragma unroll
        for (int v = 0; v < VB; v++) {
            const int idx = t + NT * v, r = idx >> 3, c = 8 * (idx & 7);
            nb[v] = *reinterpret_cast<const bf8*>(Bt + (size_t)(bn0 + r) * ldb + k0 + c);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int v = 0; v < VA; v++) {
            const int idx = t + NT * v, r = idx >> 3, c = 8 * (idx & 7);
            *reinterpret_cast<bf8*>(&sA[buf][r * PITCH + c]) = na[v];
        }
#pragma unroll
        for (int v = 0; v < VB; v++) {
            const int idx = t + NT * v, r = idx >> 3, c = 8 * (idx & 7);
            *reinterpret_cast<bf8*>(&sB[buf][r * PITCH + c]) = nb[v];
        }
    };
    f16v acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
    const int nch = K / BK;
    load(0);
    store(0);
    for (int c = 0; c < nch; c++) {
        __syncthreads();
        if (c + 1 < nch) load((c + 1) * BK);
        const __bf16* a_s = sA[c & 1];
        const __bf16* b_s = sB[c & 1];
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            bf8 a[WM], b[WN];
#pragma unroll
            for (int i = 0; i < WM; i++)
                a[i] = *reinterpret_cast<const bf8*>(&a_s[(wm0 + 32 * i + lr) * PITCH + ks + kq]);
#pragma unroll
            for (int j = 0; j < WN; j++)
                b[j] = *reinterpret_cast<const bf8*>(&b_s[(wn0 + 32 * j + lr) * PITCH + ks + kq]);
#pragma unroll
            for (int i = 0; i < WM; i++)
#pragma unroll
                for (int j = 0; j < WN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (c + 1 < nch) store((c + 1) & 1);
    }
    store_tiles<EPI, WM, WN>(acc, bm0 + wm0, bn0 + wn0, l, bias, mask, ldm, C, ldc, Ct, ldct);
}

// k-loop of one wave of gemm_wgrad over its batch slice (ns k-steps), operands D steps ahead
template <bool DOB>
__device__ __forceinline__ void wgrad_slice(const __bf16* pa, const __bf16* pb, int ns, f16v& acc, f16v& accb) {
    constexpr int D = 8;
    bf8 ones;
#pragma unroll
    for (int j = 0; j < 8; j++) ones[j] = to_bf(1.0f);
    bf8 ra[D], rb[D];
#pragma unroll
    for (int d = 0; d < D; d++)
        if (d < ns) {
            ra[d] = *reinterpret_cast<const bf8*>(pa + 16 * d);
            rb[d] = *reinterpret_cast<const bf8*>(pb + 16 * d);
        }
    for (int s0 = 0; s0 < ns; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            const int st = s0 + d;
            if (st < ns) {
                const bf8 a = ra[d], b = rb[d];
                if (st + D < ns) {
                    ra[d] = *reinterpret_cast<const bf8*>(pa + 16 * (st + D));
                    rb[d] = *reinterpret_cast<const bf8*>(pb + 16 * (st + D));
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
                if (DOB) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ones, b, accb, 0, 0, 0);
            }
        }
    }
}

// Weight (and bias) gradients: dW[M x N] (FP32) = A[M x Kd] * Bt[N x Kd]^T over the batch
// Kd, with A = the layer input transposed (features x batch) and Bt = the output gradient
// transposed -- both batch-contiguous, so the operands are 16-byte loads as in gemm_nt.  The
// batch is split gridDim.z ways (so that the few 32 x 32 tiles of a layer fill the GPU); a
// workgroup of 8 waves computes one tile of one split: wave w sums its slice, the eight
// partial tiles are added in wave order through LDS, and split z writes C + z split_stride
// (the Adam step adds the splits in order: deterministic, no atomics).  With BIAS, the
// workgroups of row tile 0 also form db[n] = sum over the batch of dY[b][n]: an MFMA with an
// all-ones A operand on the same Bt operands gives the column sums of the slice, added in
// wave order likewise.  Kd a multiple of 128 gridDim.z.
template <bool BIAS>
__global__ __launch_bounds__(512) void gemm_wgrad(const __bf16* __restrict__ A, const __bf16* __restrict__ Bt, int ld,
                                                  int Kd, float* __restrict__ C, int ldc, float* __restrict__ db,
                                                  size_t split_stride) {
    __shared__ float part[8][16][64];
    __shared__ float bpart[8][32];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    // batch split blockIdx.z of gridDim.z, wave w its eighth; each split writes its own copy
    const int ks = Kd / (8 * gridDim.z), kb = (blockIdx.z * 8 + w) * ks;
    C += blockIdx.z * split_stride;
    db += blockIdx.z * split_stride;
    const int lr = l & 31, kq = 8 * (l >> 5);
    const __bf16* pa = A + (size_t)(m0 + lr) * ld + kb + kq;
    const __bf16* pb = Bt + (size_t)(n0 + lr) * ld + kb + kq;
    const bool dob = BIAS && blockIdx.x == 0;
    f16v acc, accb;
#pragma unroll
    for (int r = 0; r < 16; r++) { acc[r] = 0.0f; accb[r] = 0.0f; }
    if (dob) wgrad_slice<true>(pa, pb, ks / 16, acc, accb);
    else wgrad_slice<false>(pa, pb, ks / 16, acc, accb);
#pragma unroll
    for (int r = 0; r < 16; r++) part[w][r][l] = acc[r];
    if (dob && l < 32) bpart[w][l] = accb[0];  // row 0 of ones * B: the column sums
    __syncthreads();
#pragma unroll
    for (int rr = 2 * w; rr < 2 * w + 2; rr++) {
        float v = part[0][rr][l];
#pragma unroll
        for (int u = 1; u < 8; u++) v += part[u][rr][l];
        const int m = m0 + 8 * (rr >> 2) + 4 * (l >> 5) + (rr & 3), n = n0 + lr;
        C[(size_t)m * ldc + n] = v;
    }
    if (dob && w == 0 && l < 32) {
        float v = bpart[0][l];
#pragma unroll
        for (int u = 1; u < 8; u++) v += bpart[u][l];
        db[n0 + l] = v;
    }
}

// rows of a batch are padded to this multiple (the gemm_nt workgroup tiles' row count)
constexpr int ROW_PAD = 128;

// genomes [n x 35] f32 -> X [n_pad x 64] bf16, genes scaled to [0, 1] by the bounds
// (Xt, optional: the transpose [64 x n_pad], the weight gradient's operand layout)
__global__ void k_embed(const float* __restrict__ g, int n, int n_pad, const float* __restrict__ lo,
                        const float* __restrict__ inv_range, __bf16* __restrict__ X, __bf16* __restrict__ Xt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad * K0) return;
    int r = i / K0, k = i - r * K0;
    float v = 0.0f;
    if (r < n && k < NIN) v = (g[(size_t)r * NIN + k] - lo[k]) * inv_range[k];
    X[i] = to_bf(v);
    if (Xt) Xt[(size_t)k * n_pad + r] = to_bf(v);
}

// loss and output gradient.  out [n_pad x 16] f32, targets [n x 4] raw (m, m, m, 0/1)
__global__ void k_loss(const float* __restrict__ out, const float* __restrict__ tgt, int n, int n_pad, float inv_n,
                       __bf16* __restrict__ dout, __bf16* __restrict__ doutT, float* __restrict__ loss_part) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;  // every lane stays for the wave reduction
    float l = 0.0f;
    float d[NOUT] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (r < n) {
        const float* o = out + (size_t)r * NOP;
        const float* t = tgt + (size_t)r * NOUT;
        const float over = t[3];
        if (over < 0.5f) {
            for (int k = 0; k < 3; k++) {
                const float e = o[k] - t[k] * 0.1f;
                l += e * e;
                d[k] = 2.0f * e * inv_n;
            }
        }
        const float z = o[3];
        const float p = 1.0f / (1.0f + expf(-z));
        l += fmaxf(z, 0.0f) - z * over + log1pf(expf(-fabsf(z)));
        d[3] = (p - over) * inv_n;
    }
    if (r < n_pad) {
        for (int k = 0; k < NOP; k++) {
            const float dk = k < NOUT ? d[k] : 0.0f;
            dout[(size_t)r * NOP + k] = to_bf(dk);
            doutT[(size_t)k * n_pad + r] = to_bf(dk);
        }
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (threadIdx.x == 0) loss_part[blockIdx.x] = l;  // one 64-row block per partial
}

// mean loss = the partials summed in a fixed order (lane-strided, then a butterfly), times 1/n
__global__ void k_loss_fin(const float* __restrict__ loss_part, int nblk, float inv_n, float* __restrict__ loss) {
    float l = 0.0f;
    for (int b = threadIdx.x; b < nblk; b += 64) l += loss_part[b];
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (threadIdx.x == 0) *loss = l * inv_n;
}

// out[i] = the nsplit copies of part (count apart) added in split order
__global__ void k_sum_splits(const float* __restrict__ part, int nsplit, int count, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = part[i];
    for (int z = 1; z < nsplit; z++) s += part[(size_t)z * count + i];
    out[i] = s;
}

// Adam step on the FP32 master parameters, writing their BF16 copy too. Padding parameters
// (W1 rows >= 35, W4 / b4 columns >= 4) take no step, so they stay zero.
// the parameter block's geometry: weights of layer L at off_w[L] ([k x n] row-major), and
// the transposed BF16 copy's element for parameter i (-1 for biases)
struct ParamLayout {
    int off_w[4], k[4], n[4];
};
__device__ __forceinline__ int transposed_index(const ParamLayout& P, int i) {
#pragma unroll
    for (int L = 0; L < 4; L++) {
        const int t = i - P.off_w[L];
        if (t >= 0 && t < P.k[L] * P.n[L]) {
            const int kk = t / P.n[L], nn = t - kk * P.n[L];
            return P.off_w[L] + nn * P.k[L] + kk;
        }
    }
    return -1;
}

// g: nsplit gradient copies n apart, added in split order
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, int nsplit, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ wb, __bf16* __restrict__ wbt, ParamLayout P, int n,
                       int off_w1, int off_w4, int off_b4, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool pad = (i >= off_w1 + NIN * HID && i < off_w1 + K0 * HID) ||
                     (i >= off_w4 && i < off_w4 + HID * NOP && (i - off_w4) % NOP >= NOUT) ||
                     (i >= off_b4 && i < off_b4 + NOP && i - off_b4 >= NOUT);
    if (pad) return;
    float gi = g[i];
    for (int z = 1; z < nsplit; z++) gi += g[(size_t)z * n + i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    wb[i] = to_bf(wi);
    const int it = transposed_index(P, i);
    if (it >= 0) wbt[it] = to_bf(wi);
}

// BF16 copies of the weights: row-major and transposed
__global__ void k_to_bf16(const float* __restrict__ x, int n, __bf16* __restrict__ y, __bf16* __restrict__ yt,
                          ParamLayout P) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = to_bf(x[i]);
    const int it = transposed_index(P, i);
    if (it >= 0) yt[it] = to_bf(x[i]);
}

// the five prediction values of one row from its four head outputs: mean, CVaR and worst
// distance (metres), oversize probability, and the robust score from the three regression
// heads (-100 when the oversize head says so)
__device__ __forceinline__ void head_values(const float (&o)[NOUT], float* __restrict__ pr) {
    const float s = 10.0f * (0.65f * o[0] + 0.25f * o[1] + 0.10f * o[2]);
    pr[0] = 10.0f * o[0];
    pr[1] = 10.0f * o[1];
    pr[2] = 10.0f * o[2];
    pr[3] = 1.0f / (1.0f + expf(-o[3]));
    pr[4] = o[3] > 0.0f ? -100.0f : s;
}

// ---- prediction in one pass: k_mlp_predict ----
// A workgroup of four waves takes 128 rows through the whole network: the genomes are
// embedded straight into LDS, each layer reads its input activations from an LDS buffer and,
// once every wave has read them, overwrites them with its output, and the heads are evaluated
// in registers, so the only HBM traffic is the genomes in and the five values per row out.
// One activation buffer (66 KB) lets two workgroups share a CU, each one's loads, epilogues
// and barriers overlapping the other's products.  Each layer is computed transposed,
// C^T = W^T H^T: the MFMA's A operand is a 32-feature tile of the layer's transposed weight
// copy (16-byte loads from L2, where every workgroup finds the 311 KB of weights), its B
// operand a 32-row tile of activations (16-byte LDS reads), and accumulator element r of lane l
// is then feature 8 (r / 4) + 4 (l / 32) + r % 4 of row l % 32 -- four consecutive features of
// one row, one 8-byte LDS write in the row-major layout the next layer reads.  Wave w computes
// features 64 w .. 64 w + 63 of all 128 rows (2 x 4 tiles); the output layer (32 features,
// four of them live) has one tile, and wave w takes rows 32 w .. 32 w + 31.  The weight stream
// runs through a ring of PD k-steps in registers that, near the end of a layer, is refilled
// from the next layer's weights.
constexpr int PB_ROWS = 128;           // rows per workgroup
constexpr int PB_XP = K0 + 8;          // LDS row pitch of the embedded inputs (bf16; 16-byte multiple)
constexpr int PB_HP = HID + 8;         // LDS row pitch of the hidden activations
constexpr int PD = 4;                  // weight k-steps in flight
// LDS copy of the biases and gene bounds: b1, b2, b3 [256 each], b4 [4], lo [35], 1/range [35]
constexpr int PB_B4 = 3 * HID, PB_LO = PB_B4 + NOUT, PB_IR = PB_LO + NIN, PB_CONST = PB_IR + NIN;

// one wave's k-loop of a layer: acc[i][j] += (W^T tile i) x (activation tile j)^T over KIN;
// ring slot s % PD holds k-step s of this layer on entry and k-step s of the next layer on exit
template <int KIN, int NA, int NB, int NA_NEXT>
__device__ __forceinline__ void fused_layer_k(f16v (&acc)[NA][NB], bf8 (&ring)[PD][2], const __bf16* __restrict__ pw,
                                              int ldw, const __bf16* __restrict__ pw_next, int ldw_next,
                                              const __bf16* __restrict__ act, int pitch, int j0, int lr, int kq) {
    constexpr int NS = KIN / 16;
    static_assert(NS % PD == 0, "k-steps a multiple of the ring depth");
#pragma unroll
    for (int i = 0; i < NA; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int d = s % PD;
        bf8 a[NA];
#pragma unroll
        for (int i = 0; i < NA; i++) a[i] = ring[d][i];
        if (s + PD < NS) {
#pragma unroll
            for (int i = 0; i < NA; i++)
                ring[d][i] = *reinterpret_cast<const bf8*>(pw + (size_t)32 * i * ldw + 16 * (s + PD));
        } else if (NA_NEXT > 0) {
#pragma unroll
            for (int i = 0; i < NA_NEXT; i++)
                ring[d][i] = *reinterpret_cast<const bf8*>(pw_next + (size_t)32 * i * ldw_next + 16 * (s + PD - NS));
        }
        bf8 b[NB];
#pragma unroll
        for (int j = 0; j < NB; j++)
            b[j] = *reinterpret_cast<const bf8*>(act + (32 * (j0 + j) + lr) * pitch + 16 * s + kq);
#pragma unroll
        for (int i = 0; i < NA; i++)
#pragma unroll
            for (int j = 0; j < NB; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

// bias + ReLU of a hidden layer's transposed tiles -> BF16 rows of the LDS activation buffer
// (features n0 .. n0 + 63 of rows 0 .. 127); bias from the LDS copy
__device__ __forceinline__ void fused_relu_to_lds(const f16v (&acc)[2][4], const float* __restrict__ bias, int n0,
                                                  __bf16* __restrict__ out, int l) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int n = n0 + 32 * i + 8 * g + 4 * (l >> 5);
            const f4 bv = *reinterpret_cast<const f4*>(bias + n);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                bf4 o;
#pragma unroll
                for (int q = 0; q < 4; q++) o[q] = to_bf(fmaxf(acc[i][j][4 * g + q] + bv[q], 0.0f));
                *reinterpret_cast<bf4*>(out + (32 * j + (l & 31)) * PB_HP + n) = o;
            }
        }
}

__global__ __launch_bounds__(256, 2) void k_mlp_predict(const float* __restrict__ genomes, int n,
                                                        const __bf16* __restrict__ wbt, ParamLayout P,
                                                        const float* __restrict__ w, const float* __restrict__ lo,
                                                        const float* __restrict__ inv_range,
                                                        float* __restrict__ pred) {
    __shared__ __attribute__((aligned(16))) __bf16 h[PB_ROWS * PB_HP];  // activations (in place)
    __shared__ __attribute__((aligned(16))) float cst[PB_CONST];
    const int t = threadIdx.x, l = t & 63, wv = t >> 6, lr = l & 31, kq = 8 * (l >> 5);
    const int m0 = blockIdx.x * PB_ROWS;
    const int n0 = 64 * wv;  // this wave's features in the hidden layers
    const __bf16* pw0 = wbt + P.off_w[0] + (size_t)(n0 + lr) * K0 + kq;
    const __bf16* pw1 = wbt + P.off_w[1] + (size_t)(n0 + lr) * HID + kq;
    const __bf16* pw2 = wbt + P.off_w[2] + (size_t)(n0 + lr) * HID + kq;
    const __bf16* pw3 = wbt + P.off_w[3] + (size_t)lr * HID + kq;
    bf8 ring[PD][2];
#pragma unroll
    for (int d = 0; d < PD; d++)
#pragma unroll
        for (int i = 0; i < 2; i++) ring[d][i] = *reinterpret_cast<const bf8*>(pw0 + (size_t)32 * i * K0 + 16 * d);

    // the block's genomes: 128 x 35 consecutive floats, coalesced, all loads issued at once
    constexpr int NE = (PB_ROWS * NIN + 255) / 256;
    const int live = min(PB_ROWS, n - m0) * NIN;
    const float* gb = genomes + (size_t)m0 * NIN;
    float gv[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = t + 256 * i;
        gv[i] = e < live ? gb[e] : 0.0f;
    }
    for (int i = t; i < PB_CONST; i += 256) {
        float v;
        if (i < PB_B4) v = w[P.off_w[i >> 8] + P.k[i >> 8] * P.n[i >> 8] + (i & (HID - 1))];
        else if (i < PB_LO) v = w[P.off_w[3] + P.k[3] * P.n[3] + (i - PB_B4)];
        else if (i < PB_IR) v = lo[i - PB_LO];
        else v = inv_range[i - PB_IR];
        cst[i] = v;
    }
    // padding columns 35..63 of the embedded inputs are zero
    for (int e = t; e < PB_ROWS * 32; e += 256) {
        const int r = e >> 5, k = 32 + (e & 31);
        if (k >= NIN) h[r * PB_XP + k] = to_bf(0.0f);
    }
    __syncthreads();
    // embedding: genes scaled to [0, 1] by the bounds; rows beyond n stay zero
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = t + 256 * i;
        if (e < PB_ROWS * NIN) {
            const int r = e / NIN, k = e - r * NIN;
            h[r * PB_XP + k] = to_bf(e < live ? (gv[i] - cst[PB_LO + k]) * cst[PB_IR + k] : 0.0f);
        }
    }
    __syncthreads();

    f16v acc[2][4];
    fused_layer_k<K0, 2, 4, 2>(acc, ring, pw0, K0, pw1, HID, h, PB_XP, 0, lr, kq);
    __syncthreads();  // every wave has read the inputs: overwrite them
    fused_relu_to_lds(acc, cst, n0, h, l);
    __syncthreads();
    fused_layer_k<HID, 2, 4, 2>(acc, ring, pw1, HID, pw2, HID, h, PB_HP, 0, lr, kq);
    __syncthreads();
    fused_relu_to_lds(acc, cst + HID, n0, h, l);
    __syncthreads();
    fused_layer_k<HID, 2, 4, 1>(acc, ring, pw2, HID, pw3, HID, h, PB_HP, 0, lr, kq);
    __syncthreads();
    fused_relu_to_lds(acc, cst + 2 * HID, n0, h, l);
    __syncthreads();
    f16v acc3[1][1];
    fused_layer_k<HID, 1, 1, 0>(acc3, ring, pw3, HID, nullptr, 0, h, PB_HP, wv, lr, kq);
    // lanes 0..31 hold the four live outputs of row 32 wv + lane in accumulator elements 0..3
    const int m = m0 + 32 * wv + l;
    if (l < 32 && m < n) {
        float o[NOUT];
#pragma unroll
        for (int q = 0; q < NOUT; q++) o[q] = acc3[0][0][q] + cst[PB_B4 + q];
        float pr[5];
        head_values(o, pr);
#pragma unroll
        for (int q = 0; q < 5; q++) pred[(size_t)m * 5 + q] = pr[q];
    }
}


thread_local std::string g_mlp_err;

}  // namespace

// parameter block layout (FP32): W1 [64x256] b1 [256] W2 [256x256] b2 W3 [256x256] b3 W4 [256x32] b4 [32]
// (rows 35..63 of W1 and columns 4..31 of W4 are padding and stay zero)
struct pr_mlp {
    int max_batch = 0;
    size_t off_w[NL], off_b[NL], n_params = 0;
    float* w = nullptr;   // master FP32 parameters
    float* g = nullptr;   // gradients (the splits summed; for pr_mlp_loss_grad)
    float* gpart = nullptr;  // [WG_SPLITS x n_params] gradient copies of the batch splits
    int nsplit = 1;          // splits the last backward pass wrote
    float* m = nullptr;   // Adam first moment
    float* v = nullptr;   // Adam second moment
    __bf16* wb = nullptr; // BF16 copy of the weights (biases stay FP32)
    __bf16* wbt = nullptr;  // BF16 copy of each layer's weights transposed ([n x k]), for gemm_nt
    ParamLayout layout = {};
    __bf16* X = nullptr;  // [B x 64]
    __bf16* H[3] = {};    // [B x 256] post-ReLU activations
    __bf16* dH[3] = {};   // [B x 256] gradients
    // transposes (features x B) of the layer inputs and output gradients: the operands of the
    // weight gradients
    __bf16* Xt = nullptr;
    __bf16* Ht[3] = {};
    __bf16* dHt[3] = {};
    __bf16* doutT = nullptr;  // [32 x B]
    float* out = nullptr; // [B x 32]
    __bf16* dout = nullptr;
    float* loss = nullptr;
    float* loss_part = nullptr;  // [max_batch / 64] per-block loss sums
    float* lo = nullptr;
    float* inv_range = nullptr;
    int step = 0;
    hipStream_t s = nullptr;
};

#define MCK(x)                                                                   \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) { g_mlp_err = std::string(#x) + ": " + hipGetErrorString(e_); return (int)e_; } \
    } while (0)

static int pad_rows(int n) { return (n + ROW_PAD - 1) / ROW_PAD * ROW_PAD; }
constexpr int WG_SPLITS = 4;  // batch splits of the weight gradients (their copies summed by the Adam step)

// deterministic He-uniform initialisation from a 32-bit hash of (seed, index)
static float init_u(unsigned seed, size_t i) {
    unsigned x = seed * 0x9e3779b9u + (unsigned)i * 0x85ebca6bu + (unsigned)(i >> 32);
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return (float)(x >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
}

// C[M x N] = A[M x K] * (Bt[N x K])^T through gemm_nt, with the workgroup tile chosen for the
// size: 128 x 128 for large batches (M >= 32768), 64 x 64 otherwise (a 4096-row batch then
// spreads over 256 workgroups), 64 x 32 for the 32-wide output layer.
template <int EPI>
static void launch_nt(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* Bt, int ldb,
                      const float* bias, const __bf16* mask, int ldm, void* C, int ldc, __bf16* Ct, int ldct) {
    if (N % 128 == 0 && M >= 32768 && M % 128 == 0 && K % 64 == 0)
        hipLaunchKernelGGL((gemm_lds<EPI, 2, 2, 2, 2>), dim3(M / 128, N / 128), dim3(256), 0, st, A, lda, Bt, ldb, K,
                           bias, mask, ldm, C, ldc, Ct, ldct);
    else if (N % 64 == 0 && M % 64 == 0 && K % 64 == 0)
        hipLaunchKernelGGL((gemm_lds<EPI, 2, 2, 1, 1>), dim3(M / 64, N / 64), dim3(256), 0, st, A, lda, Bt, ldb, K,
                           bias, mask, ldm, C, ldc, Ct, ldct);
    else if (N % 64 == 0)
        hipLaunchKernelGGL((gemm_nt<EPI, 2, 2, 1, 1>), dim3(M / 64, N / 64), dim3(256), 0, st, A, lda, Bt, ldb, K, bias,
                           mask, ldm, C, ldc, Ct, ldct);
    else
        hipLaunchKernelGGL((gemm_nt<EPI, 2, 1, 1, 1>), dim3(M / 64, N / 32), dim3(128), 0, st, A, lda, Bt, ldb, K, bias,
                           mask, ldm, C, ldc, Ct, ldct);
}

// One layer's forward GEMM: [B x K] x W[K x N] (its transposed copy) + bias (ReLU, BF16 for the
// hidden layers, FP32 for the output layer); Ht: the activations' transpose as well (training)
static void launch_layer(pr_mlp* q, int L, const __bf16* in, int B, void* out, __bf16* Ht) {
    const __bf16* Wt = q->wbt + q->off_w[L];
    const float* b = q->w + q->off_b[L];
    if (L < NL - 1)
        launch_nt<EPI_BIAS_RELU_BF16>(q->s, B, LN[L], LK[L], in, LK[L], Wt, LK[L], b, nullptr, 0, out, LN[L], Ht, B);
    else
        launch_nt<EPI_BIAS_F32>(q->s, B, LN[L], LK[L], in, LK[L], Wt, LK[L], b, nullptr, 0, out, NOP, nullptr, 0);
}

static int forward(pr_mlp* q, int B) {
    // layer 1..3: H = ReLU(X W + b) (BF16, and its transpose); layer 4: out = H3 W4 + b4 (FP32)
    const __bf16* in = q->X;
    for (int L = 0; L < NL; L++) {
        launch_layer(q, L, in, B, L < NL - 1 ? (void*)q->H[L] : (void*)q->out, L < NL - 1 ? q->Ht[L] : nullptr);
        MCK(hipGetLastError());
        if (L < NL - 1) in = q->H[L];
    }
    return 0;
}

// forward + loss + backward into q->gpart (FP32 gradients of the mean loss over n samples, one
// copy per batch split of the weight gradients: q->nsplit);
// want_loss: also reduce the mean loss into q->loss (the caller reads it)
static int loss_grad(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, bool want_loss) {
    const int B = pad_rows(n);
    hipLaunchKernelGGL(k_embed, dim3((B * K0 + 255) / 256), dim3(256), 0, q->s, d_genomes, n, B, q->lo, q->inv_range,
                       q->X, q->Xt);
    MCK(hipGetLastError());
    int rc = forward(q, B);
    if (rc) return rc;
    hipLaunchKernelGGL(k_loss, dim3((B + 63) / 64), dim3(64), 0, q->s, q->out, d_targets, n, B, 1.0f / n, q->dout,
                       q->doutT, q->loss_part);
    MCK(hipGetLastError());
    if (want_loss) {
        hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, q->s, q->loss_part, (B + 63) / 64, 1.0f / n, q->loss);
        MCK(hipGetLastError());
    }
    // backward, layer 4 .. 1; every gradient entry is written (weights and biases by gemm_wgrad)
    const __bf16* dY = q->dout;
    const __bf16* dYt = q->doutT;
    int ldy = NOP;
    for (int L = NL - 1; L >= 0; L--) {
        const __bf16* Xt = L == 0 ? q->Xt : q->Ht[L - 1];
        const int K = LK[L], N = LN[L];
        // dW [K x N] = Xin^T dY and db = column sums of dY, over the (zero-padded) batch
        const int z = B % (128 * WG_SPLITS) == 0 ? WG_SPLITS : 1;
        hipLaunchKernelGGL(gemm_wgrad<true>, dim3(K / 32, N / 32, z), dim3(512), 0, q->s, Xt, dYt, B, B,
                           q->gpart + q->off_w[L], N, q->gpart + q->off_b[L], q->n_params);
        q->nsplit = z;
        MCK(hipGetLastError());
        if (L > 0) {
            // dXin [B x K] = dY [B x N] W^T, masked by ReLU'(Xin) (Xin > 0); W [K x N] row-major is
            // the transposed operand gemm_nt takes; the transpose feeds the next weight gradient
            const __bf16* W = q->wb + q->off_w[L];
            launch_nt<EPI_MASK_BF16>(q->s, B, K, N, dY, ldy, W, N, nullptr, q->H[L - 1], K, (void*)q->dH[L - 1], K,
                                     q->dHt[L - 1], B);
            MCK(hipGetLastError());
            dY = q->dH[L - 1];
            dYt = q->dHt[L - 1];
            ldy = K;
        }
    }
    return 0;
}

static int refresh_bf16(pr_mlp* q) {
    const int n = (int)q->n_params;
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, q->s, q->w, n, q->wb, q->wbt, q->layout);
    MCK(hipGetLastError());
    return 0;
}

extern "C" {

const char* pr_mlp_last_error(void) { return g_mlp_err.c_str(); }

int pr_mlp_param_count(void) {
    int n = 0;
    for (int L = 0; L < NL; L++) n += LK[L] * LN[L] + LN[L];
    return n;
}

int pr_mlp_create(pr_mlp** out, int max_batch, unsigned seed) {
    if (!out || max_batch < 16) { g_mlp_err = "pr_mlp_create: max_batch >= 16"; return -1; }
    pr_mlp* q = new pr_mlp;
    q->max_batch = (max_batch + 31) / 32 * 32;
    size_t o = 0;
    for (int L = 0; L < NL; L++) {
        q->off_w[L] = o; o += (size_t)LK[L] * LN[L];
        q->off_b[L] = o; o += LN[L];
        q->layout.off_w[L] = (int)q->off_w[L];
        q->layout.k[L] = LK[L];
        q->layout.n[L] = LN[L];
    }
    q->n_params = o;
    const int B = pad_rows(q->max_batch);
#define MALLOC(p, bytes) MCK(hipMalloc((void**)&(p), (bytes)))
    MCK(hipStreamCreateWithFlags(&q->s, hipStreamNonBlocking));
    MALLOC(q->w, 4 * o); MALLOC(q->g, 4 * o); MALLOC(q->m, 4 * o); MALLOC(q->v, 4 * o);
    MALLOC(q->gpart, 4 * o * WG_SPLITS);
    MALLOC(q->wb, 2 * o);
    MALLOC(q->wbt, 2 * o);
    MALLOC(q->X, 2 * (size_t)B * K0);
    MALLOC(q->Xt, 2 * (size_t)B * K0);
    for (int k = 0; k < 3; k++) {
        MALLOC(q->H[k], 2 * (size_t)B * HID); MALLOC(q->dH[k], 2 * (size_t)B * HID);
        MALLOC(q->Ht[k], 2 * (size_t)B * HID); MALLOC(q->dHt[k], 2 * (size_t)B * HID);
    }
    MALLOC(q->out, 4 * (size_t)B * NOP);
    MALLOC(q->dout, 2 * (size_t)B * NOP);
    MALLOC(q->doutT, 2 * (size_t)B * NOP);
    MALLOC(q->loss, 4);
    MALLOC(q->loss_part, 4 * (size_t)(B / 64 + 1));  // B is a multiple of 128
    MALLOC(q->lo, 4 * NIN);
    MALLOC(q->inv_range, 4 * NIN);
#undef MALLOC
    float lo[NIN], hi[NIN], ir[NIN];
    pr_gene_bounds(lo, hi);
    for (int k = 0; k < NIN; k++) ir[k] = 1.0f / (hi[k] - lo[k]);
    MCK(hipMemcpy(q->lo, lo, sizeof lo, hipMemcpyHostToDevice));
    MCK(hipMemcpy(q->inv_range, ir, sizeof ir, hipMemcpyHostToDevice));
    std::vector<float> w(o, 0.0f);
    for (int L = 0; L < NL; L++) {
        const int fan_in = L == 0 ? NIN : LK[L];
        const float a = std::sqrt(6.0f / fan_in);
        for (int k = 0; k < LK[L]; k++)
            for (int nn = 0; nn < LN[L]; nn++) {
                const bool live = (L > 0 || k < NIN) && (L < NL - 1 || nn < NOUT);
                w[q->off_w[L] + (size_t)k * LN[L] + nn] = live ? a * init_u(seed + 17u * L, (size_t)k * LN[L] + nn) : 0.0f;
            }
    }
    MCK(hipMemcpy(q->w, w.data(), 4 * o, hipMemcpyHostToDevice));
    MCK(hipMemset(q->m, 0, 4 * o));
    MCK(hipMemset(q->v, 0, 4 * o));
    int rc = refresh_bf16(q);
    if (rc) return rc;
    MCK(hipStreamSynchronize(q->s));
    *out = q;
    return 0;
}

void pr_mlp_destroy(pr_mlp* q) {
    if (!q) return;
    if (q->s) (void)hipStreamSynchronize(q->s);
    void* bufs[] = {q->w, q->g, q->gpart, q->m, q->v, q->wb, q->wbt, q->X, q->Xt, q->out, q->dout, q->doutT, q->loss,
                    q->loss_part, q->lo, q->inv_range, q->H[0], q->H[1], q->H[2], q->dH[0], q->dH[1], q->dH[2],
                    q->Ht[0], q->Ht[1], q->Ht[2], q->dHt[0], q->dHt[1], q->dHt[2]};
    for (void* b : bufs) if (b) (void)hipFree(b);
    if (q->s) (void)hipStreamDestroy(q->s);
    delete q;
}

int pr_mlp_max_batch(pr_mlp* q) { return q ? q->max_batch : 0; }

int pr_mlp_get_params(pr_mlp* q, float* out, int cap) {
    if (!q || cap < (int)q->n_params) { g_mlp_err = "pr_mlp_get_params: capacity"; return -1; }
    MCK(hipStreamSynchronize(q->s));
    MCK(hipMemcpy(out, q->w, 4 * q->n_params, hipMemcpyDeviceToHost));
    return 0;
}

int pr_mlp_set_params(pr_mlp* q, const float* in, int n) {
    if (!q || n != (int)q->n_params) { g_mlp_err = "pr_mlp_set_params: size"; return -1; }
    MCK(hipMemcpy(q->w, in, 4 * q->n_params, hipMemcpyHostToDevice));
    int rc = refresh_bf16(q);
    if (rc) return rc;
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

// Raw head outputs for n genomes (device [n x 35]) -> d_pred [n x 5] = mean, cvar, worst
// (metres), oversize probability, predicted robust score (-100 if predicted oversize).
int pr_mlp_predict(pr_mlp* q, const float* d_genomes, int n, float* d_pred) {
    if (!q || n < 1) { g_mlp_err = "pr_mlp_predict: n >= 1"; return -1; }
    hipLaunchKernelGGL(k_mlp_predict, dim3((n + PB_ROWS - 1) / PB_ROWS), dim3(256), 0, q->s, d_genomes, n, q->wbt,
                       q->layout, q->w, q->lo, q->inv_range, d_pred);
    MCK(hipGetLastError());
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

// Mean loss and its gradient for one batch (n <= max_batch); grads (host, optional) in the
// parameter layout of pr_mlp_get_params.
int pr_mlp_loss_grad(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, float* loss, float* grads) {
    if (!q || n < 1 || n > q->max_batch) { g_mlp_err = "pr_mlp_loss_grad: 1 <= n <= max_batch"; return -1; }
    int rc = loss_grad(q, d_genomes, d_targets, n, true);
    if (rc) return rc;
    if (loss) {
        MCK(hipMemcpyAsync(loss, q->loss, sizeof(float), hipMemcpyDeviceToHost, q->s));
        MCK(hipStreamSynchronize(q->s));
    }
    if (grads) {
        hipLaunchKernelGGL(k_sum_splits, dim3((int)((q->n_params + 255) / 256)), dim3(256), 0, q->s, q->gpart,
                           q->nsplit, (int)q->n_params, q->g);
        MCK(hipGetLastError());
        MCK(hipStreamSynchronize(q->s));
        MCK(hipMemcpy(grads, q->g, 4 * q->n_params, hipMemcpyDeviceToHost));
    }
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

// Mini-batch Adam over n training pairs (device genomes [n x 35], targets [n x 4]) for
// `epochs` passes; batches are contiguous slices (shuffle the data on the caller side).
int pr_mlp_train(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, int epochs, int batch, float lr,
                 float* last_loss) {
    if (!q || n < 1 || batch < 32 || batch > q->max_batch) { g_mlp_err = "pr_mlp_train: 32 <= batch <= max_batch"; return -1; }
    batch = batch / 32 * 32;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
        for (int s0 = 0; s0 + 32 <= n; s0 += batch) {
            const int nb = std::min(batch, n - s0) / 32 * 32;
            const bool last = e == epochs - 1 && s0 + batch + 32 > n;  // the loss the caller reads
            int rc = loss_grad(q, d_genomes + (size_t)s0 * NIN, d_targets + (size_t)s0 * NOUT, nb, last);
            if (rc) return rc;
            q->step++;
            const float c1 = 1.0f / (1.0f - std::pow(b1, (float)q->step));
            const float c2 = 1.0f / (1.0f - std::pow(b2, (float)q->step));
            hipLaunchKernelGGL(k_adam, dim3((int)((q->n_params + 255) / 256)), dim3(256), 0, q->s, q->w, q->gpart,
                               q->nsplit, q->m,
                               q->v, q->wb, q->wbt, q->layout, (int)q->n_params, (int)q->off_w[0], (int)q->off_w[3],
                               (int)q->off_b[3], lr, b1, b2, eps, c1, c2);
            MCK(hipGetLastError());
        }
    }
    if (last_loss) {
        MCK(hipMemcpyAsync(last_loss, q->loss, sizeof(float), hipMemcpyDeviceToHost, q->s));
    }
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

}  // extern "C"
