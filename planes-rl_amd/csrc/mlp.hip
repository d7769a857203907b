// mlp.hip -- surrogate fitness model: MLP 35 -> 256 -> 256 -> 256 -> 4 (ReLU hidden layers)
// trained on the (genome, fitness) pairs the flight simulation produces.
//
// Arithmetic: BF16 operands on the matrix cores (v_mfma_f32_16x16x32_bf16, gfx950), FP32
// accumulation, FP32 master weights and Adam state; activations stored as BF16 between
// layers.  Heads: 0 mean distance, 1 CVaR of the worst quarter, 2 worst distance (all in
// units of 10 m, regression, masked for unprintable designs), 3 oversize logit (BCE).
//
// Every GEMM is one kernel, gemm_bf16<TA, TB, EPI>: a wave computes one 16x16 output tile,
// stepping K by 32 with one MFMA per step; the epilogue either applies bias + ReLU and
// stores BF16 (forward), multiplies by the ReLU mask of the layer input and stores BF16
// (input gradients), stores FP32, or stores FP32 partial products per batch slice (weight
// gradients, then summed in slice order: training is deterministic).
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

enum Epi { EPI_BIAS_RELU_BF16 = 0, EPI_BIAS_F32 = 1, EPI_MASK_BF16 = 2, EPI_PART_F32 = 3 };

// C[M x N] = A(m,k) * B(k,n) over k in [0, K) (split over blockIdx.z for EPI_PART_F32, which
// writes split z's partial product to C + z * ldc * gridDim.x * 16 for an ordered sum after).
// A(m,k) = TA ? A[k*lda + m] : A[m*lda + k];  B(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n].
// M, N multiples of 16, K a multiple of 32 (per split).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(64) void gemm_bf16(const __bf16* __restrict__ A, int lda, const __bf16* __restrict__ B,
                                                int ldb, int K, int k_per_split, const float* __restrict__ bias,
                                                const __bf16* __restrict__ mask, int ldm, void* __restrict__ C,
                                                int ldc) {
    const int l = threadIdx.x;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const int kb = blockIdx.z * k_per_split, ke = min(K, kb + k_per_split);
    const int mr = m0 + (l & 15), nc = n0 + (l & 15), kq = 8 * (l >> 4);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        bf8 a, b;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + kq + j;
            a[j] = TA ? A[(size_t)k * lda + mr] : A[(size_t)mr * lda + k];
            b[j] = TB ? B[(size_t)nc * ldb + k] : B[(size_t)k * ldb + nc];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D layout: lane l holds rows m0 + 4*(l/16) + j, column n0 + l%16
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int m = m0 + 4 * (l >> 4) + j, n = nc;
        float v = acc[j];
        if (EPI == EPI_BIAS_RELU_BF16) {
            v = fmaxf(v + bias[n], 0.0f);
            reinterpret_cast<__bf16*>(C)[(size_t)m * ldc + n] = to_bf(v);
        } else if (EPI == EPI_BIAS_F32) {
            reinterpret_cast<float*>(C)[(size_t)m * ldc + n] = v + bias[n];
        } else if (EPI == EPI_MASK_BF16) {
            const float keep = from_bf(mask[(size_t)m * ldm + n]) > 0.0f ? 1.0f : 0.0f;
            reinterpret_cast<__bf16*>(C)[(size_t)m * ldc + n] = to_bf(v * keep);
        } else {
            const size_t plane = (size_t)gridDim.x * 16 * ldc;
            reinterpret_cast<float*>(C)[blockIdx.z * plane + (size_t)m * ldc + n] = v;
        }
    }
}

// genomes [n x 35] f32 -> X [n_pad x 64] bf16, genes scaled to [0, 1] by the bounds
__global__ void k_embed(const float* __restrict__ g, int n, int n_pad, const float* __restrict__ lo,
                        const float* __restrict__ inv_range, __bf16* __restrict__ X) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad * K0) return;
    int r = i / K0, k = i - r * K0;
    float v = 0.0f;
    if (r < n && k < NIN) v = (g[(size_t)r * NIN + k] - lo[k]) * inv_range[k];
    X[i] = to_bf(v);
}

// loss and output gradient.  out [n_pad x 16] f32, targets [n x 4] raw (m, m, m, 0/1)
__global__ void k_loss(const float* __restrict__ out, const float* __restrict__ tgt, int n, int n_pad, float inv_n,
                       __bf16* __restrict__ dout, float* __restrict__ dout_f, float* __restrict__ loss_part) {
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
            if (dout_f) dout_f[(size_t)r * NOP + k] = dk;
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

// bias gradient = column sums of a bf16 [rows x cols] gradient, in two deterministic passes:
// block (x, y) sums columns 64x..64x+63 over row slice y (4 waves on interleaved rows, then
// an LDS sum in wave order) into part[y][cols]; k_colsum_fin adds the COLSUM_SLICES partials
// of each column in slice order.
constexpr int COLSUM_SLICES = 32;
constexpr int WG_SPLIT = 512;  // batch rows per weight-gradient partial product
__global__ __launch_bounds__(256) void k_colsum_part(const __bf16* __restrict__ G, int rows, int cols,
                                                     float* __restrict__ part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int w = threadIdx.x >> 6;
    const int per = (rows + COLSUM_SLICES - 1) / COLSUM_SLICES;
    const int r0 = blockIdx.y * per, r1 = min(rows, r0 + per);
    float s = 0.0f;
    if (c < cols)
        for (int r = r0 + w; r < r1; r += 4) s += from_bf(G[(size_t)r * cols + c]);
    red[w][threadIdx.x & 63] = s;
    __syncthreads();
    if (w == 0 && c < cols) {
        const int l = threadIdx.x;
        part[(size_t)blockIdx.y * cols + c] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
    }
}

__global__ void k_colsum_fin(const float* __restrict__ part, int cols, float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int y = 0; y < COLSUM_SLICES; y++) s += part[(size_t)y * cols + c];
    out[c] = s;
}

// out[i] = sum over the splits of part[z * count + i], in split order (deterministic)
__global__ void k_sum_splits(const float* __restrict__ part, int nsplit, int count, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < nsplit; z++) s += part[(size_t)z * count + i];
    out[i] = s;
}

// Adam step on the FP32 master parameters, writing their BF16 copy too. Padding parameters
// (W1 rows >= 35, W4 / b4 columns >= 4) take no step, so they stay zero.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       __bf16* __restrict__ wb, int n, int off_w1, int off_w4, int off_b4, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool pad = (i >= off_w1 + NIN * HID && i < off_w1 + K0 * HID) ||
                     (i >= off_w4 && i < off_w4 + HID * NOP && (i - off_w4) % NOP >= NOUT) ||
                     (i >= off_b4 && i < off_b4 + NOP && i - off_b4 >= NOUT);
    if (pad) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    wb[i] = to_bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, int n, __bf16* __restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf(x[i]);
}

// robust score from the three regression heads; -100 when the oversize head says so
__global__ void k_head_scores(const float* __restrict__ out, int n, float* __restrict__ pred) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* o = out + (size_t)r * NOP;
    float s = 10.0f * (0.65f * o[0] + 0.25f * o[1] + 0.10f * o[2]);
    pred[(size_t)r * 5 + 0] = 10.0f * o[0];
    pred[(size_t)r * 5 + 1] = 10.0f * o[1];
    pred[(size_t)r * 5 + 2] = 10.0f * o[2];
    pred[(size_t)r * 5 + 3] = 1.0f / (1.0f + expf(-o[3]));
    pred[(size_t)r * 5 + 4] = o[3] > 0.0f ? -100.0f : s;
}


thread_local std::string g_mlp_err;

}  // namespace

// parameter block layout (FP32): W1 [64x256] b1 [256] W2 [256x256] b2 W3 [256x256] b3 W4 [256x32] b4 [32]
// (rows 35..63 of W1 and columns 4..31 of W4 are padding and stay zero)
struct pr_mlp {
    int max_batch = 0;
    size_t off_w[NL], off_b[NL], n_params = 0;
    float* w = nullptr;   // master FP32 parameters
    float* g = nullptr;   // gradients
    float* m = nullptr;   // Adam first moment
    float* v = nullptr;   // Adam second moment
    __bf16* wb = nullptr; // BF16 copy of the weights (biases stay FP32)
    __bf16* X = nullptr;  // [B x 64]
    __bf16* H[3] = {};    // [B x 256] post-ReLU activations
    __bf16* dH[3] = {};   // [B x 256] gradients
    float* out = nullptr; // [B x 16]
    __bf16* dout = nullptr;
    float* dout_f = nullptr;
    float* loss = nullptr;
    float* loss_part = nullptr;  // [max_batch / 64] per-block loss sums
    float* lo = nullptr;
    float* inv_range = nullptr;
    float* colsum_part = nullptr;  // [COLSUM_SLICES x 256] bias-gradient partial sums
    float* wgrad_part = nullptr;   // [batch splits x 256 x 256] weight-gradient partial products
    int step = 0;
    hipStream_t s = nullptr;
};

#define MCK(x)                                                                   \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) { g_mlp_err = std::string(#x) + ": " + hipGetErrorString(e_); return (int)e_; } \
    } while (0)

static int pad32(int n) { return (n + 31) / 32 * 32; }

// deterministic He-uniform initialisation from a 32-bit hash of (seed, index)
static float init_u(unsigned seed, size_t i) {
    unsigned x = seed * 0x9e3779b9u + (unsigned)i * 0x85ebca6bu + (unsigned)(i >> 32);
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return (float)(x >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
}

static int forward(pr_mlp* q, int B, bool keep_out) {
    // layer 1..3: H = ReLU(X W + b) (BF16); layer 4: out = H3 W4 + b4 (FP32)
    const __bf16* in = q->X;
    int ldin = K0;
    for (int L = 0; L < NL; L++) {
        const __bf16* W = q->wb + q->off_w[L];
        const float* b = q->w + q->off_b[L];
        dim3 grid(B / 16, LN[L] / 16, 1);
        if (L < NL - 1) {
            hipLaunchKernelGGL((gemm_bf16<false, false, EPI_BIAS_RELU_BF16>), grid, dim3(64), 0, q->s, in, ldin, W,
                               LN[L], LK[L], LK[L], b, nullptr, 0, (void*)q->H[L], LN[L]);
            in = q->H[L];
            ldin = LN[L];
        } else {
            hipLaunchKernelGGL((gemm_bf16<false, false, EPI_BIAS_F32>), grid, dim3(64), 0, q->s, in, ldin, W, LN[L],
                               LK[L], LK[L], b, nullptr, 0, (void*)q->out, NOP);
        }
        MCK(hipGetLastError());
    }
    (void)keep_out;
    return 0;
}

// forward + loss + backward into q->g (FP32 gradients of the mean loss over n samples)
static int loss_grad(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, float* loss_out) {
    const int B = pad32(n);
    hipLaunchKernelGGL(k_embed, dim3((B * K0 + 255) / 256), dim3(256), 0, q->s, d_genomes, n, B, q->lo, q->inv_range,
                       q->X);
    MCK(hipGetLastError());
    int rc = forward(q, B, true);
    if (rc) return rc;
    // every gradient entry is written below (weights by k_sum_splits, biases by k_colsum_fin)
    hipLaunchKernelGGL(k_loss, dim3((B + 63) / 64), dim3(64), 0, q->s, q->out, d_targets, n, B, 1.0f / n, q->dout,
                       q->dout_f, q->loss_part);
    MCK(hipGetLastError());
    hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, q->s, q->loss_part, (B + 63) / 64, 1.0f / n, q->loss);
    MCK(hipGetLastError());
    // backward, layer 4 .. 1
    const __bf16* dY = q->dout;
    int ldy = NOP;
    for (int L = NL - 1; L >= 0; L--) {
        const __bf16* Xin = L == 0 ? q->X : q->H[L - 1];
        const int K = LK[L], N = LN[L];
        // dW[K x N] = Xin^T [K x B] * dY [B x N]: partial products per batch split of
        // WG_SPLIT rows, then summed in split order (deterministic, no atomics); the batch is
        // padded to a multiple of 32 with zero rows (zero activations and gradients)
        const int nsplit = (B + WG_SPLIT - 1) / WG_SPLIT;
        dim3 gw(K / 16, N / 16, nsplit);
        hipLaunchKernelGGL((gemm_bf16<true, false, EPI_PART_F32>), gw, dim3(64), 0, q->s, Xin, K, dY, ldy, B, WG_SPLIT,
                           nullptr, nullptr, 0, (void*)q->wgrad_part, N);
        MCK(hipGetLastError());
        hipLaunchKernelGGL(k_sum_splits, dim3((K * N + 255) / 256), dim3(256), 0, q->s, q->wgrad_part, nsplit, K * N,
                           q->g + q->off_w[L]);
        MCK(hipGetLastError());
        // db = column sums of dY
        hipLaunchKernelGGL(k_colsum_part, dim3((N + 63) / 64, COLSUM_SLICES), dim3(256), 0, q->s, dY, B, N,
                           q->colsum_part);
        MCK(hipGetLastError());
        hipLaunchKernelGGL(k_colsum_fin, dim3((N + 63) / 64), dim3(64), 0, q->s, q->colsum_part, N,
                           q->g + q->off_b[L]);
        MCK(hipGetLastError());
        if (L > 0) {
            // dXin [B x K] = dY [B x N] * W^T, masked by ReLU'(Xin) (Xin > 0)
            const __bf16* W = q->wb + q->off_w[L];
            dim3 gx(B / 16, K / 16, 1);
            hipLaunchKernelGGL((gemm_bf16<false, true, EPI_MASK_BF16>), gx, dim3(64), 0, q->s, dY, ldy, W, N, N, N,
                               nullptr, Xin, K, (void*)q->dH[L - 1], K);
            MCK(hipGetLastError());
            dY = q->dH[L - 1];
            ldy = K;
        }
    }
    if (loss_out) {
        MCK(hipMemcpyAsync(loss_out, q->loss, sizeof(float), hipMemcpyDeviceToHost, q->s));
        MCK(hipStreamSynchronize(q->s));
    }
    return 0;
}

static int refresh_bf16(pr_mlp* q) {
    for (int L = 0; L < NL; L++) {
        int n = LK[L] * LN[L];
        hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, q->s, q->w + q->off_w[L], n,
                           q->wb + q->off_w[L]);
        MCK(hipGetLastError());
    }
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
    }
    q->n_params = o;
    const int B = q->max_batch;
#define MALLOC(p, bytes) MCK(hipMalloc((void**)&(p), (bytes)))
    MCK(hipStreamCreateWithFlags(&q->s, hipStreamNonBlocking));
    MALLOC(q->w, 4 * o); MALLOC(q->g, 4 * o); MALLOC(q->m, 4 * o); MALLOC(q->v, 4 * o);
    MALLOC(q->wb, 2 * o);
    MALLOC(q->X, 2 * (size_t)B * K0);
    for (int k = 0; k < 3; k++) { MALLOC(q->H[k], 2 * (size_t)B * HID); MALLOC(q->dH[k], 2 * (size_t)B * HID); }
    MALLOC(q->out, 4 * (size_t)B * NOP);
    MALLOC(q->dout, 2 * (size_t)B * NOP);
    MALLOC(q->dout_f, 4 * (size_t)B * NOP);
    MALLOC(q->loss, 4);
    MALLOC(q->loss_part, 4 * (size_t)(B / 64 + 1));
    MALLOC(q->lo, 4 * NIN);
    MALLOC(q->inv_range, 4 * NIN);
    MALLOC(q->colsum_part, 4 * (size_t)COLSUM_SLICES * HID);
    MALLOC(q->wgrad_part, 4 * (size_t)((B + WG_SPLIT - 1) / WG_SPLIT) * HID * HID);
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
    void* bufs[] = {q->w, q->g, q->m, q->v, q->wb, q->X, q->out, q->dout, q->dout_f, q->loss, q->loss_part, q->lo, q->inv_range,
                    q->H[0], q->H[1], q->H[2], q->dH[0], q->dH[1], q->dH[2], q->colsum_part,
                    q->wgrad_part};
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
    for (int s0 = 0; s0 < n; s0 += q->max_batch) {
        const int nb = std::min(q->max_batch, n - s0);
        const int B = pad32(nb);
        hipLaunchKernelGGL(k_embed, dim3((B * K0 + 255) / 256), dim3(256), 0, q->s, d_genomes + (size_t)s0 * NIN, nb,
                           B, q->lo, q->inv_range, q->X);
        MCK(hipGetLastError());
        int rc = forward(q, B, false);
        if (rc) return rc;
        hipLaunchKernelGGL(k_head_scores, dim3((nb + 255) / 256), dim3(256), 0, q->s, q->out, nb,
                           d_pred + (size_t)s0 * 5);
        MCK(hipGetLastError());
    }
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

// Mean loss and its gradient for one batch (n <= max_batch); grads (host, optional) in the
// parameter layout of pr_mlp_get_params.
int pr_mlp_loss_grad(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, float* loss, float* grads) {
    if (!q || n < 1 || n > q->max_batch) { g_mlp_err = "pr_mlp_loss_grad: 1 <= n <= max_batch"; return -1; }
    int rc = loss_grad(q, d_genomes, d_targets, n, loss);
    if (rc) return rc;
    if (grads) MCK(hipMemcpy(grads, q->g, 4 * q->n_params, hipMemcpyDeviceToHost));
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
            int rc = loss_grad(q, d_genomes + (size_t)s0 * NIN, d_targets + (size_t)s0 * NOUT, nb, nullptr);
            if (rc) return rc;
            q->step++;
            const float c1 = 1.0f / (1.0f - std::pow(b1, (float)q->step));
            const float c2 = 1.0f / (1.0f - std::pow(b2, (float)q->step));
            hipLaunchKernelGGL(k_adam, dim3((int)((q->n_params + 255) / 256)), dim3(256), 0, q->s, q->w, q->g, q->m,
                               q->v, q->wb, (int)q->n_params, (int)q->off_w[0], (int)q->off_w[3], (int)q->off_b[3], lr,
                               b1, b2, eps, c1, c2);
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
