// mlp.hip -- surrogate fitness model: MLP 35 -> 256 -> 256 -> 256 -> 4 (ReLU hidden layers)
// trained on the (genome, fitness) pairs the flight simulation produces.
//
// Arithmetic: BF16 operands on the matrix cores (v_mfma_f32_32x32x16_bf16, gfx950), FP32
// accumulation, FP32 master weights and Adam state; activations stored as BF16 between
// layers.  Heads: 0 mean distance, 1 CVaR of the worst quarter, 2 worst distance (all in
// units of 10 m, regression, masked for unprintable designs), 3 oversize logit (BCE).
//
// Three kernels do the work, every MFMA operand a 16-byte load per lane:
//   k_mlp_predict  one pass per 128 rows: embedding, four layers and heads, the activations
//                  in LDS between layers;
//   k_mlp_fwd_bwd  a training batch's forward pass, loss, output gradient and the input
//                  gradients back through the hidden layers, 32 rows per workgroup, the
//                  activations (the ReLU masks) in LDS; it writes each layer input and output
//                  gradient transposed (features x batch, fragment order), the operands of
//   k_wgrad_all    dW = Xin^T dY for all four layers in one launch, one 32 x 32 tile per
//                  workgroup over the whole batch (16 waves on batch slices summed in wave
//                  order through LDS, the bias gradient from an all-ones MFMA operand
//                  alongside), each tile's parameters taking their Adam step in the same
//                  workgroup: training is deterministic.
// Weights are read from BF16 copies in MFMA fragment order (1 KB contiguous per wave load).
// A training step is two launches (three when the loss is read).
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

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

// MFMA fragment order of a [rows x red] BF16 operand (reduction index red): a fragment is one
// k-step of one 32-row tile, the 64 lanes' 16-byte operands back to back (lane l: row
// 32 t + l % 32, reduction index 16 s + 8 (l / 32) .. + 7), so that a wave's operand load is
// 1 KB contiguous; fragment (t, s) at (t * steps + s) * FRAG, steps = the reduction length / 16.
constexpr int FRAG = 512;  // bf16 elements per fragment
__device__ __forceinline__ int frag_offset(int row, int red, int steps) {
    return ((row >> 5) * steps + (red >> 4)) * FRAG + ((row & 31) + 32 * ((red >> 3) & 1)) * 8 + (red & 7);
}

// k-loop of one wave of a weight-gradient tile over its batch slice (ns k-steps), operands D steps ahead
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
            ra[d] = *reinterpret_cast<const bf8*>(pa + FRAG * d);
            rb[d] = *reinterpret_cast<const bf8*>(pb + FRAG * d);
        }
    for (int s0 = 0; s0 < ns; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            const int st = s0 + d;
            if (st < ns) {
                const bf8 a = ra[d], b = rb[d];
                if (st + D < ns) {
                    ra[d] = *reinterpret_cast<const bf8*>(pa + FRAG * (st + D));
                    rb[d] = *reinterpret_cast<const bf8*>(pb + FRAG * (st + D));
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
                if (DOB) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ones, b, accb, 0, 0, 0);
            }
        }
    }
}

// Weight (and bias) gradients of one 32 x 32 tile: dW[m][n] = sum over the batch of
// A[m][b] Bt[n][b], with A = the layer input transposed (features x batch) and Bt = the output
// gradient transposed -- both in fragment order over the batch (steps = B / 16), so every
// operand load of a wave is 1 KB contiguous.  A workgroup of WG_WAVES waves: wave w sums its
// share of the batch, the partial tiles are added in wave order through LDS (deterministic, no
// atomics).  The workgroups of row tile 0 also form db[n] = sum over the batch of dY[b][n]: an
// MFMA with an all-ones A operand on the same Bt operands gives the column sums of the slice,
// added in wave order likewise.  emit(m, n, value) receives every element (m = -1: the bias
// of column n).  B a multiple of 16 WG_WAVES.
constexpr int WG_WAVES = 16;
template <class Emit>
__device__ __forceinline__ void wgrad_tile(const __bf16* __restrict__ A, const __bf16* __restrict__ Bt, int B, int m0,
                                           int n0, float (*part)[16][64], float (*bpart)[32], Emit emit) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, lr = l & 31;
    const int ks = B / WG_WAVES, kb = w * ks;
    const __bf16* pa = A + ((size_t)(m0 >> 5) * (B >> 4) + (kb >> 4)) * FRAG + 8 * l;
    const __bf16* pb = Bt + ((size_t)(n0 >> 5) * (B >> 4) + (kb >> 4)) * FRAG + 8 * l;
    const bool dob = m0 == 0;
    f16v acc, accb;
#pragma unroll
    for (int r = 0; r < 16; r++) { acc[r] = 0.0f; accb[r] = 0.0f; }
    if (dob) wgrad_slice<true>(pa, pb, ks / 16, acc, accb);
    else wgrad_slice<false>(pa, pb, ks / 16, acc, accb);
#pragma unroll
    for (int r = 0; r < 16; r++) part[w][r][l] = acc[r];
    if (dob && l < 32) bpart[w][l] = accb[0];  // row 0 of ones * B: the column sums
    __syncthreads();
    static_assert(WG_WAVES == 16, "one accumulator row per wave in the reduction");
    {
        const int rr = w;
        float v = part[0][rr][l];
#pragma unroll
        for (int u = 1; u < WG_WAVES; u++) v += part[u][rr][l];
        emit(m0 + 8 * (rr >> 2) + 4 * (l >> 5) + (rr & 3), n0 + lr, v);
    }
    if (dob && w == 0 && l < 32) {
        float v = bpart[0][l];
#pragma unroll
        for (int u = 1; u < WG_WAVES; u++) v += bpart[u][l];
        emit(-1, n0 + l, v);
    }
}

// rows of a training batch are padded to this multiple: the weight gradients give each of
// their 16 waves a batch slice of whole 16-row k-steps
constexpr int ROW_PAD = 256;

// mean loss = the partials summed in a fixed order (lane-strided, then a butterfly), times 1/n
__global__ void k_loss_fin(const float* __restrict__ loss_part, int nblk, float inv_n, float* __restrict__ loss) {
    float l = 0.0f;
    for (int b = threadIdx.x; b < nblk; b += 64) l += loss_part[b];
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (threadIdx.x == 0) *loss = l * inv_n;
}

// The parameter block's geometry: the weights of layer L at off_w[L], [k x n] row-major.
// The kernels read the weights as MFMA A operands from two BF16 copies in fragment order: a
// fragment is one k-step of one 32-row tile, the 64 lanes' 16-byte operands back to back
// (lane l: row 32 t + l % 32, reduction index 16 s + 8 (l / 32) .. + 7), so that every weight
// load of a wave is 1 KB contiguous.  The forward copy's rows are the layer's outputs n
// (reduction over k: the transposed weights), the backward copy's rows its inputs k
// (reduction over n: the input gradient dY W^T); fragment (t, s) of layer L sits at
// off_w[L] + (t * steps + s) * FRAG, steps = the reduction length / 16.
struct ParamLayout {
    int off_w[4], k[4], n[4];
};
// positions of parameter i in the forward and backward copies (-1: a bias)
__device__ __forceinline__ void frag_index(const ParamLayout& P, int i, int& fwd, int& bwd) {
    fwd = bwd = -1;
#pragma unroll
    for (int L = 0; L < 4; L++) {
        const int t = i - P.off_w[L];
        if (t >= 0 && t < P.k[L] * P.n[L]) {
            const int kk = t / P.n[L], nn = t - kk * P.n[L];
            fwd = P.off_w[L] + frag_offset(nn, kk, P.k[L] >> 4);
            bwd = P.off_w[L] + frag_offset(kk, nn, P.n[L] >> 4);
        }
    }
}

// Adam step of parameter i with gradient gi on the FP32 master copy and moments, writing its
// BF16 fragment copies too.  Padding parameters (W1 rows >= 35, W4 / b4 columns >= 4) take no
// step, so they stay zero.
struct AdamArgs {
    float lr, b1, b2, eps, c1, c2;  // c1, c2: the bias corrections 1 / (1 - b^step)
};
__device__ __forceinline__ void adam_param(int i, float gi, float* __restrict__ w, float* __restrict__ m,
                                           float* __restrict__ v, __bf16* __restrict__ wf, __bf16* __restrict__ wbk,
                                           const ParamLayout& P, const AdamArgs& A) {
    const int off_w1 = P.off_w[0], off_w4 = P.off_w[3], off_b4 = P.off_w[3] + P.k[3] * P.n[3];
    const bool pad = (i >= off_w1 + NIN * HID && i < off_w1 + K0 * HID) ||
                     (i >= off_w4 && i < off_w4 + HID * NOP && (i - off_w4) % NOP >= NOUT) ||
                     (i >= off_b4 && i < off_b4 + NOP && i - off_b4 >= NOUT);
    if (pad) return;
    const float mi = A.b1 * m[i] + (1.0f - A.b1) * gi;
    const float vi = A.b2 * v[i] + (1.0f - A.b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - A.lr * (mi * A.c1) / (sqrtf(vi * A.c2) + A.eps);
    w[i] = wi;
    int f, b;
    frag_index(P, i, f, b);
    if (f >= 0) { wf[f] = to_bf(wi); wbk[b] = to_bf(wi); }
}

// the operands and parameter block of one layer's weight gradient
struct WgradLayer {
    const __bf16* A;   // layer input, transposed [k x batch] (fragment order)
    const __bf16* Bt;  // output gradient, transposed [n x batch] (fragment order)
    int off_w, off_b;  // the layer's weights [k x n] and biases [n] in the parameter block
    int mt, nt;        // 32 x 32 tiles along k and n
};
struct WgradAll {
    WgradLayer L[NL];  // in launch order
    int first[NL + 1]; // first workgroup of each layer (multiples of 8), and the total
    int gm[NL], gn[NL];  // the gm x gn block of tiles each XCD takes in each layer
};

// every layer's weight and bias gradients in one launch: workgroup b belongs to the layer whose
// range holds it and computes one 32 x 32 tile over the whole batch; ADAM: the tile's
// parameters take their Adam step right there, else the gradient goes to g
template <bool ADAM>
__global__ __launch_bounds__(64 * WG_WAVES) void k_wgrad_all(WgradAll W, int B, float* __restrict__ g, float* __restrict__ w,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   __bf16* __restrict__ wf, __bf16* __restrict__ wbk, ParamLayout P,
                                                   AdamArgs A) {
    __shared__ float part[WG_WAVES][16][64];
    __shared__ float bpart[WG_WAVES][32];
    int L = 0;
#pragma unroll
    for (int u = 1; u < NL; u++)
        if ((int)blockIdx.x >= W.first[u]) L = u;
    const WgradLayer& G = W.L[L];
    // workgroups are dispatched to the eight XCDs round-robin: workgroup b of the layer runs on
    // XCD b % 8, which takes a gm x gn block of tiles, so that the operand rows the tiles share
    // are read into that XCD's L2 once
    const int b = blockIdx.x - W.first[L], x = b & 7, j = b >> 3;
    const int gm = W.gm[L], gn = W.gn[L], gx = x % (G.mt / gm), gy = x / (G.mt / gm);
    const int mi = gx * gm + j % gm, ni = gy * gn + j / gm;
    const int ldc = 32 * G.nt;
    wgrad_tile(G.A, G.Bt, B, 32 * mi, 32 * ni, part, bpart, [&](int r, int c, float gv) {
        const int i = r < 0 ? G.off_b + c : G.off_w + r * ldc + c;
        if (ADAM) adam_param(i, gv, w, m, v, wf, wbk, P, A);
        else g[i] = gv;
    });
}

// the BF16 fragment copies of the weights
__global__ void k_to_bf16(const float* __restrict__ x, int n, __bf16* __restrict__ wf, __bf16* __restrict__ wbk,
                          ParamLayout P) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int f, b;
    frag_index(P, i, f, b);
    if (f >= 0) { wf[f] = to_bf(x[i]); wbk[b] = to_bf(x[i]); }
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
// ring slot s % PD holds k-step s of this layer on entry and k-step s of the next layer on exit.
// The activation operands of step s + 1 are read from LDS during step s.  A scheduling barrier
// closes every step: without it the compiler sinks the ring's loads next to their MFMAs and
// every step waits out an L2 round trip.
// pw, pw_next: the lane's operand in the layer's first fragment of this wave's first tile;
// tiles tstride apart, k-steps FRAG apart.
template <int KIN, int NA, int NB, int NA_NEXT>
__device__ __forceinline__ void fused_layer_k(f16v (&acc)[NA][NB], bf8 (&ring)[PD][2], const __bf16* __restrict__ pw,
                                              int tstride, const __bf16* __restrict__ pw_next, int tstride_next,
                                              const __bf16* __restrict__ act, int pitch, int j0, int lr, int kq) {
    constexpr int NS = KIN / 16;
    static_assert(NS % PD == 0, "k-steps a multiple of the ring depth");
#pragma unroll
    for (int i = 0; i < NA; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
    const __bf16* pa = act + (32 * j0 + lr) * pitch + kq;
    bf8 b[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) b[j] = *reinterpret_cast<const bf8*>(pa + 32 * j * pitch);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int d = s % PD;
        bf8 a[NA];
#pragma unroll
        for (int i = 0; i < NA; i++) a[i] = ring[d][i];
        if (s + PD < NS) {
#pragma unroll
            for (int i = 0; i < NA; i++)
                ring[d][i] = *reinterpret_cast<const bf8*>(pw + i * tstride + FRAG * (s + PD));
        } else if (NA_NEXT > 0) {
#pragma unroll
            for (int i = 0; i < NA_NEXT; i++)
                ring[d][i] = *reinterpret_cast<const bf8*>(pw_next + i * tstride_next + FRAG * (s + PD - NS));
        }
        bf8 bn[NB];
        if (s + 1 < NS) {
#pragma unroll
            for (int j = 0; j < NB; j++) bn[j] = *reinterpret_cast<const bf8*>(pa + 32 * j * pitch + 16 * (s + 1));
        }
#pragma unroll
        for (int i = 0; i < NA; i++)
#pragma unroll
            for (int j = 0; j < NB; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (s + 1 < NS) {
#pragma unroll
            for (int j = 0; j < NB; j++) b[j] = bn[j];
        }
        __builtin_amdgcn_sched_barrier(0);
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
                                                        const __bf16* __restrict__ wf, ParamLayout P,
                                                        const float* __restrict__ w, const float* __restrict__ lo,
                                                        const float* __restrict__ inv_range,
                                                        float* __restrict__ pred) {
    __shared__ __attribute__((aligned(16))) __bf16 h[PB_ROWS * PB_HP];  // activations (in place)
    __shared__ __attribute__((aligned(16))) float cst[PB_CONST];
    const int t = threadIdx.x, l = t & 63, wv = t >> 6, lr = l & 31, kq = 8 * (l >> 5);
    const int m0 = blockIdx.x * PB_ROWS;
    const int n0 = 64 * wv;  // this wave's features in the hidden layers
    // this wave's first fragment of each layer (forward copy; tiles of 32 outputs)
    constexpr int T0 = K0 / 16 * FRAG, TH = HID / 16 * FRAG;  // tile strides
    const __bf16* pw0 = wf + P.off_w[0] + (n0 >> 5) * T0 + 8 * l;
    const __bf16* pw1 = wf + P.off_w[1] + (n0 >> 5) * TH + 8 * l;
    const __bf16* pw2 = wf + P.off_w[2] + (n0 >> 5) * TH + 8 * l;
    const __bf16* pw3 = wf + P.off_w[3] + 8 * l;
    bf8 ring[PD][2];
#pragma unroll
    for (int d = 0; d < PD; d++)
#pragma unroll
        for (int i = 0; i < 2; i++) ring[d][i] = *reinterpret_cast<const bf8*>(pw0 + i * T0 + FRAG * d);

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
#pragma unroll
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
    fused_layer_k<K0, 2, 4, 2>(acc, ring, pw0, T0, pw1, TH, h, PB_XP, 0, lr, kq);
    __syncthreads();  // every wave has read the inputs: overwrite them
    fused_relu_to_lds(acc, cst, n0, h, l);
    __syncthreads();
    fused_layer_k<HID, 2, 4, 2>(acc, ring, pw1, TH, pw2, TH, h, PB_HP, 0, lr, kq);
    __syncthreads();
    fused_relu_to_lds(acc, cst + HID, n0, h, l);
    __syncthreads();
    fused_layer_k<HID, 2, 4, 1>(acc, ring, pw2, TH, pw3, TH, h, PB_HP, 0, lr, kq);
    __syncthreads();
    fused_relu_to_lds(acc, cst + 2 * HID, n0, h, l);
    __syncthreads();
    f16v acc3[1][1];
    fused_layer_k<HID, 1, 1, 0>(acc3, ring, pw3, TH, nullptr, 0, h, PB_HP, wv, lr, kq);
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


// ---- a training step's forward pass, loss and input gradients in one pass: k_mlp_fwd_bwd ----
// A workgroup of four waves takes 32 rows of the batch: embedding, the four layers, the loss
// and its output gradient, and the input gradients back through layers 4, 3 and 2 (ReLU
// masks from the activations, which stay in LDS), with the layers computed transposed as in
// k_mlp_predict (wave w: features 64 w .. 64 w + 63; the A operand the transposed weight copy
// in the forward pass, the row-major weights in the backward pass).  What leaves the block is
// what the weight gradients read: the layer inputs and output gradients transposed
// (features x batch), X^T, H1^T..H3^T, dY4^T, dH3^T..dH1^T, plus one loss partial.
constexpr int TB_ROWS = 32;      // rows per workgroup
constexpr int TB_DP = NOP + 8;   // LDS row pitch of the output gradient

// the transposed global copies a block writes (each [features x B], B the padded batch), in
// fragment order over the batch: element (f, m) at frag_offset(f, m, B / 16)
struct TrainOut {
    __bf16* xt;      // [64 x B]
    __bf16* ht[3];   // [256 x B] hidden activations
    __bf16* dht[3];  // [256 x B] their gradients
    __bf16* doutt;   // [32 x B] output gradient
};

// epilogue of a 32-row block's transposed tiles (features n0 .. n0 + 63): MASK false:
// ReLU(acc + bias); MASK true: acc where the activation in `mask` (LDS rows) is positive, else
// 0; the BF16 result to the LDS rows of `out`
template <bool MASK>
__device__ __forceinline__ void train_epilogue(const f16v (&acc)[2][1], const float* __restrict__ bias,
                                               const __bf16* __restrict__ mask, int n0, __bf16* __restrict__ out,
                                               int l) {
    const int m = l & 31;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int n = n0 + 32 * i + 8 * g + 4 * (l >> 5);
            bf4 o;
            if (MASK) {
                const bf4 mk = *reinterpret_cast<const bf4*>(mask + m * PB_HP + n);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float keep = from_bf(mk[q]) > 0.0f ? 1.0f : 0.0f;
                    o[q] = to_bf(acc[i][0][4 * g + q] * keep);
                }
            } else {
                const f4 bv = *reinterpret_cast<const f4*>(bias + n);
#pragma unroll
                for (int q = 0; q < 4; q++) o[q] = to_bf(fmaxf(acc[i][0][4 * g + q] + bv[q], 0.0f));
            }
            *reinterpret_cast<bf4*>(out + m * PB_HP + n) = o;
        }
}

// a block's 32 LDS rows src (features 0 .. F - 1, row pitch `pitch`) -> the global copy g
// [F x B] in fragment order over the batch, rows m0 .. m0 + 31: each thread moves 16-byte
// chunks (8 batch rows of one feature); consecutive threads take consecutive features, so
// the LDS reads are conflict-free and a wave's stores are two 512-byte runs
template <int F>
__device__ __forceinline__ void copy_out(const __bf16* __restrict__ src, int pitch, __bf16* __restrict__ g, int B,
                                         int m0, int t) {
#pragma unroll
    for (int c = t; c < 4 * F; c += 256) {
        const int f = c % F, r8 = c / F;
        bf8 v;
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = src[(8 * r8 + j) * pitch + f];
        *reinterpret_cast<bf8*>(g + frag_offset(f, m0 + 8 * r8, B >> 4)) = v;
    }
}

__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float* __restrict__ genomes,
                                                     const float* __restrict__ tgt, int n, int B, float inv_n,
                                                     const __bf16* __restrict__ wf, const __bf16* __restrict__ wbk,
                                                     ParamLayout P, const float* __restrict__ w,
                                                     const float* __restrict__ lo,
                                                     const float* __restrict__ inv_range, TrainOut T,
                                                     float* __restrict__ loss_part) {
    __shared__ __attribute__((aligned(16))) __bf16 xs[TB_ROWS * PB_XP];     // embedded inputs
    __shared__ __attribute__((aligned(16))) __bf16 hs[3][TB_ROWS * PB_HP];  // H1..H3 (post-ReLU)
    __shared__ __attribute__((aligned(16))) __bf16 ds[TB_ROWS * PB_HP];     // dH3, then dH2 (in place)
    __shared__ __attribute__((aligned(16))) __bf16 dos[TB_ROWS * TB_DP];    // output gradient
    __shared__ __attribute__((aligned(16))) float cst[PB_CONST];
    __shared__ float red[4][NOUT][32];  // the output layer's four k-quarter partial sums
    const int t = threadIdx.x, l = t & 63, wv = t >> 6, lr = l & 31, kq = 8 * (l >> 5);
    const int m0 = blockIdx.x * TB_ROWS, n0 = 64 * wv;
    // this wave's first fragment of each layer: forward copy (tiles of outputs), backward copy
    // (tiles of inputs: W3 and W2 rows)
    constexpr int T0 = K0 / 16 * FRAG, TH = HID / 16 * FRAG, T4 = NOP / 16 * FRAG;  // tile strides
    const __bf16* pf0 = wf + P.off_w[0] + (n0 >> 5) * T0 + 8 * l;
    const __bf16* pf1 = wf + P.off_w[1] + (n0 >> 5) * TH + 8 * l;
    const __bf16* pf2 = wf + P.off_w[2] + (n0 >> 5) * TH + 8 * l;
    const __bf16* pb2 = wbk + P.off_w[2] + (n0 >> 5) * TH + 8 * l;
    const __bf16* pb1 = wbk + P.off_w[1] + (n0 >> 5) * TH + 8 * l;
    bf8 ring[PD][2];
#pragma unroll
    for (int d = 0; d < PD; d++)
#pragma unroll
        for (int i = 0; i < 2; i++) ring[d][i] = *reinterpret_cast<const bf8*>(pf0 + i * T0 + FRAG * d);

    // genomes of the block (element e: row e % 32, gene e / 32), targets of wave 0's rows
    constexpr int NE = TB_ROWS * K0 / 256;
    float gv[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = t + 256 * i, m = e & 31, k = e >> 5;
        gv[i] = (k < NIN && m0 + m < n) ? genomes[(size_t)(m0 + m) * NIN + k] : 0.0f;
    }
    float tg[NOUT] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (t < 32 && m0 + t < n) {
#pragma unroll
        for (int q = 0; q < NOUT; q++) tg[q] = tgt[(size_t)(m0 + t) * NOUT + q];
    }
#pragma unroll
    for (int i = t; i < PB_CONST; i += 256) {
        float v;
        if (i < PB_B4) v = w[P.off_w[i >> 8] + P.k[i >> 8] * P.n[i >> 8] + (i & (HID - 1))];
        else if (i < PB_LO) v = w[P.off_w[3] + P.k[3] * P.n[3] + (i - PB_B4)];
        else if (i < PB_IR) v = lo[i - PB_LO];
        else v = inv_range[i - PB_IR];
        cst[i] = v;
    }
    __syncthreads();
    // embedding (genes scaled to [0, 1] by the bounds, zero beyond gene 35 and row n) -> LDS
    // rows, and X^T
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = t + 256 * i, m = e & 31, k = e >> 5;
        xs[m * PB_XP + k] = to_bf((k < NIN && m0 + m < n) ? (gv[i] - cst[PB_LO + k]) * cst[PB_IR + k] : 0.0f);
    }
    __syncthreads();
    copy_out<K0>(xs, PB_XP, T.xt, B, m0, t);

    f16v acc[2][1];
    fused_layer_k<K0, 2, 1, 2>(acc, ring, pf0, T0, pf1, TH, xs, PB_XP, 0, lr, kq);
    train_epilogue<false>(acc, cst, nullptr, n0, hs[0], l);
    __syncthreads();
    copy_out<HID>(hs[0], PB_HP, T.ht[0], B, m0, t);
    fused_layer_k<HID, 2, 1, 2>(acc, ring, pf1, TH, pf2, TH, hs[0], PB_HP, 0, lr, kq);
    train_epilogue<false>(acc, cst + HID, nullptr, n0, hs[1], l);
    __syncthreads();
    copy_out<HID>(hs[1], PB_HP, T.ht[1], B, m0, t);
    // the ring continues with the backward pass through layer 3 (W3 rows); layer 4 forward and
    // backward load their few operands directly
    fused_layer_k<HID, 2, 1, 2>(acc, ring, pf2, TH, pb2, TH, hs[1], PB_HP, 0, lr, kq);
    bf8 a4[4];  // output layer, this wave's k-quarter 64 wv .. 64 wv + 63 (W4^T rows 0..31)
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++)
        a4[s4] = *reinterpret_cast<const bf8*>(wf + P.off_w[3] + (4 * wv + s4) * FRAG + 8 * l);
    bf8 w4[2][2];  // backward through the output layer: W4 rows n0 + 32 i + lane (its 32 outputs)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int s4 = 0; s4 < 2; s4++)
            w4[i][s4] = *reinterpret_cast<const bf8*>(wbk + P.off_w[3] + ((n0 >> 5) + i) * T4 + s4 * FRAG + 8 * l);
    train_epilogue<false>(acc, cst + 2 * HID, nullptr, n0, hs[2], l);
    __syncthreads();
    copy_out<HID>(hs[2], PB_HP, T.ht[2], B, m0, t);

    // output layer: out^T = W4^T H3^T, the k range split over the four waves, summed in wave order
    f16v a3;
#pragma unroll
    for (int r = 0; r < 16; r++) a3[r] = 0.0f;
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
        const bf8 b = *reinterpret_cast<const bf8*>(hs[2] + lr * PB_HP + 64 * wv + 16 * s4 + kq);
        a3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a4[s4], b, a3, 0, 0, 0);
    }
    if (l < 32) {
#pragma unroll
        for (int q = 0; q < NOUT; q++) red[wv][q][l] = a3[q];  // features 0..3 of row l
    }
    __syncthreads();
    if (wv == 0) {
        // loss and output gradient of row m0 + l: squared error of the three distance heads
        // (targets in metres, heads in tens of metres; skipped for oversize designs) and the
        // logistic loss of the oversize head, each gradient over the batch mean
        float lossv = 0.0f;
        float d[NOUT] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (l < 32 && m0 + l < n) {
            float o[NOUT];
#pragma unroll
            for (int q = 0; q < NOUT; q++)
                o[q] = ((red[0][q][l] + red[1][q][l]) + red[2][q][l]) + red[3][q][l] + cst[PB_B4 + q];
            const float over = tg[3];
            if (over < 0.5f) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const float e = o[k] - tg[k] * 0.1f;
                    lossv += e * e;
                    d[k] = 2.0f * e * inv_n;
                }
            }
            const float z = o[3];
            const float p = 1.0f / (1.0f + expf(-z));
            lossv += fmaxf(z, 0.0f) - z * over + log1pf(expf(-fabsf(z)));
            d[3] = (p - over) * inv_n;
        }
        if (l < 32) {
            bf8 row0;
#pragma unroll
            for (int k = 0; k < 8; k++) row0[k] = to_bf(k < NOUT ? d[k] : 0.0f);
            bf8 zero;
#pragma unroll
            for (int k = 0; k < 8; k++) zero[k] = to_bf(0.0f);
            *reinterpret_cast<bf8*>(dos + l * TB_DP) = row0;
#pragma unroll
            for (int c = 1; c < NOP / 8; c++) *reinterpret_cast<bf8*>(dos + l * TB_DP + 8 * c) = zero;
        }
        for (int o = 32; o > 0; o >>= 1) lossv += __shfl_xor(lossv, o);
        if (l == 0) loss_part[blockIdx.x] = lossv;
    }
    __syncthreads();
    copy_out<NOP>(dos, TB_DP, T.doutt, B, m0, t);

    // dH3 = (dY4 W4^T) masked by H3 > 0: K = the output layer's 32 features
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[i][0][r] = 0.0f;
#pragma unroll
    for (int s4 = 0; s4 < 2; s4++) {
        const bf8 b = *reinterpret_cast<const bf8*>(dos + lr * TB_DP + 16 * s4 + kq);
#pragma unroll
        for (int i = 0; i < 2; i++) acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w4[i][s4], b, acc[i][0], 0, 0, 0);
    }
    train_epilogue<true>(acc, nullptr, hs[2], n0, ds, l);
    __syncthreads();
    copy_out<HID>(ds, PB_HP, T.dht[2], B, m0, t);
    // dH2 = (dH3 W3^T) masked by H2 > 0, written over dH3 once every wave has read it
    fused_layer_k<HID, 2, 1, 2>(acc, ring, pb2, TH, pb1, TH, ds, PB_HP, 0, lr, kq);
    __syncthreads();
    train_epilogue<true>(acc, nullptr, hs[1], n0, ds, l);
    __syncthreads();
    copy_out<HID>(ds, PB_HP, T.dht[1], B, m0, t);
    // dH1 = (dH2 W2^T) masked by H1 > 0 (the last input gradient the weight gradients need),
    // staged in the rows of H3, which no step reads any more
    fused_layer_k<HID, 2, 1, 0>(acc, ring, pb1, TH, nullptr, 0, ds, PB_HP, 0, lr, kq);
    train_epilogue<true>(acc, nullptr, hs[0], n0, hs[2], l);
    __syncthreads();
    copy_out<HID>(hs[2], PB_HP, T.dht[0], B, m0, t);
}


thread_local std::string g_mlp_err;

}  // namespace

// parameter block layout (FP32): W1 [64x256] b1 [256] W2 [256x256] b2 W3 [256x256] b3 W4 [256x32] b4 [32]
// (rows 35..63 of W1 and columns 4..31 of W4 are padding and stay zero)
struct pr_mlp {
    int max_batch = 0;
    size_t off_w[NL], off_b[NL], n_params = 0;
    float* w = nullptr;   // master FP32 parameters
    float* g = nullptr;   // gradients (pr_mlp_loss_grad)
    float* m = nullptr;   // Adam first moment
    float* v = nullptr;   // Adam second moment
    // BF16 copies of the weights in MFMA fragment order (biases stay FP32): the forward pass's
    // A operands (rows = outputs) and the input gradients' (rows = inputs)
    __bf16* wf = nullptr;
    __bf16* wbk = nullptr;
    ParamLayout layout = {};
    // transposes (features x B) of the layer inputs and output gradients of a training batch:
    // the operands of the weight gradients
    __bf16* Xt = nullptr;     // [64 x B]
    __bf16* Ht[3] = {};       // [256 x B] post-ReLU activations
    __bf16* dHt[3] = {};      // [256 x B] their gradients
    __bf16* doutT = nullptr;  // [32 x B]
    float* loss = nullptr;
    float* loss_part = nullptr;  // per-32-row-block loss sums
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

// deterministic He-uniform initialisation from a 32-bit hash of (seed, index)
static float init_u(unsigned seed, size_t i) {
    unsigned x = seed * 0x9e3779b9u + (unsigned)i * 0x85ebca6bu + (unsigned)(i >> 32);
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return (float)(x >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
}

// One training batch: the fused forward / loss / input-gradient pass, the mean loss (when
// wanted), and every layer's weight and bias gradients in one launch, which either takes the
// Adam step (adam given) or leaves the gradient in q->g
static int loss_grad(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, bool want_loss,
                     const AdamArgs* adam) {
    const int B = pad_rows(n);
    TrainOut T;
    T.xt = q->Xt;
    for (int k = 0; k < 3; k++) { T.ht[k] = q->Ht[k]; T.dht[k] = q->dHt[k]; }
    T.doutt = q->doutT;
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(B / TB_ROWS), dim3(256), 0, q->s, d_genomes, d_targets, n, B, 1.0f / n,
                       q->wf, q->wbk, q->layout, q->w, q->lo, q->inv_range, T, q->loss_part);
    MCK(hipGetLastError());
    if (want_loss) {
        hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, q->s, q->loss_part, B / TB_ROWS, 1.0f / n, q->loss);
        MCK(hipGetLastError());
    }
    // dW [K x N] = Xin^T dY and db = column sums of dY over the (zero-padded) batch; the two
    // 256 x 256 layers first
    WgradAll W;
    const int order[NL] = {1, 2, 0, 3};
    int first = 0;
    for (int u = 0; u < NL; u++) {
        const int L = order[u];
        WgradLayer& G = W.L[u];
        G.A = L == 0 ? q->Xt : q->Ht[L - 1];
        G.Bt = L == NL - 1 ? q->doutT : q->dHt[L];
        G.off_w = (int)q->off_w[L];
        G.off_b = (int)q->off_b[L];
        G.mt = LK[L] / 32;
        G.nt = LN[L] / 32;
        // per-XCD blocks: 8 x 8 tiles -> 2 x 4, 2 x 8 -> 1 x 2, 8 x 1 -> 1 x 1
        W.gm[u] = G.mt * G.nt == 64 ? 2 : 1;
        W.gn[u] = G.mt * G.nt / 8 / W.gm[u];
        W.first[u] = first;
        first += G.mt * G.nt;
    }
    W.first[NL] = first;
    if (adam)
        hipLaunchKernelGGL(k_wgrad_all<true>, dim3(first), dim3(64 * WG_WAVES), 0, q->s, W, B, q->g, q->w, q->m, q->v, q->wf,
                           q->wbk, q->layout, *adam);
    else
        hipLaunchKernelGGL(k_wgrad_all<false>, dim3(first), dim3(64 * WG_WAVES), 0, q->s, W, B, q->g, q->w, q->m, q->v, q->wf,
                           q->wbk, q->layout, AdamArgs{});
    MCK(hipGetLastError());
    return 0;
}

static int refresh_bf16(pr_mlp* q) {
    const int n = (int)q->n_params;
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, q->s, q->w, n, q->wf, q->wbk, q->layout);
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
    MALLOC(q->wf, 2 * o);
    MALLOC(q->wbk, 2 * o);
    MALLOC(q->Xt, 2 * (size_t)B * K0);
    for (int k = 0; k < 3; k++) { MALLOC(q->Ht[k], 2 * (size_t)B * HID); MALLOC(q->dHt[k], 2 * (size_t)B * HID); }
    MALLOC(q->doutT, 2 * (size_t)B * NOP);
    MALLOC(q->loss, 4);
    MALLOC(q->loss_part, 4 * (size_t)(B / TB_ROWS));
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
    void* bufs[] = {q->w,      q->g,      q->m,      q->v,      q->wf,     q->wbk,
                    q->Xt,     q->doutT,  q->loss,   q->loss_part, q->lo,  q->inv_range,
                    q->Ht[0],  q->Ht[1],  q->Ht[2],  q->dHt[0], q->dHt[1], q->dHt[2]};
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
    hipLaunchKernelGGL(k_mlp_predict, dim3((n + PB_ROWS - 1) / PB_ROWS), dim3(256), 0, q->s, d_genomes, n, q->wf,
                       q->layout, q->w, q->lo, q->inv_range, d_pred);
    MCK(hipGetLastError());
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

// Mean loss and its gradient for one batch (n <= max_batch); grads (host, optional) in the
// parameter layout of pr_mlp_get_params.
int pr_mlp_loss_grad(pr_mlp* q, const float* d_genomes, const float* d_targets, int n, float* loss, float* grads) {
    if (!q || n < 1 || n > q->max_batch) { g_mlp_err = "pr_mlp_loss_grad: 1 <= n <= max_batch"; return -1; }
    int rc = loss_grad(q, d_genomes, d_targets, n, true, nullptr);
    if (rc) return rc;
    if (loss) {
        MCK(hipMemcpyAsync(loss, q->loss, sizeof(float), hipMemcpyDeviceToHost, q->s));
        MCK(hipStreamSynchronize(q->s));
    }
    if (grads) {
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
            q->step++;
            const AdamArgs A{lr, b1, b2, eps, 1.0f / (1.0f - std::pow(b1, (float)q->step)),
                             1.0f / (1.0f - std::pow(b2, (float)q->step))};
            int rc = loss_grad(q, d_genomes + (size_t)s0 * NIN, d_targets + (size_t)s0 * NOUT, nb, last, &A);
            if (rc) return rc;
        }
    }
    if (last_loss) {
        MCK(hipMemcpyAsync(last_loss, q->loss, sizeof(float), hipMemcpyDeviceToHost, q->s));
    }
    MCK(hipStreamSynchronize(q->s));
    return 0;
}

}  // extern "C"
