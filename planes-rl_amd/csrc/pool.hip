// pool.hip -- candidate-pool GA with optional surrogate pre-screening (no reference
// counterpart; the benchmark configuration "score 1M mutants, fly the top 100k").
//
// One generation, all on the GPU except the host's bookkeeping:
//   1. pool: the elite (re-flown every generation under new seeds, like the reference's
//      GA) + Gaussian mutants of random elite parents (sigma = mutation fraction of each
//      gene's range, clamped to the bounds) -- k_mutate
//   2. fly set: the whole pool (plain), or the elite + the candidates with the best
//      surrogate-predicted robust score (pr_mlp_predict, hipCUB radix sort)
//   3. flights of the fly set under the generation's seeds + robust scores (pr_fly/pr_scores)
//   4. new elite = the best flown genomes by simulated score; top-8 re-flown on the
//      128-scenario holdout suite -> champion
//   5. surrogate mode: the flown (genome, fitness statistics) pairs join a replay buffer
//      and the network trains on it (pr_mlp_train)
// The simulation alone decides the elite and the champion.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/planes_hip.h"
#include "pr_physics.h"

namespace {

thread_local std::string g_pool_err;

__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) {
    return pr::mix32(a * 0x9e3779b9u ^ pr::mix32(b * 0x85ebca6bu ^ pr::mix32(c + 0x27d4eb2fu)));
}
__device__ __forceinline__ float unit24(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

// hash3 on the host (the parent draw of k_mutate, for the key calibration)
static inline uint32_t host_hash3(uint32_t a, uint32_t b, uint32_t c) {
    return pr::mix32(a * 0x9e3779b9u ^ pr::mix32(b * 0x85ebca6bu ^ pr::mix32(c + 0x27d4eb2fu)));
}

// pool[i] = elite[i] for i < n_elite, else a mutant of a random elite parent
__global__ void k_mutate(const float* __restrict__ elite, int n_elite, float* __restrict__ pool, int n_pool,
                         const float* __restrict__ lo, const float* __restrict__ hi, float sigma_frac, uint32_t seed,
                         uint32_t gen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pool) return;
    if (i < n_elite) {
        for (int j = 0; j < pr::NG; j++) pool[(size_t)i * pr::NG + j] = elite[(size_t)i * pr::NG + j];
        return;
    }
    const int par = (int)(hash3(seed, gen, (uint32_t)i) % (uint32_t)n_elite);
    for (int j = 0; j < pr::NG; j += 2) {
        // Box-Muller pair
        const uint32_t h1 = hash3(seed ^ 0x5bd1e995u, gen * 64u + (uint32_t)j, (uint32_t)i);
        const uint32_t h2 = pr::mix32(h1 + 0x68e31da4u);
        const float u1 = ((float)(h1 >> 8) + 1.0f) * (1.0f / 16777216.0f);
        const float u2 = unit24(h2);
        const float r = sqrtf(-2.0f * logf(u1));
        const float a = 6.2831853f * u2;
        const float z[2] = {r * cosf(a), r * sinf(a)};
        for (int q = 0; q < 2 && j + q < pr::NG; q++) {
            const int k = j + q;
            const float range = hi[k] - lo[k];
            float v = elite[(size_t)par * pr::NG + k] + sigma_frac * range * z[q];
            pool[(size_t)i * pr::NG + k] = fminf(hi[k], fmaxf(lo[k], v));
        }
    }
}

__global__ void k_gather(const float* __restrict__ src, const int* __restrict__ idx, int n, float* __restrict__ dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * pr::NG) return;
    const int i = t / pr::NG, j = t - i * pr::NG;
    dst[t] = src[(size_t)idx[i] * pr::NG + j];
}

// keys for a descending sort: -score, with the elite (i < n_keep) and an exploration sample
// (each other candidate with probability p_explore, hashed from (seed, gen, i)) forced to
// the front
// escore (non-null): the elite's simulated scores; a mutant's key is then its parent's
// simulated score plus the predicted change from parent to mutant (the network's error
// common to both cancels), with the parent drawn as k_mutate draws it
// cw: the key's weights (a, b, c) of the parent's simulated score, the mutant's predicted
// score and the parent's predicted score -- (1, 1, -1): the parent's score plus the predicted
// change; fitted: least squares of the mutants' simulated scores on the three over the last
// exploration sample
__global__ void k_sort_keys(const float* __restrict__ pred, int stride, int col, int n, int n_keep,
                            float p_explore, uint32_t seed, uint32_t gen, const float* __restrict__ escore,
                            float ca, float cb, float cc, float* __restrict__ keys, int* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool explore = i >= n_keep && unit24(hash3(seed ^ 0x2545f491u, gen, (uint32_t)i)) < p_explore;
    float s = pred[(size_t)i * stride + col];
    if (escore && i >= n_keep) {
        const int par = (int)(hash3(seed, gen, (uint32_t)i) % (uint32_t)n_keep);
        s = ca * escore[par] + (cb * s + cc * pred[(size_t)par * stride + col]);
    }
    keys[i] = i < n_keep ? -3.0e38f : explore ? -2.0e38f : -s;
    vals[i] = i;
}

// keys for a selection without the network (the surrogate's ablations): mode 1, a uniform
// random subset of the mutants (hashed from (seed, gen, i)); mode 2, the mutant's parent's
// simulated score (escore; ties -- mutants of one parent -- keep the pool order, which is
// independent of the parent because k_mutate draws parents by hash).  The elite first.
__global__ void k_sort_keys_nonet(int n, int n_keep, int mode, uint32_t seed, uint32_t gen,
                                  const float* __restrict__ escore, float* __restrict__ keys, int* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float k;
    if (i < n_keep) k = -3.0e38f;
    else if (mode == 2 && escore) k = -escore[(int)(hash3(seed, gen, (uint32_t)i) % (uint32_t)n_keep)];
    else k = -unit24(hash3(seed ^ 0x6a09e667u, gen, (uint32_t)i));
    keys[i] = k;
    vals[i] = i;
}

__global__ void k_gather_col(const float* __restrict__ pred, int stride, int col, const int* __restrict__ idx, int n,
                             float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pred[(size_t)idx[i] * stride + col];
}

// Spearman rank correlation (average ranks are not needed: ties have measure zero here)
static double spearman(const std::vector<float>& a, const std::vector<float>& b) {
    const size_t n = a.size();
    if (n < 3) return 0.0;
    auto ranks = [n](const std::vector<float>& x) {
        std::vector<int> o(n);
        for (size_t i = 0; i < n; i++) o[i] = (int)i;
        std::stable_sort(o.begin(), o.end(), [&](int p, int q) { return x[p] < x[q]; });
        std::vector<double> r(n);
        for (size_t i = 0; i < n; i++) r[o[i]] = (double)i;
        return r;
    };
    const std::vector<double> ra = ranks(a), rb = ranks(b);
    const double m = (n - 1) / 2.0;
    double sab = 0, saa = 0, sbb = 0;
    for (size_t i = 0; i < n; i++) {
        sab += (ra[i] - m) * (rb[i] - m);
        saa += (ra[i] - m) * (ra[i] - m);
        sbb += (rb[i] - m) * (rb[i] - m);
    }
    return sab / std::sqrt(saa * sbb);
}

// per-genome fitness statistics of the flown samples: mean (seed order), CVaR of the worst
// quarter, worst, oversize flag (all samples -100) -- the surrogate's training targets
__global__ void k_targets(const float* __restrict__ samples, int n, int k, float* __restrict__ tgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float row[32];
    float mean = 0.0f;
    bool over = true;
    for (int s = 0; s < k; s++) {
        const float x = samples[(size_t)i * k + s];
        mean += x / (float)k;
        over = over && x == -100.0f;
        int q = s;
        while (q > 0 && row[q - 1] > x) { row[q] = row[q - 1]; q--; }
        row[q] = x;
    }
    const int tail = (k + 3) / 4;
    float cv = 0.0f;
    for (int s = 0; s < tail; s++) cv += row[s] / (float)tail;
    tgt[(size_t)i * 4 + 0] = mean;
    tgt[(size_t)i * 4 + 1] = cv;
    tgt[(size_t)i * 4 + 2] = row[0];
    tgt[(size_t)i * 4 + 3] = over ? 1.0f : 0.0f;
}

// sort keys of one target column: the flyable rows' values, +inf for the others (sorted last)
__global__ void k_median_keys(const float* __restrict__ tgt, int n, int col, float* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = tgt[(size_t)i * 4 + 3] == 0.0f ? tgt[(size_t)i * 4 + col] : __builtin_inff();
}

// regression targets (heads 0..2) relative to their batch's offsets (target normalisation)
__global__ void k_shift_targets(float* __restrict__ tgt, int n, float o0, float o1, float o2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tgt[(size_t)i * 4 + 0] -= o0;
    tgt[(size_t)i * 4 + 1] -= o1;
    tgt[(size_t)i * 4 + 2] -= o2;
}

}  // namespace

static const float kDeltaWeights[3] = {1.0f, 1.0f, -1.0f};

struct pr_pool {
    int n_elite = 0, n_pool = 0, n_fly = 0, nseeds = 0, buf_cap = 0;
    float sigma = 0.05f;
    unsigned seed = 1;
    pr_gen* gen_eval = nullptr;
    float *d_elite = nullptr, *d_pool = nullptr, *d_fly = nullptr, *d_lo = nullptr, *d_hi = nullptr;
    float *d_pred = nullptr, *d_keys = nullptr, *d_keys2 = nullptr, *d_samples = nullptr, *d_tgt = nullptr;
    float *d_buf_g = nullptr, *d_buf_t = nullptr;
    float* d_escore = nullptr;  // [n_elite] simulated scores of the current elite
    float* d_med = nullptr;     // [2 n_pool] sort keys of the target medians
    int has_escore = 0;         // set by a generation, cleared by pr_pool_set_elite
    int parent_delta = 1;       // mutants' keys from the parent's simulated score (when known); 2: calibrated
    float cw[3] = {1.0f, 1.0f, -1.0f};  // key weights of (parent score, prediction, parent's prediction)
    int cw_fitted = 0;          // cw comes from a fit (mode 2) rather than the default
    float kw[3] = {kDeltaWeights[0], kDeltaWeights[1], kDeltaWeights[2]};  // mode 1's key weights
    std::vector<float> escore_h;  // host copy of the elite's scores the last selection used
    int *d_vals = nullptr, *d_vals2 = nullptr, *d_idx = nullptr;
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int buf_n = 0, buf_head = 0;
    int collect = 0;  // plain generations also feed the replay buffer
    int select_mode = 0;  // without a model: 0 fly the whole pool, 1 uniform subset, 2 parent-score subset
    float explore = 0.0f;  // fraction of the surrogate fly set drawn at random from the pool
    int norm_targets = 1;  // 1: each batch's regression targets relative to the batch's medians
    float tgt_off[3] = {0.0f, 0.0f, 0.0f};  // the last collected batch's offsets
    float best_holdout = -1e9f;
    std::vector<float> acc, samples, champion;
    std::vector<int> idx;
    float champion_score = -1e9f;
    const float* last_fly = nullptr;  // the last generation's fly set (d_pool or d_fly)
    int last_nf = 0, last_collected = 0;
};

#define PCK(x)                                                                                   \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) { g_pool_err = std::string(#x) + ": " + hipGetErrorString(e_); return (int)e_; } \
    } while (0)
#define PRC(x)                                                                                   \
    do {                                                                                         \
        int r_ = (x);                                                                            \
        if (r_) { g_pool_err = std::string(#x) + ": " + pr_last_error() + pr_mlp_last_error(); return r_; } \
    } while (0)

extern "C" {

const char* pr_pool_last_error(void) { return g_pool_err.c_str(); }

int pr_pool_create(pr_pool** out, int n_elite, int n_pool, int n_fly, int nseeds, int buf_cap, float sigma_frac,
                   unsigned seed) {
    if (!out || n_elite < 8 || n_pool < n_fly || n_fly < n_elite || nseeds < 1 || nseeds > 32) {
        g_pool_err = "pr_pool_create: need 8 <= n_elite <= n_fly <= n_pool and 1 <= nseeds <= 32";
        return -1;
    }
    pr_pool* p = new pr_pool;
    p->n_elite = n_elite; p->n_pool = n_pool; p->n_fly = n_fly; p->nseeds = nseeds;
    p->buf_cap = std::max(buf_cap, n_fly);
    p->sigma = sigma_frac;
    p->seed = seed;
    PRC(pr_gen_create(&p->gen_eval, n_pool, nseeds, 0, 0));
    const size_t G = sizeof(float) * pr::NG;
    PCK(hipMalloc(&p->d_elite, G * n_elite));
    PCK(hipMalloc(&p->d_escore, sizeof(float) * n_elite));
    PCK(hipMalloc(&p->d_pool, G * n_pool));
    PCK(hipMalloc(&p->d_fly, G * n_pool));
    PCK(hipMalloc(&p->d_lo, G));
    PCK(hipMalloc(&p->d_hi, G));
    PCK(hipMalloc(&p->d_pred, sizeof(float) * 5 * (size_t)n_pool));
    PCK(hipMalloc(&p->d_keys, sizeof(float) * n_pool));
    PCK(hipMalloc(&p->d_keys2, sizeof(float) * n_pool));
    PCK(hipMalloc(&p->d_vals, sizeof(int) * n_pool));
    PCK(hipMalloc(&p->d_vals2, sizeof(int) * n_pool));
    PCK(hipMalloc(&p->d_idx, sizeof(int) * n_pool));
    PCK(hipMalloc(&p->d_samples, sizeof(float) * (size_t)n_pool * nseeds));
    PCK(hipMalloc(&p->d_tgt, sizeof(float) * 4 * (size_t)n_pool));
    PCK(hipMalloc(&p->d_buf_g, G * p->buf_cap));
    PCK(hipMalloc(&p->d_buf_t, sizeof(float) * 4 * (size_t)p->buf_cap));
    PCK(hipMalloc(&p->d_med, sizeof(float) * 2 * (size_t)n_pool));
    PCK(hipcub::DeviceRadixSort::SortPairs(nullptr, p->sort_tmp_bytes, p->d_keys, p->d_keys2, p->d_vals, p->d_vals2,
                                           n_pool));
    size_t key_bytes = 0;
    PCK(hipcub::DeviceRadixSort::SortKeys(nullptr, key_bytes, p->d_med, p->d_med + n_pool, n_pool));
    p->sort_tmp_bytes = std::max(p->sort_tmp_bytes, key_bytes);
    PCK(hipMalloc(&p->d_sort_tmp, p->sort_tmp_bytes + 16));
    float lo[pr::NG], hi[pr::NG];
    pr_gene_bounds(lo, hi);
    PCK(hipMemcpy(p->d_lo, lo, G, hipMemcpyHostToDevice));
    PCK(hipMemcpy(p->d_hi, hi, G, hipMemcpyHostToDevice));
    p->champion.assign(pr::NG, 0.0f);
    *out = p;
    return 0;
}

void pr_pool_destroy(pr_pool* p) {
    if (!p) return;
    pr_gen_destroy(p->gen_eval);
    void* bufs[] = {p->d_elite, p->d_pool, p->d_fly, p->d_lo, p->d_hi, p->d_pred, p->d_keys, p->d_keys2,
                    p->d_vals, p->d_vals2, p->d_idx, p->d_samples, p->d_tgt, p->d_buf_g, p->d_buf_t, p->d_sort_tmp,
                    p->d_escore, p->d_med};
    for (void* b : bufs) if (b) (void)hipFree(b);
    delete p;
}

// initial elite (host genomes, n >= n_elite; the first n_elite are used)
int pr_pool_set_elite(pr_pool* p, const float* genomes, int n) {
    if (!p || n < p->n_elite) { g_pool_err = "pr_pool_set_elite: need n >= n_elite"; return -1; }
    PCK(hipMemcpy(p->d_elite, genomes, sizeof(float) * pr::NG * p->n_elite, hipMemcpyHostToDevice));
    p->has_escore = 0;
    p->escore_h.clear();
    p->cw_fitted = 0;
    return 0;
}

int pr_pool_get_elite(pr_pool* p, float* genomes, int n) {
    if (!p || n < p->n_elite) { g_pool_err = "pr_pool_get_elite: capacity"; return -1; }
    PCK(hipMemcpy(genomes, p->d_elite, sizeof(float) * pr::NG * p->n_elite, hipMemcpyDeviceToHost));
    return 0;
}

// Adds (genome, target) pairs (device) to the surrogate's replay buffer (ring).
static int buffer_add(pr_pool* p, const float* d_g, const float* d_t, int n) {
    for (int s = 0; s < n;) {
        const int k = std::min(n - s, p->buf_cap - p->buf_head);
        PCK(hipMemcpy(p->d_buf_g + (size_t)p->buf_head * pr::NG, d_g + (size_t)s * pr::NG, sizeof(float) * pr::NG * k,
                      hipMemcpyDeviceToDevice));
        PCK(hipMemcpy(p->d_buf_t + (size_t)p->buf_head * 4, d_t + (size_t)s * 4, sizeof(float) * 4 * k,
                      hipMemcpyDeviceToDevice));
        p->buf_head = (p->buf_head + k) % p->buf_cap;
        p->buf_n = std::min(p->buf_cap, p->buf_n + k);
        s += k;
    }
    // the network trains on its own (non-blocking) stream: the buffer must be complete first
    PCK(hipStreamSynchronize(nullptr));
    return 0;
}

// Target normalisation: every design of a batch flew the same scenarios, whose difficulty
// moves all of the batch's distances together (the same designs' robust scores differ by
// metres between scenario sets), so the regression targets (mean, CVaR, worst) are taken
// relative to the batch's medians over its flyable designs: the buffer then holds one
// frame across generations.  The selection keys use differences of predictions and rank
// within one generation, so they are unaffected by the offsets.
static int normalize_targets(pr_pool* p, float* d_t, int n, const float* samples) {
    p->tgt_off[0] = p->tgt_off[1] = p->tgt_off[2] = 0.0f;
    if (!p->norm_targets || n < 1) return 0;
    // flyable rows: not every sample -100 (the targets' oversize flag is 0 exactly there)
    int n_ok = 0;
    for (int i = 0; i < n; i++) {
        const float* r = samples + (size_t)i * p->nseeds;
        for (int k = 0; k < p->nseeds; k++)
            if (r[k] != -100.0f) { n_ok++; break; }
    }
    if (n_ok == 0) return 0;
    // each head's upper median over the flyable rows: element n_ok / 2 of its sorted values
    for (int k = 0; k < 3; k++) {
        hipLaunchKernelGGL(k_median_keys, dim3((n + 255) / 256), dim3(256), 0, 0, d_t, n, k, p->d_med);
        PCK(hipGetLastError());
        size_t tb = p->sort_tmp_bytes;
        PCK(hipcub::DeviceRadixSort::SortKeys(p->d_sort_tmp, tb, p->d_med, p->d_med + p->n_pool, n));
        PCK(hipMemcpy(&p->tgt_off[k], p->d_med + p->n_pool + n_ok / 2, sizeof(float), hipMemcpyDeviceToHost));
    }
    hipLaunchKernelGGL(k_shift_targets, dim3((n + 255) / 256), dim3(256), 0, 0, d_t, n, p->tgt_off[0], p->tgt_off[1],
                       p->tgt_off[2]);
    PCK(hipGetLastError());
    return 0;
}

// One generation.  model = null: plain (fly the whole pool); else pre-screen with the
// surrogate and train it on the flown pairs for train_epochs epochs over the buffer.
// out[0..13]: best simulated score, elite mean score, holdout score of the champion,
// wall ms, flight ms, surrogate ms (predict + train), flights flown, plane-flight-s,
// last training loss, fly-set size, best holdout score so far (the champion record),
// Spearman correlation of predicted and simulated scores over the fly set's exploration
// sample (NaN when plain; 0 when it has fewer than 3 members), the mean absolute error
// of the predicted robust score in metres over the sample's members that both the network
// and the simulation call flyable, and the accuracy of the network's oversize flag over
// the sample (NaN when plain or empty).
int pr_pool_generation(pr_pool* p, int gen, const pr_sim_params* prm, pr_mlp* model, int train_epochs, float lr,
                       double* out) {
    if (!p || !prm || !out) { g_pool_err = "pr_pool_generation: null argument"; return -1; }
    auto t0 = std::chrono::steady_clock::now();
    double mlp_ms = 0;
    hipLaunchKernelGGL(k_mutate, dim3((p->n_pool + 255) / 256), dim3(256), 0, 0, p->d_elite, p->n_elite, p->d_pool,
                       p->n_pool, p->d_lo, p->d_hi, p->sigma, p->seed, (uint32_t)gen);
    PCK(hipGetLastError());
    const float* fly = p->d_pool;
    int nf = p->n_pool;
    if (model) {
        PCK(hipDeviceSynchronize());
        auto m0 = std::chrono::steady_clock::now();
        PRC(pr_mlp_predict(model, p->d_pool, p->n_pool, p->d_pred));
        const float p_explore =
            p->n_pool > p->n_elite ? p->explore * (float)(p->n_fly - p->n_elite) / (float)(p->n_pool - p->n_elite) : 0.f;
        const float* escore = p->has_escore && p->parent_delta ? p->d_escore : nullptr;
        const float* w = p->parent_delta == 2 && p->cw_fitted ? p->cw : p->kw;
        hipLaunchKernelGGL(k_sort_keys, dim3((p->n_pool + 255) / 256), dim3(256), 0, 0, p->d_pred, 5, 4, p->n_pool,
                           p->n_elite, p_explore, p->seed, (uint32_t)gen, escore, w[0], w[1], w[2], p->d_keys,
                           p->d_vals);
        PCK(hipGetLastError());
        size_t tb = p->sort_tmp_bytes;
        PCK(hipcub::DeviceRadixSort::SortPairs(p->d_sort_tmp, tb, p->d_keys, p->d_keys2, p->d_vals, p->d_vals2,
                                               p->n_pool));
        nf = p->n_fly;
        hipLaunchKernelGGL(k_gather, dim3((nf * pr::NG + 255) / 256), dim3(256), 0, 0, p->d_pool, p->d_vals2, nf,
                           p->d_fly);
        PCK(hipGetLastError());
        PCK(hipDeviceSynchronize());
        mlp_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0).count();
        fly = p->d_fly;
    } else if (p->select_mode != 0) {
        // the fly budget chosen without the network (ablation of the pre-screen)
        const float* escore = p->has_escore ? p->d_escore : nullptr;
        hipLaunchKernelGGL(k_sort_keys_nonet, dim3((p->n_pool + 255) / 256), dim3(256), 0, 0, p->n_pool, p->n_elite,
                           p->select_mode, p->seed, (uint32_t)gen, escore, p->d_keys, p->d_vals);
        PCK(hipGetLastError());
        size_t tb = p->sort_tmp_bytes;
        PCK(hipcub::DeviceRadixSort::SortPairs(p->d_sort_tmp, tb, p->d_keys, p->d_keys2, p->d_vals, p->d_vals2,
                                               p->n_pool));
        nf = p->n_fly;
        hipLaunchKernelGGL(k_gather, dim3((nf * pr::NG + 255) / 256), dim3(256), 0, 0, p->d_pool, p->d_vals2, nf,
                           p->d_fly);
        PCK(hipGetLastError());
        fly = p->d_fly;
    }
    PCK(hipDeviceSynchronize());
    std::vector<int> seeds(p->nseeds);
    for (int s = 0; s < p->nseeds; s++) seeds[s] = gen * p->nseeds + s + 1;
    p->acc.resize(nf);
    p->idx.resize(nf);
    p->samples.resize((size_t)nf * p->nseeds);
    float holdout = 0;
    int champ = 0;
    const bool collect = model != nullptr || p->collect;
    const bool had_escore = p->has_escore != 0;  // this generation's keys used the elite's scores
    std::vector<float> escore_prev;
    PRC(pr_gen_evaluate(p->gen_eval, fly, nullptr, nf, seeds.data(), p->nseeds, 8, prm, p->acc.data(), p->idx.data(),
                        collect ? p->samples.data() : nullptr, &holdout, &champ));
    pr_stats st_main, st_hold;
    pr_gen_stats(p->gen_eval, &st_main, &st_hold);
    // new elite: best simulated scores
    PCK(hipMemcpy(p->d_idx, p->idx.data(), sizeof(int) * p->n_elite, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_gather, dim3((p->n_elite * pr::NG + 255) / 256), dim3(256), 0, 0, fly, p->d_idx, p->n_elite,
                       p->d_elite);
    PCK(hipGetLastError());
    {
        std::vector<float> es(p->n_elite);
        for (int k = 0; k < p->n_elite; k++) es[k] = p->acc[p->idx[k]];
        PCK(hipMemcpy(p->d_escore, es.data(), sizeof(float) * p->n_elite, hipMemcpyHostToDevice));
        escore_prev.swap(p->escore_h);  // the scores this generation's selection used
        p->escore_h = es;
        p->has_escore = 1;
    }
    PCK(hipMemcpy(p->champion.data(), fly + (size_t)champ * pr::NG, sizeof(float) * pr::NG, hipMemcpyDeviceToHost));
    p->champion_score = holdout;
    p->best_holdout = std::max(p->best_holdout, holdout);
    // training targets of the flown pairs (before the screening figures: the mean absolute
    // error compares predictions in the targets' frame)
    if (collect) {
        auto m0 = std::chrono::steady_clock::now();
        PCK(hipMemcpy(p->d_samples, p->samples.data(), sizeof(float) * p->samples.size(), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_targets, dim3((nf + 255) / 256), dim3(256), 0, 0, p->d_samples, nf, p->nseeds, p->d_tgt);
        PCK(hipGetLastError());
        int rc = normalize_targets(p, p->d_tgt, nf, p->samples.data());
        if (rc) return rc;
        mlp_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0).count();
    }
    // the predicted robust score's offset in that frame
    const float score_off = 0.65f * p->tgt_off[0] + 0.25f * p->tgt_off[1] + 0.10f * p->tgt_off[2];
    // screening quality: rank correlation of predicted and simulated scores over the
    // exploration sample of the fly set (uniform draws from the pool, so unbiased by the
    // selection; the surrogate-selected part over-represents over-predicted candidates)
    double rho = std::nan(""), mae = std::nan(""), acc_over = std::nan("");
    int fit_ok = 0;
    float fit_w[3] = {0.0f, 0.0f, 0.0f};
    if (model) {
        hipLaunchKernelGGL(k_gather_col, dim3((nf + 255) / 256), dim3(256), 0, 0, p->d_pred, 5, 4, p->d_vals2, nf,
                           p->d_keys);
        PCK(hipGetLastError());
        std::vector<float> pf(nf), key(nf);
        PCK(hipMemcpy(pf.data(), p->d_keys, sizeof(float) * nf, hipMemcpyDeviceToHost));
        PCK(hipMemcpy(key.data(), p->d_keys2, sizeof(float) * nf, hipMemcpyDeviceToHost));
        std::vector<float> a, b;
        for (int i = 0; i < nf; i++)
            if (key[i] == -2.0e38f) { a.push_back(pf[i]); b.push_back(p->acc[i]); }
        rho = spearman(a, b);
        // fit of the selection key (mode 2): over the exploration sample, least squares of the
        // mutants' simulated scores on (the parent's score the key used, the mutant's predicted
        // score, the parent's predicted score) with an intercept; flyable members only
        fit_ok = 0;
        if (had_escore && (int)escore_prev.size() == p->n_elite) {
            std::vector<int> pidx(nf);
            PCK(hipMemcpy(pidx.data(), p->d_vals2, sizeof(int) * nf, hipMemcpyDeviceToHost));
            std::vector<float> pp(p->n_elite);
            hipLaunchKernelGGL(k_gather_col, dim3((p->n_elite + 255) / 256), dim3(256), 0, 0, p->d_pred, 5, 4,
                               p->d_vals2, p->n_elite, p->d_keys);
            PCK(hipGetLastError());
            PCK(hipMemcpy(pp.data(), p->d_keys, sizeof(float) * p->n_elite, hipMemcpyDeviceToHost));
            double A[4][5] = {};
            long long m = 0;
            for (int i = p->n_elite; i < nf; i++) {
                if (key[i] != -2.0e38f) continue;
                const int par = (int)(host_hash3(p->seed, (uint32_t)gen, (uint32_t)pidx[i]) % (uint32_t)p->n_elite);
                const float xe = escore_prev[par], xc = pf[i], xp = pp[par], y = p->acc[i];
                if (xe == -100.0f || xc == -100.0f || xp == -100.0f || y == -100.0f) continue;
                const double v[5] = {1.0, xe, xc, xp, y};
                for (int r = 0; r < 4; r++)
                    for (int c = 0; c < 5; c++) A[r][c] += v[r] * v[c];
                m++;
            }
            // solve the 4x4 normal equations (Gaussian elimination, partial pivoting)
            if (m >= 256) {
                bool ok = true;
                for (int c = 0; c < 4 && ok; c++) {
                    int piv = c;
                    for (int r = c + 1; r < 4; r++) if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
                    if (std::fabs(A[piv][c]) < 1e-12 * (1.0 + std::fabs(A[0][0]))) { ok = false; break; }
                    for (int k = 0; k < 5; k++) std::swap(A[c][k], A[piv][k]);
                    for (int r = 0; r < 4; r++) {
                        if (r == c) continue;
                        const double f = A[r][c] / A[c][c];
                        for (int k = c; k < 5; k++) A[r][k] -= f * A[c][k];
                    }
                }
                if (ok) {
                    for (int k = 0; k < 3; k++) fit_w[k] = (float)(A[k + 1][4] / A[k + 1][k + 1]);
                    fit_ok = 1;
                    if (p->parent_delta == 2) {
                        for (int k = 0; k < 3; k++) p->cw[k] = fit_w[k];
                        p->cw_fitted = 1;
                    }
                }
            }
        }
        // the predicted score is -100 where the oversize head says so: regression error over
        // the members both call flyable, and the oversize flag's accuracy over all of them
        if (!a.empty()) {
            double e = 0;
            int n_fly = 0, agree = 0;
            for (size_t i = 0; i < a.size(); i++) {
                const bool pa = a[i] == -100.0f, pb = b[i] == -100.0f;
                agree += pa == pb;
                if (!pa && !pb) { e += std::fabs((double)a[i] + (double)score_off - (double)b[i]); n_fly++; }
            }
            if (n_fly) mae = e / n_fly;
            acc_over = (double)agree / (double)a.size();
        }
    }
    float loss = 0;
    if (collect) {
        auto m0 = std::chrono::steady_clock::now();
        int rc = buffer_add(p, fly, p->d_tgt, nf);
        if (rc) return rc;
        if (model && train_epochs > 0)
            PRC(pr_mlp_train(model, p->d_buf_g, p->d_buf_t, p->buf_n, train_epochs, pr_mlp_max_batch(model), lr, &loss));
        PCK(hipDeviceSynchronize());
        mlp_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0).count();
    }
    p->last_fly = fly;
    p->last_nf = nf;
    p->last_collected = collect ? 1 : 0;
    double mean = 0;
    for (int i = 0; i < p->n_elite; i++) mean += p->acc[p->idx[i]];
    out[0] = p->acc[p->idx[0]];
    out[1] = mean / p->n_elite;
    out[2] = holdout;
    out[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out[4] = st_main.fly_ms + st_hold.fly_ms;
    out[5] = mlp_ms;
    out[6] = (double)nf * p->nseeds + 8.0 * 128;
    out[7] = (double)(st_main.flight_steps + st_hold.flight_steps) / 600.0;
    out[8] = loss;
    out[9] = nf;
    out[10] = p->best_holdout;
    out[11] = rho;
    out[12] = mae;
    out[13] = acc_over;
    for (int k = 0; k < 3; k++) out[14 + k] = fit_ok ? (double)fit_w[k] : std::nan("");
    return 0;
}

// Seeds the surrogate's replay buffer with already-flown pairs (device genomes [n x 35],
// host samples [n x nseeds]) -- e.g. the flights of earlier generations.
int pr_pool_add_training(pr_pool* p, const float* d_genomes, const float* samples, int n) {
    if (!p || n < 1 || n > p->n_pool) { g_pool_err = "pr_pool_add_training: 1 <= n <= n_pool"; return -1; }
    PCK(hipMemcpy(p->d_samples, samples, sizeof(float) * (size_t)n * p->nseeds, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_targets, dim3((n + 255) / 256), dim3(256), 0, 0, p->d_samples, n, p->nseeds, p->d_tgt);
    PCK(hipGetLastError());
    int rc = normalize_targets(p, p->d_tgt, n, samples);
    if (rc) return rc;
    return buffer_add(p, d_genomes, p->d_tgt, n);
}

// Fraction (0..1) of the surrogate fly set beyond the elite drawn uniformly from the pool
// instead of by predicted score: exploration, and unbiased training pairs for the network.
int pr_pool_set_explore(pr_pool* p, float frac) {
    if (!p || !(frac >= 0.0f && frac <= 1.0f)) { g_pool_err = "pr_pool_set_explore: fraction in [0, 1]"; return -1; }
    p->explore = frac;
    return 0;
}

int pr_pool_set_parent_delta(pr_pool* p, int on) {
    if (!p || on < 0 || on > 2) { g_pool_err = "pr_pool_set_parent_delta: null pool or mode outside 0..2"; return -1; }
    p->parent_delta = on;
    return 0;
}

int pr_pool_set_key_weights(pr_pool* p, float w_parent, float w_pred, float w_parent_pred) {
    if (!p || !std::isfinite(w_parent) || !std::isfinite(w_pred) || !std::isfinite(w_parent_pred)) {
        g_pool_err = "pr_pool_set_key_weights: null pool or non-finite weight";
        return -1;
    }
    p->kw[0] = w_parent; p->kw[1] = w_pred; p->kw[2] = w_parent_pred;
    return 0;
}

int pr_pool_set_selection(pr_pool* p, int mode) {
    if (!p || mode < 0 || mode > 2) { g_pool_err = "pr_pool_set_selection: mode 0, 1 or 2"; return -1; }
    p->select_mode = mode;
    return 0;
}

// 1: plain generations (model = null) also add their flown pairs to the replay buffer,
// e.g. to bootstrap the surrogate before pre-screening starts.
int pr_pool_set_collect(pr_pool* p, int on) {
    if (!p) { g_pool_err = "pr_pool_set_collect: null pool"; return -1; }
    p->collect = on ? 1 : 0;
    return 0;
}

// 1 (default): each collected batch's regression targets relative to the batch's medians
// (see normalize_targets); 0: absolute targets.
int pr_pool_set_target_norm(pr_pool* p, int on) {
    if (!p) { g_pool_err = "pr_pool_set_target_norm: null pool"; return -1; }
    p->norm_targets = on ? 1 : 0;
    return 0;
}

int pr_pool_train(pr_pool* p, pr_mlp* model, int epochs, float lr, float* loss) {
    if (!p || !model || p->buf_n < 32) { g_pool_err = "pr_pool_train: model and >= 32 buffered pairs needed"; return -1; }
    PCK(hipStreamSynchronize(nullptr));  // buffer writes on the null stream, training on the network's stream
    PRC(pr_mlp_train(model, p->d_buf_g, p->d_buf_t, p->buf_n, epochs, pr_mlp_max_batch(model), lr, loss));
    return 0;
}

// The last generation's pool [n_pool x 35], fly set [n_fly x 35] (in flight order),
// simulated robust scores of the fly set [n_fly] and, when it collected training pairs,
// their targets [n_fly x 4] (mean, CVaR25, worst, oversize); null outputs are skipped.
int pr_pool_last(pr_pool* p, float* pool, float* fly, float* scores, float* targets, int* n_fly) {
    if (!p || !p->last_fly) { g_pool_err = "pr_pool_last: no generation has run"; return -1; }
    if (targets && !p->last_collected) { g_pool_err = "pr_pool_last: the last generation collected no targets"; return -1; }
    const int nf = p->last_nf;
    if (pool) PCK(hipMemcpy(pool, p->d_pool, sizeof(float) * pr::NG * (size_t)p->n_pool, hipMemcpyDeviceToHost));
    if (fly) PCK(hipMemcpy(fly, p->last_fly, sizeof(float) * pr::NG * (size_t)nf, hipMemcpyDeviceToHost));
    if (scores) std::memcpy(scores, p->acc.data(), sizeof(float) * nf);
    if (targets) PCK(hipMemcpy(targets, p->d_tgt, sizeof(float) * 4 * (size_t)nf, hipMemcpyDeviceToHost));
    if (n_fly) *n_fly = nf;
    return 0;
}

int pr_pool_champion(pr_pool* p, float* genome35, float* holdout) {
    if (!p) return -1;
    std::memcpy(genome35, p->champion.data(), sizeof(float) * pr::NG);
    if (holdout) *holdout = p->champion_score;
    return 0;
}

}  // extern "C"
