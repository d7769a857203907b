// ga.cpp -- host side of the fitness path: generation evaluation (robust scores, ranking,
// top-8 holdout re-test), the GA loop and the audit/finalist drivers, i.e. the callers
// of the GPU path in the reference's sim.cu:250-520, rebuilt over the batched C ABI.
//
// Bit-exact with the reference where it is integer/ordering work: std::sort with the
// reference's comparator, libstdc++ mt19937 / uniform_real_distribution<float> /
// normal_distribution<float> drawn in the reference's call order (sim.cu:318-321,398,
// 429-430,472-490), float accumulations in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <random>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/planes_hip.h"
#include "../../include/planes_sim.h"

namespace {

const int NG = 35;
// per-gene bounds (common.h:98-103): the genome format's clamp table
const float LO[NG] = {0.12f, 0.05f, 0.04f, 0.00f, -0.30f, -1.20f, 0.45f, -0.04f, -0.06f, -0.20f, -0.20f, 0.0f,
                      0.005f, 0.000f, 0.50f, 0.06f, 0.004f, 0.12f, 0.00f, 0.12f, 0.10f, -0.16f, -0.80f, -0.80f,
                      -0.65f, -0.65f, 0.00f, 0.00f, 0.10f, 0.15f, -0.40f, 0.15f, -0.50f, 0.00f, -0.25f};
const float HI[NG] = {0.25f, 0.20f, 1.00f, 0.90f, 0.60f, 1.20f, 0.90f, 0.10f, 0.06f, 0.08f, 0.20f, 1.5f,
                      0.050f, 0.060f, 1.00f, 0.30f, 0.065f, 0.65f, 0.85f, 0.65f, 0.45f, 0.16f, 0.80f, 0.80f,
                      1.00f, 1.00f, 1.00f, 0.75f, 1.00f, 1.00f, 1.00f, 1.00f, 0.50f, 1.00f, 0.80f};

void die(const char* what) {
    fprintf(stderr, "planes_hip: %s: %s\n", what, pr_last_error());
    exit(1);
}
#define CKP(x) do { if ((x) != 0) die(#x); } while (0)

inline void gene_set(float* g, int i, float v) { g[i] = std::fmin(HI[i], std::fmax(LO[i], v)); }

// robust score of one row of n samples (sim.cu:340-357)
float robust_row(const float* r, int n, std::vector<float>& row) {
    row.assign(r, r + n);
    float mean = 0;
    for (int s = 0; s < n; s++) mean += row[s] / n;
    std::sort(row.begin(), row.end());
    int tail = (n + 3) / 4;
    float cv = 0;
    for (int s = 0; s < tail; s++) cv += row[s] / tail;
    return 0.65f * mean + 0.25f * cv + 0.10f * row[0];
}

float percentile(const std::vector<float>& sorted, float q) {  // sim.cu:250-256
#pragma clang fp contract(off)
    if (sorted.empty()) return 0;
    float x = q * (sorted.size() - 1);
    int a = (int)floorf(x), b = (int)ceilf(x);
    float t = x - a;
    return sorted[a] * (1.0f - t) + sorted[b] * t;
}

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    void ensure(size_t bytes) {
        if (bytes <= n) return;
        if (p) pr_dev_free(p);
        CKP(pr_dev_alloc(&p, bytes));
        n = bytes;
    }
    ~DevBuf() { if (p) pr_dev_free(p); }
};

}  // namespace

// ------------------------------------------------------------------ GA RNG state
struct pr_ga {
    std::mt19937 rng;
    std::uniform_real_distribution<float> u01{0, 1};
    std::normal_distribution<float> nrm{0, 1};
};

// ------------------------------------------------------------------ generation evaluator
struct pr_gen {
    pr_ctx* main = nullptr;
    int pop = 0, nseeds_max = 0, hold_n = 128;
    DevBuf seeds, fit, steps, scores, hseeds;
    // two holdout slots, so that one generation's 8 x 128 holdout batch can fly on its own
    // stream while the next generation's population flies (pr_gen_evaluate_async)
    struct Hold {
        pr_ctx* ctx = nullptr;
        DevBuf gen, fit, steps, scores;
        int pending = 0, k = 0;
        int idx[8] = {};
    } hold[2];
    int next_slot = 0;       // slot the next holdout batch goes to
    int oldest = -1;         // oldest uncollected slot (-1: none)
    std::vector<float> samples, acc_tmp;
    pr_stats st_main = {}, st_hold = {};
    int last_pop = 0, last_nseeds = 0;  // shape of the last population batch (its samples stay in fit)
};

extern "C" {

void pr_gene_bounds(float* lo, float* hi) {
    for (int i = 0; i < NG; i++) { lo[i] = LO[i]; hi[i] = HI[i]; }
}

void* pr_ga_new(unsigned seed) {
    pr_ga* g = new pr_ga;
    g->rng.seed(seed);
    return g;
}
void pr_ga_free(void* h) { delete (pr_ga*)h; }

// grand() (sim.cu:318-321): a fresh uniform_real_distribution per genome, like the reference
void pr_ga_random(void* h, float* genomes, int n) {
    pr_ga* G = (pr_ga*)h;
    for (int i = 0; i < n; i++) {
        std::uniform_real_distribution<float> u(0, 1);
        for (int j = 0; j < NG; j++) gene_set(genomes + (size_t)NG * i, j, LO[j] + u(G->rng) * (HI[j] - LO[j]));
    }
}

void pr_ga_sanitize(float* genomes, int n) {  // sim.cu:323-325
    for (int i = 0; i < n; i++)
        for (int j = 0; j < NG; j++) gene_set(genomes + (size_t)NG * i, j, genomes[(size_t)NG * i + j]);
}

}  // extern "C"

// ------------------------------------------------------------------ ranking
// The reference ranks with std::sort of the indices by acc[a] > acc[b] (sim.cu:440-441);
// with equal scores the order is whatever libstdc++'s introsort makes of them, and breeding
// reads it (elite rows, tournament entrants by rank).  The ranking here runs the same
// introsort steps -- median of (second, middle, last) moved to the front, an unguarded
// partition around it, the right part first, ranges of at most 16 left to one final
// insertion sort, heapsort (std::partial_sort over the whole range) where the depth limit
// 2 floor(log2 n) runs out -- so it makes the same moves and returns the same permutation,
// but the two sides of a partition are disjoint, so large right parts go to other threads.
// The tests compare it with std::sort of the indices on tied and untied data.
namespace {
using RankItem = std::pair<float, int>;
inline bool rank_before(const RankItem& a, const RankItem& b) { return a.first > b.first; }

void rank_median_to_first(RankItem* r, RankItem* a, RankItem* b, RankItem* c) {
    if (rank_before(*a, *b)) {
        if (rank_before(*b, *c)) std::iter_swap(r, b);
        else if (rank_before(*a, *c)) std::iter_swap(r, c);
        else std::iter_swap(r, a);
    } else if (rank_before(*a, *c)) {
        std::iter_swap(r, a);
    } else if (rank_before(*b, *c)) {
        std::iter_swap(r, c);
    } else {
        std::iter_swap(r, b);
    }
}

RankItem* rank_partition(RankItem* first, RankItem* last) {
    rank_median_to_first(first, first + 1, first + (last - first) / 2, last - 1);
    RankItem* lo = first + 1;
    RankItem* hi = last;
    const RankItem& pivot = *first;
    for (;;) {
        while (rank_before(*lo, pivot)) ++lo;
        --hi;
        while (rank_before(pivot, *hi)) --hi;
        if (!(lo < hi)) return lo;
        std::iter_swap(lo, hi);
        ++lo;
    }
}

// Sorts [first, last).  threads: how many more threads this call may start for right parts
// (shared with the calls it starts, which receive part of it).  Every range is left sorted
// by the call that owns it: the partitions order the ranges among themselves, so the final
// insertion sort moves every element only within its leaf range (at most 16 elements, or a
// heapsorted range, where it moves nothing) and runs here leaf by leaf.
void rank_introsort(RankItem* first, RankItem* last, int depth, int threads) {
    std::vector<std::thread> kids;
    kids.reserve(threads > 0 ? threads : 0);  // each start uses at least one of them
    while (last - first > 16) {
        if (depth == 0) {
            std::partial_sort(first, last, last, rank_before);
            break;
        }
        --depth;
        RankItem* cut = rank_partition(first, last);
        bool spawned = false;
        if (threads > 0 && last - cut > 8192) {
            // the threads (this one included) split between the two sides by their sizes
            const long long all = threads + 1;
            const int right = (int)std::min(all - 1, std::max(1LL, (all * (last - cut) + (last - first) / 2) / (last - first)));
            try {
                kids.emplace_back(rank_introsort, cut, last, depth, right - 1);
                threads -= right;
                spawned = true;
            } catch (const std::system_error&) {
                threads = 0;
            }
        }
        if (!spawned) rank_introsort(cut, last, depth, 0);
        last = cut;
    }
    for (RankItem* i = first + 1; i < last; i++) {  // the leaf's insertion sort
        const RankItem v = *i;
        RankItem* j = i;
        while (j > first && rank_before(v, *(j - 1))) {
            *j = *(j - 1);
            j--;
        }
        *j = v;
    }
    for (auto& k : kids) k.join();
}

void rank_sort(RankItem* a, int n) {
    if (n < 2) return;
    int lg = 0;
    while ((2LL << lg) <= n) lg++;
    const unsigned hw = std::thread::hardware_concurrency();
    const int threads = n > 32768 ? (int)std::min(15u, hw > 1 ? hw - 1 : 0u) : 0;
    rank_introsort(a, a + n, 2 * lg, threads);
}
}  // namespace

extern "C" {

void pr_ga_rank(const float* acc, int pop, int* idx) {
    std::vector<RankItem> v((size_t)std::max(pop, 0));
    for (int i = 0; i < pop; i++) v[i] = {acc[i], i};
    rank_sort(v.data(), pop);
    for (int i = 0; i < pop; i++) idx[i] = v[i].second;
}

// breeding (sim.cu:472-490): elite copy, 5% immigrants, 2x tournament-of-4, blend, annealed
// Gaussian mutation, clamp
void pr_ga_breed(void* h, const float* gs, const float* acc, const int* idx, int pop, int gen, int gens, float* next) {
    pr_ga* G = (pr_ga*)h;
    int elite = pop / 64;
    float anneal = 1.0f - 0.6f * gen / gens;
    for (int i = 0; i < elite; i++) std::memcpy(next + (size_t)NG * i, gs + (size_t)NG * idx[i], sizeof(float) * NG);
    for (int i = elite; i < pop; i++) {
        float* out = next + (size_t)NG * i;
        if (G->u01(G->rng) < 0.05f) { pr_ga_random(h, out, 1); continue; }
        auto tourn = [&]() {
            int b = idx[(int)(G->u01(G->rng) * pop)];
            for (int k = 0; k < 3; k++) {
                int c = (int)(G->u01(G->rng) * pop);
                if (acc[c] > acc[b]) b = c;
            }
            return b;
        };
        int pa = tourn();
        int pc = tourn();
        for (int j = 0; j < NG; j++) {
            float a = -0.1f + 1.2f * G->u01(G->rng);
            float v = gs[(size_t)NG * pa + j] * a + gs[(size_t)NG * pc + j] * (1 - a);
            if (G->u01(G->rng) < 0.25f) v += G->nrm(G->rng) * 0.08f * (HI[j] - LO[j]) * anneal;
            gene_set(out, j, v);
        }
    }
}

float pr_robust_score_host(const float* row, int n) {
    std::vector<float> tmp;
    return robust_row(row, n, tmp);
}

int pr_gen_create(pr_gen** out, int pop, int nseeds_max, int max_ch, int max_sp) {
    if (!out || pop < 1 || nseeds_max < 1) return -1;
    pr_gen* e = new pr_gen;
    e->pop = pop;
    e->nseeds_max = nseeds_max;
    int rc = pr_create(&e->main, pop, 0, nseeds_max, pop * nseeds_max, max_ch, max_sp);
    for (int s = 0; s < 2 && !rc; s++) rc = pr_create(&e->hold[s].ctx, 8, 0, e->hold_n, 8 * e->hold_n, max_ch, max_sp);
    // The holdout's small late chunks keep one wave per flight: a pipelined holdout shares
    // the GPU with the next generation's population flights, where the small-chunk mapping's
    // extra instructions (its rigid-body step runs in every wave of a flight) slowed the
    // population more than they sped the holdout up (100k x 12 generation -3 %; 10k x 8 +1 %).
    for (int s = 0; s < 2 && !rc; s++) rc = pr_set_split(e->hold[s].ctx, 0);
    if (rc) { pr_gen_destroy(e); return rc; }
    e->seeds.ensure(sizeof(int) * nseeds_max);
    e->fit.ensure(sizeof(float) * pop * nseeds_max);
    e->steps.ensure(sizeof(int) * pop * nseeds_max);
    e->scores.ensure(sizeof(float) * pop);
    e->hseeds.ensure(sizeof(int) * e->hold_n);
    for (auto& h : e->hold) {
        h.gen.ensure(sizeof(float) * NG * 8);
        h.fit.ensure(sizeof(float) * 8 * e->hold_n);
        h.steps.ensure(sizeof(int) * 8 * e->hold_n);
        h.scores.ensure(sizeof(float) * 8);
    }
    std::vector<int> hs(e->hold_n);
    for (int p = 0; p < e->hold_n; p++) hs[p] = 100003 + p * 7919;  // holdout suite (sim.cu:64,452)
    if ((rc = pr_h2d(e->hseeds.p, hs.data(), sizeof(int) * e->hold_n))) { pr_gen_destroy(e); return rc; }
    *out = e;
    return 0;
}

// The evaluator's flight contexts (0: population flights, 1 and 2: the two holdout slots),
// for the pr_set_* options (fused kernel, tail mapping, buckets, graph, prepare precision).
pr_ctx* pr_gen_context(pr_gen* e, int which) {
    return !e || which < 0 || which > 2 ? nullptr : which == 0 ? e->main : e->hold[which - 1].ctx;
}

void pr_gen_destroy(pr_gen* e) {
    if (!e) return;
    pr_destroy(e->main);
    for (auto& h : e->hold) pr_destroy(h.ctx);
    delete e;
}

// Population part of a generation (sim.cu:432-441): every genome x every seed in one batch,
// robust scores, std::sort ranking on the host.
static int gen_main(pr_gen* e, const float* d_genomes, const int* d_res, int pop, const int* seeds, int nseeds,
                    const pr_sim_params* p, float* acc, int* idx, float* samples) {
    int rc;
    if ((rc = pr_h2d(e->seeds.p, seeds, sizeof(int) * nseeds))) return rc;
    if ((rc = pr_prepare(e->main, d_genomes, d_res, pop))) return rc;
    if ((rc = pr_scenarios(e->main, (const int*)e->seeds.p, nseeds, p))) return rc;
    if ((rc = pr_fly(e->main, 0, pop, nseeds, p, (float*)e->fit.p, (int*)e->steps.p))) return rc;
    pr_stats_get(e->main, &e->st_main);
    if (nseeds <= 256) {
        if ((rc = pr_scores(e->main, (const float*)e->fit.p, pop, nseeds, (float*)e->scores.p))) return rc;
        if ((rc = pr_d2h(acc, e->scores.p, sizeof(float) * pop))) return rc;
        if (samples && (rc = pr_d2h(samples, e->fit.p, sizeof(float) * pop * nseeds))) return rc;
    } else {
        e->samples.resize((size_t)pop * nseeds);
        if ((rc = pr_d2h(e->samples.data(), e->fit.p, sizeof(float) * pop * nseeds))) return rc;
        std::vector<float> tmp;
        for (int i = 0; i < pop; i++) acc[i] = robust_row(e->samples.data() + (size_t)i * nseeds, nseeds, tmp);
        if (samples) std::memcpy(samples, e->samples.data(), sizeof(float) * pop * nseeds);
    }
    e->last_pop = pop;
    e->last_nseeds = nseeds;
    return 0;
}

// Holdout re-test (sim.cu:446-455) of the top min(holdout_k, 8) genomes on the fixed
// 128-scenario suite (seed 100003 with the vary rule), launched on slot s's own stream.
// The genomes are copied out before returning, so the caller may overwrite d_genomes.
static int hold_launch(pr_gen* e, int s, const float* d_genomes, const int* idx, int k, const pr_sim_params* p) {
    pr_gen::Hold& h = e->hold[s];
    int rc;
    for (int c = 0; c < k; c++)
        if ((rc = hipMemcpyAsync((char*)h.gen.p + sizeof(float) * NG * c, d_genomes + (size_t)NG * idx[c],
                                 sizeof(float) * NG, hipMemcpyDeviceToDevice, 0)))
            return rc;
    if ((rc = hipStreamSynchronize(0))) return rc;
    pr_sim_params hp = *p;
    hp.spread = 0;
    if ((rc = pr_prepare(h.ctx, (const float*)h.gen.p, nullptr, k))) return rc;
    if ((rc = pr_scenarios(h.ctx, (const int*)e->hseeds.p, e->hold_n, &hp))) return rc;
    if ((rc = pr_fly_async(h.ctx, 0, k, e->hold_n, &hp, (float*)h.fit.p, (int*)h.steps.p))) return rc;
    h.pending = 1;
    h.k = k;
    for (int c = 0; c < k; c++) h.idx[c] = idx[c];
    return 0;
}

// The first k (<= 8) places of the ranking without the full sort, when they are forced: if
// the k + 1 largest scores are strictly decreasing (no tie among them) and no score is NaN,
// every sort by acc[a] > acc[b] puts exactly these genomes, in this order, first.  false
// otherwise (ties, NaN, pop <= k): then only the full ranking decides.
static bool forced_top(const float* acc, int pop, int k, int* top) {
    if (k < 1 || k > 8 || pop <= k) return false;
    float v[9];
    int ix[9];
    int m = 0;
    for (int i = 0; i < pop; i++) {
        const float x = acc[i];
        if (x != x) return false;
        if (m == k + 1 && !(x > v[k])) continue;
        int j = m < k + 1 ? m++ : k;  // a new slot, or the smallest kept one when full
        while (j > 0 && x > v[j - 1]) {
            v[j] = v[j - 1];
            ix[j] = ix[j - 1];
            j--;
        }
        v[j] = x;
        ix[j] = i;
    }
    for (int j = 0; j < k; j++)
        if (!(v[j] > v[j + 1])) return false;
    for (int j = 0; j < k; j++) top[j] = ix[j];
    return true;
}

// Ranking (sim.cu:440-441) and the holdout launch of its top k.  When forced_top fixes the
// top k, the full std::sort ranking runs on a worker thread (host memory only) while this
// thread launches the holdout, so the host's launch work is off the ranking's path.
static int rank_and_hold(pr_gen* e, int s, const float* d_genomes, const float* acc, int pop, int* idx, int k,
                         const pr_sim_params* p) {
    int top[8];
    if (!forced_top(acc, pop, k, top)) {
        pr_ga_rank(acc, pop, idx);
        return k > 0 ? hold_launch(e, s, d_genomes, idx, k, p) : 0;
    }
    std::thread w;
    try {
        w = std::thread(pr_ga_rank, acc, pop, idx);
    } catch (const std::system_error&) {
        pr_ga_rank(acc, pop, idx);
    }
    const int rc = hold_launch(e, s, d_genomes, top, k, p);
    if (w.joinable()) w.join();
    return rc;
}

// Waits for slot s's holdout batch; champion = the first strict maximum of the holdout
// robust scores (sim.cu:447-450), as a genome index of that generation.
static int hold_collect(pr_gen* e, int s, float* holdout_out, int* champion_out) {
    pr_gen::Hold& h = e->hold[s];
    int rc;
    if ((rc = pr_fly_wait(h.ctx))) return rc;
    h.pending = 0;
    pr_stats_get(h.ctx, &e->st_hold);
    if ((rc = pr_scores(h.ctx, (const float*)h.fit.p, h.k, e->hold_n, (float*)h.scores.p))) return rc;
    float hsc[8];
    if ((rc = pr_d2h(hsc, h.scores.p, sizeof(float) * h.k))) return rc;
    float best = -1e9f;
    int champ = h.idx[0];
    for (int c = 0; c < h.k; c++)
        if (hsc[c] > best) { best = hsc[c]; champ = h.idx[c]; }
    if (holdout_out) *holdout_out = best;
    if (champion_out) *champion_out = champ;
    return 0;
}

int pr_gen_evaluate(pr_gen* e, const float* d_genomes, const int* d_res, int pop, const int* seeds, int nseeds,
                    int holdout_k, const pr_sim_params* p, float* acc, int* idx, float* samples,
                    float* holdout_out, int* champion_out) {
    if (!e || pop < 1 || pop > e->pop || nseeds < 1 || nseeds > e->nseeds_max || !acc || !idx || !p) return -1;
    if (e->oldest >= 0) return -2;  // an asynchronous holdout batch is still uncollected
    int rc = gen_main(e, d_genomes, d_res, pop, seeds, nseeds, p, acc, idx, samples);
    if (rc) return rc;
    const int k = std::min(std::min(holdout_k, pop), 8);
    if ((rc = rank_and_hold(e, 0, d_genomes, acc, pop, idx, k, p))) return rc;
    float holdout = -1e9f;
    int champion = idx[0];
    if (k > 0 && (rc = hold_collect(e, 0, &holdout, &champion))) return rc;
    if (holdout_out) *holdout_out = holdout;
    if (champion_out) *champion_out = champion;
    return 0;
}

// Pipelined variant: the population part as pr_gen_evaluate, then the holdout batch is
// launched on its own stream and the call returns; pr_gen_collect returns its result. At
// most two holdout batches may be outstanding, so generation g+1's population flights can
// overlap generation g's holdout. Returns -2 when both slots are still uncollected.
int pr_gen_evaluate_async(pr_gen* e, const float* d_genomes, const int* d_res, int pop, const int* seeds,
                          int nseeds, int holdout_k, const pr_sim_params* p, float* acc, int* idx, float* samples) {
    if (!e || pop < 1 || pop > e->pop || nseeds < 1 || nseeds > e->nseeds_max || !acc || !idx || !p) return -1;
    const int s = e->next_slot;
    if (e->hold[s].pending) return -2;
    int rc = gen_main(e, d_genomes, d_res, pop, seeds, nseeds, p, acc, idx, samples);
    if (rc) return rc;
    const int k = std::min(std::min(holdout_k, pop), 8);
    e->hold[s].k = 0;
    if ((rc = rank_and_hold(e, s, d_genomes, acc, pop, idx, k, p))) return rc;
    if (k <= 0) {  // nothing to fly: collect returns -1e9 and idx[0]
        e->hold[s].idx[0] = idx[0];
        e->hold[s].pending = 2;
    }
    if (e->oldest < 0) e->oldest = s;
    e->next_slot = s ^ 1;
    return 0;
}

// Samples of genomes [first, first + count) of the last population batch, row-major by seed;
// they stay on the device until the next evaluation, so reading them costs the evaluation nothing.
int pr_gen_read_samples(pr_gen* e, int first, int count, float* out) {
    if (!e || !out || first < 0 || count < 0 || e->last_pop == 0 || first + count > e->last_pop) return -1;
    const size_t ns = (size_t)e->last_nseeds;
    return pr_d2h(out, (const float*)e->fit.p + (size_t)first * ns, sizeof(float) * (size_t)count * ns);
}

// Result of the oldest outstanding pr_gen_evaluate_async holdout batch (waits for it).
int pr_gen_collect(pr_gen* e, float* holdout_out, int* champion_out) {
    if (!e || e->oldest < 0) return -1;
    const int s = e->oldest;
    pr_gen::Hold& h = e->hold[s];
    int rc = 0;
    if (h.pending == 2) {
        h.pending = 0;
        if (holdout_out) *holdout_out = -1e9f;
        if (champion_out) *champion_out = h.idx[0];
    } else if ((rc = hold_collect(e, s, holdout_out, champion_out))) {
        return rc;
    }
    e->oldest = e->hold[s ^ 1].pending ? (s ^ 1) : -1;
    return 0;
}

int pr_gen_stats(pr_gen* e, pr_stats* main_st, pr_stats* hold_st) {
    if (!e) return -1;
    if (main_st) *main_st = e->st_main;
    if (hold_st) *hold_st = e->st_hold;
    return 0;
}

}  // extern "C"

// ================================================================== reference call surface
static pr_sim_params reference_params() {
    pr_sim_params p;
    p.preset = 0; p.gust_rms_max = 0; p.warp_max = 0; p.wind_sd = 0;
    p.max_steps = 7200; p.chunk = 600; p.spread = 0;
    return p;
}

static void save_best(const char* out_dir, const Genome* g) {  // sim.cu:327-331
    std::string path = std::string(out_dir) + "/best.bin";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); return; }
    fwrite(g, sizeof(Genome), 1, f);
    fclose(f);
}

static void save_genomes(const char* out_dir, const char* name, const Genome* gs, int n) {  // sim.cu:333-338
    std::string path = std::string(out_dir) + "/" + name;
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); return; }
    fwrite(&n, 4, 1, f);
    fwrite(gs, sizeof(Genome), n, f);
    fclose(f);
}

void run_ga(int pop, int gens, int nseeds, const char* out_dir, const char* init_path) {
    if (pop < 2 || gens < 1 || nseeds < 1) {
        fprintf(stderr, "evolve requires pop>=2, gens>=1, seeds>=1\n");
        return;
    }
    std::filesystem::create_directories(out_dir);
    void* ga = pr_ga_new(42);
    std::vector<Genome> gs(pop), next(pop);
    pr_ga_random(ga, (float*)gs.data(), pop);
    if (init_path) {
        FILE* f = fopen(init_path, "rb");
        if (f) {
            int n = 0;
            if (fread(&n, 4, 1, f) != 1) n = 0;
            if (n > pop) n = pop;
            if (n < 0) n = 0;
            size_t got = fread(gs.data(), sizeof(Genome), n, f);
            fclose(f);
            (void)got;
            pr_ga_sanitize((float*)gs.data(), n);
            printf("warm start: %d genomes from %s (clamped to printable bounds)\n", n, init_path);
        }
    }
    pr_gen* E = nullptr;
    CKP(pr_gen_create(&E, pop, nseeds, 0, 0));
    DevBuf dg;
    dg.ensure(sizeof(Genome) * pop);
    pr_sim_params prm = reference_params();
    std::vector<float> acc(pop);
    std::vector<int> idx(pop), seeds(nseeds);
    Genome best_ever;
    std::memset(&best_ever, 0, sizeof best_ever);
    float best_ever_f = -1e9f;
    std::vector<Genome> archive;
    std::string hpath = std::string(out_dir) + "/history.csv";
    FILE* hist = fopen(hpath.c_str(), "w");
    if (!hist) { fprintf(stderr, "cannot write %s\n", hpath.c_str()); pr_gen_destroy(E); pr_ga_free(ga); return; }
    fprintf(hist, "gen,train_best,holdout128,mean,p90\n");
    for (int gen = 0; gen < gens; gen++) {
        CKP(pr_h2d(dg.p, gs.data(), sizeof(Genome) * pop));
        for (int s = 0; s < nseeds; s++) seeds[s] = gen * nseeds + s + 1;
        float holdout;
        int champion_i;
        CKP(pr_gen_evaluate(E, (const float*)dg.p, nullptr, pop, seeds.data(), nseeds, 8, &prm, acc.data(),
                            idx.data(), nullptr, &holdout, &champion_i));
        float mean = 0;
        for (int i = 0; i < pop; i++) mean += acc[i] / pop;
        fprintf(hist, "%d,%.3f,%.3f,%.3f,%.3f\n", gen, acc[idx[0]], holdout, mean, acc[idx[pop / 10]]);
        fflush(hist);
        if (gen % 10 == 0 || gen == gens - 1)
            printf("gen %4d train %7.2fm holdout %7.2fm mean %7.2fm p90 %7.2fm\n", gen, acc[idx[0]], holdout, mean,
                   acc[idx[pop / 10]]);
        if (holdout > best_ever_f) { best_ever_f = holdout; best_ever = gs[champion_i]; }
        archive.push_back(gs[champion_i]);
        if (gen % 25 == 24) {  // checkpoint (sim.cu:464-470)
            save_best(out_dir, &best_ever);
            std::vector<Genome> srt(pop);
            for (int i = 0; i < pop; i++) srt[i] = gs[idx[i]];
            save_genomes(out_dir, "pop.bin", srt.data(), pop);
            save_genomes(out_dir, "archive.bin", archive.data(), (int)archive.size());
        }
        pr_ga_breed(ga, (const float*)gs.data(), acc.data(), idx.data(), pop, gen, gens, (float*)next.data());
        gs.swap(next);
    }
    fclose(hist);
    save_best(out_dir, &best_ever);
    // re-evaluate the final (bred, unevaluated) population on seeds 777+s (sim.cu:494-508)
    CKP(pr_h2d(dg.p, gs.data(), sizeof(Genome) * pop));
    for (int s = 0; s < nseeds; s++) seeds[s] = 777 + s;
    CKP(pr_gen_evaluate(E, (const float*)dg.p, nullptr, pop, seeds.data(), nseeds, 0, &prm, acc.data(), idx.data(),
                        nullptr, nullptr, nullptr));
    std::vector<Genome> srt(pop);
    for (int i = 0; i < pop; i++) srt[i] = gs[idx[i]];
    save_genomes(out_dir, "pop.bin", srt.data(), pop);
    save_genomes(out_dir, "archive.bin", archive.data(), (int)archive.size());
    printf("best holdout score: %.2fm (saved %s/best.bin)\n", best_ever_f, out_dir);
    const Genome* bg = &best_ever;
    printf("genome: span=%.3f chord=%.3f taper=%.2f sweep=%.2f dih_in=%.2f dih_out=%.2f kink=%.2f\n"
           "        camber=(%.3f,%.3f) washout=%.3f elev=%.3f nose=%.2f keel=%.3f deck=(%.3f,%.2f) body=%.2f\n"
           "        fin=(h %.3f, chord %.2f, sweep %.2f) tail=(span %.2f, chord %.2f, angle %.3f)\n",
           bg->span, bg->chord, bg->taper, bg->sweep, bg->dihedral_in, bg->dihedral_out, bg->kink, bg->camber1,
           bg->camber2, bg->washout, bg->elevator, bg->nose_mass, bg->keel_h, bg->deck_gap, bg->deck_scale,
           bg->body_frac, bg->fin_height, bg->fin_chord, bg->fin_sweep, bg->tail_span, bg->tail_chord,
           bg->tail_angle);
    pr_gen_destroy(E);
    pr_ga_free(ga);
}

// one genome under `scenarios` seeds base + i*7919 (layout 0, one genome)
// Robustness statistics of one design's scenario distances (sim.cu:291-309).  The moments
// are accumulated over the samples in ascending order with a division per term, as there,
// because that order fixes the float result; order statistics interpolate linearly
// (sim.cu:250-256); CVaR10 averages the worst max(1, n/10); failures count x < 0.
int pr_design_stats(const float* samples, int n, float* out) {
#pragma clang fp contract(off)
    if (!samples || !out || n < 1) return -1;
    std::vector<float> v(samples, samples + n);
    std::sort(v.begin(), v.end());
    const float fn = (float)n;
    float m1 = 0, m2 = 0;
    for (int i = 0; i < n; i++) {
        m1 += v[i] / fn;
        m2 += v[i] * v[i] / fn;
    }
    const int worst = std::max(1, n / 10);
    float cv = 0;
    for (int i = 0; i < worst; i++) cv += v[i] / (float)worst;
    int neg = 0;
    while (neg < n && v[neg] < 0.0f) neg++;  // sorted: negatives come first
    out[PR_ST_MIN] = v.front();
    out[PR_ST_P05] = percentile(v, .05f);
    out[PR_ST_P10] = percentile(v, .10f);
    out[PR_ST_P25] = percentile(v, .25f);
    out[PR_ST_MEDIAN] = percentile(v, .50f);
    out[PR_ST_MEAN] = m1;
    out[PR_ST_SD] = sqrtf(fmaxf(0.0f, m2 - m1 * m1));
    out[PR_ST_P75] = percentile(v, .75f);
    out[PR_ST_P90] = percentile(v, .90f);
    out[PR_ST_P95] = percentile(v, .95f);
    out[PR_ST_MAX] = v.back();
    out[PR_ST_CVAR10] = cv;
    out[PR_ST_FAILURES] = (float)neg;
    return 0;
}

// Batched robustness audit: evaluate_design (sim.cu:258-312) for many designs at once.
// Every genome flies the same n_scen scenarios, seed base_seed + s * 7919 indexed by the
// scenario (32-bit wrap, as the reference's seed arithmetic), in one pr_fly; the distances
// land in d_fitness [n_genomes x n_scen] (device) and the per-genome statistics in
// stats [n_genomes x PR_NSTAT] (host), computed on nthreads host threads (0: all cores).
int pr_audit(pr_ctx* c, const float* d_genomes, const int* d_res, int n_genomes, int n_scen, int base_seed,
             const pr_sim_params* p, float* d_fitness, int* d_steps, float* stats, int nthreads) {
    if (!c || !d_genomes || !p || !d_fitness || !d_steps || !stats || n_genomes < 1 || n_scen < 1) return -1;
    std::vector<int> seeds(n_scen);
    for (int s = 0; s < n_scen; s++) seeds[s] = (int)((unsigned)base_seed + (unsigned)s * 7919u);
    DevBuf ds;
    ds.ensure(sizeof(int) * n_scen);
    int rc = pr_h2d(ds.p, seeds.data(), sizeof(int) * n_scen);
    if (!rc) rc = pr_prepare(c, d_genomes, d_res, n_genomes);
    if (!rc) rc = pr_scenarios(c, (const int*)ds.p, n_scen, p);
    if (!rc) rc = pr_fly(c, 0, n_genomes, n_scen, p, d_fitness, d_steps);
    if (rc) return rc;
    std::vector<float> dist((size_t)n_genomes * n_scen);
    rc = pr_d2h(dist.data(), d_fitness, sizeof(float) * dist.size());
    if (rc) return rc;
    int nth = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    nth = std::min(nth, n_genomes);
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; t++)
        pool.emplace_back([&]() {
            for (int g; (g = next.fetch_add(1)) < n_genomes;)
                pr_design_stats(dist.data() + (size_t)g * n_scen, n_scen, stats + (size_t)g * PR_NSTAT);
        });
    for (auto& th : pool) th.join();
    return 0;
}

static void fly_one_genome(pr_ctx* c, DevBuf& dg, DevBuf& ds, DevBuf& df, DevBuf& dst, const Genome* g, int n,
                           int base_seed, std::vector<float>& out) {
    std::vector<int> seeds(n);
    for (int i = 0; i < n; i++) seeds[i] = (int)((unsigned)base_seed + (unsigned)i * 7919u);
    pr_sim_params prm = reference_params();
    CKP(pr_h2d(dg.p, g, sizeof(Genome)));
    CKP(pr_h2d(ds.p, seeds.data(), sizeof(int) * n));
    CKP(pr_prepare(c, (const float*)dg.p, nullptr, 1));
    CKP(pr_scenarios(c, (const int*)ds.p, n, &prm));
    CKP(pr_fly(c, 0, 1, n, &prm, (float*)df.p, (int*)dst.p));
    out.resize(n);
    CKP(pr_d2h(out.data(), df.p, sizeof(float) * n));
}

void evaluate_design(const Genome* g, int scenarios, int base_seed, const char* csv_path) {  // sim.cu:258-312
    if (scenarios < 8) scenarios = 8;
    pr_ctx* c = nullptr;
    CKP(pr_create(&c, 1, 0, scenarios, scenarios, 0, 0));
    DevBuf dg, ds, df, dst;
    dg.ensure(sizeof(Genome)); ds.ensure(sizeof(int) * scenarios); df.ensure(sizeof(float) * scenarios);
    dst.ensure(sizeof(int) * scenarios);
    std::vector<float> distances;
    fly_one_genome(c, dg, ds, df, dst, g, scenarios, base_seed, distances);
    std::vector<float> sc((size_t)scenarios * 26);
    CKP(pr_read_scenarios(c, scenarios, sc.data()));
    std::filesystem::path out(csv_path);
    if (out.has_parent_path()) std::filesystem::create_directories(out.parent_path());
    FILE* f = fopen(csv_path, "w");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", csv_path);
    } else {
        fprintf(f, "scenario,seed,distance_m,speed_mps,height_m,pitch_deg,yaw_deg,roll_deg,force_pitch_deg,force_yaw_deg,"
                   "grip_x_mm,grip_y_mm,grip_z_mm,roll_rate,pitch_rate,yaw_rate,air_density,skin_cf,gust_scale,"
                   "base_wind_x,base_wind_y,material_scale,infill_ratio,print_bias_x_mm,print_bias_y_mm,"
                   "print_bias_z_mm,warp_pitch_deg,warp_roll_deg\n");
        for (int i = 0; i < scenarios; i++) {
            int seed = (int)((unsigned)base_seed + (unsigned)i * 7919u);
            const float* d = sc.data() + (size_t)26 * i;
            const float r2d = 57.2957795131f;
            fprintf(f, "%d,%d,%.6f,%.5f,%.5f,%.4f,%.4f,%.4f,%.4f,%.4f,%.3f,%.3f,%.3f,%.5f,%.5f,%.5f,%.6f,%.6f,%.6f,"
                       "%.6f,%.6f,%.5f,%.5f,%.3f,%.3f,%.3f,%.4f,%.4f\n",
                    i, seed, distances[i], d[0], d[4], d[1] * r2d, d[2] * r2d, d[3] * r2d, d[13] * r2d, d[14] * r2d,
                    d[15] * 1000, d[16] * 1000, d[17] * 1000, d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12],
                    d[18], d[19], d[20] * 1000, d[21] * 1000, d[22] * 1000, d[23] * r2d, d[24] * r2d);
        }
        fclose(f);
    }
    float st[PR_NSTAT];
    pr_design_stats(distances.data(), scenarios, st);
    float mass20[20], ext[3];
    int nt, nv, off[2];
    CKP(pr_read_prepared(c, 1, mass20, &nt, &nv, off));
    CKP(pr_read_extent(c, 1, ext));
    printf("robustness audit: %d GPU scenarios -> %s\n", scenarios, csv_path);
    printf("  min %.2fm  p05 %.2fm  p10 %.2fm  p25 %.2fm  median %.2fm\n", st[PR_ST_MIN], st[PR_ST_P05], st[PR_ST_P10],
           st[PR_ST_P25], st[PR_ST_MEDIAN]);
    printf("  mean %.2fm  sd %.2fm  p75 %.2fm  p90 %.2fm  p95 %.2fm  max %.2fm\n", st[PR_ST_MEAN], st[PR_ST_SD],
           st[PR_ST_P75], st[PR_ST_P90], st[PR_ST_P95], st[PR_ST_MAX]);
    printf("  CVaR10 %.2fm  failures %d/%d  PLA %.1fg  bounds %.0fx%.0fx%.0fmm\n", st[PR_ST_CVAR10],
           (int)st[PR_ST_FAILURES], scenarios, mass20[0] * 1000, ext[0] * 1000, ext[1] * 1000, ext[2] * 1000);
    pr_destroy(c);
}

void select_finalist(const Genome* candidates, int count, int scenarios, int base_seed, const char* out_dir) {
    if (count < 1 || scenarios < 128) {  // sim.cu:360
        fprintf(stderr, "select-best needs candidates>=1 and scenarios>=128\n");
        return;
    }
    std::filesystem::create_directories(out_dir);
    // candidates flown in batches: batch x scenarios flights per pr_fly (layout 0)
    const long long max_flights = 1 << 22;
    int batch = (int)std::max(1LL, std::min<long long>(count, max_flights / scenarios));
    pr_ctx* c = nullptr;
    CKP(pr_create(&c, batch, 0, scenarios, batch * scenarios, 0, 0));
    DevBuf dg, ds, df, dst;
    dg.ensure(sizeof(Genome) * batch); ds.ensure(sizeof(int) * scenarios);
    df.ensure(sizeof(float) * batch * scenarios); dst.ensure(sizeof(int) * batch * scenarios);
    std::vector<int> seeds(scenarios);
    for (int i = 0; i < scenarios; i++) seeds[i] = (int)((unsigned)base_seed + (unsigned)i * 7919u);
    CKP(pr_h2d(ds.p, seeds.data(), sizeof(int) * scenarios));
    pr_sim_params prm = reference_params();
    CKP(pr_scenarios(c, (const int*)ds.p, scenarios, &prm));
    std::string csv_path = std::string(out_dir) + "/finalist-selection.csv";
    FILE* csv = fopen(csv_path.c_str(), "w");
    if (!csv) { fprintf(stderr, "cannot write %s\n", csv_path.c_str()); pr_destroy(c); return; }
    fprintf(csv, "candidate,robust_score,mean,cvar25,worst\n");
    float best_score = -1e9f;
    int best_i = 0;
    std::vector<float> fit, sorted, tmp;
    for (int c0 = 0; c0 < count; c0 += batch) {
        int nb = std::min(batch, count - c0);
        CKP(pr_h2d(dg.p, candidates + c0, sizeof(Genome) * nb));
        CKP(pr_prepare(c, (const float*)dg.p, nullptr, nb));
        CKP(pr_fly(c, 0, nb, scenarios, &prm, (float*)df.p, (int*)dst.p));
        fit.resize((size_t)nb * scenarios);
        CKP(pr_d2h(fit.data(), df.p, sizeof(float) * fit.size()));
        for (int k = 0; k < nb; k++) {
            int ci = c0 + k;
            const float* row = fit.data() + (size_t)k * scenarios;
            float score = robust_row(row, scenarios, tmp);
            sorted.assign(row, row + scenarios);
            std::sort(sorted.begin(), sorted.end());
            float mean = 0, cvar = 0;
            int tail = (scenarios + 3) / 4;
            for (float x : sorted) mean += x / scenarios;
            for (int i = 0; i < tail; i++) cvar += sorted[i] / tail;
            fprintf(csv, "%d,%.6f,%.6f,%.6f,%.6f\n", ci, score, mean, cvar, sorted[0]);
            if (score > best_score) { best_score = score; best_i = ci; }
            if (ci % 25 == 0 || ci == count - 1) {
                printf("finalist %d/%d best %.3fm (candidate %d)\n", ci + 1, count, best_score, best_i);
                fflush(stdout);
            }
        }
    }
    fclose(csv);
    std::string best_path = std::string(out_dir) + "/best-final.bin";
    FILE* f = fopen(best_path.c_str(), "wb");
    if (f) { fwrite(&candidates[best_i], sizeof(Genome), 1, f); fclose(f); }
    printf("independent finalist winner: candidate %d score %.3fm -> %s\n", best_i, best_score, best_path.c_str());
    pr_destroy(c);
}

// ------------------------------------------------------------------ PanelBuf-style shims (sim.h:9-13)
struct RefPanelBuf {
    pr_ctx* ctx;
    int pop;
};

void* panelbuf_alloc(int pop) {
    RefPanelBuf* pb = new RefPanelBuf;
    pb->pop = pop;
    CKP(pr_create(&pb->ctx, pop, 0, pop, pop, 0, 0));
    return pb;
}

void panelbuf_free(void* h) {
    if (!h) return;
    RefPanelBuf* pb = (RefPanelBuf*)h;
    pr_destroy(pb->ctx);
    delete pb;
}

void gpu_prepare(const Genome* d_gs, void* h, int pop) {
    RefPanelBuf* pb = (RefPanelBuf*)h;
    CKP(pr_prepare(pb->ctx, (const float*)d_gs, nullptr, pop));
}

void gpu_sim(void* h, int pop, int seed, float* d_fitness, float8* d_rec, int rec_every, int rec_cap, float spread,
             int vary_scenarios) {
    RefPanelBuf* pb = (RefPanelBuf*)h;
    pr_sim_params prm = reference_params();
    CKP(pr_sim(pb->ctx, pop, seed, d_fitness, (float*)d_rec, rec_every, rec_cap, spread, vary_scenarios, &prm));
}
