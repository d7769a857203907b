// shim_driver.cpp -- a caller of the sim.h entry points (include/planes_sim.h), used the way
// the reference's run_ga uses them for one generation's evaluation (sim.cu:411-418, 431-456):
//
//   panelbuf_alloc(pop) and a 128-plane holdout buffer;
//   upload the genomes, gpu_prepare;
//   for each seed s: gpu_sim(pb, pop, gen*nseeds + s + 1, d_fit, null, 1, 0, 0, 0), download;
//   robust scores, std::sort ranking by score (descending);
//   for each of the top min(8, pop): 128 copies of the genome, gpu_prepare,
//   gpu_sim(hold, 128, 100003, d_hfit, null, 1, 0, 0, 1), robust score of the 128 distances;
//   champion = first strict maximum of the holdout scores.
//
// Test infrastructure (tests/test_gpu_shims.py): reads pop genomes (the best.bin record, 140 B
// each) from a file and writes, as little-endian binary: int32 pop, nseeds, k; float32
// samples[pop * nseeds] (genome-major); int32 idx[pop]; float32 holdout[k]; int32 champion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "planes_sim.h"

#define HCK(x)                                                                              \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            fprintf(stderr, "HIP %s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_));   \
            exit(1);                                                                        \
        }                                                                                   \
    } while (0)

// robust score of one row (sim.cu:340-357): 0.65 mean + 0.25 CVaR of the worst quarter + 0.10 worst
static float robust(const float* r, int n) {
    std::vector<float> row(r, r + n);
    float mean = 0;
    for (int s = 0; s < n; s++) mean += row[s] / n;
    std::sort(row.begin(), row.end());
    int tail = (n + 3) / 4;
    float cvar = 0;
    for (int s = 0; s < tail; s++) cvar += row[s] / tail;
    return 0.65f * mean + 0.25f * cvar + 0.10f * row[0];
}

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: shim_driver genomes.bin pop nseeds gen out.bin\n");
        return 2;
    }
    const int pop = atoi(argv[2]), nseeds = atoi(argv[3]), gen = atoi(argv[4]);
    if (pop < 1 || nseeds < 1 || gen < 0) return 2;
    std::vector<Genome> gs(pop);
    FILE* f = fopen(argv[1], "rb");
    if (!f || fread(gs.data(), sizeof(Genome), pop, f) != (size_t)pop) {
        fprintf(stderr, "cannot read %d genomes from %s\n", pop, argv[1]);
        return 1;
    }
    fclose(f);

    const int HOLD = 128;
    Genome* d_gs = nullptr;
    Genome* d_hold = nullptr;
    float *d_fit = nullptr, *d_hfit = nullptr;
    HCK(hipMalloc(&d_gs, sizeof(Genome) * pop));
    HCK(hipMalloc(&d_hold, sizeof(Genome) * HOLD));
    HCK(hipMalloc(&d_fit, sizeof(float) * pop));
    HCK(hipMalloc(&d_hfit, sizeof(float) * HOLD));
    void* pb = panelbuf_alloc(pop);
    void* hpb = panelbuf_alloc(HOLD);

    HCK(hipMemcpy(d_gs, gs.data(), sizeof(Genome) * pop, hipMemcpyHostToDevice));
    gpu_prepare(d_gs, pb, pop);
    std::vector<float> fit(pop), samples((size_t)pop * nseeds), acc(pop);
    for (int s = 0; s < nseeds; s++) {
        gpu_sim(pb, pop, gen * nseeds + s + 1, d_fit, nullptr, 1, 0, 0.0f, 0);
        HCK(hipMemcpy(fit.data(), d_fit, sizeof(float) * pop, hipMemcpyDeviceToHost));
        for (int i = 0; i < pop; i++) samples[(size_t)i * nseeds + s] = fit[i];
    }
    for (int i = 0; i < pop; i++) acc[i] = robust(&samples[(size_t)i * nseeds], nseeds);
    std::vector<int> idx(pop);
    for (int i = 0; i < pop; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return acc[a] > acc[b]; });

    const int k = std::min(8, pop);
    std::vector<float> holdout(k), hfit(HOLD);
    std::vector<Genome> copies(HOLD);
    int champion = idx[0];
    float best = -1e9f;
    for (int h = 0; h < k; h++) {
        for (int j = 0; j < HOLD; j++) copies[j] = gs[idx[h]];
        HCK(hipMemcpy(d_hold, copies.data(), sizeof(Genome) * HOLD, hipMemcpyHostToDevice));
        gpu_prepare(d_hold, hpb, HOLD);
        gpu_sim(hpb, HOLD, 100003, d_hfit, nullptr, 1, 0, 0.0f, 1);
        HCK(hipMemcpy(hfit.data(), d_hfit, sizeof(float) * HOLD, hipMemcpyDeviceToHost));
        holdout[h] = robust(hfit.data(), HOLD);
        if (holdout[h] > best) { best = holdout[h]; champion = idx[h]; }
    }
    panelbuf_free(pb);
    panelbuf_free(hpb);
    HCK(hipFree(d_gs));
    HCK(hipFree(d_hold));
    HCK(hipFree(d_fit));
    HCK(hipFree(d_hfit));

    FILE* o = fopen(argv[5], "wb");
    if (!o) return 1;
    int hdr[3] = {pop, nseeds, k};
    fwrite(hdr, sizeof(int), 3, o);
    fwrite(samples.data(), sizeof(float), samples.size(), o);
    fwrite(idx.data(), sizeof(int), pop, o);
    fwrite(holdout.data(), sizeof(float), k, o);
    fwrite(&champion, sizeof(int), 1, o);
    fclose(o);
    printf("shim driver: %d genomes x %d seeds, holdout of %d, champion %d (%.3f m)\n", pop, nseeds, k, champion,
           best);
    return 0;
}
