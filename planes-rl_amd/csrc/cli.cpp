// cli.cpp -- command-line front end with the reference's modes, flags and defaults
// (main.cu:36-105) for the modes on the fitness-evaluation path: evolve, evaluate,
// select-best, inspect.  The video, STL and preset modes belong to the reference's
// renderer/exporter and are not part of this build.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/planes_hip.h"
#include "../../include/planes_sim.h"

static const char* opt(int argc, char** argv, const char* key, const char* dflt) {
    for (int i = 2; i < argc - 1; i++)
        if (!strcmp(argv[i], key)) return argv[i + 1];
    return dflt;
}

static Genome load_best(const char* p) {  // best.bin: one raw 140-byte record
    FILE* f = fopen(p, "rb");
    if (!f) { fprintf(stderr, "missing %s\n", p); exit(1); }
    Genome g;
    if (fread(&g, sizeof g, 1, f) != 1) { fprintf(stderr, "short read %s\n", p); exit(1); }
    fclose(f);
    return g;
}

static std::vector<Genome> load_pop(const char* p) {  // pop.bin/archive.bin: int32 n + n records
    FILE* f = fopen(p, "rb");
    if (!f) { fprintf(stderr, "missing %s\n", p); exit(1); }
    int n = 0;
    if (fread(&n, 4, 1, f) != 1 || n < 0) { fprintf(stderr, "bad header %s\n", p); exit(1); }
    std::vector<Genome> gs(n);
    if (fread(gs.data(), sizeof(Genome), n, f) != (size_t)n) { fprintf(stderr, "short read %s\n", p); exit(1); }
    fclose(f);
    return gs;
}

// designs.cu:45-54 describe(): mass, bounding box, CG, printability -- computed on the GPU
static void inspect(const Genome* g, const char* name) {
    pr_ctx* c = nullptr;
    void* d = nullptr;
    if (pr_create(&c, 1, 0, 1, 1, 0, 0) || pr_dev_alloc(&d, sizeof(Genome)) || pr_h2d(d, g, sizeof(Genome)) ||
        pr_prepare(c, (const float*)d, nullptr, 1)) {
        fprintf(stderr, "inspect: %s\n", pr_last_error());
        exit(1);
    }
    float m[20], e[3];
    int nt, nv, off[2];
    if (pr_read_prepared(c, 1, m, &nt, &nv, off) || pr_read_extent(c, 1, e)) {
        fprintf(stderr, "inspect: %s\n", pr_last_error());
        exit(1);
    }
    printf("%-18s %6.1fg  %3.0fx%3.0fx%3.0f mm  CG=(%5.1f,%5.1f,%5.1f) mm  %s\n", name, m[0] * 1000, e[0] * 1000,
           e[1] * 1000, e[2] * 1000, m[2] * 1000, m[3] * 1000, m[4] * 1000, nt > 0 ? "printable" : "TOO LARGE");
    pr_dev_free(d);
    pr_destroy(c);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        printf("usage: planes_amd <evolve|evaluate|select-best|inspect> [opts]\n"
               "  evolve    --pop 4096 --gens 400 --seeds 8 --out out [--init out/pop.bin]\n"
               "  evaluate  --best out/best.bin --n 512 --seed 9001 --o out/robustness.csv\n"
               "  select-best --pop-file out/archive.bin --n 8192 --seed 1900001 --out out/run\n"
               "  inspect   --best out/best.bin\n");
        return 1;
    }
    const char* mode = argv[1];
    if (!strcmp(mode, "evolve")) {
        run_ga(atoi(opt(argc, argv, "--pop", "4096")), atoi(opt(argc, argv, "--gens", "400")),
               atoi(opt(argc, argv, "--seeds", "3")), opt(argc, argv, "--out", "out"), opt(argc, argv, "--init", nullptr));
    } else if (!strcmp(mode, "evaluate")) {
        Genome g = load_best(opt(argc, argv, "--best", "out/best.bin"));
        evaluate_design(&g, atoi(opt(argc, argv, "--n", "512")), atoi(opt(argc, argv, "--seed", "9001")),
                        opt(argc, argv, "--o", "out/robustness.csv"));
    } else if (!strcmp(mode, "select-best")) {
        auto gs = load_pop(opt(argc, argv, "--pop-file", "out/archive.bin"));
        select_finalist(gs.data(), (int)gs.size(), atoi(opt(argc, argv, "--n", "8192")),
                        atoi(opt(argc, argv, "--seed", "1900001")), opt(argc, argv, "--out", "out"));
    } else if (!strcmp(mode, "inspect")) {
        Genome g = load_best(opt(argc, argv, "--best", "out/best.bin"));
        inspect(&g, opt(argc, argv, "--best", "out/best.bin"));
    } else if (!strcmp(mode, "showcase") || !strcmp(mode, "hero") || !strcmp(mode, "robust-video") ||
               !strcmp(mode, "evolution-video") || !strcmp(mode, "turntable") || !strcmp(mode, "stl") ||
               !strcmp(mode, "presets")) {
        fprintf(stderr, "mode %s belongs to the reference's renderer/exporter, which this library does not include\n",
                mode);
        return 1;
    } else {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 1;
    }
    return 0;
}
