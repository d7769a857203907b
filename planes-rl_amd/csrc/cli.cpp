// cli.cpp -- command-line front end with the reference's modes, flags and defaults
// (main.cu:36-105) for the modes on the fitness-evaluation path: evolve, evaluate,
// select-best, inspect, presets (seed population and descriptions; its STL export is the
// reference's exporter).  The video and STL modes belong to the reference's
// renderer/exporter and are not part of this library.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
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

// The five preset designs (designs.cu:35-41, built by its design() rule, designs.cu:10-33):
// their genomes as the reference's host build computes them, bit for bit, in gene order.
static const struct {
    const char* name;
    float g[35];
} kPresets[5] = {
    {"stable_gull",
     {0x1.e147aep-3f, 0x1.ae147ap-4f, 0x1.28f5c2p-1f, 0x1.eb851ep-3f, 0x1.0a3d7p-3f, 0x1.28f5c2p-1f,
      0x1.70a3d8p-1f, 0x1.99999ap-6f, 0x1.47ae14p-7f, -0x1.c28f5cp-5f, 0x1.1eb852p-5f, 0x1.ae147ap-2f,
      0x1.ba5e36p-6f, 0.0f, 0x1.99999ap-1f, 0x1.eb851ep-4f, 0x1.374bc6p-5f, 0x1.666666p-2f,
      0x1.333334p-2f, 0x1.70a3d8p-2f, 0x1.c28f5cp-3f, -0x1.99999ap-6f, -0x1.8p-2f, 0x1.23d70ap-2f,
      0x1.051ebap-2f, -0x1.851eb8p-4f, 0x1.99999ap-4f, 0x1.51eb84p-3f, 0x1.970a3ep-1f, 0x1.9eb852p-2f,
      -0x1.999998p-6f, 0x1.ca3d7p-1f, -0x1.70a3d6p-4f, 0x1.cccccep-3f, 0x1.947ae2p-2f}},
    {"fast_dart",
     {0x1.a3d70ap-3f, 0x1.28f5c2p-3f, 0x1.47ae14p-3f, 0x1.0a3d7p-1f, 0x1.eb851ep-5f, 0x1.333334p-2f,
      0x1.8f5c28p-1f, 0x1.89374cp-7f, -0x1.47ae14p-8f, -0x1.1eb852p-5f, 0x1.99999ap-6f, 0x1.70a3d8p-1f,
      0x1.78d4fep-6f, 0.0f, 0x1.8p-1f, 0x1.99999ap-4f, 0x1.6872bp-6f, 0x1.c28f5cp-3f,
      0x1.333334p-1f, 0x1.70a3d8p-3f, 0x1.1eb852p-3f, 0x1.eb851ep-7f, -0x1.cccccap-3f, 0x1.3d70a2p-3f,
      0x1.51eb86p-3f, 0x1.eb852p-7f, 0x1.333334p-2f, 0x1.051eb8p-2f, 0x1.5eb852p-1f, 0x1.07ae14p-1f,
      0x1p-3f, 0x1.91eb84p-1f, -0x1.eb8518p-6f, 0x1.8p-2f, 0x1.23d70ap-2f}},
    {"efficient_plank",
     {0x1.f5c29p-3f, 0x1.4fdf3cp-4f, 0x1.a3d70ap-1f, 0x1.eb851ep-4f, 0x1.99999ap-4f, 0x1.ae147ap-2f,
      0x1.5c28f6p-1f, 0x1.eb851ep-6f, 0x1.0624dep-8f, -0x1.1eb852p-4f, 0x1.eb851ep-5f, 0x1.333334p-2f,
      0x1.0624dep-5f, 0.0f, 0x1.ccccccp-1f, 0x1.47ae14p-3f, 0x1.eb851ep-6f, 0x1.ae147ap-2f,
      0x1.70a3d8p-3f, 0x1.eb851ep-2f, 0x1.70a3d8p-3f, -0x1.70a3d8p-5f, -0x1.33332p-5f, -0x1.eb85p-8f,
      0x1.ae148p-5f, 0x1.3851ecp-3f, 0x1.19999ap-1f, 0x1.7851ecp-2f, 0x1.1851ecp-1f, 0x1.4e147cp-1f,
      0x1.400002p-2f, 0x1.4b851ep-1f, 0x1.70a3dcp-5f, 0x1.2p-1f, 0x1.2e147cp-3f}},
    {"boxy_biplane",
     {0x1.a3d70ap-3f, 0x1.70a3d8p-4f, 0x1.666666p-1f, 0x1.47ae14p-3f, 0x1.47ae14p-4f, 0x1.47ae14p-2f,
      0x1.666666p-1f, 0x1.6872bp-6f, 0x1.47ae14p-8f, -0x1.47ae14p-5f, 0x1.1eb852p-5f, 0x1.851eb8p-2f,
      0x1.89374cp-6f, 0x1.16872cp-5f, 0x1.a3d70ap-1f, 0x1.1eb852p-3f, 0x1.70a3d8p-5f, 0x1.eb851ep-2f,
      0x1.c28f5cp-3f, 0x1.ae147ap-2f, 0x1.1eb852p-2f, -0x1.47ae14p-6f, 0x1.70a3ep-4f, -0x1.e353f8p-4f,
      -0x1.89374p-6f, 0x1.f7cedcp-3f, 0x1.70a3d8p-1f, 0x1.c6a7fp-2f, 0x1.d0e56p-2f, 0x1.7df3b8p-1f,
      0x1.c28f5ep-2f, 0x1.1ba5e2p-1f, 0x1.893752p-4f, 0x1.6147bp-1f, 0x1.ba5e2cp-5f}},
    {"compact_trainer",
     {0x1.7ae148p-3f, 0x1.99999ap-4f, 0x1.5c28f6p-1f, 0x1.99999ap-3f, 0x1.70a3d8p-3f, 0x1.70a3d8p-1f,
      0x1.3d70a4p-1f, 0x1.1eb852p-5f, 0x1.89374cp-7f, -0x1.0a3d7p-4f, 0x1.99999ap-5f, 0x1.eb851ep-2f,
      0x1.26e978p-5f, 0.0f, 0x1.99999ap-1f, 0x1.70a3d8p-3f, 0x1.99999ap-5f, 0x1.19999ap-1f,
      0x1.eb851ep-4f, 0x1.28f5c2p-1f, 0x1.47ae14p-2f, -0x1.eb851ep-5f, 0x1.eb852p-3f, -0x1.fbe76cp-3f,
      -0x1.d2f1a8p-4f, 0x1.6c8b46p-2f, 0x1.d70a3ep-1f, 0x1.116872p-1f, 0x1.60418ap-2f, 0x1.b645a4p-1f,
      0x1.2e147ap-1f, 0x1.c6a7ecp-2f, 0x1.3f7cfp-3f, 0x1.ae147cp-1f, -0x1.cac088p-5f}},
};

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

// designs.cu:58-73 write_presets(): the presets as a seed population (pop.bin layout, for
// evolve --init) and one description line each; STL files are the exporter's part.
static int write_presets(const char* out_dir) {
    std::error_code ec;
    std::filesystem::create_directories(out_dir, ec);
    const std::string path = (std::filesystem::path(out_dir) / "seeds.bin").string();
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); return 1; }
    const int n = 5;
    bool ok = fwrite(&n, sizeof n, 1, f) == 1;
    for (const auto& d : kPresets) ok = ok && fwrite(d.g, sizeof d.g, 1, f) == 1;
    if (fclose(f) != 0 || !ok) { fprintf(stderr, "cannot write %s\n", path.c_str()); return 1; }
    for (const auto& d : kPresets) {
        Genome g;
        static_assert(sizeof(Genome) == sizeof d.g, "Genome is 35 floats");
        memcpy(&g, d.g, sizeof g);
        inspect(&g, d.name);
    }
    printf("preset seed population -> %s\n", path.c_str());
    printf("(STL export of the presets is not part of this library)\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        printf("usage: planes_amd <evolve|evaluate|select-best|inspect|presets> [opts]\n"
               "  evolve    --pop 4096 --gens 400 --seeds 8 --out out [--init out/pop.bin]\n"
               "  evaluate  --best out/best.bin --n 512 --seed 9001 --o out/robustness.csv\n"
               "  select-best --pop-file out/archive.bin --n 8192 --seed 1900001 --out out/run\n"
               "  inspect   --best out/best.bin\n"
               "  presets   --out out/presets\n");
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
    } else if (!strcmp(mode, "presets")) {
        return write_presets(opt(argc, argv, "--out", "out/presets"));
    } else if (!strcmp(mode, "showcase") || !strcmp(mode, "hero") || !strcmp(mode, "robust-video") ||
               !strcmp(mode, "evolution-video") || !strcmp(mode, "turntable") || !strcmp(mode, "stl")) {
        fprintf(stderr, "mode %s belongs to the reference's renderer/exporter, which this library does not include\n",
                mode);
        return 1;
    } else {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 1;
    }
    return 0;
}
