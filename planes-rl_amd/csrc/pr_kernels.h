// pr_kernels.h -- argument blocks and host launchers of kernels.hip (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "pr_physics.h"

namespace pr {

struct PrepareArgs {
    const float* genomes;  // [n*35]
    const int* res;        // [n*2] (ch, sp) per genome, or null -> def_ch/def_sp
    int def_ch, def_sp;
    int nv_max;            // LDS vertex capacity
    const int* off;        // [n] exclusive prefix of the closed-form triangle counts
    Mass* mass;            // [n]
    int* nt;               // [n] triangles flown (0 if outside the print volume)
    int* nv;               // [n]
    float* ext;            // [n*3] bounding-box extent of the mesh
    float *pcx, *pcy, *pcz, *pA, *pnx, *pny, *pnz;  // panel SoA at off[g] + t
};

struct QueueArgs {
    int n_flights;
    int layout;            // 0: flight f = (genome f / nseeds, scenario f % nseeds); 1: (f, f)
    int nseeds;
    const int* nt;
    const Mass* mass;
    const float* scen;     // [n_scen*26]
    const int* seeds;
    float spread;
    FlightConst* fc;
    int n_buckets;
    const int* bucket_nt;  // device [n_buckets] upper triangle count of each bucket
    int* count;            // bucket b's list count at count[b * count_stride]
    int count_stride;
    int* list;             // [n_buckets * list_stride]
    int list_stride;
    FlightState* state;
    float* fitness;
    int* steps;
};

struct FlyArgs {
    const int* nt;
    const int* off;
    const float *pcx, *pcy, *pcz, *pA, *pnx, *pny, *pnz;
    const FlightConst* fc;
    FlightState* state;
    float* fitness;
    int* steps;
    const int* list_in;
    const int* count_in;
    int* list_out;
    int* count_out;
    int* head;
    int chunk, max_steps;
    int run_lo, run_hi;    // the launch works only if run_lo <= flights in list_in <= run_hi
    float spread;
    unsigned long long* stats;  // [0] flight-steps, [1] panel-steps
    unsigned long long* cstats; // this chunk's [0] flight-steps, [1] panel-steps (null: none)
};

// Per-bucket chunk counters (PR_CNT ints per bucket): [0..2] list counts, rotating with the
// chunk index mod 3 (in, out, cleared for the next chunk), [3..4] queue heads, rotating mod 2.
#define PR_CNT 5

// All buckets' chunk in one persistent launch (k_fly_all).  Bucket b's lists are
// lists{0,1} + b * list_stride (list_in = lists0 when in_list == 0); its counters are
// counts + PR_CNT * b.  vid_p / vid_w: its fused-variant ids (many flights per wave / one
// per wave, -1 none: the bucket is flown by its own k_fly launch); vid_w takes over once
// the bucket's count is <= thr.  Block 0 clears the counters of the next chunk for every
// bucket (c_zero, h_zero), so a chunk is one launch.
struct FlyAllArgs {
    FlyArgs base;
    int* lists0;
    int* lists1;
    int list_stride;
    int in_list;
    int* counts;
    int c_in, c_out, c_zero, h_run, h_zero;
    int nb;
    int vid_p[8], vid_w[8], thr[8];
    unsigned long long* stats;  // slot 2b + (tail ? 1 : 0): flight-steps, panel-steps
    int* chunk_n;               // [nb] flights entering this chunk per bucket (null: none)
};

int fused_variant_id(int L, int P);
hipError_t launch_fly_all(const FlyAllArgs& a, int grid, hipStream_t s);
hipError_t fly_all_occupancy(int* blocks_per_cu);
hipError_t launch_zero_chunk(int* counts, int nb, int c_in, int c_zero, int h_zero, int* chunk_n, hipStream_t s);

hipError_t launch_layout(const float* genomes, const int* res, int ch, int sp, int n, int* nt_raw, hipStream_t s);
hipError_t launch_prepare(const PrepareArgs& a, int n, int fp64, hipStream_t s);
hipError_t launch_scenarios(const int* seeds, int n, const ScenarioParams& sp, float* out, hipStream_t s);
hipError_t launch_queue(const QueueArgs& a, hipStream_t s);
hipError_t launch_zero(int* p, int n, hipStream_t s);
hipError_t launch_fly(int L, int P, const FlyArgs& a, int grid, hipStream_t s);
hipError_t fly_occupancy(int L, int P, int* blocks_per_cu);
hipError_t launch_fly_rec(int L, int P, const FlyArgs& a, int n_flights, float* rec, int rec_every, int rec_cap,
                          hipStream_t s);
int fly_variant_count();
int fly_variant_info(int i, int* L, int* P);
hipError_t launch_scores(const float* samples, int pop, int nseeds, float* scores, hipStream_t s);

}  // namespace pr
