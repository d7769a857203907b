// pr_kernels.h -- argument blocks and host launchers of kernels.hip (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "pr_physics.h"

namespace pr {

struct PrepareArgs {
    const float* genomes;  // [n*35]
    const int* res;        // [n*2] (ch, sp) per genome, or null -> def_ch/def_sp
    int def_ch, def_sp;
    int nv_max;            // LDS vertex capacity of this launch
    int nv_lo, nv_hi;      // this launch prepares the genomes with nv_lo <= vertices <= nv_hi
    int max_ch, max_sp;    // the context's resolution bounds (others: not meshed, unprintable)
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
    int* bucket;           // [n_flights] bucket of each flight, -1 when it ends at launch
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
    const int* list_in;    // flights of this chunk (one bucket), in list order
    const int* count_in;
    int* marks;            // marks[k] = list_in[k] if that flight goes on to the next chunk, else -1
    int* head;             // claim counter of list_in
    int chunk, max_steps;
    int run_lo, run_hi;    // the launch works only if run_lo <= flights in list_in <= run_hi
    float spread;
    unsigned long long* stats;  // [0] flight-steps, [1] panel-steps
    unsigned long long* cstats; // this chunk's [0] flight-steps, [1] panel-steps (null: none)
    // split_max > 0: a chunk whose flights, summed over the nb_all buckets' counters
    // counts_all[PR_CNT b + c_all], number at most split_max is flown by k_fly_split
    // (several waves per flight), and k_fly / k_fly_all launches of that chunk do nothing
    const int* counts_all;
    int nb_all, c_all, split_max;
};

// Per-bucket chunk counters (PR_CNT ints per bucket): [0..1] list counts (chunk ck reads
// slot ck & 1, its compaction writes slot (ck + 1) & 1), [2] claim counter of a bucket's
// own k_fly many-flights-per-wave launch, [3] of its one-flight-per-wave launch,
// [4..11] the fused kernel's many-flights-per-wave claim counters, one per XCD (its
// share of the list positions), [12] that pass's single counter (small chunks), [13] the
// fused one-flight-per-wave pass's counter.  The compaction after each chunk clears
// [2..19].
#define PR_CNT 20
#define PR_CNT_CLAIM0 2
#define PR_XCDS 8
static_assert(4 + PR_XCDS + 2 <= PR_CNT, "counter layout");

// All buckets' chunk in one persistent launch (k_fly_all).  Bucket b's list is lists + b *
// list_stride, its marks marks + b * list_stride, its counters counts + PR_CNT * b.
// vid_p / vid_w: its fused-variant ids (many flights per wave / one per wave, -1 none: the
// bucket is flown by its own k_fly launch); vid_w flies list positions [n - thr, n).
struct FlyAllArgs {
    FlyArgs base;
    const int* lists;
    int* marks;
    int list_stride;
    int* counts;
    int c_in;
    int nb;
    int vid_p[8], vid_w[8], thr[8];
    int tail_when_short;        // 1: vid_w only where the later-drained buckets hold < thr flights
    // widen = k > 0: a chunk whose flights would fill fewer than k x the grid's waves at a
    // bucket's width lanes[b] flies that bucket with its 2x / 4x wider variant vid_2 / vid_4
    // (-1: none); 0: never
    int widen;
    int lanes[8], vid_2[8], vid_4[8];
    unsigned long long* stats;  // slot 2b + (tail ? 1 : 0): flight-steps, panel-steps
    int split_lo, split_hi;     // k_fly_split: flies the chunk if split_lo <= flights <= split_hi
    unsigned long long* split_stats;  // k_fly_split's flight-steps, panel-steps
};

// Stable compaction of every bucket's marks into its next list (k_compact_*): keeps the
// list order, so which flights a chunk's passes fly is a function of the data alone.
// by_bucket = 0: bucket b's source is src + b * src_stride, n = counts[PR_CNT b + c_src]
// entries, kept if >= 0 (the value is written); by_bucket = 1: one source for all buckets
// (k_init's bucket of each flight), n = n_fixed, entry i kept for bucket b if src[i] == b
// (i is written).  Writes dst + b * dst_stride and counts[PR_CNT b + c_dst]; clears the
// claim counters; records the source lengths in chunk_n (null: none).
struct CompactArgs {
    const int* src;
    long long src_stride;
    int by_bucket;
    int c_src, n_fixed;
    int* dst;
    long long dst_stride;
    int c_dst;
    int* counts;
    int* tcount;           // [nb * max_tiles] survivors per tile
    int max_tiles;
    int nb;
    int* chunk_n;
    // by_bucket only: slot i of the source stands for flight perm[i / perm_div] * perm_div +
    // i % perm_div (genomes in the order of perm, each with its perm_div scenarios); that
    // flight's bucket is read and its id written.  Null perm: slot i is flight i.
    const int* perm;
    int perm_div;
};
constexpr int PR_COMPACT_TILE = 4096;

int fused_variant_id(int L, int P);
// fused variant with L lanes and the fewest panel slots that hold nt_max triangles (-1: none)
int fused_variant_for(int L, int nt_max);
hipError_t launch_fly_all(const FlyAllArgs& a, int grid, hipStream_t s);
// Small chunks: one flight per workgroup of `waves` (2 or 4) waves, every bucket's list
// (positions numbered across the buckets in bucket order); triangles <= PR_SPLIT_NT.
hipError_t launch_fly_split(const FlyAllArgs& a, int waves, int grid, hipStream_t s);
constexpr int PR_SPLIT_NT = 2048;
hipError_t fly_all_occupancy(int* blocks_per_cu);
hipError_t launch_compact(const CompactArgs& a, long long n_total, hipStream_t s);
// perm[0..n) = genome indices stably sorted by panel-slot key (triangle count / 8); work:
// sort_work_ints(n) ints of device memory (per-tile key counts)
hipError_t launch_sort_genomes(const int* nt, int n, int* perm, int* work, hipStream_t s);
size_t sort_work_ints(int n);

hipError_t launch_layout(const float* genomes, const int* res, int ch, int sp, int max_ch, int max_sp, int n,
                         int* nt_raw, hipStream_t s);
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
