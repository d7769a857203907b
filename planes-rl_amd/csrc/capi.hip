// capi.hip -- host runtime behind include/planes_hip.h.
//
// One pr_ctx owns the HBM buffers of one evaluation slot (the reference's PanelBuf,
// sim.cu:15-21, without its MAX_T padding) and one HIP stream.  pr_fly schedules the
// flights as 1-s chunks: each chunk flies every triangle-count bucket's work list with
// the persistent flight kernel, which marks the flights that go on; a stable compaction
// turns the marks into the next chunk's lists on the device, so the host never
// synchronises between chunks and the lists keep a data-determined order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/planes_hip.h"
#include "pr_kernels.h"

using namespace pr;

static thread_local std::string g_err;

static int fail(int code, const char* what) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s (code %d: %s)", what, code,
             code > 0 ? hipGetErrorString((hipError_t)code) : "invalid argument");
    g_err = buf;
    return code ? code : -1;
}
#define HCK(x)                                                        \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) return fail((int)e_, #x);               \
    } while (0)
#define ARG(cond, msg)                          \
    do {                                        \
        if (!(cond)) { g_err = msg; return -1; } \
    } while (0)

// Flight-kernel mapping of one triangle-count bucket: (L, P) while many flights remain,
// (Lw, Pw) = one flight per wave once they fit one per wave (Lw = 0: none).
struct Bucket { int nt_max, L, P, Lw, Pw; };

struct pr_ctx {
    hipStream_t stream = nullptr;
    int cap_genomes = 0, cap_scen = 0, cap_flights = 0;
    long long cap_panels = 0;
    int max_ch = 5, max_sp = 8, nv_max = 0;
    // genome side
    int n_prepared = 0;
    int *d_ntraw = nullptr, *d_off = nullptr, *d_nt = nullptr, *d_nv = nullptr;
    float* d_ext = nullptr;
    Mass* d_mass = nullptr;
    float* d_panel[7] = {};
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    // scenarios
    int n_scen = 0;
    int* d_seeds = nullptr;
    float* d_scen = nullptr;
    // flights
    FlightState* d_state = nullptr;
    FlightConst* d_fc = nullptr;
    int* d_lists[2] = {};
    int* d_counts = nullptr;  // PR_CNT per bucket: 2 alternating list counts, 2 claim counters
    int* d_tcount = nullptr;  // [NB_MAX][max_tiles] survivors per compaction tile
    int* d_perm = nullptr;    // [max_genomes] genomes in chunk 0's list order (sorted by mesh size)
    int* d_sortwork = nullptr;  // [sort_work_ints(max_genomes)] the size sort's per-tile key counts
    int max_tiles = 0;
    int* d_bucket_nt = nullptr;
    unsigned long long* d_stats = nullptr;
    unsigned long long* d_cstats = nullptr;  // [PR_MAX_CHUNKS][2] per-chunk flight-/panel-steps
    int* d_steps_own = nullptr;              // step counts for callers that do not ask for them (pr_sim)
    int* d_chunk_n = nullptr;                // [PR_MAX_CHUNKS][NB_MAX] flights entering each chunk
    std::vector<Bucket> buckets;
    std::vector<int> grid;    // persistent grid per bucket
    std::vector<int> grid_w;  // persistent grid of the bucket's one-flight-per-wave mapping
    std::vector<int> vid_p, vid_w;  // fused-kernel variant ids per bucket (-1: own k_fly launch)
    int grid_f = 0;                 // persistent grid of the fused kernel
    int simds = 0;                  // SIMDs on the device (4 per CU)
    hipEvent_t ev[2 * 512 + 2] = {};
    int ev_slot[512] = {};          // slot of a timed launch; -1 = the fused kernel
    int ev_chunk[512] = {};         // chunk of a timed launch
    hipEvent_t g0 = nullptr, g1 = nullptr;
    // hipGraph of the whole flight sequence (optional), keyed by its launch parameters
    int use_graph = 0;
    int prep_fp64 = 0;
    int tail_wide = 2;
    int fused = 1;
    int sort_lists = 1;
    int split = 1;  // small chunks by k_fly_split (several waves per flight)
    int pending = 0, pend_events = 0, pend_chunks = 0;  // an asynchronous pr_fly in flight
    hipGraphExec_t gexec = nullptr;
    long long gkey[10] = {};
    pr_stats stats = {};
    hipEvent_t prep0 = nullptr, prep1 = nullptr;
};

static const int NB_MAX = 8;
// work counters: [bucket][many per wave, one per wave][flight-steps, panel-steps], then the
// small-chunk kernel's pair
static const int PR_STATS_N = 4 * NB_MAX + 2;

static int mesh_bounds(int ch, int sp, int* nv, int* nt) {
    int nw = (2 * sp + 1) * (ch + 1), tw = 4 * sp * ch;
    *nv = 2 * nw + 9 + 12 + 8 + 8;
    *nt = 2 * tw + 8 + 6 + 6 + 4;
    return 0;
}

static int wide_for(int nt_max, int* Lw, int* Pw) {
    static const int Ps[] = {8, 16, 24, 32};
    for (int p : Ps)
        if (64 * p >= nt_max) { *Lw = 64; *Pw = p; return 0; }
    *Lw = 0; *Pw = 0;
    return 0;
}

static int default_buckets(pr_ctx* c) {
    // smallest group whose lanes hold the flight's panels: up to 24 slots per lane in VGPRs,
    // 33 with the LDS slots (P = 33: 22 in VGPRs + 11 in LDS) -- narrower groups put more
    // flights in a wave and share its per-step rigid-body work among them (measured +5.8 %
    // on the 100k-genome 200..900-triangle mix).  A wave runs as many slots as its fullest
    // group needs, so the VGPR-only L = 16 and 32 ranges are split in two (split = 1): the
    // flights sharing a wave then differ less in triangle count.  Split entries are dropped
    // first when the resolution needs more than NB_MAX buckets.
    static const int table[][4] = {{192, 8, 24, 0},  {264, 8, 33, 0},  {320, 16, 24, 1},  {384, 16, 24, 0},
                                   {528, 16, 33, 0}, {640, 32, 24, 1}, {768, 32, 24, 0},  {1024, 32, 33, 0},
                                   {1536, 64, 24, 0}, {2048, 64, 32, 0}};
    int nv, nt;
    mesh_bounds(c->max_ch, c->max_sp, &nv, &nt);
    std::vector<const int*> rows;
    for (const auto& t : table) {
        rows.push_back(t);
        if (t[0] >= nt) break;
    }
    if (rows.back()[0] < nt) return fail(-1, "mesh resolution exceeds the largest flight-kernel bucket (2048 triangles)");
    for (size_t i = 0; i < rows.size() && rows.size() > (size_t)NB_MAX;)
        if (rows[i][3]) rows.erase(rows.begin() + i);
        else i++;
    c->buckets.clear();
    for (const int* t : rows) {
        Bucket b = {t[0], t[1], t[2], 0, 0};
        wide_for(b.nt_max, &b.Lw, &b.Pw);
        if (b.Lw == b.L && b.Pw == b.P) b.Lw = 0;
        c->buckets.push_back(b);
    }
    return 0;
}

static int upload_buckets(pr_ctx* c) {
    if (c->gexec) {  // a captured flight graph holds the old mapping
        HCK(hipGraphExecDestroy(c->gexec));
        c->gexec = nullptr;
    }
    int nb = (int)c->buckets.size();
    std::vector<int> h(nb);
    for (int i = 0; i < nb; i++) h[i] = c->buckets[i].nt_max;
    HCK(hipMemcpyAsync(c->d_bucket_nt, h.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    HCK(hipStreamSynchronize(c->stream));
    int dev = 0, ncu = 0;
    HCK(hipGetDevice(&dev));
    HCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    c->grid.assign(nb, 0);
    c->grid_w.assign(nb, 0);
    c->vid_p.assign(nb, -1);
    c->vid_w.assign(nb, -1);
    for (int i = 0; i < nb; i++) {
        int occ = 0;
        HCK(fly_occupancy(c->buckets[i].L, c->buckets[i].P, &occ));
        c->grid[i] = (occ < 1 ? 1 : occ) * ncu;
        if (c->buckets[i].Lw) {
            HCK(fly_occupancy(c->buckets[i].Lw, c->buckets[i].Pw, &occ));
            c->grid_w[i] = (occ < 1 ? 1 : occ) * ncu;
        }
        // a bucket joins the fused kernel only if both of its mappings are compiled into it
        const int vp = fused_variant_id(c->buckets[i].L, c->buckets[i].P);
        const int vw = c->buckets[i].Lw ? fused_variant_id(c->buckets[i].Lw, c->buckets[i].Pw) : -1;
        if (c->fused && vp >= 0 && (c->buckets[i].Lw == 0 || vw >= 0)) {
            c->vid_p[i] = vp;
            c->vid_w[i] = vw;
        }
    }
    int occ = 0;
    HCK(fly_all_occupancy(&occ));
    c->grid_f = (occ < 1 ? 1 : occ) * ncu;
    c->simds = 4 * ncu;
    return 0;
}

extern "C" {

const char* pr_last_error(void) { return g_err.c_str(); }

int pr_create(pr_ctx** out, int max_genomes, long long max_panels, int max_scenarios, int max_flights, int max_ch,
              int max_sp) {
    ARG(out, "pr_create: null output");
    ARG(max_genomes > 0 && max_scenarios > 0 && max_flights > 0, "pr_create: capacities must be positive");
    pr_ctx* c = new pr_ctx;
    *out = nullptr;
    c->max_ch = max_ch > 0 ? max_ch : 5;
    c->max_sp = max_sp > 0 ? max_sp : 8;
    int nv, nt;
    mesh_bounds(c->max_ch, c->max_sp, &nv, &nt);
    c->nv_max = nv;
    if (max_panels <= 0) max_panels = (long long)max_genomes * nt;
    c->cap_genomes = max_genomes;
    c->cap_panels = max_panels;
    c->cap_scen = max_scenarios;
    c->cap_flights = max_flights;
    int rc;
    if ((rc = default_buckets(c))) { delete c; return rc; }
#define ALLOC(p, bytes)                                              \
    do {                                                             \
        hipError_t e_ = hipMalloc((void**)&(p), (bytes));           \
        if (e_ != hipSuccess) { pr_destroy(c); return fail((int)e_, "hipMalloc " #p); } \
    } while (0)
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return fail(-1, "stream"); }
    ALLOC(c->d_ntraw, sizeof(int) * (max_genomes + 1));
    ALLOC(c->d_off, sizeof(int) * (max_genomes + 1));
    ALLOC(c->d_nt, sizeof(int) * max_genomes);
    ALLOC(c->d_nv, sizeof(int) * max_genomes);
    ALLOC(c->d_ext, sizeof(float) * 3 * max_genomes);
    ALLOC(c->d_mass, sizeof(Mass) * max_genomes);
    for (int k = 0; k < 7; k++) ALLOC(c->d_panel[k], sizeof(float) * max_panels);
    ALLOC(c->d_seeds, sizeof(int) * max_scenarios);
    ALLOC(c->d_scen, sizeof(float) * NSF * max_scenarios);
    ALLOC(c->d_state, sizeof(FlightState) * max_flights);
    ALLOC(c->d_steps_own, sizeof(int) * max_flights);
    ALLOC(c->d_fc, sizeof(FlightConst) * max_flights);
    for (int k = 0; k < 2; k++) ALLOC(c->d_lists[k], sizeof(int) * (size_t)NB_MAX * max_flights);
    ALLOC(c->d_counts, sizeof(int) * PR_CNT * NB_MAX);
    c->max_tiles = (int)((max_flights + PR_COMPACT_TILE - 1) / PR_COMPACT_TILE);
    ALLOC(c->d_tcount, sizeof(int) * NB_MAX * (size_t)c->max_tiles);
    ALLOC(c->d_perm, sizeof(int) * max_genomes);
    ALLOC(c->d_sortwork, sizeof(int) * sort_work_ints(max_genomes));
    ALLOC(c->d_bucket_nt, sizeof(int) * NB_MAX);
    ALLOC(c->d_stats, sizeof(unsigned long long) * PR_STATS_N);
    ALLOC(c->d_cstats, sizeof(unsigned long long) * 2 * PR_MAX_CHUNKS);
    ALLOC(c->d_chunk_n, sizeof(int) * PR_MAX_CHUNKS * NB_MAX);
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, c->scan_tmp_bytes, c->d_ntraw, c->d_off, max_genomes + 1, c->stream) !=
        hipSuccess) { pr_destroy(c); return fail(-1, "scan workspace query"); }
    ALLOC(c->d_scan_tmp, c->scan_tmp_bytes + 16);
#undef ALLOC
    for (auto& e : c->ev) if (hipEventCreate(&e) != hipSuccess) { pr_destroy(c); return fail(-1, "event"); }
    if (hipEventCreate(&c->g0) != hipSuccess || hipEventCreate(&c->g1) != hipSuccess) { pr_destroy(c); return fail(-1, "event"); }
    if (hipEventCreate(&c->prep0) != hipSuccess || hipEventCreate(&c->prep1) != hipSuccess) { pr_destroy(c); return fail(-1, "event"); }
    if ((rc = upload_buckets(c))) { pr_destroy(c); return rc; }
    *out = c;
    return 0;
}

void pr_destroy(pr_ctx* c) {
    if (!c) return;
    // teardown: release everything; failures here have nowhere to be reported
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* bufs[] = {c->d_ntraw, c->d_off, c->d_nt, c->d_nv, c->d_ext, c->d_mass, c->d_seeds, c->d_scen, c->d_state,
                    c->d_fc, c->d_lists[0], c->d_lists[1], c->d_counts, c->d_bucket_nt, c->d_stats, c->d_scan_tmp,
                    c->d_cstats, c->d_chunk_n, c->d_steps_own, c->d_tcount, c->d_perm, c->d_sortwork};
    for (void* b : bufs) if (b) (void)hipFree(b);
    for (int k = 0; k < 7; k++) if (c->d_panel[k]) (void)hipFree(c->d_panel[k]);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->prep0) (void)hipEventDestroy(c->prep0);
    if (c->prep1) (void)hipEventDestroy(c->prep1);
    if (c->g0) (void)hipEventDestroy(c->g0);
    if (c->g1) (void)hipEventDestroy(c->g1);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int pr_set_buckets(pr_ctx* c, int n, const int* nt_max, const int* L, const int* P) {
    ARG(c && n > 0 && n <= NB_MAX, "pr_set_buckets: 1..8 buckets");
    std::vector<Bucket> b(n);
    for (int i = 0; i < n; i++) {
        b[i] = {nt_max[i], L[i], P[i], 0, 0};
        ARG(L[i] * P[i] >= nt_max[i], "pr_set_buckets: L*P below the bucket's triangle count");
        ARG(i == 0 || nt_max[i] > nt_max[i - 1], "pr_set_buckets: nt_max must increase");
        int occ = 0;
        ARG(fly_occupancy(L[i], P[i], &occ) == hipSuccess, "pr_set_buckets: (L, P) is not a compiled variant");
        if (c->tail_wide) {
            wide_for(nt_max[i], &b[i].Lw, &b[i].Pw);
            if (b[i].Lw == L[i] && b[i].Pw == P[i]) b[i].Lw = 0;
        }
    }
    int nv, nt;
    mesh_bounds(c->max_ch, c->max_sp, &nv, &nt);
    ARG(b.back().nt_max >= nt, "pr_set_buckets: last bucket below the context's largest triangle count");
    c->buckets = b;
    return upload_buckets(c);
}

// 0: one mapping per bucket; 1: every bucket flies its last flights of a chunk one per
// wave; 2 (default): in the fused kernel a bucket does so only when the buckets drained
// after it hold fewer flights than there are waves (else its last flights overlap their
// flights); own-launch buckets as in 1
int pr_set_tail_mapping(pr_ctx* c, int on) {
    ARG(c && on >= 0 && on <= 2, "pr_set_tail_mapping: null context or mode outside 0..2");
    c->tail_wide = on;
    for (Bucket& b : c->buckets) {
        b.Lw = 0; b.Pw = 0;
        if (c->tail_wide) {
            wide_for(b.nt_max, &b.Lw, &b.Pw);
            if (b.Lw == b.L && b.Pw == b.P) b.Lw = 0;
        }
    }
    return upload_buckets(c);
}

int pr_fly_variants(int* n, int* L, int* P, int cap) {
    int k = fly_variant_count();
    *n = k;
    for (int i = 0; i < k && i < cap; i++) fly_variant_info(i, L + i, P + i);
    return 0;
}

int pr_prepare(pr_ctx* c, const float* d_genomes, const int* d_res, int n) {
    ARG(c && d_genomes, "pr_prepare: null argument");
    ARG(n > 0 && n <= c->cap_genomes, "pr_prepare: n outside 1..max_genomes");
    if (c->prep0) HCK(hipEventRecord(c->prep0, c->stream));  // untimed in the allocation-free path
    HCK(launch_layout(d_genomes, d_res, 5, 8, c->max_ch, c->max_sp, n, c->d_ntraw, c->stream));
    HCK(launch_zero(c->d_ntraw + n, 1, c->stream));
    size_t tb = c->scan_tmp_bytes;
    HCK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, c->d_ntraw, c->d_off, n + 1, c->stream));
    int nv, nt;
    mesh_bounds(c->max_ch, c->max_sp, &nv, &nt);
    if ((long long)n * nt > c->cap_panels) {  // the packed total may still fit: check it exactly
        int total = 0;
        HCK(hipMemcpyAsync(&total, c->d_off + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HCK(hipStreamSynchronize(c->stream));
        ARG(total <= c->cap_panels, "pr_prepare: packed panel total exceeds max_panels");
    }
    PrepareArgs a;
    a.genomes = d_genomes; a.res = d_res; a.def_ch = 5; a.def_sp = 8; a.nv_max = c->nv_max;
    a.max_ch = c->max_ch; a.max_sp = c->max_sp;
    a.off = c->d_off; a.mass = c->d_mass; a.nt = c->d_nt; a.nv = c->d_nv; a.ext = c->d_ext;
    a.pcx = c->d_panel[0]; a.pcy = c->d_panel[1]; a.pcz = c->d_panel[2]; a.pA = c->d_panel[3];
    a.pnx = c->d_panel[4]; a.pny = c->d_panel[5]; a.pnz = c->d_panel[6];
    HCK(launch_prepare(a, n, c->prep_fp64, c->stream));
    if (c->prep1) HCK(hipEventRecord(c->prep1, c->stream));
    c->n_prepared = n;
    return 0;
}

int pr_scenarios(pr_ctx* c, const int* d_seeds, int n, const pr_sim_params* p) {
    ARG(c && d_seeds && p, "pr_scenarios: null argument");
    ARG(n > 0 && n <= c->cap_scen, "pr_scenarios: n outside 1..max_scenarios");
    ARG(p->preset == 0 || p->preset == 1, "pr_scenarios: preset must be 0 or 1");
    HCK(hipMemcpyAsync(c->d_seeds, d_seeds, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    ScenarioParams sp;
    sp.preset = p->preset; sp.gust_rms_max = p->gust_rms_max; sp.warp_max = p->warp_max; sp.wind_sd = p->wind_sd;
    HCK(launch_scenarios(c->d_seeds, n, sp, c->d_scen, c->stream));
    c->n_scen = n;
    return 0;
}

static void fill_fly_args(pr_ctx* c, const pr_sim_params* p, float* d_fitness, int* d_steps, FlyArgs& a) {
    a.nt = c->d_nt; a.off = c->d_off; a.fc = c->d_fc;
    a.pcx = c->d_panel[0]; a.pcy = c->d_panel[1]; a.pcz = c->d_panel[2]; a.pA = c->d_panel[3];
    a.pnx = c->d_panel[4]; a.pny = c->d_panel[5]; a.pnz = c->d_panel[6];
    a.state = c->d_state; a.fitness = d_fitness; a.steps = d_steps;
    a.chunk = p->chunk; a.max_steps = p->max_steps; a.spread = p->spread; a.stats = c->d_stats;
    a.list_in = nullptr; a.count_in = nullptr; a.marks = nullptr; a.head = nullptr;
    a.run_lo = 0; a.run_hi = 0x7fffffff;
    a.counts_all = nullptr; a.nb_all = 0; a.c_all = 0; a.split_max = 0;
}

static CompactArgs compact_args(pr_ctx* c, int NB) {
    CompactArgs ca = {};
    ca.dst = c->d_lists[0];
    ca.dst_stride = c->cap_flights;
    ca.counts = c->d_counts;
    ca.tcount = c->d_tcount;
    ca.max_tiles = c->max_tiles;
    ca.nb = NB;
    ca.chunk_n = nullptr;
    return ca;
}

static int init_flights(pr_ctx* c, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                        int* d_steps, long long* nf_out) {
    ARG(c && p && d_fitness && d_steps, "pr_fly: null argument");
    ARG(layout == 0 || layout == 1, "pr_fly: layout must be 0 or 1");
    ARG(n_genomes > 0 && n_genomes <= c->n_prepared, "pr_fly: n_genomes exceeds the prepared genomes");
    ARG(n_scen > 0 && n_scen <= c->n_scen, "pr_fly: n_scen exceeds the prepared scenarios");
    ARG(layout == 0 || n_genomes == n_scen, "pr_fly: layout 1 needs n_genomes == n_scen");
    ARG(p->max_steps > 0 && p->chunk > 0, "pr_fly: max_steps and chunk must be positive");
    long long nf = layout == 0 ? (long long)n_genomes * n_scen : n_genomes;
    ARG(nf <= c->cap_flights, "pr_fly: flights exceed max_flights");
    const int NB = (int)c->buckets.size();
    HCK(launch_zero(c->d_counts, PR_CNT * NB_MAX, c->stream));
    HCK(launch_zero(reinterpret_cast<int*>(c->d_stats), 2 * PR_STATS_N, c->stream));
    HCK(launch_zero(reinterpret_cast<int*>(c->d_cstats), 4 * PR_MAX_CHUNKS, c->stream));
    HCK(launch_zero(c->d_chunk_n, PR_MAX_CHUNKS * NB_MAX, c->stream));
    QueueArgs q;
    q.n_flights = (int)nf; q.layout = layout; q.nseeds = n_scen; q.nt = c->d_nt;
    q.mass = c->d_mass; q.scen = c->d_scen; q.seeds = c->d_seeds; q.spread = p->spread; q.fc = c->d_fc;
    q.n_buckets = NB; q.bucket_nt = c->d_bucket_nt; q.bucket = c->d_lists[1];
    q.state = c->d_state; q.fitness = d_fitness; q.steps = d_steps;
    HCK(launch_queue(q, c->stream));
    // chunk 0's lists (count slot 0): each bucket's flights with their genomes in order of
    // mesh size (a stable sort, so the order is a function of the data), or in flight order
    CompactArgs ca = compact_args(c, NB);
    ca.src = c->d_lists[1]; ca.src_stride = 0; ca.by_bucket = 1; ca.n_fixed = (int)nf; ca.c_dst = 0;
    if (c->sort_lists) {
        HCK(launch_sort_genomes(c->d_nt, n_genomes, c->d_perm, c->d_sortwork, c->stream));
        ca.perm = c->d_perm;
        ca.perm_div = layout == 0 ? n_scen : 1;
    }
    HCK(launch_compact(ca, nf * NB, c->stream));
    *nf_out = nf;
    return 0;
}

int pr_init_flights(pr_ctx* c, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                    int* d_steps) {
    long long nf;
    int rc = init_flights(c, layout, n_genomes, n_scen, p, d_fitness, d_steps, &nf);
    if (rc) return rc;
    HCK(hipStreamSynchronize(c->stream));
    return 0;
}

// Issue the flight sequence after k_init on the context stream, one 1-s chunk at a time:
// the fused kernel flies every bucket compiled into it (one launch per chunk); a bucket
// outside it gets its own k_fly launches; then one compaction builds every bucket's list
// for the next chunk from the marks (list order kept).  Chunk ck reads count slot ck & 1
// and its compaction writes slot (ck + 1) & 1.
static int issue_chunks(pr_ctx* c, const pr_sim_params* p, FlyArgs a, long long nf, bool timed, int* n_events,
                        bool capturing) {
    const int NB = (int)c->buckets.size();
    // a batch that fits one flight per SIMD (e.g. the 8 x 128 holdout) launches one wave per
    // SIMD: every flight then has a SIMD to itself and its steps run at single-wave latency
    const int grid = nf <= c->simds ? std::min(c->simds, c->grid_f) : c->grid_f;
    const int n_chunks = (p->max_steps + p->chunk - 1) / p->chunk;
    const int n_ev = (int)(sizeof(c->ev) / sizeof(c->ev[0]));
    bool any_fused = false, any_own = false;
    for (int b = 0; b < NB; b++) (c->vid_p[b] >= 0 ? any_fused : any_own) = true;
    // a bucket outside the fused kernel that received no flights needs no launches (lists
    // only shrink between chunks); known from k_init's counts, except under graph capture
    bool own_used[NB_MAX] = {};
    if (any_own) {
        int cnt[PR_CNT * NB_MAX];
        if (!capturing) {
            HCK(hipMemcpyAsync(cnt, c->d_counts, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
            HCK(hipStreamSynchronize(c->stream));
        }
        for (int b = 0; b < NB; b++) own_used[b] = c->vid_p[b] < 0 && (capturing || cnt[PR_CNT * b] > 0);
        any_own = false;
        for (int b = 0; b < NB; b++) any_own = any_own || own_used[b];
    }
    FlyAllArgs F;
    F.base = a;
    F.lists = c->d_lists[0];
    F.marks = c->d_lists[1];
    F.list_stride = c->cap_flights;
    F.counts = c->d_counts;
    F.nb = NB;
    F.stats = c->d_stats;
    F.split_stats = c->d_stats + 4 * NB_MAX;
    for (int b = 0; b < 8; b++) {
        F.vid_p[b] = b < NB ? c->vid_p[b] : -1;
        F.vid_w[b] = b < NB ? c->vid_w[b] : -1;
        F.thr[b] = grid;  // one flight per wave once they fit one per wave
        F.lanes[b] = b < NB ? c->buckets[b].L : 64;
        F.vid_2[b] = b < NB && c->buckets[b].L < 64 ? fused_variant_for(2 * c->buckets[b].L, c->buckets[b].nt_max) : -1;
        F.vid_4[b] = b < NB && c->buckets[b].L < 32 ? fused_variant_for(4 * c->buckets[b].L, c->buckets[b].nt_max) : -1;
    }
    F.tail_when_short = c->tail_wide == 2;
    // widening factor 2 (measured against 1, 4 and 8: C3 within noise of each other up to 4,
    // C2 +1 % over 1; 8 slows C3)
    F.widen = c->tail_wide == 2 ? 2 : 0;
    // Small chunks: at most one flight per 4 (2) SIMDs are flown with the flight's panels
    // spread over a workgroup of 4 (2) waves, one wave per SIMD; the other launches of such
    // a chunk return at once.  Decided on the device from the chunk's flight count, so a
    // captured graph holds all the launches.
    const int s4 = c->simds / 4, s2 = c->simds / 2;
    const bool split = c->split && s4 > 0 && c->buckets.back().nt_max <= PR_SPLIT_NT;
    a.counts_all = c->d_counts;
    a.nb_all = NB;
    a.split_max = split ? s2 : 0;
    F.base.counts_all = a.counts_all;
    F.base.nb_all = NB;
    F.base.split_max = a.split_max;
    CompactArgs ca = compact_args(c, NB);
    ca.src = c->d_lists[1];
    ca.src_stride = c->cap_flights;
    ca.by_bucket = 0;
    int ne = 0;
    for (int ck = 0; ck < n_chunks; ck++) {
        const int cs = ck < PR_MAX_CHUNKS ? ck : PR_MAX_CHUNKS - 1;  // later chunks share the last slot
        F.c_in = ck & 1;
        F.base.c_all = a.c_all = F.c_in;
        F.base.cstats = a.cstats = c->d_cstats + 2 * cs;
        // the fused kernel's timed bracket (when it has one) includes the small-chunk launches
        const bool tev = any_fused && timed && ne + 2 <= n_ev;
        if (tev) HCK(hipEventRecord(c->ev[ne], c->stream));
        if (split) {
            F.split_lo = 1;
            F.split_hi = s4;
            HCK(launch_fly_split(F, 4, s4, c->stream));
            F.split_lo = s4 + 1;
            F.split_hi = s2;
            HCK(launch_fly_split(F, 2, s2, c->stream));
        }
        if (any_fused) {
            HCK(launch_fly_all(F, grid, c->stream));
            if (tev) { HCK(hipEventRecord(c->ev[ne + 1], c->stream)); c->ev_slot[ne / 2] = -1; c->ev_chunk[ne / 2] = cs; ne += 2; }
        }
        for (int b = 0; any_own && b < NB; b++) {
            if (!own_used[b]) continue;
            int* cb = c->d_counts + PR_CNT * b;
            a.list_in = c->d_lists[0] + (size_t)b * c->cap_flights;
            a.marks = c->d_lists[1] + (size_t)b * c->cap_flights;
            a.count_in = cb + F.c_in;
            a.head = cb + 2;
            const Bucket& bk = c->buckets[b];
            for (int w = 0; w < (bk.Lw ? 2 : 1); w++) {
                // w = 0: many flights per wave while more than grid_w flights remain;
                // w = 1: one flight per wave (lowest per-step latency) for the tail
                a.run_lo = (w == 0 && bk.Lw) ? c->grid_w[b] + 1 : 0;
                a.run_hi = w == 0 ? 0x7fffffff : c->grid_w[b];
                a.stats = c->d_stats + 2 * (2 * b + w);
                const bool tev = timed && ne + 2 <= n_ev;
                if (tev) HCK(hipEventRecord(c->ev[ne], c->stream));
                if (w == 0) HCK(launch_fly(bk.L, bk.P, a, c->grid[b], c->stream));
                else HCK(launch_fly(bk.Lw, bk.Pw, a, c->grid_w[b], c->stream));
                if (tev) {
                    HCK(hipEventRecord(c->ev[ne + 1], c->stream));
                    c->ev_slot[ne / 2] = 2 * b + w;
                    c->ev_chunk[ne / 2] = cs;
                    ne += 2;
                }
            }
        }
        ca.c_src = ck & 1;
        ca.c_dst = (ck + 1) & 1;
        ca.chunk_n = c->d_chunk_n + NB_MAX * cs;
        HCK(launch_compact(ca, nf, c->stream));
    }
    *n_events = ne;
    return 0;
}

// 1 (default): buckets whose mappings are compiled into the fused flight kernel are flown
// by it, one launch per chunk; 0: every bucket by its own k_fly launches (per-mapping timing)
int pr_set_fused(pr_ctx* c, int on) {
    ARG(c, "pr_set_fused: null context");
    c->fused = on ? 1 : 0;
    return upload_buckets(c);
}

int pr_set_split(pr_ctx* c, int on) {
    ARG(c, "pr_set_split: null context");
    c->split = on ? 1 : 0;
    return 0;
}

int pr_set_list_order(pr_ctx* c, int sorted) {
    ARG(c, "pr_set_list_order: null context");
    c->sort_lists = sorted ? 1 : 0;
    if (c->gexec) {  // a captured flight graph holds the old order
        HCK(hipGraphExecDestroy(c->gexec));
        c->gexec = nullptr;
    }
    return 0;
}

int pr_set_prepare_precision(pr_ctx* c, int bits) {
    ARG(c && (bits == 32 || bits == 64), "pr_set_prepare_precision: 32 or 64");
    c->prep_fp64 = bits == 64;
    return 0;
}

int pr_set_graph(pr_ctx* c, int on) {
    ARG(c, "pr_set_graph: null context");
    c->use_graph = on ? 1 : 0;
    return 0;
}

int pr_fly_async(pr_ctx* c, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                 int* d_steps) {
    ARG(c && p, "pr_fly: null argument");
    ARG(!c->pending, "pr_fly: the previous asynchronous flight batch has not been waited for");
    FlyArgs a;
    fill_fly_args(c, p, d_fitness, d_steps, a);
    int ne = 0;
    long long nf;
    if (c->use_graph) {
        long long key[10] = {layout, n_genomes, n_scen, p->max_steps, p->chunk, (long long)(size_t)d_fitness,
                             (long long)(size_t)d_steps, p->preset, (long long)c->buckets.size() + 64 * c->fused + 128 * c->split,
                             c->n_prepared};
        if (!c->gexec || memcmp(key, c->gkey, sizeof key) != 0) {
            if (c->gexec) { HCK(hipGraphExecDestroy(c->gexec)); c->gexec = nullptr; }
            HCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = init_flights(c, layout, n_genomes, n_scen, p, d_fitness, d_steps, &nf);
            int unused = 0;
            if (rc == 0) rc = issue_chunks(c, p, a, nf, false, &unused, true);
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            HCK(ec);
            hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HCK(ei);
            memcpy(c->gkey, key, sizeof key);
        }
        HCK(hipEventRecord(c->g0, c->stream));
        HCK(hipGraphLaunch(c->gexec, c->stream));
        HCK(hipEventRecord(c->g1, c->stream));
    } else {
        int rc = init_flights(c, layout, n_genomes, n_scen, p, d_fitness, d_steps, &nf);
        if (rc) return rc;
        rc = issue_chunks(c, p, a, nf, true, &ne, false);
        if (rc) return rc;
    }
    c->pending = 1;
    c->pend_events = ne;
    c->pend_chunks = (p->max_steps + p->chunk - 1) / p->chunk;
    return 0;
}

// Waits for the context's flight batch and gathers its statistics.
int pr_fly_wait(pr_ctx* c) {
    ARG(c, "pr_fly_wait: null context");
    if (!c->pending) return 0;
    c->pending = 0;
    HCK(hipStreamSynchronize(c->stream));
    const int NB = (int)c->buckets.size();
    const int ne = c->pend_events;
    pr_stats& S = c->stats;
    S = pr_stats();
    S.n_slots = 2 * NB;
    for (int b = 0; b < NB; b++) {
        const Bucket& bk = c->buckets[b];
        S.s_nt_max[2 * b] = S.s_nt_max[2 * b + 1] = bk.nt_max;
        S.s_L[2 * b] = bk.L; S.s_P[2 * b] = bk.P;
        S.s_L[2 * b + 1] = bk.Lw; S.s_P[2 * b + 1] = bk.Pw;
    }
    for (int k = 0; k < ne; k += 2) {
        float t = 0;
        HCK(hipEventElapsedTime(&t, c->ev[k], c->ev[k + 1]));
        S.fly_ms += t;
        const int sl = c->ev_slot[k / 2];
        if (sl < 0) { S.fused_ms += t; S.fused_launches += 1; }
        else { S.s_fly_ms[sl] += t; S.s_launches[sl] += 1; }
        S.c_ms[c->ev_chunk[k / 2]] += t;
    }
    {
        const int n_chunks = c->pend_chunks;
        S.n_chunks = n_chunks < PR_MAX_CHUNKS ? n_chunks : PR_MAX_CHUNKS;
        unsigned long long cst[2 * PR_MAX_CHUNKS];
        int cn[PR_MAX_CHUNKS * NB_MAX];
        HCK(hipMemcpy(cst, c->d_cstats, sizeof cst, hipMemcpyDeviceToHost));
        HCK(hipMemcpy(cn, c->d_chunk_n, sizeof cn, hipMemcpyDeviceToHost));
        for (int k = 0; k < S.n_chunks; k++) {
            S.c_flight_steps[k] = cst[2 * k];
            S.c_panel_steps[k] = cst[2 * k + 1];
            for (int b = 0; b < NB; b++) S.c_flights[k] += cn[NB_MAX * k + b];
        }
    }
    for (int b = 0; b < NB; b++) S.s_fused[2 * b] = S.s_fused[2 * b + 1] = c->vid_p[b] >= 0;
    if (c->use_graph) {
        float t = 0;
        HCK(hipEventElapsedTime(&t, c->g0, c->g1));
        S.fly_ms = t;
    }
    unsigned long long st[PR_STATS_N];
    HCK(hipMemcpy(st, c->d_stats, sizeof st, hipMemcpyDeviceToHost));
    S.split_flight_steps = st[4 * NB_MAX];
    S.split_panel_steps = st[4 * NB_MAX + 1];
    S.flight_steps += S.split_flight_steps;
    S.panel_steps += S.split_panel_steps;
    for (int k = 0; k < 2 * NB; k++) {
        S.s_flight_steps[k] = st[2 * k];
        S.s_panel_steps[k] = st[2 * k + 1];
        S.flight_steps += st[2 * k];
        S.panel_steps += st[2 * k + 1];
    }
    S.fly_launches = ne / 2;
    float pm = 0;
    if (hipEventElapsedTime(&pm, c->prep0, c->prep1) == hipSuccess) S.prep_ms = pm;
    return 0;
}

int pr_fly(pr_ctx* c, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness, int* d_steps) {
    const int rc = pr_fly_async(c, layout, n_genomes, n_scen, p, d_fitness, d_steps);
    return rc ? rc : pr_fly_wait(c);
}

int pr_fly_record(pr_ctx* c, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                  int* d_steps, float* d_rec, int rec_every, int rec_cap) {
    ARG(d_rec && rec_every >= 1 && rec_cap >= 0, "pr_fly_record: d_rec, rec_every >= 1, rec_cap >= 0");
    long long nf;
    int rc = init_flights(c, layout, n_genomes, n_scen, p, d_fitness, d_steps, &nf);
    if (rc) return rc;
    FlyArgs a;
    fill_fly_args(c, p, d_fitness, d_steps, a);
    a.run_lo = 0;
    a.run_hi = 0x7fffffff;
    // every flight flown by the widest-bucket mapping (recording is not throughput work)
    const Bucket& b = c->buckets.back();
    HCK(launch_fly_rec(b.L, b.P, a, (int)nf, d_rec, rec_every, rec_cap, c->stream));
    HCK(hipStreamSynchronize(c->stream));
    return 0;
}

// gpu_sim (sim.h:12, sim.cu:238-248) as a C entry point: flies the prepared genomes from
// seed `seed` (vary = 0: every plane the same scenario; 1: plane p flies seed + p * 7919),
// launch grid offset `spread`, optional pose recording into d_rec [pop * rec_cap * 8]
// every rec_every steps (>= 1 as in the reference, whose step0 / rec_every runs always).
int pr_sim(pr_ctx* c, int pop, int seed, float* d_fitness, float* d_rec, int rec_every, int rec_cap, float spread,
           int vary, const pr_sim_params* p) {
    ARG(c && d_fitness && p, "pr_sim: null argument");
    ARG(rec_every >= 1, "pr_sim: rec_every must be >= 1");
    ARG(pop > 0 && pop <= c->n_prepared, "pr_sim: pop exceeds the prepared genomes");
    const int ns = vary ? pop : 1;
    ARG(ns <= c->cap_scen, "pr_sim: vary = 1 needs max_scenarios >= pop");
    std::vector<int> seeds(ns);
    for (int i = 0; i < ns; i++) seeds[i] = (int)((unsigned)seed + (unsigned)i * 7919u);  // 32-bit wrap
    HCK(hipMemcpyAsync(c->d_seeds, seeds.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->stream));
    ScenarioParams sp;
    sp.preset = p->preset; sp.gust_rms_max = p->gust_rms_max; sp.warp_max = p->warp_max; sp.wind_sd = p->wind_sd;
    HCK(launch_scenarios(c->d_seeds, ns, sp, c->d_scen, c->stream));
    c->n_scen = ns;
    pr_sim_params q = *p;
    q.spread = spread;
    if (d_rec) return pr_fly_record(c, vary ? 1 : 0, pop, ns, &q, d_fitness, c->d_steps_own, d_rec, rec_every, rec_cap);
    return pr_fly(c, vary ? 1 : 0, pop, ns, &q, d_fitness, c->d_steps_own);
}

int pr_read_extent(pr_ctx* c, int n, float* ext3) {
    ARG(c && ext3 && n >= 0 && n <= c->n_prepared, "pr_read_extent: n exceeds the prepared genomes");
    HCK(hipStreamSynchronize(c->stream));
    HCK(hipMemcpy(ext3, c->d_ext, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    return 0;
}

int pr_scores(pr_ctx* c, const float* d_samples, int pop, int nseeds, float* d_scores) {
    ARG(c && d_samples && d_scores, "pr_scores: null argument");
    ARG(pop > 0 && nseeds > 0 && nseeds <= 256, "pr_scores: pop > 0 and 1 <= nseeds <= 256");
    HCK(launch_scores(d_samples, pop, nseeds, d_scores, c->stream));
    HCK(hipStreamSynchronize(c->stream));
    return 0;
}

int pr_stats_get(pr_ctx* c, pr_stats* out) {
    ARG(c && out, "pr_stats_get: null argument");
    *out = c->stats;
    return 0;
}

int pr_read_prepared(pr_ctx* c, int n, float* mass20, int* nt, int* nv, int* off) {
    ARG(c && n >= 0 && n <= c->n_prepared, "pr_read_prepared: n exceeds the prepared genomes");
    HCK(hipStreamSynchronize(c->stream));
    if (mass20) HCK(hipMemcpy(mass20, c->d_mass, sizeof(Mass) * n, hipMemcpyDeviceToHost));
    if (nt) HCK(hipMemcpy(nt, c->d_nt, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (nv) HCK(hipMemcpy(nv, c->d_nv, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (off) HCK(hipMemcpy(off, c->d_off, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    return 0;
}

int pr_read_panels(pr_ctx* c, long long first, int count, float* out7) {
    ARG(c && out7 && first >= 0 && count >= 0 && first + count <= c->cap_panels, "pr_read_panels: range");
    HCK(hipStreamSynchronize(c->stream));
    std::vector<float> tmp(count);
    for (int k = 0; k < 7; k++) {
        HCK(hipMemcpy(tmp.data(), c->d_panel[k] + first, sizeof(float) * count, hipMemcpyDeviceToHost));
        for (int i = 0; i < count; i++) out7[7 * i + k] = tmp[i];
    }
    return 0;
}

int pr_read_scenarios(pr_ctx* c, int n, float* out26) {
    ARG(c && out26 && n >= 0 && n <= c->n_scen, "pr_read_scenarios: n exceeds the prepared scenarios");
    HCK(hipStreamSynchronize(c->stream));
    HCK(hipMemcpy(out26, c->d_scen, sizeof(float) * NSF * n, hipMemcpyDeviceToHost));
    return 0;
}

int pr_read_genome_order(pr_ctx* c, int n, int* out) {
    ARG(c && out && n >= 0 && n <= c->n_prepared, "pr_read_genome_order: n exceeds the prepared genomes");
    HCK(hipStreamSynchronize(c->stream));
    HCK(hipMemcpy(out, c->d_perm, sizeof(int) * n, hipMemcpyDeviceToHost));
    return 0;
}

int pr_read_states(pr_ctx* c, int first, int n, float* out16) {
    ARG(c && out16 && first >= 0 && n >= 0 && first + n <= c->cap_flights, "pr_read_states: range");
    HCK(hipStreamSynchronize(c->stream));
    HCK(hipMemcpy(out16, c->d_state + first, sizeof(FlightState) * n, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------- allocation-free evaluation
// pr_eval_noalloc: pr_prepare + pr_scenarios + pr_fly (layout 0) for callers that cannot
// allocate device memory (code under stream capture, tight serving loops).  Every buffer the
// flight path needs beyond the caller's inputs and outputs -- mesh counts and offsets, mass
// properties, packed panels, scenario descriptors, per-flight constants and states, work
// lists, counters, the scan's workspace -- is carved from one file-scope device array.  Each
// device holds its own copy of the array (at its own address, found once per device with
// hipGetSymbolAddress) until the process ends.  Calls that overlap on one device race for it.
static const int PR_NA_GENOMES = 2048;   // genomes per call (reference resolution CH=5, SP=8)
static const int PR_NA_SEEDS = 16;       // scenarios per call
static const int PR_NA_FLIGHTS = PR_NA_GENOMES * PR_NA_SEEDS;
static const size_t PR_NA_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[PR_NA_BYTES];

static const int PR_NA_MAX_DEVICES = 16;
struct NoAllocShell {
    int ready = 0;
    pr_ctx ctx;
};
static NoAllocShell g_na[PR_NA_MAX_DEVICES];

// carves the context's buffers out of the device's scratch array; 256-byte aligned pieces
static int noalloc_setup(int dev, pr_ctx* c) {
    unsigned char* base = nullptr;
    HCK(hipGetSymbolAddress((void**)&base, HIP_SYMBOL(scratch)));
    c->max_ch = 5;
    c->max_sp = 8;
    int nv, nt;
    mesh_bounds(c->max_ch, c->max_sp, &nv, &nt);
    c->nv_max = nv;
    c->cap_genomes = PR_NA_GENOMES;
    c->cap_panels = (long long)PR_NA_GENOMES * nt;
    c->cap_scen = PR_NA_SEEDS;
    c->cap_flights = PR_NA_FLIGHTS;
    c->max_tiles = (PR_NA_FLIGHTS + PR_COMPACT_TILE - 1) / PR_COMPACT_TILE;
    int rc;
    if ((rc = default_buckets(c))) return rc;
    size_t used = 0;
    bool fits = true;
    auto take = [&](size_t bytes) -> void* {
        void* p = base + used;
        used += (bytes + 255) & ~(size_t)255;
        if (used > PR_NA_BYTES) fits = false;
        return p;
    };
    const int G = PR_NA_GENOMES, F = PR_NA_FLIGHTS;
    c->d_ntraw = (int*)take(sizeof(int) * (G + 1));
    c->d_off = (int*)take(sizeof(int) * (G + 1));
    c->d_nt = (int*)take(sizeof(int) * G);
    c->d_nv = (int*)take(sizeof(int) * G);
    c->d_ext = (float*)take(sizeof(float) * 3 * G);
    c->d_mass = (Mass*)take(sizeof(Mass) * G);
    for (int k = 0; k < 7; k++) c->d_panel[k] = (float*)take(sizeof(float) * c->cap_panels);
    c->d_seeds = (int*)take(sizeof(int) * PR_NA_SEEDS);
    c->d_scen = (float*)take(sizeof(float) * NSF * PR_NA_SEEDS);
    c->d_state = (FlightState*)take(sizeof(FlightState) * F);
    c->d_steps_own = (int*)take(sizeof(int) * F);
    c->d_fc = (FlightConst*)take(sizeof(FlightConst) * F);
    for (int k = 0; k < 2; k++) c->d_lists[k] = (int*)take(sizeof(int) * (size_t)NB_MAX * F);
    c->d_counts = (int*)take(sizeof(int) * PR_CNT * NB_MAX);
    c->d_tcount = (int*)take(sizeof(int) * NB_MAX * (size_t)c->max_tiles);
    c->d_perm = (int*)take(sizeof(int) * G);
    c->d_sortwork = (int*)take(sizeof(int) * sort_work_ints(G));
    c->d_bucket_nt = (int*)take(sizeof(int) * NB_MAX);
    c->d_stats = (unsigned long long*)take(sizeof(unsigned long long) * PR_STATS_N);
    c->d_cstats = (unsigned long long*)take(sizeof(unsigned long long) * 2 * PR_MAX_CHUNKS);
    c->d_chunk_n = (int*)take(sizeof(int) * PR_MAX_CHUNKS * NB_MAX);
    HCK(hipcub::DeviceScan::ExclusiveSum(nullptr, c->scan_tmp_bytes, c->d_ntraw, c->d_off, G + 1, (hipStream_t)0));
    c->d_scan_tmp = take(c->scan_tmp_bytes + 16);
    if (!fits) return fail(-1, "pr_eval_noalloc: scratch layout exceeds the array (build error)");
    c->stream = nullptr;  // set-up on the null stream; each call then runs on the caller's stream
    if ((rc = upload_buckets(c))) return rc;
    (void)dev;
    return 0;
}

int pr_eval_noalloc(const float* d_genomes, int n_genomes, const int* d_seeds, int n_seeds, const pr_sim_params* p,
                    float* d_fitness, int* d_steps, void* stream) {
    ARG(d_genomes && d_seeds && p && d_fitness, "pr_eval_noalloc: null argument");
    ARG(n_genomes >= 1 && n_genomes <= PR_NA_GENOMES,
        "pr_eval_noalloc: n_genomes outside 1..2048 (the scratch array's capacity)");
    ARG(n_seeds >= 1 && n_seeds <= PR_NA_SEEDS, "pr_eval_noalloc: n_seeds outside 1..16 (the scratch array's capacity)");
    ARG(p->max_steps > 0 && p->chunk > 0, "pr_eval_noalloc: max_steps and chunk must be positive");
    int dev = 0;
    HCK(hipGetDevice(&dev));
    ARG(dev >= 0 && dev < PR_NA_MAX_DEVICES, "pr_eval_noalloc: device index beyond 15");
    NoAllocShell& sh = g_na[dev];
    pr_ctx* c = &sh.ctx;
    if (!sh.ready) {
        int rc = noalloc_setup(dev, c);
        if (rc) return rc;
        sh.ready = 1;
    }
    c->stream = (hipStream_t)stream;
    int rc;
    if ((rc = pr_prepare(c, d_genomes, nullptr, n_genomes))) return rc;
    if ((rc = pr_scenarios(c, d_seeds, n_seeds, p))) return rc;
    FlyArgs a;
    int* steps = d_steps ? d_steps : c->d_steps_own;
    fill_fly_args(c, p, d_fitness, steps, a);
    long long nf = 0;
    if ((rc = init_flights(c, 0, n_genomes, n_seeds, p, d_fitness, steps, &nf))) return rc;
    int ne = 0;
    return issue_chunks(c, p, a, nf, false, &ne, false);
}

int pr_dev_alloc(void** p, size_t bytes) { HCK(hipMalloc(p, bytes)); return 0; }
int pr_dev_free(void* p) { HCK(hipFree(p)); return 0; }
int pr_h2d(void* d, const void* h, size_t bytes) { HCK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice)); return 0; }
int pr_d2h(void* h, const void* d, size_t bytes) { HCK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost)); return 0; }
int pr_device_sync(void) { HCK(hipDeviceSynchronize()); return 0; }
int pr_device_name(char* buf, int cap) {
    hipDeviceProp_t prop;
    int dev = 0;
    HCK(hipGetDevice(&dev));
    HCK(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return 0;
}

}  // extern "C"
