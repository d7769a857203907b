// kernels.hip -- gfx950 kernels of the fitness-evaluation path.
//
//   k_layout     genome -> closed-form triangle count (for the exact panel prefix ranges)
//   k_prepare    one wave64 per genome: lane-parallel meshing into LDS, print-volume
//                check, mass properties in the reference's float summation order,
//                packed SoA panel records at exact prefix offsets (no MAX_T padding)
//   k_scenarios  scenario descriptors per seed (reference or config preset)
//   k_init       (genome, scenario) flights -> constants, launch state, bucket of each
//                flight; unprintable -> -100
//   k_fly<L,P>   persistent flight kernel: each group of L lanes flies one flight with
//                P panels per lane held in VGPRs; groups refill from the work list
//   k_fly_all    one chunk of every bucket in one launch (the variants of PR_FUSED)
//   k_compact_*  stable compaction: per-bucket work lists from k_init's buckets and from
//                each chunk's marks, in a data-determined order
//   k_scores     tail-aware robust score per genome
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pr_kernels.h"
#include "pr_physics.h"

namespace pr {

// ------------------------------------------------------------------ cross-lane helpers
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over each aligned group of L lanes; every lane of the group receives the same bits
// (each butterfly stage adds a symmetric partner, and IEEE addition commutes).
template <int L>
__device__ __forceinline__ float group_sum(float v) {
    if (L >= 2) v += dpp_mov<0xB1>(v);    // quad_perm [1,0,3,2]: lane ^ 1
    if (L >= 4) v += dpp_mov<0x4E>(v);    // quad_perm [2,3,0,1]: lane ^ 2
    if (L >= 8) v += dpp_mov<0x141>(v);   // row_half_mirror: quad 0 <-> quad 1 within 8
    if (L >= 16) v += dpp_mov<0x140>(v);  // row_mirror: half-row 0 <-> 1 within 16
    if (L >= 32) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));  // lane ^ 16
    if (L >= 64) v += __shfl_xor(v, 32);
    return v;
}

template <int L>
__device__ __forceinline__ float group_max(float v) {
    if (L >= 2) v = fmaxf(v, dpp_mov<0xB1>(v));
    if (L >= 4) v = fmaxf(v, dpp_mov<0x4E>(v));
    if (L >= 8) v = fmaxf(v, dpp_mov<0x141>(v));
    if (L >= 16) v = fmaxf(v, dpp_mov<0x140>(v));
    if (L >= 32) v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F)));
    if (L >= 64) v = fmaxf(v, __shfl_xor(v, 32));
    return v;
}

// Largest panel-slot count over the wave's groups, as a wave-uniform (scalar) value; every
// lane of the wave must be active.
template <int P>
__device__ __forceinline__ int wave_slots(int np) {
    const int m = (int)group_max<64>((float)np);
    return __builtin_amdgcn_readfirstlane(m < P ? m : P);
}

// ------------------------------------------------------------------ k_layout
// A mesh resolution outside 1..(max_ch, max_sp) -- the bounds the context's buffers and LDS
// are sized for -- is not meshed: the genome gets no triangles and flies as unprintable.
__device__ __forceinline__ bool res_ok(int ch, int sp, int max_ch, int max_sp) {
    return ch >= 1 && sp >= 1 && ch <= max_ch && sp <= max_sp;
}

__global__ void k_layout(const float* __restrict__ genomes, const int* __restrict__ res, int def_ch, int def_sp,
                         int max_ch, int max_sp, int n, int* __restrict__ nt_raw) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g[NG];
    for (int k = 0; k < NG; k++) g[k] = genomes[(size_t)i * NG + k];
    int ch = res ? res[2 * i] : def_ch, sp = res ? res[2 * i + 1] : def_sp;
    nt_raw[i] = res_ok(ch, sp, max_ch, max_sp) ? layout_of(g, ch, sp).nt : 0;
}

// ------------------------------------------------------------------ k_prepare
// block = one wave64 = one genome.  LDS: vertices [nv_max] + 12 x 64 triangle terms, in T.
// T = float reproduces the reference's float computation bit for bit; T = double is the
// FP64 prepare mode (every step in double, results rounded once to float).
template <class T>
__device__ __forceinline__ T wave_min_t(T v) {
    for (int o = 32; o > 0; o >>= 1) v = tmin(v, __shfl_xor(v, o));
    return v;
}
template <class T>
__device__ __forceinline__ T wave_max_t(T v) {
    for (int o = 32; o > 0; o >>= 1) v = tmax(v, __shfl_xor(v, o));
    return v;
}

template <class T>
__global__ __launch_bounds__(64) void k_prepare(PrepareArgs a) {
    extern __shared__ double lds_raw[];
    __shared__ T g[NG];
    const int gi = blockIdx.x, lane = threadIdx.x;
    float gf[2];
    gf[0] = a.genomes[(size_t)gi * NG + G_FIN_LAYOUT];
    gf[1] = a.genomes[(size_t)gi * NG + G_DECK_GAP];
    float gl[NG];  // topology decisions are taken on the float genes
    for (int k = 0; k < NG; k++) gl[k] = 0.0f;
    gl[G_FIN_LAYOUT] = gf[0];
    gl[G_DECK_GAP] = gf[1];
    const int ch = a.res ? a.res[2 * gi] : a.def_ch, sp = a.res ? a.res[2 * gi + 1] : a.def_sp;
    if (!res_ok(ch, sp, a.max_ch, a.max_sp)) {  // not meshed (the first launch records it)
        if (a.nv_lo == 0) {
            float* mo = reinterpret_cast<float*>(a.mass + gi);
            if (lane < 20) mo[lane] = 0.0f;
            if (lane < 3) a.ext[3 * gi + lane] = 0.0f;
            if (lane == 0) {
                a.nt[gi] = 0;
                a.nv[gi] = 0;
            }
        }
        return;
    }
    const MeshLayout L = layout_of(gl, ch, sp);
    if (L.nv < a.nv_lo || L.nv > a.nv_hi) return;  // the other launch's genome (the whole block)
    if (lane < NG) g[lane] = T(a.genomes[(size_t)gi * NG + lane]);
    __syncthreads();
    V3T<T>* Vx = reinterpret_cast<V3T<T>*>(lds_raw);
    T* term = reinterpret_cast<T*>(lds_raw) + 3 * (size_t)a.nv_max;  // [12][64]

    // vertices, lane-parallel; extent by exact min/max reductions
    T lo[3] = {T(3.4e38f), T(3.4e38f), T(3.4e38f)}, hi[3] = {T(-3.4e38f), T(-3.4e38f), T(-3.4e38f)};
    for (int v = lane; v < L.nv; v += 64) {
        V3T<T> p = mesh_vertex(g, L, v);
        Vx[v] = p;
        lo[0] = tmin(lo[0], p.x); lo[1] = tmin(lo[1], p.y); lo[2] = tmin(lo[2], p.z);
        hi[0] = tmax(hi[0], p.x); hi[1] = tmax(hi[1], p.y); hi[2] = tmax(hi[2], p.z);
    }
    bool fits = true;
    for (int k = 0; k < 3; k++) {
        T e;
        {
#pragma clang fp contract(off)
            e = wave_max_t(hi[k]) - wave_min_t(lo[k]);
        }
        fits = fits && (e <= T(PR_PRINT_LIMIT));
        if (lane == 0) a.ext[3 * gi + k] = (float)e;
    }
    __syncthreads();

    // pass 1: mass and first moment; lanes 0..3 each own one serial accumulator chain
    T acc = T(0);
    for (int base = 0; base < L.nt; base += 64) {
        int t = base + lane;
        if (t < L.nt) {
            int id[3];
            tri_of(L, t, id);
            T mi;
            V3T<T> mct;
            tri_mass_terms(Vx[id[0]], Vx[id[1]], Vx[id[2]], &mi, &mct);
            term[0 * 64 + lane] = mi; term[1 * 64 + lane] = mct.x; term[2 * 64 + lane] = mct.y; term[3 * 64 + lane] = mct.z;
        }
        __syncthreads();
        if (lane < 4) {
            int cnt = min(64, L.nt - base);
            for (int k = 0; k < cnt; k++) {
#pragma clang fp contract(off)
                acc = acc + term[lane * 64 + k];
            }
        }
        __syncthreads();
    }
    V3T<T> c;
    T m, mn;
    {
#pragma clang fp contract(off)
        m = __shfl(acc, 0);
        c = VT<T>(__shfl(acc, 1), __shfl(acc, 2), __shfl(acc, 3));
        T shell = m;
        V3T<T> np = VT<T>(T(0.5f) * g[G_CHORD], T(0), T(0));
        mn = g[G_NOSE_MASS] * shell;
        m += mn;
        c = vadd(c, vmul(np, mn));
        c = vmul(c, T(1) / m);
    }
    // pass 2: inertia about c; lanes 0..5 own I[0..5]
    T Ik = T(0);
    for (int base = 0; base < L.nt; base += 64) {
        int t = base + lane;
        if (t < L.nt) {
            int id[3];
            tri_of(L, t, id);
            T o[12];
            tri_inertia_terms(Vx[id[0]], Vx[id[1]], Vx[id[2]], c, o);
            for (int k = 0; k < 12; k++) term[k * 64 + lane] = o[k];
        }
        __syncthreads();
        if (lane < 6) {
            int cnt = min(64, L.nt - base);
            const T* ta = term + lane * 64;
            const T* tb = term + (lane + 6) * 64;
            for (int k = 0; k < cnt; k++) {
#pragma clang fp contract(off)
                if (lane < 3) { Ik = Ik + ta[k]; Ik = Ik + tb[k]; }
                else { Ik = Ik - ta[k]; Ik = Ik - tb[k]; }
            }
        }
        __syncthreads();
    }
    float out[20];
    {
#pragma clang fp contract(off)
        T I[6];
        for (int k = 0; k < 6; k++) I[k] = __shfl(Ik, k);
        V3T<T> r = vsub(VT<T>(T(0.5f) * g[G_CHORD], T(0), T(0)), c);
        I[0] += mn * (r.y * r.y + r.z * r.z); I[1] += mn * (r.x * r.x + r.z * r.z);
        I[2] += mn * (r.x * r.x + r.y * r.y); I[3] -= mn * r.x * r.y;
        I[4] -= mn * r.x * r.z; I[5] -= mn * r.y * r.z;
        const T reg = T(1e-9f);
        T xx = I[0] + reg, yy = I[1] + reg, zz = I[2] + reg;
        T inv[9];
        invert_sym3_t(xx, yy, zz, I[3], I[4], I[5], false, inv);
        out[0] = (float)m;
        out[1] = (float)(T(1) / m);
        out[2] = (float)c.x; out[3] = (float)c.y; out[4] = (float)c.z;
        out[5] = (float)xx; out[6] = (float)yy; out[7] = (float)zz;
        out[8] = (float)I[3]; out[9] = (float)I[4]; out[10] = (float)I[5];
        for (int k = 0; k < 9; k++) out[11 + k] = (float)inv[k];
    }
    const int nt_eff = fits ? L.nt : 0;
    float* mo = reinterpret_cast<float*>(a.mass + gi);
#pragma unroll
    for (int k = 0; k < 20; k++)
        if (lane == k) mo[k] = out[k];
    if (lane == 0) {
        a.nt[gi] = nt_eff;
        a.nv[gi] = L.nv;
    }
    // panel records relative to the shell COM (sim.cu:40-48), exact prefix range
    const int off = a.off[gi];
    for (int t = lane; t < nt_eff; t += 64) {
        int id[3];
        tri_of(L, t, id);
        V3T<T> cen, nrm;
        T area;
        panel_of(Vx[id[0]], Vx[id[1]], Vx[id[2]], c, &cen, &area, &nrm);
        size_t o = (size_t)off + t;
        a.pcx[o] = (float)cen.x; a.pcy[o] = (float)cen.y; a.pcz[o] = (float)cen.z; a.pA[o] = (float)area;
        a.pnx[o] = (float)nrm.x; a.pny[o] = (float)nrm.y; a.pnz[o] = (float)nrm.z;
    }
}

// ------------------------------------------------------------------ k_scenarios
__global__ void k_scenarios(const int* __restrict__ seeds, int n, ScenarioParams sp, float* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float d[NSF];
    if (sp.preset == 0) scenario_reference(seeds[i], d);
    else scenario_config(seeds[i], sp, d);
    for (int k = 0; k < NSF; k++) out[(size_t)i * NSF + k] = d[k];
}

// ------------------------------------------------------------------ k_init
// Flight f -> (genome, scenario) by layout.  Writes the flight's constants (printed mass,
// scenario terms) and launch state (sim.cu:63-109), or fitness -100 for an unprintable
// genome (sim.cu:35-36,109), and the triangle-count bucket of its genome (-1 when the
// flight ends at launch), from which k_compact_* builds the first chunk's lists.
__global__ void k_init(QueueArgs a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;
    if (f < a.n_flights) {
        const int gi = a.layout == 0 ? f / a.nseeds : f;
        const int si = a.layout == 0 ? f - gi * a.nseeds : f;
        const int nt = a.nt[gi];
        FlightState st;
        FlightConst fc;
        float d[NSF];
        for (int k = 0; k < NSF; k++) d[k] = a.scen[(size_t)si * NSF + k];
        Mass m0, mp;
        for (int k = 0; k < 20; k++) reinterpret_cast<float*>(&m0)[k] = reinterpret_cast<const float*>(a.mass + gi)[k];
        vec3 shift;
        printed_mass(m0, d, mp, shift);
        launch_state(d, mp, shift, a.layout == 0 ? gi : f, a.spread, st);
        fc.mass = mp.mass; fc.invmass = mp.invmass;
        for (int k = 0; k < 6; k++) fc.I[k] = mp.I[k];
        fc.iv[0] = mp.invI[0]; fc.iv[1] = mp.invI[1]; fc.iv[2] = mp.invI[2];
        fc.iv[3] = mp.invI[4]; fc.iv[4] = mp.invI[5]; fc.iv[5] = mp.invI[8];
        fc.cf = d[S_CF]; fc.half_rho = 0.5f * d[S_RHO];
        fc.base[0] = d[S_BASE_X]; fc.base[1] = d[S_BASE_Y]; fc.base[2] = d[S_BASE_Z];
        fc.gust = d[S_GUST]; fc.phase = (float)phase_of(a.seeds[si]);
        fc.shift[0] = shift.x; fc.shift[1] = shift.y; fc.shift[2] = shift.z;
        fc.warp_p = d[S_WARP_P]; fc.warp_r = d[S_WARP_R];
        fc.genome = gi; fc.scen = si;
        fc.pad[0] = fc.pad[1] = fc.pad[2] = 0.0f;
        if (nt <= 0) {
            st.done = 1; st.fitness = -100.0f;
            a.fitness[f] = -100.0f;
            a.steps[f] = 0;
        } else {
            b = 0;
            while (b < a.n_buckets - 1 && nt > a.bucket_nt[b]) b++;
        }
        a.fc[f] = fc;
        a.state[f] = st;
        a.bucket[f] = b;
    }
}

__global__ void k_zero_ints(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ k_fly
// Per-lane copy of one group's flight.  All lanes of a group hold bitwise-identical
// values (they compute the same expressions on the same inputs).
struct Body {
    float px, py, pz, qw, qx, qy, qz, vx, vy, vz, wx, wy, wz, t;
    float mass, ks, I0, I1, I2, I3, I4, I5, v0, v1, v2, v4, v5, v8;  // ks = invmass * dt, v* = invI * dt
    float g, bx, by, bz;  // g = -2 / cf (cf: skin friction)
    int step, end, fid, nt, s0;  // s0: step at which this chunk's flying started
};

// wind terms (common.h:462-465): term k = weight * sin(a*p_axis + b*t + c*phase)
__device__ __forceinline__ void wind_term(int k, int& axis, float& a, float& b, float& c, float& w, int& comp) {
    switch (k) {
        case 0: axis = 0; a = 0.9f; b = 1.3f; c = 1.0f; w = 0.55f; comp = 0; break;
        case 1: axis = 2; a = 2.3f; b = 2.1f; c = 1.7f; w = 0.35f; comp = 0; break;
        case 2: axis = 1; a = 4.1f; b = 3.3f; c = 0.4f; w = 0.25f; comp = 0; break;
        case 3: axis = 0; a = 1.1f; b = 1.7f; c = 2.2f; w = 0.50f; comp = 1; break;
        case 4: axis = 2; a = 3.1f; b = 2.9f; c = 0.9f; w = 0.30f; comp = 1; break;
        case 5: axis = 0; a = 1.4f; b = 2.3f; c = 3.1f; w = 0.30f; comp = 2; break;
        case 6: axis = 1; a = 2.7f; b = 1.9f; c = 1.2f; w = 0.20f; comp = 2; break;
        default: axis = 0; a = 0; b = 0; c = 0; w = 0; comp = 3; break;
    }
}

// Gust terms by lane: every aligned 8 lanes hold the 7 terms, x on lanes 0-2 (lane 3
// none), y on 4-5, z on 6-7, so one quad sum gives x and one pair sum gives y and z (the
// addition order of a butterfly over the group, without its selects and extra stages).
__device__ __forceinline__ int wind_term_of_lane(int j8) {
    return j8 < 3 ? j8 : j8 == 3 ? -1 : j8 - 1;
}

// 8-lane broadcast through the LDS crossbar (ds_swizzle bitmask mode: lane & 0x18 | or):
// no VALU issue slot, which is the flight kernel's bottleneck.
template <int OR>
__device__ __forceinline__ float bcast8(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x18 | (OR << 5)));
}

// Rounded terms w * sin, summed without contraction: the same value in every lane of the 8.
__device__ __forceinline__ void gust_sums(float w, float sn, float& gx, float& gy, float& gz) {
#pragma clang fp contract(off)
    const float sv = w * sn;
    const float s1 = sv + dpp_mov<0xB1>(sv);  // lane ^ 1
    const float s2 = s1 + dpp_mov<0x4E>(s1);  // lane ^ 2
    gx = bcast8<0>(s2);
    gy = bcast8<4>(s1);
    gz = bcast8<6>(s1);
}

struct WindLane {  // this lane's gust term (zero weight on lane 3 of each 8)
    float ax, ay, az, bt, cs, wg;
};

// Panel slots per lane beyond which the extra slots live in LDS: P <= 24 slots sit in VGPRs
// (7 per slot; 24 is what 256 VGPRs hold beside the flight state at two waves per SIMD);
// a variant with P > 24 keeps 22 slots in VGPRs (room for the LDS loads in flight) and the
// other P - 22 in the wave's LDS, one column per lane: float [P - 22][7][64].
constexpr int PR_LDS_PV = 22;
__host__ __device__ constexpr int lds_slots(int P) { return P > 24 ? P - PR_LDS_PV : 0; }
// LDS bytes of one wave's panel slots; 11 slots (19.7 KB) keep eight waves per CU within
// the 160 KB of LDS, the same occupancy the VGPRs allow
__host__ __device__ constexpr int lds_bytes(int P) { return lds_slots(P) * 7 * 64 * 4; }

// One group's flight: setup from the flight's constants/state/panels, then 600 Hz steps.
// NW > 1 (L = 64): the flight's panels are spread over the NW waves of a workgroup (lane
// threadIdx.x holds triangles threadIdx.x + s * 64 NW); each step the waves' partial force
// and torque sums meet in LDS behind one barrier and every wave adds them in the same fixed
// order, so all NW waves carry bitwise-identical flight states and take the same number of
// steps (every barrier is reached by all of them).
template <int L, int P, int NW = 1>
struct Flyer {
    static_assert(L >= 8, "the gust terms need 8 lanes per group");
    static_assert(NW == 1 || (L == 64 && lds_slots(P) == 0 && (NW == 2 || NW == 4)),
                  "a flight over several waves: 2 or 4 full waves, panels in VGPRs");
    static constexpr int PQ = lds_slots(P);  // slots in LDS
    static constexpr int PV = P - PQ;        // slots in VGPRs
    static_assert(PQ == 0 || PV % 2 == 0, "the LDS slots continue the VGPR slots' pairs");
    Body B;
    WindLane W;
    float rx[PV], ry[PV], rz[PV], nx[PV], ny[PV], nz[PV], kk[PV];  // kk = cf rho A / 2
    float* lds;  // this lane's LDS column (PQ x 7 rows, 64 floats apart); only when PQ > 0
    float* xsum;  // NW > 1: the workgroup's exchange block, float [2][NW][8]
    int np;  // panel slots this group uses per lane: ceil(nt / (L NW)) <= P
    bool bad;

    // smem: the wave's LDS panel block (lds_bytes(P)); unused when PQ == 0.  NW > 1: the
    // workgroup's exchange block instead
    __device__ __forceinline__ void setup(const FlyArgs& a, int fid, int j, float* smem) {
        lds = smem + (threadIdx.x & 63);
        xsum = smem;
        // this lane's first triangle and the stride between its slots
        const int jt = NW > 1 ? (int)threadIdx.x : j;
        constexpr int TS = L * NW;
        const FlightConst& fc = a.fc[fid];
        const FlightState& st = a.state[fid];
        const int gi = fc.genome;
        B.fid = fid;
        B.nt = a.nt[gi];
        np = (B.nt + TS - 1) / TS;
        const int off = a.off[gi];
        B.px = st.pos[0]; B.py = st.pos[1]; B.pz = st.pos[2];
        B.qw = st.q[0]; B.qx = st.q[1]; B.qy = st.q[2]; B.qz = st.q[3];
        B.vx = st.vel[0]; B.vy = st.vel[1]; B.vz = st.vel[2];
        B.wx = st.wb[0]; B.wy = st.wb[1]; B.wz = st.wb[2];
        B.step = st.step;
        B.s0 = st.step;
        B.t = B.step == 0 ? 0.0f : (float)B.step * PR_DT;  // chunk clock rule (sim.cu:113)
        B.end = min(B.step + a.chunk, a.max_steps);
        B.mass = fc.mass; B.ks = fc.invmass * PR_DT;
        B.I0 = fc.I[0]; B.I1 = fc.I[1]; B.I2 = fc.I[2]; B.I3 = fc.I[3]; B.I4 = fc.I[4]; B.I5 = fc.I[5];
        B.v0 = fc.iv[0] * PR_DT; B.v1 = fc.iv[1] * PR_DT; B.v2 = fc.iv[2] * PR_DT;
        B.v4 = fc.iv[3] * PR_DT; B.v5 = fc.iv[4] * PR_DT; B.v8 = fc.iv[5] * PR_DT;
        // cf >= 0.0135 in both scenario presets; the floor only keeps g finite
        const float cf = fmaxf(fc.cf, 1e-12f);
        B.g = -2.0f / cf;
        B.bx = fc.base[0]; B.by = fc.base[1]; B.bz = fc.base[2];
        const float phase = fc.phase, gust = fc.gust, hr = fc.half_rho;
        const vec3 shift = V(fc.shift[0], fc.shift[1], fc.shift[2]);
        const float wp = fc.warp_p, wr = fc.warp_r;
        {
            int axis, comp;
            float ca, cb, cc, cw;
            wind_term(wind_term_of_lane(j & 7), axis, ca, cb, cc, cw, comp);
            W.ax = axis == 0 ? ca : 0.0f; W.ay = axis == 1 ? ca : 0.0f; W.az = axis == 2 ? ca : 0.0f;
            W.bt = cb; W.cs = cc * phase; W.wg = cw * gust;
        }
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int t = jt + p * TS;
            float v[7] = {0, 0, 0, 0, 0, 1, 0};  // padding: zero area, so zero force
            if (t < B.nt) {
                const size_t o = (size_t)off + t;
                vec3 rb = vsub(V(a.pcx[o], a.pcy[o], a.pcz[o]), shift);
                vec3 nb = warped_normal(rb, V(a.pnx[o], a.pny[o], a.pnz[o]), wp, wr);
                v[0] = rb.x; v[1] = rb.y; v[2] = rb.z;
                v[3] = nb.x; v[4] = nb.y; v[5] = nb.z;
                v[6] = (hr * a.pA[o]) * cf;
            }
            if (p < PV) {
                rx[p] = v[0]; ry[p] = v[1]; rz[p] = v[2]; nx[p] = v[3]; ny[p] = v[4]; nz[p] = v[5]; kk[p] = v[6];
            } else {
#pragma unroll
                for (int k = 0; k < 7; k++) lds[((p - PV) * 7 + k) * 64] = v[k];
            }
        }
    }

    // Force and torque of panel p in body axes (sim.cu:144-158 rotated into the body frame):
    // v_r = v_b + w x r, s = v_r.n/|v_r|, q = k |v_r|^2 (k = rho A / 2), and
    // f = -q cn n - q cf (v_r/|v_r| - s n) = c1 n + c2 v_r with cn = 2 s h,
    // h = 0.95 sqrt(1 - s^2) + |s|.  With d = v_r.n (so s = d/|v_r|) and the panel constant
    // kk = cf k, the |v_r| factors cancel:
    //   c1 = q s (cf - 2h) = kk d (|v_r| + g e),  e = 0.95 sqrt(|v_r|^2 - d^2) + |d|,  g = -2/cf
    //   c2 = -q cf / |v_r| = -kk |v_r|
    // two square roots and no reciprocal: 35 VALU instructions (37 issue slots) per panel.
    // |v_r|^2 - d^2 >= 0 up to rounding (|n| = 1); its magnitude is taken (abs modifier, no
    // instruction), which differs from clamping at zero only when the flow is normal to
    // the panel to within rounding.
    // The reference skips panels with |v_r|^2 < 1e-6 (its NaN guard); here v_r = 0 gives
    // zero force without a guard (both square roots are 0), and 0 < |v_r|^2 < 1e-6
    // contributes < 1e-9 N.
    // FIRST: the lane's first panel initialises the sums instead of adding to zero.
    template <bool FIRST = false>
    __device__ __forceinline__ static void panel_force_at(float prx, float pry, float prz, float pnx, float pny,
                                                          float pnz, float pk, float vbx, float vby, float vbz,
                                                          float wx, float wy, float wz, float g, float& Fx,
                                                          float& Fy, float& Fz, float& Tx, float& Ty, float& Tz) {
        const float vrx = __builtin_fmaf(wy, prz, __builtin_fmaf(-wz, pry, vbx));
        const float vry = __builtin_fmaf(wz, prx, __builtin_fmaf(-wx, prz, vby));
        const float vrz = __builtin_fmaf(wx, pry, __builtin_fmaf(-wy, prx, vbz));
        const float V2 = __builtin_fmaf(vrz, vrz, __builtin_fmaf(vry, vry, vrx * vrx));
        const float d = __builtin_fmaf(vrz, pnz, __builtin_fmaf(vry, pny, vrx * pnx));
        const float vm = __builtin_amdgcn_sqrtf(V2);
        const float sq = __builtin_amdgcn_sqrtf(fabsf(__builtin_fmaf(-d, d, V2)));
        const float e = __builtin_fmaf(0.95f, sq, fabsf(d));
        const float c1 = (pk * d) * __builtin_fmaf(g, e, vm);
        const float c2 = -(pk * vm);
        const float fx = c1 * pnx + c2 * vrx;
        const float fy = c1 * pny + c2 * vry;
        const float fz = c1 * pnz + c2 * vrz;
        if (FIRST) {
            Fx = fx; Fy = fy; Fz = fz;
            Tx = __builtin_fmaf(pry, fz, -prz * fy);
            Ty = __builtin_fmaf(prz, fx, -prx * fz);
            Tz = __builtin_fmaf(prx, fy, -pry * fx);
        } else {
            Fx += fx; Fy += fy; Fz += fz;
            Tx = __builtin_fmaf(pry, fz, __builtin_fmaf(-prz, fy, Tx));
            Ty = __builtin_fmaf(prz, fx, __builtin_fmaf(-prx, fz, Ty));
            Tz = __builtin_fmaf(prx, fy, __builtin_fmaf(-pry, fx, Tz));
        }
    }

    // slot p < PV: the panel held in VGPRs
    template <bool FIRST = false>
    __device__ __forceinline__ void panel_force(int p, float vbx, float vby, float vbz, float wx, float wy, float wz,
                                                float g, float& Fx, float& Fy, float& Fz, float& Tx, float& Ty,
                                                float& Tz) const {
        panel_force_at<FIRST>(rx[p], ry[p], rz[p], nx[p], ny[p], nz[p], kk[p], vbx, vby, vbz, wx, wy, wz, g, Fx, Fy,
                              Fz, Tx, Ty, Tz);
    }

    // slot PV + q: the panel read from this lane's LDS column
    __device__ __forceinline__ void panel_force_lds(int q, float vbx, float vby, float vbz, float wx, float wy,
                                                    float wz, float g, float& Fx, float& Fy, float& Fz, float& Tx,
                                                    float& Ty, float& Tz) const {
        const float* c = lds + q * 7 * 64;
        panel_force_at<false>(c[0], c[64], c[128], c[192], c[256], c[320], c[384], vbx, vby, vbz, wx, wy, wz, g, Fx,
                              Fy, Fz, Tx, Ty, Tz);
    }

    // One step (sim.cu:129-205) for the group; returns true when the flight ended (ground,
    // NaN, out of bounds or the step cap).  npw: wave-uniform panel-slot count (the most any
    // group of the wave uses); slots beyond it hold only zero-area padding and are skipped
    // by a scalar branch -- their force is exactly zero, so skipping changes no bit.
    __device__ __forceinline__ bool step(int max_steps, int npw) {
        const float qw = B.qw, qx = B.qx, qy = B.qy, qz = B.qz;
        const float r00 = 1 - 2 * (qy * qy + qz * qz), r01 = 2 * (qx * qy - qw * qz), r02 = 2 * (qx * qz + qw * qy);
        const float r10 = 2 * (qx * qy + qw * qz), r11 = 1 - 2 * (qx * qx + qz * qz), r12 = 2 * (qy * qz - qw * qx);
        const float r20 = 2 * (qx * qz - qw * qy), r21 = 2 * (qy * qz + qw * qx), r22 = 1 - 2 * (qx * qx + qy * qy);
        // gust field at the CG: this lane's term, summed per component within its 8 lanes
        // ((t0 + t1) + (t2 + 0), t3 + t4, t5 + t6) and broadcast to them
        const float arg = W.ax * B.px + W.ay * B.py + W.az * B.pz + W.bt * B.t + W.cs;
        float gx, gy, gz;
        gust_sums(W.wg, sin_gust(arg), gx, gy, gz);
        // air velocity relative to the CG, in body axes
        const float ux = B.vx - (B.bx + gx), uy = B.vy - (B.by + gy), uz = B.vz - (B.bz + gz);
        const float vbx = r00 * ux + r10 * uy + r20 * uz;
        const float vby = r01 * ux + r11 * uy + r21 * uz;
        const float vbz = r02 * ux + r12 * uy + r22 * uz;
        const float wx = B.wx, wy = B.wy, wz = B.wz, g = B.g;
        float Fx, Fy, Fz, Tx, Ty, Tz;
        // slots in pairs: two independent panel evaluations per branch give the scheduler
        // instruction-level parallelism to interleave (at most one padded slot evaluated);
        // the first pair always runs (a slot past the lane's triangles adds exactly zero)
        panel_force<true>(0, vbx, vby, vbz, wx, wy, wz, g, Fx, Fy, Fz, Tx, Ty, Tz);
        if (PV > 1) panel_force(1, vbx, vby, vbz, wx, wy, wz, g, Fx, Fy, Fz, Tx, Ty, Tz);
#pragma unroll
        for (int p = 2; p < PV; p += 2) {
            if (p < npw) {
                panel_force(p, vbx, vby, vbz, wx, wy, wz, g, Fx, Fy, Fz, Tx, Ty, Tz);
                if (p + 1 < PV) panel_force(p + 1, vbx, vby, vbz, wx, wy, wz, g, Fx, Fy, Fz, Tx, Ty, Tz);
            }
        }
        // LDS slots (PV is even, so the pairs continue the VGPR slots' pairing)
#pragma unroll
        for (int q = 0; q < PQ; q += 2) {
            if (PV + q < npw) {
                panel_force_lds(q, vbx, vby, vbz, wx, wy, wz, g, Fx, Fy, Fz, Tx, Ty, Tz);
                if (q + 1 < PQ) panel_force_lds(q + 1, vbx, vby, vbz, wx, wy, wz, g, Fx, Fy, Fz, Tx, Ty, Tz);
            }
        }
        Fx = group_sum<L>(Fx); Fy = group_sum<L>(Fy); Fz = group_sum<L>(Fz);
        Tx = group_sum<L>(Tx); Ty = group_sum<L>(Ty); Tz = group_sum<L>(Tz);
        // the sums are complete here: without this (empty) statement the compiler sinks the
        // force sums' last additions past the spin clamp's branch, where their DPP moves can
        // no longer fold into them (two extra VALU instructions per sum)
        asm volatile("" : "+v"(Fx), "+v"(Fy), "+v"(Fz));
        if constexpr (NW > 1) {
            // the waves' partial sums, double-buffered by step parity: a wave writes block
            // (s & 1) at step s after passing step s - 1's barrier, which every wave reaches
            // only after its reads of that block at step s - 2; this step's barrier orders
            // the writes before the reads
            float* xb = xsum + (B.step & 1) * (8 * NW);
            if ((threadIdx.x & 63) == 0) {
                float* o = xb + 8 * (threadIdx.x >> 6);
                o[0] = Fx; o[1] = Fy; o[2] = Fz; o[3] = Tx; o[4] = Ty; o[5] = Tz;
            }
            __syncthreads();
            float sm[6];
#pragma unroll
            for (int k = 0; k < 6; k++) {
                if constexpr (NW == 2) sm[k] = xb[k] + xb[8 + k];
                else sm[k] = (xb[k] + xb[8 + k]) + (xb[16 + k] + xb[24 + k]);
            }
            Fx = sm[0]; Fy = sm[1]; Fz = sm[2]; Tx = sm[3]; Ty = sm[4]; Tz = sm[5];
        }
        // rigid-body update (sim.cu:177-197); torque is already in body axes
        float Fwx = r00 * Fx + r01 * Fy + r02 * Fz;
        float Fwy = r10 * Fx + r11 * Fy + r12 * Fz;
        float Fwz = r20 * Fx + r21 * Fy + r22 * Fz;
        Fwz -= B.mass * PR_GRAV;
        const float Iwx = B.I0 * wx + B.I3 * wy + B.I4 * wz;
        const float Iwy = B.I3 * wx + B.I1 * wy + B.I5 * wz;
        const float Iwz = B.I4 * wx + B.I5 * wy + B.I2 * wz;
        // tq = T - w x (I w), then w += dt invI tq (dt folded into the inverse inertia)
        const float tqx = __builtin_fmaf(-wy, Iwz, __builtin_fmaf(wz, Iwy, Tx));
        const float tqy = __builtin_fmaf(-wz, Iwx, __builtin_fmaf(wx, Iwz, Ty));
        const float tqz = __builtin_fmaf(-wx, Iwy, __builtin_fmaf(wy, Iwx, Tz));
        float nwx = wx + (B.v0 * tqx + B.v1 * tqy + B.v2 * tqz);
        float nwy = wy + (B.v1 * tqx + B.v4 * tqy + B.v5 * tqz);
        float nwz = wz + (B.v2 * tqx + B.v5 * tqy + B.v8 * tqz);
        // magnitude clamps compared on squares (the square roots only where a clamp applies:
        // v_sqrt / v_rsq issue at half the FMA rate); hardware v_sqrt / v_rsq (<= 1 ulp) for
        // the clamps and the renormalisation, as the reference's GPU build (-use_fast_math)
        const float wl2 = nwx * nwx + nwy * nwy + nwz * nwz;
        if (wl2 > 150.f * 150.f) {
            const float s = 150.f / __builtin_amdgcn_sqrtf(wl2);
            nwx *= s; nwy *= s; nwz *= s;
        }
        const float ks = B.ks;
        float nvx = B.vx + Fwx * ks, nvy = B.vy + Fwy * ks, nvz = B.vz + Fwz * ks;
        const float vl2 = nvx * nvx + nvy * nvy + nvz * nvz;
        if (vl2 > 60.f * 60.f) {
            const float s = 60.f / __builtin_amdgcn_sqrtf(vl2);
            nvx *= s; nvy *= s; nvz *= s;
        }
        const float npx = B.px + nvx * PR_DT, npy = B.py + nvy * PR_DT, npz = B.pz + nvz * PR_DT;
        // q += q (x) (0, w dt / 2)
        const float hd = 0.5f * PR_DT;
        const float hx = nwx * hd, hy = nwy * hd, hz = nwz * hd;
        float nqw = qw - qx * hx - qy * hy - qz * hz;
        float nqx = qx + qw * hx + qy * hz - qz * hy;
        float nqy = qy + qw * hy - qx * hz + qz * hx;
        float nqz = qz + qw * hz + qx * hy - qy * hx;
        const float il = __builtin_amdgcn_rsqf(nqw * nqw + nqx * nqx + nqy * nqy + nqz * nqz);
        B.qw = nqw * il; B.qx = nqx * il; B.qy = nqy * il; B.qz = nqz * il;
        B.vx = nvx; B.vy = nvy; B.vz = nvz;
        B.wx = nwx; B.wy = nwy; B.wz = nwz;
        B.px = npx; B.py = npy; B.pz = npz;
        B.t += PR_DT;
        B.step += 1;
        // termination (sim.cu:198-204)
        bad = !(npx == npx && nvx == nvx);
        const bool ended = npz <= 0.0f || bad || fabsf(npy) > 300.f || npx < -50.f;
        return ended || B.step >= max_steps;
    }

    __device__ __forceinline__ FlightState state(bool fin) const {
        FlightState st;
        st.pos[0] = B.px; st.pos[1] = B.py; st.pos[2] = B.pz;
        st.q[0] = B.qw; st.q[1] = B.qx; st.q[2] = B.qy; st.q[3] = B.qz;
        st.vel[0] = B.vx; st.vel[1] = B.vy; st.vel[2] = B.vz;
        st.wb[0] = B.wx; st.wb[1] = B.wy; st.wb[2] = B.wz;
        st.step = B.step; st.done = fin ? 1 : 0;
        st.fitness = fin ? (bad ? -10.0f : B.px) : 0.0f;
        return st;
    }
};

// Persistent loop over one work list: groups of L lanes take flights with one atomic each
// and fly them to the end of the chunk (or of the flight); the function returns when every
// group of the wave has stopped.
// Flies list positions [lo, hi) of one bucket's chunk: each group claims the next position
// from `head` (positions lo + 0, 1, 2, ...), and at the chunk's end marks it with the flight
// id if the flight goes on, else -1.  Which code path flies a position depends only on the
// position, so the results do not depend on the order in which groups claim.
// XCD-local claiming: list positions [lo, hi) in PR_XCDS equal shares, share x claimed
// through counter xh[x]; a workgroup takes positions from its own XCD's share first
// (workgroups are dispatched to the XCDs round-robin by index, so blockIdx % 8 is its
// XCD) and then from the others'.  A genome's flights sit next to each other in the list,
// so its panels are mostly fetched into one XCD's L2 instead of up to eight.  Every
// position is taken exactly once (each share's counter hands out its positions once);
// hi when all shares are spent.  Which code path flies a position still depends only on
// the position.
__device__ __forceinline__ int claim_xcd(int* xh, int lo, int hi) {
    const int n = hi - lo;  // < 2^28 (list capacity), so n * PR_XCDS fits an int
    const int me = (int)(blockIdx.x % PR_XCDS);
    // not unrolled: the eight shares' bounds would otherwise be kept in scalar registers at
    // every call site of the fused kernel's variants (hundreds of scalar spills)
#pragma unroll 1
    for (int r = 0; r < PR_XCDS; r++) {
        const int x = (me + r) % PR_XCDS;
        const int a = lo + n * x / PR_XCDS, b = lo + n * (x + 1) / PR_XCDS;
        if (a + __atomic_load_n(xh + x, __ATOMIC_RELAXED) < b) {
            const int k = a + atomicAdd(xh + x, 1);
            if (k < b) return k;
        }
    }
    return hi;
}

// End of a flight's chunk (one lane per flight): work counters, state, and the list mark
// (the flight id if the flight goes on, -1 when it ended, with its fitness and step count)
template <class FL>
__device__ __forceinline__ void store_flight(const FlyArgs& a, const FL& F, int pos, bool fin) {
    // algorithmic work counters: integrated steps and panel-steps of this chunk
    const unsigned long long fs = (unsigned long long)(F.B.step - F.B.s0);
    const unsigned long long ps = fs * (unsigned long long)F.B.nt;
    atomicAdd(a.stats + 0, fs);
    atomicAdd(a.stats + 1, ps);
    if (a.cstats) {
        atomicAdd(a.cstats + 0, fs);
        atomicAdd(a.cstats + 1, ps);
    }
    const FlightState st = F.state(fin);
    a.state[F.B.fid] = st;
    if (fin) {
        a.fitness[F.B.fid] = st.fitness;
        a.steps[F.B.fid] = F.B.step;
    }
    a.marks[pos] = fin ? -1 : F.B.fid;
}

// xheads: PR_XCDS claim counters (XCD-local claiming) or null (one counter, head)
template <int L, int P>
__device__ __forceinline__ void fly_list(const FlyArgs& a, int lo, int hi, int* head, float* smem,
                                         int* xheads = nullptr) {
    const int lane = threadIdx.x & 63;
    const int j = lane & (L - 1);
    const int gbase = lane & ~(L - 1);
    Flyer<L, P> F;
    bool active = false, exhausted = false;
    F.B.nt = 0;
    F.np = 0;
    int npw = 0, pos = 0;

    for (;;) {
        const bool need = !active && !exhausted;
        if (__ballot(need)) {
            if (need) {
                int k = hi;
                if (j == 0) {
                    if (xheads) {
                        k = claim_xcd(xheads, lo, hi);
                    } else {
                        // look before claiming: once the range is spent, no more atomics
                        const int h = __atomic_load_n(head, __ATOMIC_RELAXED);
                        if (lo + h < hi) k = lo + atomicAdd(head, 1);
                    }
                }
                k = __shfl(k, gbase);
                if (k >= hi) exhausted = true;
                else { pos = k; F.setup(a, a.list_in[k], j, smem); active = true; }
            }
            npw = wave_slots<P>(active ? F.np : 0);
        }
        if (!__ballot(active)) break;
        if (!active) continue;
        const bool fin = F.step(a.max_steps, npw);
        if (fin || F.B.step >= F.B.end) {
            if (j == 0) store_flight(a, F, pos, fin);
            active = false;
        }
    }
}

// Flights over all buckets' lists of the chunk (counters slot c)
__device__ __forceinline__ int chunk_total(const int* counts, int nb, int c) {
    int n = 0;
    for (int b = 0; b < nb; b++) n += counts[PR_CNT * b + c];
    return n;
}

// true when this chunk belongs to k_fly_split, so the k_fly / k_fly_all launches skip it
__device__ __forceinline__ bool chunk_is_split(const FlyArgs& a) {
    return a.split_max > 0 && chunk_total(a.counts_all, a.nb_all, a.c_all) <= a.split_max;
}

template <int L, int P>
__global__ __launch_bounds__(64) void k_fly(FlyArgs a) {
    const int n_in = *a.count_in;
    // two mappings are queued per chunk (many-flights-per-wave, one-flight-per-wave); the
    // flight count decides which one works, so the choice needs no host synchronisation
    if (n_in < a.run_lo || n_in > a.run_hi || chunk_is_split(a)) return;
    __shared__ float smem[lds_bytes(P) > 0 ? lds_bytes(P) / 4 : 1];
    fly_list<L, P>(a, 0, n_in, a.head, smem);
}

// A small chunk (at most one flight per W SIMDs): one flight per workgroup of W waves, its
// panels spread over all 64 W lanes, so a step's panel work is W times shorter than with one
// wave per flight.  Positions are numbered across the buckets in bucket order; which
// workgroup flies a position, and how, depends on the position and the chunk's flight count
// only.  The whole workgroup takes the same branches: the position loop and the flight's
// step loop are uniform (identical flight states in every wave).
template <int W, int P>
__global__ __launch_bounds__(64 * W) void k_fly_split(FlyAllArgs A) {
    static_assert(64 * W * P >= PR_SPLIT_NT, "every mesh must fit the workgroup's panel slots");
    __shared__ float xs[2 * 8 * W];
    const int n = chunk_total(A.counts, A.nb, A.c_in);
    if (n < A.split_lo || n > A.split_hi) return;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        int b = 0, k0 = 0;
        while (b < A.nb - 1 && k >= k0 + A.counts[PR_CNT * b + A.c_in]) {
            k0 += A.counts[PR_CNT * b + A.c_in];
            b++;
        }
        const int pos = k - k0;
        FlyArgs a = A.base;
        a.list_in = A.lists + (size_t)b * A.list_stride;
        a.marks = A.marks + (size_t)b * A.list_stride;
        a.stats = A.split_stats;
        Flyer<64, P, W> F;
        F.setup(a, a.list_in[pos], threadIdx.x & 63, xs);
        const int npw = __builtin_amdgcn_readfirstlane(F.np);
        bool fin;
        do {
            fin = F.step(a.max_steps, npw);
        } while (!fin && F.B.step < F.B.end);
        if (threadIdx.x == 0) store_flight(a, F, pos, fin);
        __syncthreads();  // the exchange block is reused by the next flight
    }
}

// Variants compiled into the fused flight kernel (all fit 2 waves per SIMD): id, L, P.
#define PR_FUSED(X)                                                                               \
    X(0, 8, 12) X(1, 8, 24) X(2, 16, 12) X(3, 16, 24) X(4, 32, 12) X(5, 32, 24) X(6, 64, 8) X(7, 64, 16) \
    X(8, 8, 33) X(9, 16, 33) X(10, 32, 33)

// LDS of the fused kernel: the largest panel block among its variants
constexpr int fused_lds_bytes() {
    int m = 0;
#define LDSMAX(ID, LL, PP) m = lds_bytes(PP) > m ? lds_bytes(PP) : m;
    PR_FUSED(LDSMAX)
#undef LDSMAX
    return m;
}

// One chunk of every bucket in one launch: each wave drains the buckets' lists, largest
// triangle counts first, flying each bucket with its own (L, P) code path: list positions
// [0, n - thr[b]) with the many-flights-per-wave mapping, the last thr[b] positions (one
// flight per wave) with the one-flight-per-wave mapping -- the last flights of a chunk run
// at the lowest step latency instead of leaving most lanes of a partial round idle.  With
// tail_when_short, a bucket does so only when the buckets drained after it hold fewer
// flights than thr[b]: otherwise its last flights run beside those buckets' flights
// anyway, and the one-per-wave mapping costs more instructions per flight-step.
__global__ __launch_bounds__(64, 2) void k_fly_all(FlyAllArgs A) {
    __shared__ float smem[fused_lds_bytes() > 0 ? fused_lds_bytes() / 4 : 1];
    if (chunk_is_split(A.base)) return;
    // flights of the buckets drained after the current one (in this launch)
    int below = 0;
    for (int b = 0; b < A.nb; b++)
        if (A.vid_p[b] >= 0) below += A.counts[PR_CNT * b + A.c_in];
    // widest group width at which this chunk's flights still fill widen x the grid's waves:
    // a bucket whose own width is at most half (a quarter) of it flies 2x (4x) wider --
    // fewer flights per wave, so the few long flights of a late chunk spread over more
    // SIMDs instead of crowding part of the waves (the choice depends on the counts only)
    const int lwant = A.widen && below > 0 ? (A.widen * 64 * (int)gridDim.x) / below : 0;
    // largest triangle counts first: the longest chunk jobs start first, small ones fill in
    for (int b = A.nb - 1; b >= 0; b--) {
        if (A.vid_p[b] < 0) continue;
        int* cb = A.counts + PR_CNT * b;
        const int n_in = cb[A.c_in];
        below -= n_in;
        if (n_in == 0) continue;
        FlyArgs a = A.base;
        a.list_in = A.lists + (size_t)b * A.list_stride;
        a.marks = A.marks + (size_t)b * A.list_stride;
        a.count_in = cb + A.c_in;
        const int vw = A.tail_when_short && below >= A.thr[b] ? -1 : A.vid_w[b];
        const int limit = vw >= 0 ? max(n_in - A.thr[b], 0) : n_in;
        int vp = A.vid_p[b];
        if (lwant >= 4 * A.lanes[b] && A.vid_4[b] >= 0) vp = A.vid_4[b];
        else if (lwant >= 2 * A.lanes[b] && A.vid_2[b] >= 0) vp = A.vid_2[b];
        // XCD-local claiming where the pass holds many flights per wave of the grid (the
        // large early chunks, which read most of the panels); a small chunk's flights are
        // taken from one counter, so they spread over the grid in claim order
        int* xh = limit >= 8 * (int)gridDim.x ? cb + 4 : nullptr;
        if (limit > 0) {
            a.stats = A.stats + 2 * (2 * b);
            switch (vp) {
#define CASE(ID, LL, PP) case ID: fly_list<LL, PP>(a, 0, limit, cb + 4 + PR_XCDS, smem, xh); break;
                PR_FUSED(CASE)
#undef CASE
                default: break;
            }
        }
        if (vw >= 0 && limit < n_in) {
            a.stats = A.stats + 2 * (2 * b + 1);
            switch (vw) {
#define CASE(ID, LL, PP) case ID: fly_list<LL, PP>(a, limit, n_in, cb + 4 + PR_XCDS + 1, smem); break;
                PR_FUSED(CASE)
#undef CASE
                default: break;
            }
        }
    }
}

// ------------------------------------------------------------------ k_compact_* (list upkeep)
// Tiles of PR_COMPACT_TILE source entries, numbered across the buckets in bucket order;
// every block works out the numbering from the counters, then strides over the tiles.
struct TileMap {
    int n[8], t0[9];
    __device__ void load(const CompactArgs& A) {
        t0[0] = 0;
        for (int b = 0; b < A.nb; b++) {
            n[b] = A.by_bucket ? A.n_fixed : A.counts[PR_CNT * b + A.c_src];
            t0[b + 1] = t0[b] + (n[b] + PR_COMPACT_TILE - 1) / PR_COMPACT_TILE;
        }
    }
    __device__ int bucket_of(int g, int nb) const {
        int b = 0;
        while (b < nb - 1 && g >= t0[b + 1]) b++;
        return b;
    }
};

constexpr int CT_THREADS = 256, CT_ITEMS = PR_COMPACT_TILE / CT_THREADS;

__device__ __forceinline__ bool ct_keep(const CompactArgs& A, int v, int b) { return A.by_bucket ? v == b : v >= 0; }

// by_bucket: the flight that source slot i stands for
__device__ __forceinline__ int ct_flight(const CompactArgs& A, int i) {
    return A.perm ? A.perm[i / A.perm_div] * A.perm_div + i % A.perm_div : i;
}

// block-wide sum over CT_THREADS threads; every thread receives it
__device__ __forceinline__ int ct_block_sum(int v, int* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int s = 0;
    for (int w = 0; w < CT_THREADS / 64; w++) s += red[w];
    return s;
}

__global__ __launch_bounds__(CT_THREADS) void k_compact_count(CompactArgs A) {
    __shared__ TileMap M;
    __shared__ int red[CT_THREADS / 64];
    if (threadIdx.x == 0) M.load(A);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x < A.nb) {
        for (int k = PR_CNT_CLAIM0; k < PR_CNT; k++) A.counts[PR_CNT * threadIdx.x + k] = 0;
        if (A.chunk_n) A.chunk_n[threadIdx.x] = M.n[threadIdx.x];
    }
    const int total = M.t0[A.nb];
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int b = M.bucket_of(g, A.nb);
        const int t = g - M.t0[b];
        const int* src = A.src + (A.by_bucket ? 0 : (size_t)b * A.src_stride);
        const int base = t * PR_COMPACT_TILE;
        int c = 0;
        for (int i = 0; i < CT_ITEMS; i++) {
            const int x = base + i * CT_THREADS + (int)threadIdx.x;
            if (x < M.n[b] && ct_keep(A, src[A.by_bucket ? ct_flight(A, x) : x], b)) c++;
        }
        c = ct_block_sum(c, red);
        if (threadIdx.x == 0) A.tcount[(size_t)b * A.max_tiles + t] = c;
    }
}

__global__ __launch_bounds__(CT_THREADS) void k_compact_scatter(CompactArgs A) {
    __shared__ TileMap M;
    __shared__ int red[CT_THREADS / 64];
    __shared__ int tile[PR_COMPACT_TILE];
    if (threadIdx.x == 0) M.load(A);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x < A.nb && M.n[threadIdx.x] == 0)
        A.counts[PR_CNT * threadIdx.x + A.c_dst] = 0;
    const int total = M.t0[A.nb];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int b = M.bucket_of(g, A.nb);
        const int t = g - M.t0[b];
        const int* src = A.src + (A.by_bucket ? 0 : (size_t)b * A.src_stride);
        int* dst = A.dst + (size_t)b * A.dst_stride;
        const int base = t * PR_COMPACT_TILE;
        const int n = min(M.n[b] - base, PR_COMPACT_TILE);
        // survivors of the bucket's earlier tiles
        int pre = 0;
        for (int u = threadIdx.x; u < t; u += CT_THREADS) pre += A.tcount[(size_t)b * A.max_tiles + u];
        pre = ct_block_sum(pre, red);
        // the tile through LDS (coalesced), then CT_ITEMS consecutive entries per thread
        for (int i = 0; i < CT_ITEMS; i++) {
            const int x = i * CT_THREADS + (int)threadIdx.x;
            tile[x] = x < n ? src[A.by_bucket ? ct_flight(A, base + x) : base + x] : (A.by_bucket ? -2 : -1);
        }
        __syncthreads();
        const int i0 = threadIdx.x * CT_ITEMS;
        int c = 0;
        for (int i = 0; i < CT_ITEMS; i++) c += ct_keep(A, tile[i0 + i], b) ? 1 : 0;
        // exclusive scan of c over the block: wave scan, then the earlier waves' totals
        int inc = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        __syncthreads();
        if (lane == 63) red[wave] = inc;
        __syncthreads();
        int off = pre + inc - c;
        int tile_total = 0;
        for (int w = 0; w < CT_THREADS / 64; w++) {
            if (w < wave) off += red[w];
            tile_total += red[w];
        }
        for (int i = 0; i < CT_ITEMS; i++) {
            const int v = tile[i0 + i];
            if (ct_keep(A, v, b)) dst[off++] = A.by_bucket ? ct_flight(A, base + i0 + i) : v;
        }
        if (t == M.t0[b + 1] - M.t0[b] - 1 && threadIdx.x == 0)
            A.counts[PR_CNT * b + A.c_dst] = pre + tile_total;
        __syncthreads();  // tile[] and red[] are reused by the next tile
    }
}

hipError_t launch_compact(const CompactArgs& a, long long n_total, hipStream_t s) {
    // every bucket's tiles together: at most n_total / tile + one partial tile per bucket
    const long long tiles = (n_total + PR_COMPACT_TILE - 1) / PR_COMPACT_TILE + a.nb;
    const int grid = (int)std::min<long long>(std::max<long long>(tiles, 1), 1024);
    hipLaunchKernelGGL(k_compact_count, dim3(grid), dim3(CT_THREADS), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_compact_scatter, dim3(grid), dim3(CT_THREADS), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ k_sort_* (chunk 0's list order)
// Stable counting sort of genomes by panel-slot key min(nt, 2047) / 8 (256 keys), in tiles of
// SORT_THREADS genomes: k_sort_count takes each tile's histogram; k_sort_offsets turns the
// histograms into each (tile, key)'s first output position (all genomes of smaller keys,
// then those of the same key in earlier tiles); k_sort_scatter ranks each tile's genomes
// among equal keys -- within a wave by a ballot per distinct key, in lane order, then
// across the tile's waves in wave order -- and places every genome after all earlier genomes
// of its key.  The result depends on the data alone.  Built into chunk 0's lists (through
// CompactArgs::perm), it lets the groups of a wave -- which take neighbouring list positions
// -- fly meshes of nearly equal size, so a wave's panel-slot count (its fullest group's)
// wastes little on the others.
constexpr int SORT_THREADS = 1024, SORT_KEYS = 256;

__device__ __forceinline__ int sort_key(int nt) { return min(nt, 2047) >> 3; }

size_t sort_work_ints(int n) { return (size_t)((n + SORT_THREADS - 1) / SORT_THREADS) * SORT_KEYS; }

// cnt[tile][key]: genomes of the tile with that key
__global__ __launch_bounds__(SORT_THREADS) void k_sort_count(const int* __restrict__ nt, int n, int* __restrict__ cnt) {
    __shared__ int h[SORT_KEYS];
    const int tid = threadIdx.x;
    for (int k = tid; k < SORT_KEYS; k += SORT_THREADS) h[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * SORT_THREADS + tid;
    if (i < n) atomicAdd(&h[sort_key(nt[i])], 1);
    __syncthreads();
    for (int k = tid; k < SORT_KEYS; k += SORT_THREADS) cnt[(size_t)blockIdx.x * SORT_KEYS + k] = h[k];
}

// One block, one thread per key: cnt[tile][key] is replaced by the first output position of
// that tile's genomes of that key.
__global__ __launch_bounds__(SORT_KEYS) void k_sort_offsets(int* __restrict__ cnt, int tiles) {
    __shared__ int start[SORT_KEYS];
    const int k = threadIdx.x;
    int tot = 0;
    for (int t = 0; t < tiles; t++) tot += cnt[(size_t)t * SORT_KEYS + k];
    start[k] = tot;
    __syncthreads();
    if (k == 0) {
        int s = 0;
        for (int j = 0; j < SORT_KEYS; j++) { const int c = start[j]; start[j] = s; s += c; }
    }
    __syncthreads();
    int base = start[k];
    for (int t = 0; t < tiles; t++) {
        const int c = cnt[(size_t)t * SORT_KEYS + k];
        cnt[(size_t)t * SORT_KEYS + k] = base;
        base += c;
    }
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(const int* __restrict__ nt, int n,
                                                               const int* __restrict__ off, int* __restrict__ perm) {
    __shared__ int wcnt[SORT_THREADS / 64][SORT_KEYS];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int i = blockIdx.x * SORT_THREADS + tid;
    const bool valid = i < n;
    const int key = valid ? sort_key(nt[i]) : -1;
    for (int k = lane; k < SORT_KEYS; k += 64) wcnt[w][k] = 0;
    __syncthreads();
    // rank among this wave's lanes of equal key, in lane order (every lane loops until the
    // wave has placed all its valid lanes)
    const unsigned long long lt = (1ull << lane) - 1ull;
    int rank = 0;
    bool left = valid;
    for (unsigned long long m = __ballot(left); m; m = __ballot(left)) {
        const int first = __builtin_ctzll(m);
        const int kk = __shfl(key, first);
        const unsigned long long same = __ballot(left && key == kk);
        if (left && key == kk) { rank = __popcll(same & lt); left = false; }
        if (lane == first) wcnt[w][kk] = __popcll(same);
    }
    __syncthreads();
    for (int k = tid; k < SORT_KEYS; k += SORT_THREADS) {  // waves' offsets per key
        int s = 0;
        for (int ww = 0; ww < SORT_THREADS / 64; ww++) { const int c = wcnt[ww][k]; wcnt[ww][k] = s; s += c; }
    }
    __syncthreads();
    if (valid) perm[off[(size_t)blockIdx.x * SORT_KEYS + key] + wcnt[w][key] + rank] = i;
}

hipError_t launch_sort_genomes(const int* nt, int n, int* perm, int* work, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int tiles = (n + SORT_THREADS - 1) / SORT_THREADS;
    hipLaunchKernelGGL(k_sort_count, dim3(tiles), dim3(SORT_THREADS), 0, s, nt, n, work);
    hipLaunchKernelGGL(k_sort_offsets, dim3(1), dim3(SORT_KEYS), 0, s, work, tiles);
    hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(SORT_THREADS), 0, s, nt, n, (const int*)work, perm);
    return hipGetLastError();
}

int fused_variant_id(int L, int P) {
#define ID(I, LL, PP) if (L == LL && P == PP) return I;
    PR_FUSED(ID)
#undef ID
    return -1;
}

int fused_variant_for(int L, int nt_max) {
    int best = -1, bp = 1 << 30;
#define FOR(I, LL, PP) if (L == LL && LL * PP >= nt_max && PP < bp) { best = I; bp = PP; }
    PR_FUSED(FOR)
#undef FOR
    return best;
}

hipError_t launch_fly_all(const FlyAllArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_fly_all, dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fly_split(const FlyAllArgs& a, int waves, int grid, hipStream_t s) {
    if (waves == 4) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fly_split<4, 8>), dim3(grid), dim3(256), 0, s, a);
    else if (waves == 2) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fly_split<2, 16>), dim3(grid), dim3(128), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t fly_all_occupancy(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_fly_all, 64, 0);
}

// Pose recording (sim.cu:117-128): one group per flight over the whole horizon.  Frames
// {pos, q, alive} at steps with step % rec_every == 0, index step0/rec_every + k within
// each chunk (the reference's per-chunk index rule), capped at rec_cap; ended flights keep
// writing their final pose with alive = 0.
template <int L, int P>
__global__ __launch_bounds__(64) void k_fly_rec(FlyArgs a, int n_flights, float* rec, int rec_every, int rec_cap) {
    const int lane = threadIdx.x & 63;
    const int j = lane & (L - 1);
    const int fid = blockIdx.x * (64 / L) + lane / L;
    if (fid >= n_flights) return;  // whole groups leave together
    __shared__ float smem[lds_bytes(P) > 0 ? lds_bytes(P) / 4 : 1];
    Flyer<L, P> F;
    F.setup(a, fid, j, smem);
    bool done = a.state[fid].done != 0;
    int rec_i = 0;
    for (int st = 0; st < a.max_steps; st++) {
        if (st % a.chunk == 0) { F.B.t = (float)st * PR_DT; rec_i = st / rec_every; }
        if (st % rec_every == 0 && rec_i < rec_cap) {
            if (j == 0) {
                float* r = rec + ((size_t)fid * rec_cap + rec_i) * 8;
                r[0] = F.B.px; r[1] = F.B.py; r[2] = F.B.pz;
                r[3] = F.B.qw; r[4] = F.B.qx; r[5] = F.B.qy; r[6] = F.B.qz;
                r[7] = done ? 0.0f : 1.0f;
            }
            rec_i++;
        }
        if (done) continue;
        const bool fin = F.step(a.max_steps, P);
        if (fin) {
            done = true;
            if (j == 0) {
                const FlightState s = F.state(true);
                a.state[fid] = s;
                a.fitness[fid] = s.fitness;
                a.steps[fid] = F.B.step;
            }
        }
    }
}

// ------------------------------------------------------------------ k_scores (sim.cu:340-357)
template <int MAXS>
__global__ void k_scores(const float* __restrict__ samples, int pop, int nseeds, float* __restrict__ scores) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pop) return;
    float row[MAXS];
    float mean = 0.0f;
    for (int s = 0; s < nseeds; s++) {
        float x = samples[(size_t)i * nseeds + s];
        mean += x / (float)nseeds;
        int k = s;  // insertion sort, ascending
        while (k > 0 && row[k - 1] > x) { row[k] = row[k - 1]; k--; }
        row[k] = x;
    }
    int tail = (nseeds + 3) / 4;
    float cv = 0.0f;
    for (int s = 0; s < tail; s++) cv += row[s] / (float)tail;
    scores[i] = 0.65f * mean + 0.25f * cv + 0.10f * row[0];
}

// Wide rows (the 128-scenario holdout, evaluate_design-sized batches): one wave per genome.
// Each lane places its samples at their stable ascending rank (the position the insertion
// sort above gives them: count of smaller samples plus earlier equal ones), then one lane
// sums the mean in seed order and the CVaR over the sorted head -- the same float
// operations in the same order as k_scores, so the same bits.
constexpr int SCORES_WAVE_MAX = 256;
__global__ __launch_bounds__(64) void k_scores_wave(const float* __restrict__ samples, int pop, int nseeds,
                                                    float* __restrict__ scores) {
#pragma clang fp contract(off)
    __shared__ float row[SCORES_WAVE_MAX];
    __shared__ float sorted[SCORES_WAVE_MAX];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= pop) return;  // whole block
    for (int s = lane; s < nseeds; s += 64) row[s] = samples[(size_t)i * nseeds + s];
    __syncthreads();
    for (int s = lane; s < nseeds; s += 64) {
        const float x = row[s];
        int r = 0;
        for (int j = 0; j < nseeds; j++) {
            const float y = row[j];
            r += (y < x || (y == x && j < s)) ? 1 : 0;
        }
        sorted[r] = x;
    }
    __syncthreads();
    if (lane == 0) {
        float mean = 0.0f;
        for (int s = 0; s < nseeds; s++) mean += row[s] / (float)nseeds;
        const int tail = (nseeds + 3) / 4;
        float cv = 0.0f;
        for (int s = 0; s < tail; s++) cv += sorted[s] / (float)tail;
        scores[i] = 0.65f * mean + 0.25f * cv + 0.10f * sorted[0];
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launch_layout(const float* genomes, const int* res, int ch, int sp, int max_ch, int max_sp, int n,
                         int* nt_raw, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_layout, dim3((n + 255) / 256), dim3(256), 0, s, genomes, res, ch, sp, max_ch, max_sp, n,
                       nt_raw);
    return hipGetLastError();
}

// One block per genome, LDS for its vertices and the 12 x 64 triangle terms.  The LDS a
// block needs sets how many blocks share a CU, and the context's vertex bound (from its
// largest chordwise and spanwise resolutions taken separately) is often twice what any of
// its genomes has: genomes with at most nv_small vertices go to a launch whose blocks fit
// 16 to a CU (10 KB each), the rest to one sized for the bound; each block of either launch
// returns at once when its genome belongs to the other.
template <class T>
static void launch_prepare_t(PrepareArgs a, int n, hipStream_t s) {
    const int nv_small = (10240 / (int)sizeof(T) - 12 * 64) / 3;
    const int nv_max = a.nv_max;
    auto go = [&](int lo, int hi) {
        a.nv_lo = lo;
        a.nv_hi = hi;
        a.nv_max = hi;
        const size_t lds = sizeof(T) * (3 * (size_t)hi + 12 * 64);
        hipLaunchKernelGGL(k_prepare<T>, dim3(n), dim3(64), lds, s, a);
    };
    if (nv_max <= nv_small) {
        go(0, nv_max);
    } else {
        go(0, nv_small);
        go(nv_small + 1, nv_max);
    }
}

hipError_t launch_prepare(const PrepareArgs& a, int n, int fp64, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (fp64) launch_prepare_t<double>(a, n, s);
    else launch_prepare_t<float>(a, n, s);
    return hipGetLastError();
}

hipError_t launch_scenarios(const int* seeds, int n, const ScenarioParams& sp, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scenarios, dim3((n + 127) / 128), dim3(128), 0, s, seeds, n, sp, out);
    return hipGetLastError();
}

hipError_t launch_queue(const QueueArgs& a, hipStream_t s) {
    if (a.n_flights <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_init, dim3((a.n_flights + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zero(int* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_ints, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

template <int L, int P>
static hipError_t fly_one(const FlyArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_fly<L, P>), dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

// compiled (L, P) variants; capacity L*P triangles per flight
#define PR_VARIANTS(X) \
    X(8, 12) X(8, 24) X(16, 12) X(16, 24) X(32, 12) X(32, 24) X(64, 8) X(64, 16) X(64, 24) X(64, 32) \
    X(8, 33) X(16, 33) X(32, 33)

int fly_variant_count() {
    int n = 0;
#define CNT(L, P) n++;
    PR_VARIANTS(CNT)
#undef CNT
    return n;
}

int fly_variant_info(int i, int* L, int* P) {
    int k = 0;
#define INF(LL, PP) if (k++ == i) { *L = LL; *P = PP; return 0; }
    PR_VARIANTS(INF)
#undef INF
    return -1;
}

hipError_t launch_fly(int L, int P, const FlyArgs& a, int grid, hipStream_t s) {
#define DISP(LL, PP) if (L == LL && P == PP) return fly_one<LL, PP>(a, grid, s);
    PR_VARIANTS(DISP)
#undef DISP
    return hipErrorInvalidValue;
}

hipError_t launch_fly_rec(int L, int P, const FlyArgs& a, int n_flights, float* rec, int rec_every, int rec_cap,
                          hipStream_t s) {
    const int grid = (n_flights * L + 63) / 64;
#define DISPR(LL, PP) \
    if (L == LL && P == PP) { hipLaunchKernelGGL((k_fly_rec<LL, PP>), dim3(grid), dim3(64), 0, s, a, n_flights, rec, rec_every, rec_cap); return hipGetLastError(); }
    PR_VARIANTS(DISPR)
#undef DISPR
    return hipErrorInvalidValue;
}

hipError_t fly_occupancy(int L, int P, int* blocks_per_cu) {
#define OCC(LL, PP) if (L == LL && P == PP) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_fly<LL, PP>, 64, 0);
    PR_VARIANTS(OCC)
#undef OCC
    return hipErrorInvalidValue;
}

hipError_t launch_scores(const float* samples, int pop, int nseeds, float* scores, hipStream_t s) {
    if (pop <= 0) return hipSuccess;
    dim3 grid((pop + 127) / 128), blk(128);
    if (nseeds <= 16) hipLaunchKernelGGL(k_scores<16>, grid, blk, 0, s, samples, pop, nseeds, scores);
    else if (nseeds <= SCORES_WAVE_MAX) hipLaunchKernelGGL(k_scores_wave, dim3(pop), dim3(64), 0, s, samples, pop, nseeds, scores);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace pr
