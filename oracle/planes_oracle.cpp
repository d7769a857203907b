// planes_oracle.cpp -- CPU restatement of the planes-rl fitness-evaluation path.
//
// THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it, and only to check the GPU path's results
// or as the timed CPU baseline.  The product (planes-rl_amd/, libplanes_hip.so) never links it.
//
// Every function is templated on the arithmetic type R (float = the FP32
// transcription, double = the FP64 oracle).  Model constants are the reference's
// float literals promoted to R, so FP64 and FP32 describe the same discrete model
// and differ only in rounding.  Operation order inside each expression follows
// the reference's source order so that the FP32 instantiation reproduces the
// reference's host-compiled float results (checked against oracle/_ref).
//
// Transcendentals: sin/cos/tan are evaluated in double and rounded once to R
// ("correctly rounded float" in practice).  sqrt and division are IEEE.
// Build with -ffp-contract=off: no FMA contraction anywhere.
//
// How each function is pinned (tests/test_oracle_golden.py, test_oracle_physics.py):
//   mesh (build, surf, empennage, fin, deck): bit-exact against oracle/_ref, the
//     reference's own common.h/designs.cu/stl.cu compiled with g++; closed-form counts,
//     mirror symmetry, the champion STL's triangle count.
//   lamina_mass, extent: bit-exact against oracle/_ref; geometry_test.cpp properties.
//   scenario_ref, hash32, printed, gust_wind: bit-exact against oracle/_ref.
//   osin/ocos/otan: correctly rounded against an arbitrary-precision evaluation;
//     libm-float mode bit-exact against oracle/_ref.
//   launch, step, aero: closed forms (ballistic motion at rho = 0, angular-momentum
//     conservation without torque, flat-plate special cases, torque = r x F, normal-flip
//     invariance) and FP32 <-> FP64 convergence.
//   robust_scores, ga_rank, ga2_breed: hand vectors and the reference's score weights.
//   scenario_cfg, lane_normal: no counterpart in the reference (the configurations'
//     scenario preset).  The fields drawn as in the reference's scenario_desc (speed, skin
//     friction, grip offsets, print biases, infill) bit-exact against the reference's
//     scenario vectors; the rest by the distributions the configurations state (KS and
//     moment tests per field, pairwise independence of all 26, the gust-field RMS constant);
//     planted slips in both fail those tests.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- math
template <class R> struct V3 { R x, y, z; };
template <class R> struct Qt { R w, x, y, z; };

template <class R> static inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return mk<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return mk<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> static inline V3<R> mul(V3<R> a, R s) { return mk<R>(a.x * s, a.y * s, a.z * s); }
template <class R> static inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> static inline V3<R> crs(V3<R> a, V3<R> b) {
    return mk<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> static inline R len(V3<R> a) { return std::sqrt(dot(a, a)); }
// common.h:22-23 -- unit vector with +z fallback for (near) zero length
template <class R> static inline V3<R> unit(V3<R> a) {
    R l = len(a);
    return l > R(1e-12f) ? mul(a, R(1) / l) : mk<R>(0, 0, 1);
}

// Float transcendentals: by default the double result rounded once (correctly rounded in
// practice; pinned against mpmath in the tests).  With g_libm_float set they call libm's
// float functions instead, which is what the reference's host build (g++, glibc) does --
// used only to pin the restatement bit-for-bit against oracle/_ref.
static int g_libm_float = 0;
static inline float osin(float x) { return g_libm_float ? ::sinf(x) : (float)std::sin((double)x); }
static inline float ocos(float x) { return g_libm_float ? ::cosf(x) : (float)std::cos((double)x); }
static inline float otan(float x) { return g_libm_float ? ::tanf(x) : (float)std::tan((double)x); }
static inline double osin(double x) { return std::sin(x); }
static inline double ocos(double x) { return std::cos(x); }
static inline double otan(double x) { return std::tan(x); }
template <class R> static inline R omin(R a, R b) { return std::fmin(a, b); }
template <class R> static inline R omax(R a, R b) { return std::fmax(a, b); }

// common.h:27-32 Hamilton product
template <class R> static inline Qt<R> qprod(Qt<R> a, Qt<R> b) {
    return Qt<R>{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
                 a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                 a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                 a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
// common.h:33-36
template <class R> static inline Qt<R> qunit(Qt<R> q) {
    R l = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    R il = R(1) / l;
    return Qt<R>{q.w * il, q.x * il, q.y * il, q.z * il};
}
// common.h:38-43 rows of the body->world rotation
template <class R> struct M3 { V3<R> r0, r1, r2; };
template <class R> static inline M3<R> qrows(Qt<R> q) {
    R w = q.w, x = q.x, y = q.y, z = q.z;
    M3<R> m;
    m.r0 = mk<R>(R(1) - R(2) * (y * y + z * z), R(2) * (x * y - w * z), R(2) * (x * z + w * y));
    m.r1 = mk<R>(R(2) * (x * y + w * z), R(1) - R(2) * (x * x + z * z), R(2) * (y * z - w * x));
    m.r2 = mk<R>(R(2) * (x * z - w * y), R(2) * (y * z + w * x), R(1) - R(2) * (x * x + y * y));
    return m;
}
template <class R> static inline V3<R> apply(const M3<R>& m, V3<R> v) { return mk<R>(dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)); }
// common.h:48-53 transpose apply
template <class R> static inline V3<R> apply_t(const M3<R>& m, V3<R> v) {
    return mk<R>(m.r0.x * v.x + m.r1.x * v.y + m.r2.x * v.z,
                 m.r0.y * v.x + m.r1.y * v.y + m.r2.y * v.z,
                 m.r0.z * v.x + m.r1.z * v.y + m.r2.z * v.z);
}
// common.h:54-57 axis-angle quaternion
template <class R> static inline Qt<R> qfrom(V3<R> axis, R ang) {
    R s = osin(ang * R(0.5f));
    V3<R> a = unit(axis);
    return Qt<R>{ocos(ang * R(0.5f)), a.x * s, a.y * s, a.z * s};
}

// ---------------------------------------------------------------- genome
enum Gene {
    G_SPAN, G_CHORD, G_TAPER, G_SWEEP, G_DIH_IN, G_DIH_OUT, G_KINK, G_CAMBER1, G_CAMBER2,
    G_WASHOUT, G_ELEVATOR, G_NOSE_MASS, G_KEEL_H, G_DECK_GAP, G_DECK_SCALE, G_BODY_FRAC,
    G_FIN_HEIGHT, G_FIN_CHORD, G_FIN_SWEEP, G_TAIL_SPAN, G_TAIL_CHORD, G_TAIL_ANGLE,
    G_LE_B1, G_LE_B2, G_CHORD_B1, G_CHORD_B2, G_FIN_LAYOUT, G_FIN_CANT, G_FIN_TAPER,
    G_FIN_SPAN_POS, G_TAIL_SWEEP, G_TAIL_TAPER, G_TAIL_DIHEDRAL, G_TIP_ROUND, G_CENTER_CHORD,
    N_GENES
};
// common.h:98-103 per-gene bounds (same table: it defines the genome format)
static const float LO[N_GENES] = {0.12f, 0.05f, 0.04f, 0.00f, -0.30f, -1.20f, 0.45f, -0.04f, -0.06f, -0.20f,
                                  -0.20f, 0.0f, 0.005f, 0.000f, 0.50f, 0.06f, 0.004f, 0.12f, 0.00f, 0.12f,
                                  0.10f, -0.16f, -0.80f, -0.80f, -0.65f, -0.65f, 0.00f, 0.00f, 0.10f, 0.15f,
                                  -0.40f, 0.15f, -0.50f, 0.00f, -0.25f};
static const float HI[N_GENES] = {0.25f, 0.20f, 1.00f, 0.90f, 0.60f, 1.20f, 0.90f, 0.10f, 0.06f, 0.08f,
                                  0.20f, 1.5f, 0.050f, 0.060f, 1.00f, 0.30f, 0.065f, 0.65f, 0.85f, 0.65f,
                                  0.45f, 0.16f, 0.80f, 0.80f, 1.00f, 1.00f, 1.00f, 0.75f, 1.00f, 1.00f,
                                  1.00f, 1.00f, 0.50f, 1.00f, 0.80f};

// physical constants (common.h:112-115, 470-478); float literals as the model defines them
static const float AREAL = 1240.0f * 0.00045f;   // PLA density x wall
static const float WALL = 0.00045f;
static const float PRINT_LIMIT = 0.250f - 0.00045f;  // sim.cu:35
static const float DT = 1.0f / 600.0f;
static const float GRAV = 9.81f;
static const float CN_SLOPE = 1.9f;

template <class R> struct Gn {  // genome promoted to R
    R v[N_GENES];
    explicit Gn(const float* g) { for (int i = 0; i < N_GENES; i++) v[i] = R(g[i]); }
    R operator[](int i) const { return v[i]; }
};

// ---------------------------------------------------------------- mesh (common.h:105-292)
struct Res { int ch = 5, sp = 8; };

template <class R> static R camber(const Gn<R>& g, R xc) {  // common.h:124-129
    R te = g[G_ELEVATOR] * R(0.22f);
    R u = xc, v = R(1) - u;
    return R(3) * v * v * u * g[G_CAMBER1] + R(3) * v * u * u * g[G_CAMBER2] + u * u * u * te;
}

template <class R> static void dihedral_yz(const Gn<R>& g, R t, R& y, R& z) {  // common.h:132-141
    R half = g[G_SPAN] * R(0.5f);
    R b = g[G_BODY_FRAC] * half, k = g[G_KINK] * half;
    if (k < b) k = b;
    R yy = 0, zz = 0, seg;
    seg = omin(t, b);
    yy += seg;
    seg = omin(t, k) - b;
    if (seg > 0) { yy += seg * ocos(g[G_DIH_IN]); zz += seg * osin(g[G_DIH_IN]); }
    seg = t - k;
    if (seg > 0) { yy += seg * ocos(g[G_DIH_OUT]); zz += seg * osin(g[G_DIH_OUT]); }
    y = yy; z = zz;
}

template <class R> static R bez(R v, R tf, R p0, R p1, R p2, R p3) {
    // cubic Bernstein blend in the reference's association order (common.h:156,165)
    return v * v * v * p0 + R(3) * v * v * tf * p1 + R(3) * v * tf * tf * p2 + tf * tf * tf * p3;
}

template <class R> static V3<R> surf(const Gn<R>& g, R u, R xc, R zoff, R scale) {  // common.h:144-172
    R half = g[G_SPAN] * R(0.5f) * scale;
    R t = std::fabs(u) * half;
    R y, z;
    dihedral_yz(g, t, y, z);
    R tf = std::fabs(u);
    R v = R(1) - tf;
    R cr = R(1) + g[G_CENTER_CHORD];
    R ct = g[G_TAPER];
    R c1 = R(1) + g[G_CHORD_B1], c2 = ct + g[G_CHORD_B2];
    R ratio = bez(v, tf, cr, c1, c2, ct);
    R ell = std::sqrt(omax(R(0.08f), R(1) - R(0.92f) * tf * tf));
    ratio *= R(1) - g[G_TIP_ROUND] + g[G_TIP_ROUND] * ell;
    ratio = omax(R(0.035f), omin(R(1.9f), ratio));
    R c = g[G_CHORD] * scale * ratio;
    R xr = R(0.5f) * g[G_CHORD] * scale * cr;
    R xt = R(0.5f) * g[G_CHORD] * scale - otan(g[G_SWEEP]) * half;
    R xb1 = xr + g[G_LE_B1] * g[G_CHORD] * scale;
    R xb2 = xt + g[G_LE_B2] * g[G_CHORD] * scale;
    R xle = bez(v, tf, xr, xb1, xb2, xt);
    R x = xle - xc * c;
    R zc = z * scale + camber(g, xc) * c;
    R xq = xle - R(0.25f) * c;
    R tw = g[G_WASHOUT] * tf;
    zc += tw * (x - xq);
    return mk<R>(x, (u < 0 ? -y : y), zc + zoff);
}

template <class R> struct Mesh {
    std::vector<V3<R>> V;
    std::vector<int> T;  // 3 per triangle
    void tri(int a, int b, int c) { T.push_back(a); T.push_back(b); T.push_back(c); }
    int nv() const { return (int)V.size(); }
    int nt() const { return (int)T.size() / 3; }
};

template <class R> static void deck(const Gn<R>& g, Mesh<R>& m, R zoff, R scale, Res rs) {  // common.h:177-192
    int base = m.nv();
    for (int i = 0; i <= 2 * rs.sp; i++) {
        R u = R((float)(i - rs.sp) / (float)rs.sp);
        for (int j = 0; j <= rs.ch; j++) {
            R xc = R((float)j / (float)rs.ch);
            m.V.push_back(surf(g, u, xc, zoff, scale));
        }
    }
    for (int i = 0; i < 2 * rs.sp; i++)
        for (int j = 0; j < rs.ch; j++) {
            int a = base + i * (rs.ch + 1) + j, b = a + rs.ch + 1;
            m.tri(a, b, a + 1);
            m.tri(b, b + 1, a + 1);
        }
}

template <class R> static void fin(const Gn<R>& g, Mesh<R>& m, R y0, R z0, int side) {  // common.h:197-213
    R fc = g[G_CHORD] * g[G_FIN_CHORD];
    R sw = fc * g[G_FIN_SWEEP];
    R xr = R(-0.32f) * g[G_CHORD] - g[G_TAIL_CHORD] * g[G_CHORD] * R(0.82f);
    R tipc = fc * g[G_FIN_TAPER];
    R cant = side == 0 ? R(0) : g[G_FIN_CANT];
    R ya = y0 + R(side) * osin(cant) * g[G_FIN_HEIGHT];
    R za = z0 + ocos(cant) * g[G_FIN_HEIGHT];
    int b = m.nv();
    m.V.push_back(mk<R>(xr + R(0.5f) * fc, y0, z0));
    m.V.push_back(mk<R>(xr - R(0.5f) * fc, y0, z0));
    m.V.push_back(mk<R>(xr - sw + R(0.5f) * tipc, ya, za));
    m.V.push_back(mk<R>(xr - sw - R(0.5f) * tipc, ya, za));
    m.tri(b, b + 1, b + 2);
    m.tri(b + 1, b + 3, b + 2);
}

template <class R> static void empennage(const Gn<R>& g, Mesh<R>& m) {  // common.h:215-255
    R hs = R(0.5f) * g[G_SPAN] * g[G_TAIL_SPAN];
    R hc = g[G_CHORD] * g[G_TAIL_CHORD];
    R x0 = R(-0.32f) * g[G_CHORD];
    R z0 = camber(g, R(0.82f)) * g[G_CHORD];
    int base = m.nv();
    for (int i = 0; i < 3; i++) {
        R u = R((float)(i - 1)), au = std::fabs(u);
        R tc = hc * (R(1) + (g[G_TAIL_TAPER] - R(1)) * au);
        R xle = x0 - g[G_TAIL_SWEEP] * hc * au;
        for (int j = 0; j < 3; j++) {
            R xc = R(0.5f * (float)j);
            R x = xle - xc * tc;
            R z = z0 + g[G_TAIL_ANGLE] * (x - xle) + au * hs * osin(g[G_TAIL_DIHEDRAL]);
            m.V.push_back(mk<R>(x, u * hs * ocos(g[G_TAIL_DIHEDRAL]), z));
        }
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            int a = base + i * 3 + j, b = a + 3;
            m.tri(a, b, a + 1);
            m.tri(b, b + 1, a + 1);
        }
    R yp = hs * g[G_FIN_SPAN_POS] * ocos(g[G_TAIL_DIHEDRAL]);
    R zp = z0 + hs * g[G_FIN_SPAN_POS] * osin(g[G_TAIL_DIHEDRAL]);
    R lay = g[G_FIN_LAYOUT];
    if (lay < R(0.25f)) {
        fin(g, m, R(0), z0, 0);
    } else if (lay < R(0.50f)) {
        fin(g, m, -yp, zp, -1);
        fin(g, m, yp, zp, 1);
    } else if (lay < R(0.75f)) {
        fin(g, m, R(0), z0, 0);
        fin(g, m, -yp, zp, -1);
        fin(g, m, yp, zp, 1);
    } else {
        R yo = hs * R(0.94f) * ocos(g[G_TAIL_DIHEDRAL]);
        R zo = z0 + hs * R(0.94f) * osin(g[G_TAIL_DIHEDRAL]);
        fin(g, m, -yo, zo, -1);
        fin(g, m, yo, zo, 1);
    }
}

template <class R> static Mesh<R> build(const float* gf, Res rs) {  // common.h:257-292
    Gn<R> g(gf);
    Mesh<R> m;
    deck(g, m, R(0), R(1), rs);
    empennage(g, m);
    bool bi = g[G_DECK_GAP] > R(0.02f);
    if (bi) deck(g, m, g[G_DECK_GAP], g[G_DECK_SCALE], rs);
    int base = m.nv();
    for (int j = 0; j <= 3; j++) {  // keel sheet
        R xc = R((float)j / 3.0f);
        V3<R> top = surf(g, R(0), xc, R(0), R(1));
        m.V.push_back(top);
        m.V.push_back(mk<R>(top.x, R(0), top.z - g[G_KEEL_H] * (R(0.35f) + R(0.65f) * xc)));
    }
    for (int j = 0; j < 3; j++) {
        int a = base + j * 2;
        m.tri(a, a + 1, a + 2);
        m.tri(a + 1, a + 3, a + 2);
    }
    if (bi) {  // struts at the kink stations
        for (int s = -1; s <= 1; s += 2) {
            R u = R(s) * g[G_KINK];
            int b = m.nv();
            m.V.push_back(surf(g, u, R(0.25f), R(0), R(1)));
            m.V.push_back(surf(g, u, R(0.55f), R(0), R(1)));
            m.V.push_back(surf(g, u, R(0.25f), g[G_DECK_GAP], g[G_DECK_SCALE]));
            m.V.push_back(surf(g, u, R(0.55f), g[G_DECK_GAP], g[G_DECK_SCALE]));
            m.tri(b, b + 1, b + 2);
            m.tri(b + 1, b + 3, b + 2);
        }
    }
    return m;
}

template <class R> static V3<R> extent(const Mesh<R>& m) {  // common.h:294-302
    if (m.V.empty()) return mk<R>(0, 0, 0);
    V3<R> lo = m.V[0], hi = m.V[0];
    for (size_t i = 1; i < m.V.size(); i++) {
        lo.x = omin(lo.x, m.V[i].x); lo.y = omin(lo.y, m.V[i].y); lo.z = omin(lo.z, m.V[i].z);
        hi.x = omax(hi.x, m.V[i].x); hi.y = omax(hi.y, m.V[i].y); hi.z = omax(hi.z, m.V[i].z);
    }
    return hi - lo;
}

// ---------------------------------------------------------------- mass properties (common.h:309-365)
template <class R> struct Mass {
    R mass, invmass;
    V3<R> com;
    R I[6];     // xx yy zz xy xz yz (xy.. are -sum m x y)
    R invI[9];
};

template <class R> static void inv3(Mass<R>& mp, R xx, R yy, R zz, R xy, R xz, R yz, bool floor_det) {
    R det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    R id = R(1) / (floor_det ? omax(det, R(1e-24f)) : det);
    mp.invI[0] = (yy * zz - yz * yz) * id; mp.invI[1] = (xz * yz - xy * zz) * id; mp.invI[2] = (xy * yz - xz * yy) * id;
    mp.invI[3] = mp.invI[1]; mp.invI[4] = (xx * zz - xz * xz) * id; mp.invI[5] = (xy * xz - xx * yz) * id;
    mp.invI[6] = mp.invI[2]; mp.invI[7] = mp.invI[5]; mp.invI[8] = (xx * yy - xy * xy) * id;
}

template <class R> static Mass<R> lamina_mass(const float* gf, const Mesh<R>& msh) {
    Gn<R> g(gf);
    const int nt = msh.nt();
    auto vert = [&](int t, int k) { return msh.V[msh.T[3 * t + k]]; };
    R m = 0;
    V3<R> c = mk<R>(0, 0, 0);
    const R third = R(1.0f / 3);
    for (int i = 0; i < nt; i++) {
        V3<R> a = vert(i, 0), b = vert(i, 1), d = vert(i, 2);
        R A = R(0.5f) * len(crs(b - a, d - a));
        R mi = A * R(AREAL);
        V3<R> ct = mul((a + b) + d, third);
        m += mi;
        c = c + mul(ct, mi);
    }
    R shell = m;
    V3<R> np = mk<R>(R(0.5f) * g[G_CHORD], R(0), R(0));  // nose ballast point
    R mn = g[G_NOSE_MASS] * shell;
    m += mn;
    c = c + mul(np, mn);
    c = mul(c, R(1) / m);
    R I[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < nt; i++) {
        V3<R> a = vert(i, 0), b = vert(i, 1), d = vert(i, 2);
        R A = R(0.5f) * len(crs(b - a, d - a));
        R mi = A * R(AREAL);
        V3<R> ra = a - c, rb = b - c, rd = d - c;
        // lamina covariance: (sum r r^T + s s^T)/12 written per component
        R exx = (ra.x * ra.x + rb.x * rb.x + rd.x * rd.x + ra.x * rb.x + ra.x * rd.x + rb.x * rd.x) / R(6);
        R eyy = (ra.y * ra.y + rb.y * rb.y + rd.y * rd.y + ra.y * rb.y + ra.y * rd.y + rb.y * rd.y) / R(6);
        R ezz = (ra.z * ra.z + rb.z * rb.z + rd.z * rd.z + ra.z * rb.z + ra.z * rd.z + rb.z * rd.z) / R(6);
        R exy = (R(2) * (ra.x * ra.y + rb.x * rb.y + rd.x * rd.y) + ra.x * rb.y + rb.x * ra.y + ra.x * rd.y +
                 rd.x * ra.y + rb.x * rd.y + rd.x * rb.y) / R(12);
        R exz = (R(2) * (ra.x * ra.z + rb.x * rb.z + rd.x * rd.z) + ra.x * rb.z + rb.x * ra.z + ra.x * rd.z +
                 rd.x * ra.z + rb.x * rd.z + rd.x * rb.z) / R(12);
        R eyz = (R(2) * (ra.y * ra.z + rb.y * rb.z + rd.y * rd.z) + ra.y * rb.z + rb.y * ra.z + ra.y * rd.z +
                 rd.y * ra.z + rb.y * rd.z + rd.y * rb.z) / R(12);
        I[0] += mi * (eyy + ezz); I[1] += mi * (exx + ezz); I[2] += mi * (exx + eyy);
        I[3] -= mi * exy; I[4] -= mi * exz; I[5] -= mi * eyz;
        V3<R> nn = unit(crs(b - a, d - a));
        R th = mi * R(WALL) * R(WALL) / R(12);
        I[0] += th * (R(1) - nn.x * nn.x); I[1] += th * (R(1) - nn.y * nn.y); I[2] += th * (R(1) - nn.z * nn.z);
        I[3] -= th * nn.x * nn.y; I[4] -= th * nn.x * nn.z; I[5] -= th * nn.y * nn.z;
    }
    {
        V3<R> r = np - c;
        I[0] += mn * (r.y * r.y + r.z * r.z); I[1] += mn * (r.x * r.x + r.z * r.z);
        I[2] += mn * (r.x * r.x + r.y * r.y); I[3] -= mn * r.x * r.y;
        I[4] -= mn * r.x * r.z; I[5] -= mn * r.y * r.z;
    }
    const R reg = R(1e-9f);
    Mass<R> mp;
    R xx = I[0] + reg, yy = I[1] + reg, zz = I[2] + reg;
    inv3(mp, xx, yy, zz, I[3], I[4], I[5], false);
    mp.I[0] = xx; mp.I[1] = yy; mp.I[2] = zz; mp.I[3] = I[3]; mp.I[4] = I[4]; mp.I[5] = I[5];
    mp.mass = m; mp.invmass = R(1) / m; mp.com = c;
    return mp;
}

// ---------------------------------------------------------------- scenarios (common.h:367-421)
static inline uint32_t hash32(uint32_t x) {  // lowbias32 (C. Wellons): xorshift-multiply
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
static inline float lane_u01(int seed, int lane) {
    uint32_t key = (uint32_t)seed * 0x9e3779b9u + (uint32_t)lane * 0x85ebca6bu;
    return (float)(hash32(key) >> 8) * (1.0f / 16777216.0f);
}
static inline float lane_pm1(int seed, int lane) { return lane_u01(seed, lane) * 2.0f - 1.0f; }
// (seed*37) % 97 with 32-bit wrap (the reference's int product overflows for large seeds)
static inline int phase_of(int seed) { return (int)((uint32_t)seed * 37u) % 97; }

// 26 fields: the reference's ScenarioDesc order (common.h:377-387) + base_z (0 in the reference preset)
enum SF {
    S_SPEED, S_PITCH, S_YAW, S_ROLL, S_HEIGHT, S_ROLL_RATE, S_PITCH_RATE, S_YAW_RATE, S_RHO, S_CF, S_GUST,
    S_BASE_X, S_BASE_Y, S_FPITCH, S_FYAW, S_GRIP_X, S_GRIP_Y, S_GRIP_Z, S_MATERIAL, S_INFILL, S_BIAS_X,
    S_BIAS_Y, S_BIAS_Z, S_WARP_P, S_WARP_R, S_BASE_Z, N_SF
};

struct ScenarioCfg {
    int preset = 0;            // 0 reference (hash lanes, common.h:389-421), 1 parametric config preset
    float gust_rms_max = 4.0f; // config: 3-D RMS of the gust field, drawn U(0, max)
    float warp_max = 0.0174532925f;  // config: warp pitch/roll U(-max, max) [rad]
    float wind_sd = 2.0f;      // config: steady wind N(0, sd) per axis
};

static void scenario_ref(int seed, float* d) {  // common.h:389-421 (float arithmetic, no contraction)
    d[S_SPEED] = 9.2f + 2.0f * lane_pm1(seed, 0);
    d[S_PITCH] = 0.10f + 0.20f * lane_pm1(seed, 1);
    d[S_YAW] = 0.12f * lane_pm1(seed, 2);
    d[S_ROLL] = 0.16f * lane_pm1(seed, 3);
    d[S_HEIGHT] = 1.75f + 0.35f * lane_pm1(seed, 4);
    d[S_ROLL_RATE] = 1.8f * lane_pm1(seed, 5);
    d[S_PITCH_RATE] = 1.2f * lane_pm1(seed, 6);
    d[S_YAW_RATE] = 1.5f * lane_pm1(seed, 7);
    d[S_RHO] = 1.225f * (0.88f + 0.22f * lane_u01(seed, 8));
    d[S_CF] = 0.02f * (0.75f + 0.55f * lane_u01(seed, 9));
    d[S_GUST] = 0.45f + 1.10f * (0.5f + 0.5f * osin((float)seed * 1.61803f));
    float ph = (float)phase_of(seed);
    d[S_BASE_X] = -0.25f + 1.15f * osin(ph * 0.71f);
    d[S_BASE_Y] = 1.00f * osin(ph * 1.13f);
    d[S_FPITCH] = 0.075f * lane_pm1(seed, 10);
    d[S_FYAW] = 0.060f * lane_pm1(seed, 11);
    d[S_GRIP_X] = 0.018f * lane_pm1(seed, 12);
    d[S_GRIP_Y] = 0.022f * lane_pm1(seed, 13);
    d[S_GRIP_Z] = 0.010f * lane_pm1(seed, 14);
    d[S_MATERIAL] = 0.88f + 0.24f * lane_u01(seed, 15);
    d[S_INFILL] = 0.06f + 0.34f * lane_u01(seed, 16);
    d[S_BIAS_X] = 0.018f * lane_pm1(seed, 17);
    d[S_BIAS_Y] = 0.010f * lane_pm1(seed, 18);
    d[S_BIAS_Z] = 0.006f * lane_pm1(seed, 19);
    d[S_CF] *= 0.90f + 0.25f * lane_u01(seed, 20);
    d[S_WARP_P] = 0.025f * lane_pm1(seed, 21);
    d[S_WARP_R] = 0.035f * lane_pm1(seed, 22);
    d[S_BASE_Z] = 0.0f;
}

// Box-Muller (Box & Muller 1958) standard normal from hash lanes (lane, lane+1), evaluated in
// double and rounded once.  u1 is taken in (0,1] so the log is finite.
static float lane_normal(int seed, int lane) {
    uint32_t k1 = (uint32_t)seed * 0x9e3779b9u + (uint32_t)lane * 0x85ebca6bu;
    uint32_t k2 = (uint32_t)seed * 0x9e3779b9u + (uint32_t)(lane + 1) * 0x85ebca6bu;
    double u1 = (double)((hash32(k1) >> 8) + 1u) * (1.0 / 16777216.0);
    double u2 = (double)(hash32(k2) >> 8) * (1.0 / 16777216.0);
    double r = std::sqrt(-2.0 * std::log(u1));
    return (float)(r * std::cos(6.283185307179586 * u2));
}

// Parametric config preset (no reference counterpart): the distributions of the benchmark
// configurations.  Lanes 0-22 keep their reference meaning where the distribution is unchanged;
// normals use lane pairs from 32 upward.
static void scenario_cfg(int seed, const ScenarioCfg& cf, float* d) {
    const float deg = 0.0174532925f;
    d[S_SPEED] = 9.2f + 2.0f * lane_pm1(seed, 0);                  // U(7.2, 11.2)
    d[S_PITCH] = 0.10f + 5.0f * deg * lane_normal(seed, 32);       // nominal 0.10 rad + N(0, 5 deg)
    d[S_YAW] = 5.0f * deg * lane_normal(seed, 34);
    d[S_ROLL] = 5.0f * deg * lane_normal(seed, 36);
    d[S_HEIGHT] = 1.75f + 0.25f * lane_pm1(seed, 4);               // U(1.5, 2.0)
    d[S_ROLL_RATE] = 30.0f * deg * lane_normal(seed, 38);          // N(0, 30 deg/s)
    d[S_PITCH_RATE] = 30.0f * deg * lane_normal(seed, 40);
    d[S_YAW_RATE] = 30.0f * deg * lane_normal(seed, 42);
    d[S_RHO] = 1.15f + 0.10f * lane_u01(seed, 8);                  // U(1.15, 1.25)
    d[S_CF] = 0.02f * (0.75f + 0.55f * lane_u01(seed, 9));
    // gust field 3-D RMS = 0.6919 * amplitude (7 sinusoids, weights .55/.35/.25/.50/.30/.30/.20)
    d[S_GUST] = cf.gust_rms_max * lane_u01(seed, 23) * (1.0f / 0.69191763f);
    d[S_BASE_X] = cf.wind_sd * lane_normal(seed, 44);
    d[S_BASE_Y] = cf.wind_sd * lane_normal(seed, 46);
    d[S_FPITCH] = 5.0f * deg * lane_normal(seed, 48);
    d[S_FYAW] = 5.0f * deg * lane_normal(seed, 50);
    d[S_GRIP_X] = 0.018f * lane_pm1(seed, 12);
    d[S_GRIP_Y] = 0.022f * lane_pm1(seed, 13);
    d[S_GRIP_Z] = 0.010f * lane_pm1(seed, 14);
    d[S_MATERIAL] = 0.95f + 0.10f * lane_u01(seed, 15);            // +-5 %
    d[S_INFILL] = 0.06f + 0.34f * lane_u01(seed, 16);              // 6-40 %
    d[S_BIAS_X] = 0.018f * lane_pm1(seed, 17);
    d[S_BIAS_Y] = 0.010f * lane_pm1(seed, 18);
    d[S_BIAS_Z] = 0.006f * lane_pm1(seed, 19);
    d[S_CF] *= 0.90f + 0.25f * lane_u01(seed, 20);
    d[S_WARP_P] = cf.warp_max * lane_pm1(seed, 21);
    d[S_WARP_R] = cf.warp_max * lane_pm1(seed, 22);
    d[S_BASE_Z] = cf.wind_sd * lane_normal(seed, 52);
}

static void scenario(int seed, const ScenarioCfg& cf, float* d) {
    if (cf.preset == 0) scenario_ref(seed, d); else scenario_cfg(seed, cf, d);
}

// common.h:423-457 printed (per flight) mass properties
template <class R> static Mass<R> printed(Mass<R> src, const float* d, V3<R>& shift) {
    R m_shell = src.mass * R(d[S_MATERIAL]);
    R m_fill = m_shell * R(d[S_INFILL]);
    R m = m_shell + m_fill;
    V3<R> bias = mk<R>(R(d[S_BIAS_X]), R(d[S_BIAS_Y]), R(d[S_BIAS_Z]));
    shift = mul(bias, m_fill / m);
    R sc = m / src.mass;
    for (int k = 0; k < 6; k++) src.I[k] *= sc;
    auto pt = [&](V3<R> r, R pm) {
        src.I[0] += pm * (r.y * r.y + r.z * r.z); src.I[1] += pm * (r.x * r.x + r.z * r.z);
        src.I[2] += pm * (r.x * r.x + r.y * r.y); src.I[3] -= pm * r.x * r.y;
        src.I[4] -= pm * r.x * r.z; src.I[5] -= pm * r.y * r.z;
    };
    pt(mul(shift, R(-1.0f)), m_shell);
    pt(bias - shift, m_fill);
    src.mass = m; src.invmass = R(1) / m;
    src.com = src.com + shift;
    inv3(src, src.I[0], src.I[1], src.I[2], src.I[3], src.I[4], src.I[5], true);
    return src;
}

// common.h:459-467 gust field at the CG
template <class R> static V3<R> gust_wind(V3<R> p, R t, int seed, const float* d) {
    R s = R((float)phase_of(seed));
    R gu = R(d[S_GUST]);
    R gx = gu * (R(0.55f) * osin(R(0.9f) * p.x + R(1.3f) * t + s) + R(0.35f) * osin(R(2.3f) * p.z + R(2.1f) * t + R(1.7f) * s) +
                 R(0.25f) * osin(R(4.1f) * p.y + R(3.3f) * t + R(0.4f) * s));
    R gy = gu * (R(0.50f) * osin(R(1.1f) * p.x + R(1.7f) * t + R(2.2f) * s) + R(0.30f) * osin(R(3.1f) * p.z + R(2.9f) * t + R(0.9f) * s));
    R gz = gu * (R(0.30f) * osin(R(1.4f) * p.x + R(2.3f) * t + R(3.1f) * s) + R(0.20f) * osin(R(2.7f) * p.y + R(1.9f) * t + R(1.2f) * s));
    return mk<R>(R(d[S_BASE_X]) + gx, R(d[S_BASE_Y]) + gy, R(d[S_BASE_Z]) + gz);
}

// ---------------------------------------------------------------- panels (sim.cu:25-49)
template <class R> struct Panels {
    std::vector<V3<R>> c;  // centroid - com
    std::vector<R> A;
    std::vector<V3<R>> n;
};

template <class R> static Panels<R> panels_of(const Mesh<R>& m, const Mass<R>& mp, int nt) {
    Panels<R> P;
    const R third = R(1.0f / 3);
    for (int t = 0; t < nt; t++) {
        V3<R> a = m.V[m.T[3 * t]] - mp.com, b = m.V[m.T[3 * t + 1]] - mp.com, c = m.V[m.T[3 * t + 2]] - mp.com;
        V3<R> nrm = crs(b - a, c - a);
        R A = R(0.5f) * len(nrm);
        P.c.push_back(mul((a + b) + c, third));
        P.A.push_back(A);
        P.n.push_back(A > R(1e-12f) ? mul(nrm, R(0.5f) / A) : mk<R>(0, 0, 1));
    }
    return P;
}

// ---------------------------------------------------------------- flight (sim.cu:55-212)
struct SimCfg {
    int max_steps = 7200;  // flight cap (SIM_MAX_STEPS)
    int chunk = 600;       // clock reset period (sim.cu:113,241)
    float spread = 0.0f;   // showcase launch grid (sim.cu:71-78)
};

template <class R> struct State {
    V3<R> pos; Qt<R> q; V3<R> vel, wb;
    bool done; R fitness; int steps;  // steps = integrated steps (plane-flight-seconds * 600)
    R t;
};

template <class R> struct Flight {
    Panels<R> P;
    Mass<R> mp;       // printed
    V3<R> shift;
    float d[N_SF];
    int seed, nt;
    R rho, cf;
    bool flip = false;
};

template <class R> static void launch(Flight<R>& F, State<R>& s, int plane_index, const SimCfg& sc) {  // sim.cu:70-109
    const float* d = F.d;
    R ox = 0, oy = 0, oz = 0;
    if (sc.spread > 0) {
        int cols = 25;
        oy = R((float)((plane_index % cols) - cols / 2) * sc.spread);
        ox = R(-(float)(plane_index / cols) * sc.spread * 1.4f);
        oz = R(0.25f * osin((float)plane_index * 12.9898f));
    }
    R speed = R(d[S_SPEED]), pitch = R(d[S_PITCH]), yaw = R(d[S_YAW]), roll = R(d[S_ROLL]), height = R(d[S_HEIGHT]);
    s.pos = mk<R>(ox, oy, height + oz);
    Qt<R> qy = qfrom(mk<R>(0, 0, 1), yaw), qp = qfrom(mk<R>(0, 1, 0), -pitch), qr = qfrom(mk<R>(1, 0, 0), roll);
    s.q = qunit(qprod(qy, qprod(qp, qr)));
    Qt<R> fy = qfrom(mk<R>(0, 0, 1), yaw + R(d[S_FYAW]));
    Qt<R> fp = qfrom(mk<R>(0, 1, 0), -(pitch + R(d[S_FPITCH])));
    s.vel = apply(qrows(qprod(fy, fp)), mk<R>(speed, 0, 0));
    s.wb = mk<R>(R(d[S_ROLL_RATE]), R(d[S_PITCH_RATE]), R(d[S_YAW_RATE]));
    V3<R> grip = mk<R>(R(d[S_GRIP_X]), R(d[S_GRIP_Y]), R(d[S_GRIP_Z]));
    V3<R> Jb = apply_t(qrows(s.q), mul(s.vel, F.mp.mass));
    V3<R> Lm = mul(crs(grip - F.shift, Jb), R(0.12f));
    const R* iv = F.mp.invI;
    V3<R> dw = mk<R>(iv[0] * Lm.x + iv[1] * Lm.y + iv[2] * Lm.z, iv[3] * Lm.x + iv[4] * Lm.y + iv[5] * Lm.z,
                     iv[6] * Lm.x + iv[7] * Lm.y + iv[8] * Lm.z);
    s.wb = s.wb + dw;
    R w0 = len(s.wb);
    if (w0 > R(35.0f)) s.wb = mul(s.wb, R(35.0f) / w0);
    s.done = F.nt <= 0;
    s.t = 0;
    s.steps = 0;
    s.fitness = s.done ? R(-100.0f) : R(0);
}

// Panel aerodynamics of one step (sim.cu:133-159): world-frame force and torque about the CG.
// flip: negate every panel normal (a test hook; the force model is invariant under it).
template <class R> static void aero(const Panels<R>& P, int nt, V3<R> shift, const M3<R>& Rm, V3<R> vel, V3<R> ww,
                                    V3<R> wind, R rho, R cf, R wp, R wr, bool flip, V3<R>& Fs, V3<R>& Ts) {
    Fs = mk<R>(0, 0, 0); Ts = mk<R>(0, 0, 0);
    for (int i = 0; i < nt; i++) {
        V3<R> rb = P.c[i] - shift;
        R A = P.A[i];
        V3<R> nb = flip ? mk<R>(-P.n[i].x, -P.n[i].y, -P.n[i].z) : P.n[i];
        R ar = wr * omax(R(-1.f), omin(R(1.f), rb.y / R(0.125f)));
        R ap = wp * omax(R(-1.f), omin(R(1.f), rb.x / R(0.125f)));
        V3<R> nr = mk<R>(nb.x, nb.y - ar * nb.z, nb.z + ar * nb.y);
        nb = unit(mk<R>(nr.x + ap * nr.z, nr.y, nr.z - ap * nr.x));
        V3<R> rw = apply(Rm, rb), nw = apply(Rm, nb);
        V3<R> vr = (vel + crs(ww, rw)) - wind;
        R V2 = dot(vr, vr);
        if (V2 < R(1e-6f)) continue;
        R Vm = std::sqrt(V2);
        V3<R> vh = mul(vr, R(1) / Vm);
        R sn = dot(vh, nw);
        R qd = R(0.5f) * rho * V2 * A;
        R cn = R(CN_SLOPE) * sn * std::sqrt(omax(R(0), R(1) - sn * sn)) + R(2) * sn * std::fabs(sn);
        V3<R> f = mul(nw, -qd * cn);
        V3<R> tang = vh - mul(nw, sn);
        f = f - mul(tang, qd * cf);
        Fs = Fs + f;
        Ts = Ts + crs(rw, f);
    }
}

// One integration step (sim.cu:129-205); sets s.done / s.fitness when the flight ends.
template <class R> static void step(Flight<R>& F, State<R>& s, int stepno, const SimCfg& sc) {
    M3<R> Rm = qrows(s.q);
    V3<R> vel = s.vel, wb = s.wb, pos = s.pos;
    V3<R> ww = apply(Rm, wb);
    V3<R> wind = gust_wind(pos, s.t, F.seed, F.d);
    V3<R> Fs, Ts;
    aero(F.P, F.nt, F.shift, Rm, vel, ww, wind, F.rho, F.cf, R(F.d[S_WARP_P]), R(F.d[S_WARP_R]), F.flip, Fs, Ts);
    const Mass<R>& mp = F.mp;
    Fs.z -= mp.mass * R(GRAV);
    V3<R> tb = apply_t(Rm, Ts);
    V3<R> Iw = mk<R>(mp.I[0] * wb.x + mp.I[3] * wb.y + mp.I[4] * wb.z, mp.I[3] * wb.x + mp.I[1] * wb.y + mp.I[5] * wb.z,
                     mp.I[4] * wb.x + mp.I[5] * wb.y + mp.I[2] * wb.z);
    V3<R> tq = tb - crs(wb, Iw);
    const R* iv = mp.invI;
    V3<R> dw = mk<R>(iv[0] * tq.x + iv[1] * tq.y + iv[2] * tq.z, iv[3] * tq.x + iv[4] * tq.y + iv[5] * tq.z,
                     iv[6] * tq.x + iv[7] * tq.y + iv[8] * tq.z);
    const R dt = R(DT);
    wb = wb + mul(dw, dt);
    R wl = len(wb);
    if (wl > R(150.f)) wb = mul(wb, R(150.f) / wl);
    vel = vel + mul(Fs, mp.invmass * dt);
    R vl = len(vel);
    if (vl > R(60.f)) vel = mul(vel, R(60.f) / vl);
    pos = pos + mul(vel, dt);
    Qt<R> dq = qprod(s.q, Qt<R>{0, wb.x, wb.y, wb.z});
    s.q = qunit(Qt<R>{s.q.w + R(0.5f) * dq.w * dt, s.q.x + R(0.5f) * dq.x * dt, s.q.y + R(0.5f) * dq.y * dt,
                      s.q.z + R(0.5f) * dq.z * dt});
    s.vel = vel; s.wb = wb; s.pos = pos; s.t += dt;
    s.steps++;
    bool bad = !(pos.x == pos.x && vel.x == vel.x);
    if (pos.z <= R(0) || bad || std::fabs(pos.y) > R(300.f) || pos.x < R(-50.f)) {
        s.done = true;
        s.fitness = bad ? R(-10.0f) : pos.x;
    } else if (stepno == sc.max_steps - 1) {
        s.done = true;
        s.fitness = pos.x;
    }
}

// Build one flight: genome -> mesh -> mass -> panels; scenario -> printed mass.
template <class R> struct Prepared {
    Mesh<R> mesh;
    Mass<R> mp;
    int nt;  // 0 if not printable
    int nt_raw;
    int nv;
};

template <class R> static Prepared<R> prepare(const float* g, Res rs) {
    Prepared<R> p;
    p.mesh = build<R>(g, rs);
    p.mp = lamina_mass<R>(g, p.mesh);
    V3<R> e = extent(p.mesh);
    bool ok = e.x <= R(PRINT_LIMIT) && e.y <= R(PRINT_LIMIT) && e.z <= R(PRINT_LIMIT);
    p.nt_raw = p.mesh.nt();
    p.nv = p.mesh.nv();
    p.nt = ok ? p.nt_raw : 0;
    return p;
}

template <class R> static Flight<R> make_flight(const Prepared<R>& p, int seed, const ScenarioCfg& scfg,
                                                const float* sc_override = nullptr, int flags = 0) {
    Flight<R> F;
    F.seed = seed;
    F.nt = p.nt;
    scenario(seed, scfg, F.d);
    if (sc_override) std::memcpy(F.d, sc_override, sizeof(F.d));
    F.flip = (flags & 1) != 0;
    F.P = panels_of(p.mesh, p.mp, p.nt);
    F.mp = printed(p.mp, F.d, F.shift);
    F.rho = R(F.d[S_RHO]);
    F.cf = R(F.d[S_CF]);
    return F;
}

// clock at a chunk start: step0 * dt in R (sim.cu:113)
template <class R> static R chunk_clock(int step0) {
    return std::is_same<R, double>::value ? R(step0) * R(DT) : R((float)step0 * DT);
}

// Full flight with the chunk clock rule: t = step0*dt at each chunk start (sim.cu:113).
// rec (optional): frames of {pos, q, alive} every rec_every steps (sim.cu:119-128).
template <class R> static State<R> fly(Flight<R>& F, int plane_index, const SimCfg& sc, float* rec = nullptr,
                                       int rec_every = 1, int rec_cap = 0, std::vector<State<R>>* traj = nullptr) {
    State<R> s;
    launch(F, s, plane_index, sc);
    for (int step0 = 0; step0 < sc.max_steps; step0 += sc.chunk) {
        if (step0 > 0) s.t = chunk_clock<R>(step0);
        int rec_i = step0 / rec_every;  // recomputed per chunk like sim.cu:117
        for (int st = step0; st < step0 + sc.chunk && st < sc.max_steps; st++) {
            if (rec && st % rec_every == 0 && rec_i < rec_cap) {
                float* r = rec + 8 * rec_i;
                r[0] = (float)s.pos.x; r[1] = (float)s.pos.y; r[2] = (float)s.pos.z;
                r[3] = (float)s.q.w; r[4] = (float)s.q.x; r[5] = (float)s.q.y; r[6] = (float)s.q.z;
                r[7] = s.done ? 0.0f : 1.0f;
                rec_i++;
            }
            if (s.done) { if (!rec) break; continue; }
            step(F, s, st, sc);
            if (traj) traj->push_back(s);
        }
        if (s.done && !rec) break;
    }
    return s;
}

}  // namespace orc

// =============================================================== C ABI for the tests
using namespace orc;

static Res mkres(int ch, int sp) { Res r; if (ch > 0) r.ch = ch; if (sp > 0) r.sp = sp; return r; }
static ScenarioCfg mkscfg(int preset, float gust_rms_max, float warp_max, float wind_sd) {
    ScenarioCfg c; c.preset = preset;
    if (gust_rms_max >= 0) c.gust_rms_max = gust_rms_max;
    if (warp_max >= 0) c.warp_max = warp_max;
    if (wind_sd >= 0) c.wind_sd = wind_sd;
    return c;
}

template <class R> static void mass_out(const Mass<R>& m, R* o) {
    o[0] = m.mass; o[1] = m.invmass; o[2] = m.com.x; o[3] = m.com.y; o[4] = m.com.z;
    for (int k = 0; k < 6; k++) o[5 + k] = m.I[k];
    for (int k = 0; k < 9; k++) o[11 + k] = m.invI[k];
}
template <class R> static Mass<R> mass_in(const R* o) {
    Mass<R> m;
    m.mass = o[0]; m.invmass = o[1]; m.com = mk<R>(o[2], o[3], o[4]);
    for (int k = 0; k < 6; k++) m.I[k] = o[5 + k];
    for (int k = 0; k < 9; k++) m.invI[k] = o[11 + k];
    return m;
}

// Flight state vector: pos3 q4 vel3 wb3 = 13 values
template <class R> static void state_out(const State<R>& s, R* o) {
    o[0] = s.pos.x; o[1] = s.pos.y; o[2] = s.pos.z; o[3] = s.q.w; o[4] = s.q.x; o[5] = s.q.y; o[6] = s.q.z;
    o[7] = s.vel.x; o[8] = s.vel.y; o[9] = s.vel.z; o[10] = s.wb.x; o[11] = s.wb.y; o[12] = s.wb.z;
}

extern "C" {

int or_n_genes() { return N_GENES; }
void or_set_libm_float(int on) { g_libm_float = on; }
void or_bounds(float* lo, float* hi) { for (int i = 0; i < N_GENES; i++) { lo[i] = LO[i]; hi[i] = HI[i]; } }

// mesh: V [cap_v*3], T [cap_t*3]; returns 0 or -1 when capacity is short
#define MESH_FN(NAME, R)                                                                          \
    int NAME(const float* g, int ch, int sp, R* V, int cap_v, int* T, int cap_t, int* nv, int* nt) { \
        Mesh<R> m = build<R>(g, mkres(ch, sp));                                                   \
        *nv = m.nv(); *nt = m.nt();                                                               \
        if (m.nv() > cap_v || m.nt() > cap_t) return -1;                                          \
        for (int i = 0; i < m.nv(); i++) { V[3 * i] = m.V[i].x; V[3 * i + 1] = m.V[i].y; V[3 * i + 2] = m.V[i].z; } \
        for (size_t i = 0; i < m.T.size(); i++) T[i] = m.T[i];                                   \
        return 0;                                                                                 \
    }
MESH_FN(or_mesh_f32, float)
MESH_FN(or_mesh_f64, double)

// mass properties of the unprinted shell: out[20] = mass, invmass, com3, I6, invI9
// printable: 1 if within the print volume; nt: triangle count
#define MASS_FN(NAME, R)                                                             \
    void NAME(const float* g, int ch, int sp, R* out, int* printable, int* nt, int* nv) { \
        Prepared<R> p = prepare<R>(g, mkres(ch, sp));                                \
        mass_out(p.mp, out);                                                         \
        *printable = p.nt > 0; *nt = p.nt_raw; *nv = p.nv;                          \
    }
MASS_FN(or_mass_f32, float)
MASS_FN(or_mass_f64, double)

// panels: c [nt*3], A [nt], n [nt*3] relative to the shell COM (k_prepare output)
#define PANEL_FN(NAME, R)                                                              \
    int NAME(const float* g, int ch, int sp, R* c, R* A, R* n, int cap) {            \
        Prepared<R> p = prepare<R>(g, mkres(ch, sp));                                  \
        if (p.nt_raw > cap) return -1;                                                 \
        Panels<R> P = panels_of(p.mesh, p.mp, p.nt_raw);                               \
        for (int t = 0; t < p.nt_raw; t++) {                                           \
            c[3 * t] = P.c[t].x; c[3 * t + 1] = P.c[t].y; c[3 * t + 2] = P.c[t].z;     \
            A[t] = P.A[t];                                                             \
            n[3 * t] = P.n[t].x; n[3 * t + 1] = P.n[t].y; n[3 * t + 2] = P.n[t].z;     \
        }                                                                              \
        return p.nt;                                                                   \
    }
PANEL_FN(or_panels_f32, float)
PANEL_FN(or_panels_f64, double)

uint32_t or_hash32(uint32_t x) { return hash32(x); }
float or_lane_u01(int seed, int lane) { return lane_u01(seed, lane); }
float or_lane_normal(int seed, int lane) { return lane_normal(seed, lane); }
int or_phase(int seed) { return phase_of(seed); }

void or_scenario(int seed, int preset, float gust_rms_max, float warp_max, float wind_sd, float* out26) {
    scenario(seed, mkscfg(preset, gust_rms_max, warp_max, wind_sd), out26);
}
void or_scenario_batch(const int* seeds, int n, int preset, float gust_rms_max, float warp_max, float wind_sd, float* out) {
    ScenarioCfg c = mkscfg(preset, gust_rms_max, warp_max, wind_sd);
    for (int i = 0; i < n; i++) scenario(seeds[i], c, out + (size_t)N_SF * i);
}

#define PRINTED_FN(NAME, R)                                          \
    void NAME(const R* mp20, const float* sc26, R* out20, R* shift3) { \
        V3<R> sh; Mass<R> m = printed(mass_in(mp20), sc26, sh);     \
        mass_out(m, out20); shift3[0] = sh.x; shift3[1] = sh.y; shift3[2] = sh.z; \
    }
PRINTED_FN(or_printed_f32, float)
PRINTED_FN(or_printed_f64, double)

void or_wind_f32(float px, float py, float pz, float t, int seed, const float* sc26, float* out3) {
    V3<float> w = gust_wind(mk<float>(px, py, pz), t, seed, sc26);
    out3[0] = w.x; out3[1] = w.y; out3[2] = w.z;
}
void or_wind_f64(double px, double py, double pz, double t, int seed, const float* sc26, double* out3) {
    V3<double> w = gust_wind(mk<double>(px, py, pz), t, seed, sc26);
    out3[0] = w.x; out3[1] = w.y; out3[2] = w.z;
}


// Launch state of one flight (after sim.cu:70-109): out13 + printed mass20 + shift3.
// sc_override (optional, 26 floats): the scenario descriptor to launch under instead of seed's.
#define LAUNCH_FN(NAME, R)                                                                               \
    void NAME(const float* g, int ch, int sp, int seed, int preset, float gust, float warp, float wsd,    \
              int plane_index, float spread, R* st13, R* mp20, R* shift3, int* nt,                     \
              const float* sc_override) {                                                               \
        Prepared<R> p = prepare<R>(g, mkres(ch, sp));                                                   \
        Flight<R> F = make_flight(p, seed, mkscfg(preset, gust, warp, wsd), sc_override);               \
        SimCfg sc; sc.spread = spread;                                                                  \
        State<R> s; launch(F, s, plane_index, sc);                                                      \
        state_out(s, st13); mass_out(F.mp, mp20);                                                       \
        shift3[0] = F.shift.x; shift3[1] = F.shift.y; shift3[2] = F.shift.z; *nt = F.nt;              \
    }
LAUNCH_FN(or_launch_f32, float)
LAUNCH_FN(or_launch_f64, double)

// Trajectory of one flight: states after each of the first nsteps steps -> out[nsteps*13].
// Returns the number of steps integrated (< nsteps if the flight ended).
#define TRAJ_FN(NAME, R)                                                                               \
    int NAME(const float* g, int ch, int sp, int seed, int preset, float gust, float warp, float wsd,  \
             int max_steps, int nsteps, R* out, const float* sc_override, int flags) {                                                      \
        Prepared<R> p = prepare<R>(g, mkres(ch, sp));                                                 \
        Flight<R> F = make_flight(p, seed, mkscfg(preset, gust, warp, wsd), sc_override, flags);                          \
        SimCfg sc; if (max_steps > 0) sc.max_steps = max_steps;                                       \
        State<R> s; launch(F, s, 0, sc);                                                              \
        int n = 0;                                                                                    \
        for (int st = 0; st < sc.max_steps && st < nsteps && !s.done; st++) {                         \
            if (st > 0 && st % sc.chunk == 0) s.t = chunk_clock<R>(st);                               \
            step(F, s, st, sc);                                                                       \
            state_out(s, out + 13 * n); n++;                                                          \
        }                                                                                             \
        return n;                                                                                     \
    }
TRAJ_FN(or_traj_f32, float)
TRAJ_FN(or_traj_f64, double)

// Aero force/torque (world frame) of explicit panels: c,n [nt*3], A [nt]; q = (w,x,y,z)
void or_aero_f64(int nt, const double* c, const double* A, const double* n, const double* shift3, const double* q4,
                 const double* vel3, const double* wb3, const double* wind3, double rho, double cf, double wp,
                 double wr, int flip, double* F3, double* T3) {
    Panels<double> P;
    for (int i = 0; i < nt; i++) {
        P.c.push_back(mk<double>(c[3 * i], c[3 * i + 1], c[3 * i + 2]));
        P.A.push_back(A[i]);
        P.n.push_back(mk<double>(n[3 * i], n[3 * i + 1], n[3 * i + 2]));
    }
    M3<double> Rm = qrows(Qt<double>{q4[0], q4[1], q4[2], q4[3]});
    V3<double> wb = mk<double>(wb3[0], wb3[1], wb3[2]);
    V3<double> Fs, Ts;
    aero(P, nt, mk<double>(shift3[0], shift3[1], shift3[2]), Rm, mk<double>(vel3[0], vel3[1], vel3[2]), apply(Rm, wb),
         mk<double>(wind3[0], wind3[1], wind3[2]), rho, cf, wp, wr, flip != 0, Fs, Ts);
    F3[0] = Fs.x; F3[1] = Fs.y; F3[2] = Fs.z; T3[0] = Ts.x; T3[1] = Ts.y; T3[2] = Ts.z;
}

// Batched fitness: flight i flies genome gi[i] under env seed seeds[i]; multithreaded with an
// atomic flight counter.  fitness[i], steps[i] (integrated steps).  Returns wall seconds.
#define FLY_FN(NAME, R)                                                                                      \
    double NAME(const float* genomes, const int* gi, const int* seeds, int n, int ch, int sp, int preset,      \
                float gust, float warp, float wsd, int max_steps, int nthreads, R* fitness, int* steps) {      \
        auto t0 = std::chrono::steady_clock::now();                                                         \
        ScenarioCfg scfg = mkscfg(preset, gust, warp, wsd);                                                 \
        SimCfg sc; if (max_steps > 0) sc.max_steps = max_steps;                                             \
        Res rs = mkres(ch, sp);                                                                             \
        std::atomic<int> next(0);                                                                           \
        auto work = [&]() {                                                                                 \
            int cached = -1; Prepared<R> p;                                                                 \
            for (;;) {                                                                                      \
                int i = next.fetch_add(1);                                                                  \
                if (i >= n) break;                                                                          \
                if (gi[i] != cached) { p = prepare<R>(genomes + (size_t)N_GENES * gi[i], rs); cached = gi[i]; } \
                Flight<R> F = make_flight(p, seeds[i], scfg);                                               \
                State<R> s = fly(F, 0, sc);                                                                 \
                fitness[i] = s.fitness; steps[i] = s.steps;                                                 \
            }                                                                                               \
        };                                                                                                  \
        if (nthreads <= 1) work();                                                                          \
        else { std::vector<std::thread> th; for (int k = 0; k < nthreads; k++) th.emplace_back(work);       \
               for (auto& x : th) x.join(); }                                                               \
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();                \
    }
FLY_FN(or_fly_f32, float)
FLY_FN(or_fly_f64, double)

// The same with a mesh resolution per genome: res[2 * g] = ch, res[2 * g + 1] = sp.
#define FLY_RES_FN(NAME, R)                                                                                  \
    double NAME(const float* genomes, const int* res, const int* gi, const int* seeds, int n, int preset,      \
                float gust, float warp, float wsd, int max_steps, int nthreads, R* fitness, int* steps) {      \
        auto t0 = std::chrono::steady_clock::now();                                                         \
        ScenarioCfg scfg = mkscfg(preset, gust, warp, wsd);                                                 \
        SimCfg sc; if (max_steps > 0) sc.max_steps = max_steps;                                             \
        std::atomic<int> next(0);                                                                           \
        auto work = [&]() {                                                                                 \
            int cached = -1; Prepared<R> p;                                                                 \
            for (;;) {                                                                                      \
                int i = next.fetch_add(1);                                                                  \
                if (i >= n) break;                                                                          \
                if (gi[i] != cached) {                                                                      \
                    p = prepare<R>(genomes + (size_t)N_GENES * gi[i], mkres(res[2 * gi[i]], res[2 * gi[i] + 1])); \
                    cached = gi[i];                                                                         \
                }                                                                                           \
                Flight<R> F = make_flight(p, seeds[i], scfg);                                               \
                State<R> s = fly(F, 0, sc);                                                                 \
                fitness[i] = s.fitness; steps[i] = s.steps;                                                 \
            }                                                                                               \
        };                                                                                                  \
        if (nthreads <= 1) work();                                                                          \
        else { std::vector<std::thread> th; for (int k = 0; k < nthreads; k++) th.emplace_back(work);       \
               for (auto& x : th) x.join(); }                                                               \
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();                \
    }
FLY_RES_FN(or_fly_res_f32, float)
FLY_RES_FN(or_fly_res_f64, double)

// Pose recording of one flight (sim.cu:119-128): rec[rec_cap*8]; returns fitness
float or_record_f32(const float* g, int seed, int plane_index, float spread, int max_steps, int rec_every,
                    int rec_cap, float* rec) {
    Prepared<float> p = prepare<float>(g, Res());
    Flight<float> F = make_flight(p, seed, ScenarioCfg());
    SimCfg sc; sc.spread = spread; if (max_steps > 0) sc.max_steps = max_steps;
    State<float> s = fly(F, plane_index, sc, rec, rec_every, rec_cap);
    return s.fitness;
}

// ------------------------------------------------------------ fitness aggregation + GA (sim.cu:315-490)
// robust score (sim.cu:340-357): 0.65 mean + 0.25 CVaR(worst quarter) + 0.10 worst
void or_robust_scores(const float* samples, int pop, int nseeds, float* scores) {
    std::vector<float> row(nseeds);
    int tail = (nseeds + 3) / 4;
    for (int i = 0; i < pop; i++) {
        float mean = 0;
        for (int s = 0; s < nseeds; s++) { row[s] = samples[(size_t)i * nseeds + s]; mean += row[s] / nseeds; }
        std::sort(row.begin(), row.end());
        float cv = 0;
        for (int s = 0; s < tail; s++) cv += row[s] / tail;
        scores[i] = 0.65f * mean + 0.25f * cv + 0.10f * row[0];
    }
}

static void clampset(float* g, int i, float v) { g[i] = std::fmin(HI[i], std::fmax(LO[i], v)); }
static void random_genome(float* g, std::mt19937& rng) {  // sim.cu:318-321
    std::uniform_real_distribution<float> u(0, 1);
    for (int i = 0; i < N_GENES; i++) clampset(g, i, LO[i] + u(rng) * (HI[i] - LO[i]));
}
void or_ga_sanitize(float* genomes, int n) {
    for (int i = 0; i < n; i++) for (int j = 0; j < N_GENES; j++) clampset(genomes + (size_t)N_GENES * i, j, genomes[(size_t)N_GENES * i + j]);
}
// ranking (sim.cu:440-441): idx sorted by score, descending, std::sort (libstdc++ order on ties)
void or_ga_rank(const float* acc, int pop, int* idx) {
    for (int i = 0; i < pop; i++) idx[i] = i;
    std::sort(idx, idx + pop, [&](int a, int b) { return acc[a] > acc[b]; });
}
// GA state and breeding (sim.cu:398, 429-430, 472-490): gs sorted by idx rank -> next.  The
// engine, the uniform and the normal distribution objects persist across generations as in
// the reference (sim.cu:398, 429-430): std::normal_distribution<float> caches its second
// Marsaglia-polar draw inside the object, so a per-generation object would lose it.
struct OrGA2 { std::mt19937 rng; std::uniform_real_distribution<float> u01{0, 1}; std::normal_distribution<float> nrm{0, 1}; };
void* or_ga2_new(unsigned seed) { OrGA2* g = new OrGA2; g->rng.seed(seed); return g; }
void or_ga2_free(void* h) { delete (OrGA2*)h; }
void or_ga2_random(void* h, float* genomes, int n) {
    OrGA2* g = (OrGA2*)h;
    for (int i = 0; i < n; i++) random_genome(genomes + (size_t)N_GENES * i, g->rng);
}
void or_ga2_breed(void* h, const float* gs, const float* acc, const int* idx, int pop, int gen, int gens, float* next) {
    OrGA2* G = (OrGA2*)h;
    std::mt19937& rng = G->rng;
    int elite = pop / 64;
    float anneal = 1.0f - 0.6f * gen / gens;
    for (int i = 0; i < elite; i++) std::memcpy(next + (size_t)N_GENES * i, gs + (size_t)N_GENES * idx[i], 4 * N_GENES);
    for (int i = elite; i < pop; i++) {
        float* out = next + (size_t)N_GENES * i;
        if (G->u01(rng) < 0.05f) { random_genome(out, rng); continue; }
        auto tourn = [&]() {
            int b = idx[(int)(G->u01(rng) * pop)];
            for (int k = 0; k < 3; k++) { int c = (int)(G->u01(rng) * pop); if (acc[c] > acc[b]) b = c; }
            return b;
        };
        int pa = tourn();
        int pc = tourn();
        for (int j = 0; j < N_GENES; j++) {
            float a = -0.1f + 1.2f * G->u01(rng);
            float v = gs[(size_t)N_GENES * pa + j] * a + gs[(size_t)N_GENES * pc + j] * (1 - a);
            if (G->u01(rng) < 0.25f) v += G->nrm(rng) * 0.08f * (HI[j] - LO[j]) * anneal;
            clampset(out, j, v);
        }
    }
}

// Raw 32-bit outputs of std::mt19937(seed) (the engine the GA draws from, sim.cu:398), so that
// tests can check it against the standard's definition and its check value.
void or_mt19937(unsigned seed, int n, uint32_t* out) {
    std::mt19937 rng(seed);
    for (int i = 0; i < n; i++) out[i] = (uint32_t)rng();
}

// percentile with linear interpolation (sim.cu:250-256)
float or_percentile(const float* sorted, int n, float q) {
    if (n <= 0) return 0;
    float x = q * (n - 1);
    int a = (int)std::floor(x), b = (int)std::ceil(x);
    float t = x - a;
    return sorted[a] * (1.0f - t) + sorted[b] * t;
}

int or_hw_threads() { return (int)std::thread::hardware_concurrency(); }
float or_sinf(float x) { return osin(x); }
float or_cosf(float x) { return ocos(x); }
float or_tanf(float x) { return otan(x); }

}  // extern "C"
