// pr_physics.h -- per-lane physics building blocks of the planes-rl fitness path.
//
// Everything here is pure arithmetic on registers; the kernels in kernels.hip compose
// these pieces with wave64 cross-lane operations.  Functions marked "exact" are
// evaluated with FMA contraction off so that they reproduce the reference's float
// rounding (common.h equations in source order); the flight step is free to contract.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PR_HD __host__ __device__ __forceinline__

namespace pr {

constexpr int NG = 35;   // genes per genome (140 B, common.h:61-97)
constexpr int NSF = 26;  // scenario fields (common.h:377-387 order + base_z)

// gene indices (Genome field order, common.h:62-96)
enum Gene {
    G_SPAN, G_CHORD, G_TAPER, G_SWEEP, G_DIH_IN, G_DIH_OUT, G_KINK, G_CAMBER1, G_CAMBER2,
    G_WASHOUT, G_ELEVATOR, G_NOSE_MASS, G_KEEL_H, G_DECK_GAP, G_DECK_SCALE, G_BODY_FRAC,
    G_FIN_HEIGHT, G_FIN_CHORD, G_FIN_SWEEP, G_TAIL_SPAN, G_TAIL_CHORD, G_TAIL_ANGLE,
    G_LE_B1, G_LE_B2, G_CHORD_B1, G_CHORD_B2, G_FIN_LAYOUT, G_FIN_CANT, G_FIN_TAPER,
    G_FIN_SPAN_POS, G_TAIL_SWEEP, G_TAIL_TAPER, G_TAIL_DIHEDRAL, G_TIP_ROUND, G_CENTER_CHORD
};
enum Field {
    S_SPEED, S_PITCH, S_YAW, S_ROLL, S_HEIGHT, S_ROLL_RATE, S_PITCH_RATE, S_YAW_RATE, S_RHO, S_CF, S_GUST,
    S_BASE_X, S_BASE_Y, S_FPITCH, S_FYAW, S_GRIP_X, S_GRIP_Y, S_GRIP_Z, S_MATERIAL, S_INFILL, S_BIAS_X,
    S_BIAS_Y, S_BIAS_Z, S_WARP_P, S_WARP_R, S_BASE_Z
};

// model constants (float literals of common.h:112-115,470-478 and sim.cu:35)
#define PR_AREAL (1240.0f * 0.00045f)
#define PR_WALL 0.00045f
#define PR_PRINT_LIMIT (0.250f - 0.00045f)
#define PR_DT (1.0f / 600.0f)
#define PR_GRAV 9.81f

template <class T>
struct V3T { T x, y, z; };
typedef V3T<float> vec3;
PR_HD vec3 V(float x, float y, float z) { vec3 a; a.x = x; a.y = y; a.z = z; return a; }
template <class T>
PR_HD V3T<T> VT(T x, T y, T z) { V3T<T> a; a.x = x; a.y = y; a.z = z; return a; }
// Contraction is decided per function body, so every helper used by an exact path
// carries its own pragma (the library is built with -ffp-contract=fast-honor-pragmas).
template <class T>
PR_HD V3T<T> vadd(V3T<T> a, V3T<T> b) {
#pragma clang fp contract(off)
    return VT<T>(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class T>
PR_HD V3T<T> vsub(V3T<T> a, V3T<T> b) {
#pragma clang fp contract(off)
    return VT<T>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class T>
PR_HD V3T<T> vmul(V3T<T> a, T s) {
#pragma clang fp contract(off)
    return VT<T>(a.x * s, a.y * s, a.z * s);
}
template <class T>
PR_HD T vdot(V3T<T> a, V3T<T> b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
PR_HD V3T<T> vcross(V3T<T> a, V3T<T> b) {
#pragma clang fp contract(off)
    return VT<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

struct quat { float w, x, y, z; };

// ------------------------------------------------------------------ double-evaluated sin/cos
// sin/cos of a double by Cody-Waite reduction against pi/2 split into three doubles (with
// FMA, so each partial product is exact before its single rounding), then Taylor
// polynomials on |r| <= pi/4 (truncation < 1e-19).  Rounded once to float this gives the
// correctly rounded float sin/cos/tan for all but vanishingly rare arguments.
PR_HD double poly_sin(double r) {
#pragma clang fp contract(off)
    double z = r * r;
    double p = 2.8114572543455206e-15;          // 1/17!
    p = p * z - 7.6471637318198164e-13;         // -1/15!
    p = p * z + 1.6059043836821613e-10;         // 1/13!
    p = p * z - 2.5052108385441720e-08;         // -1/11!
    p = p * z + 2.7557319223985893e-06;         // 1/9!
    p = p * z - 1.9841269841269841e-04;         // -1/7!
    p = p * z + 8.3333333333333332e-03;         // 1/5!
    p = p * z - 1.6666666666666666e-01;         // -1/3!
    return r + r * (z * p);
}
PR_HD double poly_cos(double r) {
#pragma clang fp contract(off)
    double z = r * r;
    double p = 1.5619206968586225e-16;          // 1/18!
    p = p * z - 4.7794773323873853e-14;         // -1/16!
    p = p * z + 1.1470745597729725e-11;         // 1/14!
    p = p * z - 2.0876756987868100e-09;         // -1/12!
    p = p * z + 2.7557319223985888e-07;         // 1/10!
    p = p * z - 2.4801587301587302e-05;         // -1/8!
    p = p * z + 1.3888888888888889e-03;         // 1/6!
    p = p * z - 4.1666666666666664e-02;         // -1/4!
    p = p * z + 0.5;                            // 1/2!
    return 1.0 - z * p;
}
PR_HD void sincos_d(double x, double* s, double* c) {
#pragma clang fp contract(off)
    const double inv_half_pi = 0.63661977236758138;
    const double P1 = 1.5707963267948966e+00, P2 = 6.1232339957367660e-17, P3 = -1.4973849048591698e-33;
    double k = __builtin_rint(x * inv_half_pi);
    double r = __builtin_fma(-k, P1, x);
    r = __builtin_fma(-k, P2, r);
    r = __builtin_fma(-k, P3, r);
    long long q = (long long)k;
    double ps = poly_sin(r), pc = poly_cos(r);
    switch ((int)(q & 3)) {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}
PR_HD float sinf_cr(float x) { double s, c; sincos_d((double)x, &s, &c); return (float)s; }
PR_HD float cosf_cr(float x) { double s, c; sincos_d((double)x, &s, &c); return (float)c; }
PR_HD float tanf_cr(float x) {
#pragma clang fp contract(off)
    double s, c;
    sincos_d((double)x, &s, &c);
    return (float)(s / c);
}

// natural log of a positive normal double: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(z), z = (m-1)/(m+1), series to z^27 (|z| <= 0.1716).
PR_HD double log_d(double x) {
#pragma clang fp contract(off)
    uint64_t b;
    __builtin_memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    __builtin_memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    double z = (m - 1.0) / (m + 1.0), z2 = z * z;
    double p = 1.0 / 27;
    p = p * z2 + 1.0 / 25; p = p * z2 + 1.0 / 23; p = p * z2 + 1.0 / 21; p = p * z2 + 1.0 / 19;
    p = p * z2 + 1.0 / 17; p = p * z2 + 1.0 / 15; p = p * z2 + 1.0 / 13; p = p * z2 + 1.0 / 11;
    p = p * z2 + 1.0 / 9;  p = p * z2 + 1.0 / 7;  p = p * z2 + 1.0 / 5;  p = p * z2 + 1.0 / 3;
    double lm = 2.0 * z + 2.0 * z * (z2 * p);
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    return (double)e * ln2_hi + (lm + (double)e * ln2_lo);
}

// ------------------------------------------------------------------ single-precision sin (flight loop)
// Cody-Waite against pi/2 in three float parts (FMA-exact products), odd/even Taylor
// polynomials on |r| <= pi/4 (truncation < 2e-9).  ~1 ulp for |x| up to ~1e4 rad,
// which covers every gust-field argument of a flight that has not been terminated.
PR_HD float sin_fast(float x) {
    const float inv_half_pi = 0.636619772f;
    const float C1 = 1.57079637e+00f, C2 = -4.37113883e-08f, C3 = -1.71512451e-15f;
    float k = __builtin_rintf(x * inv_half_pi);
    float r = __builtin_fmaf(-k, C1, x);
    r = __builtin_fmaf(-k, C2, r);
    r = __builtin_fmaf(-k, C3, r);
    int q = (int)k;
    float z = r * r;
    float s = __builtin_fmaf(z, 2.75573192e-06f, -1.98412698e-04f);
    s = __builtin_fmaf(s, z, 8.33333333e-03f);
    s = __builtin_fmaf(s, z, -1.66666667e-01f);
    s = __builtin_fmaf(s * z, r, r);
    float c = __builtin_fmaf(z, -2.75573192e-07f, 2.48015873e-05f);
    c = __builtin_fmaf(c, z, -1.38888889e-03f);
    c = __builtin_fmaf(c, z, 4.16666667e-02f);
    c = __builtin_fmaf(c, z, -0.5f);
    c = __builtin_fmaf(c, z, 1.0f);
    float v = (q & 1) ? c : s;
    return (q & 2) ? -v : v;
}

// Gust-term sine for the flight loop: Cody-Waite reduction against 2 pi in three float parts
// (the split of 2 pi into float head, middle and tail) to r in [-pi, pi], then the hardware
// v_sin_f32, whose argument is in revolutions (measured on gfx950: <= 8e-7 absolute error on
// [-0.5, 0.5] revolutions). 7 instructions instead of sin_fast's ~22; the gust amplitude
// (<= ~3 m/s) turns that into < 3e-6 m/s of wind.
__device__ __forceinline__ float sin_gust(float x) {
    const float inv_2pi = 0.159154943f;
    const float D1 = 6.28318548e+00f, D2 = -1.74845553e-07f, D3 = -6.86049804e-15f;
    const float k = __builtin_rintf(x * inv_2pi);
    float r = __builtin_fmaf(-k, D1, x);
    r = __builtin_fmaf(-k, D2, r);
    r = __builtin_fmaf(-k, D3, r);
    return __builtin_amdgcn_sinf(r * inv_2pi);
}

// ------------------------------------------------------------------ scenario hash (common.h:368-375)
PR_HD uint32_t mix32(uint32_t x) {  // Wellons' lowbias32
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
PR_HD uint32_t lane_key(int seed, int lane) { return (uint32_t)seed * 0x9e3779b9u + (uint32_t)lane * 0x85ebca6bu; }
PR_HD float u01(int seed, int lane) {
#pragma clang fp contract(off)
    return (float)(mix32(lane_key(seed, lane)) >> 8) * (1.0f / 16777216.0f);
}
PR_HD float pm1(int seed, int lane) {
#pragma clang fp contract(off)
    return u01(seed, lane) * 2.0f - 1.0f;
}
PR_HD int phase_of(int seed) { return (int)((uint32_t)seed * 37u) % 97; }  // (seed*37)%97, 32-bit wrap

// Box-Muller standard normal from lanes (lane, lane+1), evaluated in double, rounded once.
PR_HD float normal01(int seed, int lane) {
#pragma clang fp contract(off)
    double a = (double)((mix32(lane_key(seed, lane)) >> 8) + 1u) * (1.0 / 16777216.0);
    double b = (double)(mix32(lane_key(seed, lane + 1)) >> 8) * (1.0 / 16777216.0);
    double r = __builtin_sqrt(-2.0 * log_d(a));
    double s, c;
    sincos_d(6.283185307179586 * b, &s, &c);
    return (float)(r * c);
}

struct ScenarioParams {
    int preset;          // 0 = reference hash preset (common.h:389-421), 1 = parametric config preset
    float gust_rms_max;  // config: gust field 3-D RMS ~ U(0, max) [m/s]
    float warp_max;      // config: warp pitch/roll ~ U(-max, max) [rad]
    float wind_sd;       // config: steady wind N(0, sd) per axis [m/s]
};

// exact: the reference preset, field for field (common.h:389-421)
PR_HD void scenario_reference(int seed, float* d) {
#pragma clang fp contract(off)
    d[S_SPEED] = 9.2f + 2.0f * pm1(seed, 0);
    d[S_PITCH] = 0.10f + 0.20f * pm1(seed, 1);
    d[S_YAW] = 0.12f * pm1(seed, 2);
    d[S_ROLL] = 0.16f * pm1(seed, 3);
    d[S_HEIGHT] = 1.75f + 0.35f * pm1(seed, 4);
    d[S_ROLL_RATE] = 1.8f * pm1(seed, 5);
    d[S_PITCH_RATE] = 1.2f * pm1(seed, 6);
    d[S_YAW_RATE] = 1.5f * pm1(seed, 7);
    d[S_RHO] = 1.225f * (0.88f + 0.22f * u01(seed, 8));
    float cf = 0.02f * (0.75f + 0.55f * u01(seed, 9));
    d[S_GUST] = 0.45f + 1.10f * (0.5f + 0.5f * sinf_cr((float)seed * 1.61803f));
    float ph = (float)phase_of(seed);
    d[S_BASE_X] = -0.25f + 1.15f * sinf_cr(ph * 0.71f);
    d[S_BASE_Y] = 1.00f * sinf_cr(ph * 1.13f);
    d[S_FPITCH] = 0.075f * pm1(seed, 10);
    d[S_FYAW] = 0.060f * pm1(seed, 11);
    d[S_GRIP_X] = 0.018f * pm1(seed, 12);
    d[S_GRIP_Y] = 0.022f * pm1(seed, 13);
    d[S_GRIP_Z] = 0.010f * pm1(seed, 14);
    d[S_MATERIAL] = 0.88f + 0.24f * u01(seed, 15);
    d[S_INFILL] = 0.06f + 0.34f * u01(seed, 16);
    d[S_BIAS_X] = 0.018f * pm1(seed, 17);
    d[S_BIAS_Y] = 0.010f * pm1(seed, 18);
    d[S_BIAS_Z] = 0.006f * pm1(seed, 19);
    d[S_CF] = cf * (0.90f + 0.25f * u01(seed, 20));
    d[S_WARP_P] = 0.025f * pm1(seed, 21);
    d[S_WARP_R] = 0.035f * pm1(seed, 22);
    d[S_BASE_Z] = 0.0f;
}

// exact: parametric preset of the benchmark configurations (no reference counterpart)
PR_HD void scenario_config(int seed, const ScenarioParams& sp, float* d) {
#pragma clang fp contract(off)
    const float deg = 0.0174532925f;
    d[S_SPEED] = 9.2f + 2.0f * pm1(seed, 0);
    d[S_PITCH] = 0.10f + 5.0f * deg * normal01(seed, 32);
    d[S_YAW] = 5.0f * deg * normal01(seed, 34);
    d[S_ROLL] = 5.0f * deg * normal01(seed, 36);
    d[S_HEIGHT] = 1.75f + 0.25f * pm1(seed, 4);
    d[S_ROLL_RATE] = 30.0f * deg * normal01(seed, 38);
    d[S_PITCH_RATE] = 30.0f * deg * normal01(seed, 40);
    d[S_YAW_RATE] = 30.0f * deg * normal01(seed, 42);
    d[S_RHO] = 1.15f + 0.10f * u01(seed, 8);
    float cf = 0.02f * (0.75f + 0.55f * u01(seed, 9));
    d[S_GUST] = sp.gust_rms_max * u01(seed, 23) * (1.0f / 0.69191763f);
    d[S_BASE_X] = sp.wind_sd * normal01(seed, 44);
    d[S_BASE_Y] = sp.wind_sd * normal01(seed, 46);
    d[S_FPITCH] = 5.0f * deg * normal01(seed, 48);
    d[S_FYAW] = 5.0f * deg * normal01(seed, 50);
    d[S_GRIP_X] = 0.018f * pm1(seed, 12);
    d[S_GRIP_Y] = 0.022f * pm1(seed, 13);
    d[S_GRIP_Z] = 0.010f * pm1(seed, 14);
    d[S_MATERIAL] = 0.95f + 0.10f * u01(seed, 15);
    d[S_INFILL] = 0.06f + 0.34f * u01(seed, 16);
    d[S_BIAS_X] = 0.018f * pm1(seed, 17);
    d[S_BIAS_Y] = 0.010f * pm1(seed, 18);
    d[S_BIAS_Z] = 0.006f * pm1(seed, 19);
    d[S_CF] = cf * (0.90f + 0.25f * u01(seed, 20));
    d[S_WARP_P] = sp.warp_max * pm1(seed, 21);
    d[S_WARP_R] = sp.warp_max * pm1(seed, 22);
    d[S_BASE_Z] = sp.wind_sd * normal01(seed, 52);
}

// ------------------------------------------------------------------ mass properties
struct Mass {           // 80 B, the reference's MassProps layout (common.h:117-122)
    float mass, invmass;
    float com[3];
    float I[6];          // xx yy zz xy xz yz  (xy.. hold -sum m x y)
    float invI[9];
};

PR_HD void invert_sym3(float xx, float yy, float zz, float xy, float xz, float yz, bool floor_det, float* inv) {
#pragma clang fp contract(off)
    float det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    float id = 1.0f / (floor_det ? fmaxf(det, 1e-24f) : det);
    inv[0] = (yy * zz - yz * yz) * id; inv[1] = (xz * yz - xy * zz) * id; inv[2] = (xy * yz - xz * yy) * id;
    inv[3] = inv[1]; inv[4] = (xx * zz - xz * xz) * id; inv[5] = (xy * xz - xx * yz) * id;
    inv[6] = inv[2]; inv[7] = inv[5]; inv[8] = (xx * yy - xy * xy) * id;
}

// exact: per-flight printed mass properties (common.h:441-457)
PR_HD void printed_mass(const Mass& src, const float* d, Mass& out, vec3& shift) {
#pragma clang fp contract(off)
    float m_shell = src.mass * d[S_MATERIAL];
    float m_fill = m_shell * d[S_INFILL];
    float m = m_shell + m_fill;
    vec3 bias = V(d[S_BIAS_X], d[S_BIAS_Y], d[S_BIAS_Z]);
    shift = vmul(bias, m_fill / m);
    float sc = m / src.mass;
    float I[6];
    for (int k = 0; k < 6; k++) I[k] = src.I[k] * sc;
    vec3 r = vmul(shift, -1.0f);
    I[0] += m_shell * (r.y * r.y + r.z * r.z); I[1] += m_shell * (r.x * r.x + r.z * r.z);
    I[2] += m_shell * (r.x * r.x + r.y * r.y); I[3] -= m_shell * r.x * r.y;
    I[4] -= m_shell * r.x * r.z; I[5] -= m_shell * r.y * r.z;
    r = vsub(bias, shift);
    I[0] += m_fill * (r.y * r.y + r.z * r.z); I[1] += m_fill * (r.x * r.x + r.z * r.z);
    I[2] += m_fill * (r.x * r.x + r.y * r.y); I[3] -= m_fill * r.x * r.y;
    I[4] -= m_fill * r.x * r.z; I[5] -= m_fill * r.y * r.z;
    out.mass = m;
    out.invmass = 1.0f / m;
    out.com[0] = src.com[0] + shift.x; out.com[1] = src.com[1] + shift.y; out.com[2] = src.com[2] + shift.z;
    for (int k = 0; k < 6; k++) out.I[k] = I[k];
    invert_sym3(I[0], I[1], I[2], I[3], I[4], I[5], true, out.invI);
}

// ------------------------------------------------------------------ quaternion helpers (exact)
PR_HD quat qmul(quat a, quat b) {
#pragma clang fp contract(off)
    quat r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return r;
}
PR_HD quat qnormalize(quat q) {
#pragma clang fp contract(off)
    float l = sqrtf(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    float il = 1.0f / l;
    quat r; r.w = q.w * il; r.x = q.x * il; r.y = q.y * il; r.z = q.z * il;
    return r;
}
// rotation of a unit coordinate axis (axis 0=x,1=y,2=z) by ang: common.h:54-57 with norm3(axis)=axis
PR_HD quat qaxis_unit(int axis, float ang) {
#pragma clang fp contract(off)
    float s = sinf_cr(ang * 0.5f);
    quat q; q.w = cosf_cr(ang * 0.5f);
    q.x = axis == 0 ? s : 0.0f * s; q.y = axis == 1 ? s : 0.0f * s; q.z = axis == 2 ? s : 0.0f * s;
    return q;
}
struct Rot { vec3 r0, r1, r2; };  // rows of the body->world matrix (common.h:38-43)
PR_HD Rot rot_of(quat q) {
#pragma clang fp contract(off)
    float w = q.w, x = q.x, y = q.y, z = q.z;
    Rot m;
    m.r0 = V(1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y));
    m.r1 = V(2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x));
    m.r2 = V(2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y));
    return m;
}

// Per-flight constants (128 B): printed mass properties and the scenario terms the
// flight loop needs; written once per pr_fly by k_init.
struct FlightConst {
    float mass, invmass;
    float I[6];           // printed inertia xx yy zz xy xz yz
    float iv[6];          // symmetric inverse: [0] [1] [2] [4] [5] [8] of the 3x3
    float cf, half_rho;   // skin friction, 0.5 * air density
    float base[3], gust, phase;
    float shift[3];       // printed CG shift (panel centroids are relative to the shell COM)
    float warp_p, warp_r;
    int genome, scen;
    float pad[3];
};

struct FlightState {  // 64 B per flight in HBM
    float pos[3], q[4], vel[3], wb[3];
    int step;         // integrated steps so far
    int done;
    float fitness;
};

// exact: launch of one flight (sim.cu:70-109).  spread > 0 offsets the showcase grid.
PR_HD void launch_state(const float* d, const Mass& mp, vec3 shift, int plane_index, float spread,
                        FlightState& st) {
#pragma clang fp contract(off)
    float ox = 0, oy = 0, oz = 0;
    if (spread > 0) {
        int cols = 25;
        oy = (float)((plane_index % cols) - cols / 2) * spread;
        ox = (float)(-(plane_index / cols)) * spread * 1.4f;
        oz = 0.25f * sinf_cr((float)plane_index * 12.9898f);
    }
    float speed = d[S_SPEED], pitch = d[S_PITCH], yaw = d[S_YAW], roll = d[S_ROLL], height = d[S_HEIGHT];
    st.pos[0] = ox; st.pos[1] = oy; st.pos[2] = height + oz;
    quat qy = qaxis_unit(2, yaw), qp = qaxis_unit(1, -pitch), qr = qaxis_unit(0, roll);
    quat q = qnormalize(qmul(qy, qmul(qp, qr)));
    quat fy = qaxis_unit(2, yaw + d[S_FYAW]);
    quat fp = qaxis_unit(1, -(pitch + d[S_FPITCH]));
    Rot rf = rot_of(qmul(fy, fp));
    vec3 v0 = V(speed, 0, 0);
    vec3 vel = V(vdot(rf.r0, v0), vdot(rf.r1, v0), vdot(rf.r2, v0));
    vec3 wb = V(d[S_ROLL_RATE], d[S_PITCH_RATE], d[S_YAW_RATE]);
    vec3 grip = V(d[S_GRIP_X], d[S_GRIP_Y], d[S_GRIP_Z]);
    Rot rq = rot_of(q);
    vec3 j = vmul(vel, mp.mass);
    vec3 Jb = V(rq.r0.x * j.x + rq.r1.x * j.y + rq.r2.x * j.z, rq.r0.y * j.x + rq.r1.y * j.y + rq.r2.y * j.z,
                rq.r0.z * j.x + rq.r1.z * j.y + rq.r2.z * j.z);
    vec3 L = vmul(vcross(vsub(grip, shift), Jb), 0.12f);
    const float* iv = mp.invI;
    vec3 dw = V(iv[0] * L.x + iv[1] * L.y + iv[2] * L.z, iv[3] * L.x + iv[4] * L.y + iv[5] * L.z,
                iv[6] * L.x + iv[7] * L.y + iv[8] * L.z);
    wb = vadd(wb, dw);
    float w0 = sqrtf(vdot(wb, wb));
    if (w0 > 35.0f) wb = vmul(wb, 35.0f / w0);
    st.q[0] = q.w; st.q[1] = q.x; st.q[2] = q.y; st.q[3] = q.z;
    st.vel[0] = vel.x; st.vel[1] = vel.y; st.vel[2] = vel.z;
    st.wb[0] = wb.x; st.wb[1] = wb.y; st.wb[2] = wb.z;
    st.step = 0;
    st.done = 0;
    st.fitness = 0.0f;
}

// exact: print-warped unit normal of one panel (sim.cu:140-143)
PR_HD vec3 warped_normal(vec3 rb, vec3 nb, float warp_p, float warp_r) {
#pragma clang fp contract(off)
    float ar = warp_r * fmaxf(-1.f, fminf(1.f, rb.y / 0.125f));
    float ap = warp_p * fmaxf(-1.f, fminf(1.f, rb.x / 0.125f));
    vec3 nr = V(nb.x, nb.y - ar * nb.z, nb.z + ar * nb.y);
    vec3 v = V(nr.x + ap * nr.z, nr.y, nr.z - ap * nr.x);
    float l = sqrtf(vdot(v, v));
    return l > 1e-12f ? vmul(v, 1.0f / l) : V(0, 0, 1);
}

// ------------------------------------------------------------------ mesh topology (common.h:177-292)
struct MeshLayout {
    int ch, sp;        // chordwise segments, spanwise stations per half
    int nw, tw;        // vertices / triangles of one wing deck
    int nfin, fin_kind;// fins: count and layout (0 centre, 1 twin, 2 triple, 3 outboard)
    int bi;            // biplane
    int nv, nt;
    int v_tail, v_fin, v_deck2, v_keel, v_strut;  // vertex range starts
    int t_tail, t_fin, t_deck2, t_keel, t_strut;  // triangle range starts
};

PR_HD MeshLayout layout_of(const float* g, int ch, int sp) {
    MeshLayout L;
    L.ch = ch; L.sp = sp;
    L.nw = (2 * sp + 1) * (ch + 1);
    L.tw = 2 * (2 * sp) * ch;
    float lay = g[G_FIN_LAYOUT];
    L.fin_kind = lay < 0.25f ? 0 : (lay < 0.50f ? 1 : (lay < 0.75f ? 2 : 3));
    L.nfin = L.fin_kind == 0 ? 1 : (L.fin_kind == 2 ? 3 : 2);
    L.bi = g[G_DECK_GAP] > 0.02f ? 1 : 0;
    L.v_tail = L.nw; L.v_fin = L.v_tail + 9; L.v_deck2 = L.v_fin + 4 * L.nfin;
    L.v_keel = L.v_deck2 + L.bi * L.nw; L.v_strut = L.v_keel + 8; L.nv = L.v_strut + 8 * L.bi;
    L.t_tail = L.tw; L.t_fin = L.t_tail + 8; L.t_deck2 = L.t_fin + 2 * L.nfin;
    L.t_keel = L.t_deck2 + L.bi * L.tw; L.t_strut = L.t_keel + 6; L.nt = L.t_strut + 4 * L.bi;
    return L;
}

// triangle t -> vertex indices
PR_HD void tri_of(const MeshLayout& L, int t, int* a3) {
    int base, k, a, b;
    if (t < L.t_tail || (L.bi && t >= L.t_deck2 && t < L.t_keel)) {  // wing grid cell
        base = t < L.t_tail ? 0 : L.v_deck2;
        k = t < L.t_tail ? t : t - L.t_deck2;
        int cell = k >> 1, i = cell / L.ch, j = cell - i * L.ch;
        a = base + i * (L.ch + 1) + j; b = a + L.ch + 1;
        if (k & 1) { a3[0] = b; a3[1] = b + 1; a3[2] = a + 1; } else { a3[0] = a; a3[1] = b; a3[2] = a + 1; }
    } else if (t < L.t_fin) {  // tail 2x2 cells
        k = t - L.t_tail;
        int cell = k >> 1, i = cell >> 1, j = cell & 1;
        a = L.v_tail + i * 3 + j; b = a + 3;
        if (k & 1) { a3[0] = b; a3[1] = b + 1; a3[2] = a + 1; } else { a3[0] = a; a3[1] = b; a3[2] = a + 1; }
    } else if (t < L.t_deck2) {  // fins: quads (b, b+1, b+2) (b+1, b+3, b+2)
        k = t - L.t_fin;
        base = L.v_fin + 4 * (k >> 1);
        if (k & 1) { a3[0] = base + 1; a3[1] = base + 3; a3[2] = base + 2; } else { a3[0] = base; a3[1] = base + 1; a3[2] = base + 2; }
    } else if (t < L.t_strut) {  // keel strip
        k = t - L.t_keel;
        a = L.v_keel + (k >> 1) * 2;
        if (k & 1) { a3[0] = a + 1; a3[1] = a + 3; a3[2] = a + 2; } else { a3[0] = a; a3[1] = a + 1; a3[2] = a + 2; }
    } else {  // struts
        k = t - L.t_strut;
        base = L.v_strut + 4 * (k >> 1);
        if (k & 1) { a3[0] = base + 1; a3[1] = base + 3; a3[2] = base + 2; } else { a3[0] = base; a3[1] = base + 1; a3[2] = base + 2; }
    }
}

// ------------------------------------------------------------------ geometry in T (float or double)
// The float instantiation reproduces the reference's float rounding; the double one is
// the FP64 prepare mode (model constants stay the reference's float literals).
PR_HD float tsin(float x) { return sinf_cr(x); }
PR_HD float tcos(float x) { return cosf_cr(x); }
PR_HD float ttan(float x) { return tanf_cr(x); }
PR_HD double tsin(double x) { double s, c; sincos_d(x, &s, &c); return s; }
PR_HD double tcos(double x) { double s, c; sincos_d(x, &s, &c); return c; }
PR_HD double ttan(double x) {
#pragma clang fp contract(off)
    double s, c;
    sincos_d(x, &s, &c);
    return s / c;
}
PR_HD float tsqrt(float x) { return sqrtf(x); }
PR_HD double tsqrt(double x) { return sqrt(x); }
PR_HD float tmin(float a, float b) { return fminf(a, b); }
PR_HD double tmin(double a, double b) { return fmin(a, b); }
PR_HD float tmax(float a, float b) { return fmaxf(a, b); }
PR_HD double tmax(double a, double b) { return fmax(a, b); }
PR_HD float tabs(float a) { return fabsf(a); }
PR_HD double tabs(double a) { return fabs(a); }

// exact: camber line (common.h:124-129)
template <class T>
PR_HD T camber_line(const T* g, T xc) {
#pragma clang fp contract(off)
    T te = g[G_ELEVATOR] * T(0.22f);
    T u = xc, v = T(1) - u;
    return T(3) * v * v * u * g[G_CAMBER1] + T(3) * v * u * u * g[G_CAMBER2] + u * u * u * te;
}

// exact: wing surface point (common.h:132-172)
template <class T>
PR_HD V3T<T> wing_point(const T* g, T u, T xc, T zoff, T scale) {
#pragma clang fp contract(off)
    T half = g[G_SPAN] * T(0.5f) * scale;
    T t = tabs(u) * half;
    // dihedral polyline (common.h:132-141)
    T hs = g[G_SPAN] * T(0.5f);
    T b = g[G_BODY_FRAC] * hs, k = g[G_KINK] * hs;
    if (k < b) k = b;
    T yy = 0, zz = 0, seg;
    seg = tmin(t, b); yy += seg;
    seg = tmin(t, k) - b;
    if (seg > 0) { yy += seg * tcos(g[G_DIH_IN]); zz += seg * tsin(g[G_DIH_IN]); }
    seg = t - k;
    if (seg > 0) { yy += seg * tcos(g[G_DIH_OUT]); zz += seg * tsin(g[G_DIH_OUT]); }
    T tf = tabs(u);
    T v = T(1) - tf;
    T cr = T(1) + g[G_CENTER_CHORD];
    T ct = g[G_TAPER];
    T c1 = T(1) + g[G_CHORD_B1], c2 = ct + g[G_CHORD_B2];
    T ratio = v * v * v * cr + T(3) * v * v * tf * c1 + T(3) * v * tf * tf * c2 + tf * tf * tf * ct;
    T ell = tsqrt(tmax(T(0.08f), T(1) - T(0.92f) * tf * tf));
    ratio *= T(1) - g[G_TIP_ROUND] + g[G_TIP_ROUND] * ell;
    ratio = tmax(T(0.035f), tmin(T(1.9f), ratio));
    T c = g[G_CHORD] * scale * ratio;
    T xr = T(0.5f) * g[G_CHORD] * scale * cr;
    T xt = T(0.5f) * g[G_CHORD] * scale - ttan(g[G_SWEEP]) * half;
    T xb1 = xr + g[G_LE_B1] * g[G_CHORD] * scale;
    T xb2 = xt + g[G_LE_B2] * g[G_CHORD] * scale;
    T xle = v * v * v * xr + T(3) * v * v * tf * xb1 + T(3) * v * tf * tf * xb2 + tf * tf * tf * xt;
    T x = xle - xc * c;
    T zc = zz * scale + camber_line(g, xc) * c;
    T xq = xle - T(0.25f) * c;
    T tw = g[G_WASHOUT] * tf;
    zc += tw * (x - xq);
    return VT<T>(x, (u < 0 ? -yy : yy), zc + zoff);
}

// exact: vertex k (0..3) of a fin rooted at (y0, z0) (common.h:197-213)
template <class T>
PR_HD V3T<T> fin_vertex(const T* g, T y0, T z0, int side, int k) {
#pragma clang fp contract(off)
    T fc = g[G_CHORD] * g[G_FIN_CHORD];
    T sw = fc * g[G_FIN_SWEEP];
    T xr = T(-0.32f) * g[G_CHORD] - g[G_TAIL_CHORD] * g[G_CHORD] * T(0.82f);
    T tipc = fc * g[G_FIN_TAPER];
    T cant = side == 0 ? T(0) : g[G_FIN_CANT];
    T ya = y0 + T(side) * tsin(cant) * g[G_FIN_HEIGHT];
    T za = z0 + tcos(cant) * g[G_FIN_HEIGHT];
    if (k == 0) return VT<T>(xr + T(0.5f) * fc, y0, z0);
    if (k == 1) return VT<T>(xr - T(0.5f) * fc, y0, z0);
    if (k == 2) return VT<T>(xr - sw + T(0.5f) * tipc, ya, za);
    return VT<T>(xr - sw - T(0.5f) * tipc, ya, za);
}

// exact: vertex v of the mesh (build_mesh emission order, common.h:257-292)
template <class T>
PR_HD V3T<T> mesh_vertex(const T* g, const MeshLayout& L, int v) {
#pragma clang fp contract(off)
    if (v < L.v_tail || (L.bi && v >= L.v_deck2 && v < L.v_keel)) {
        bool top = v >= L.v_deck2;
        int k = top ? v - L.v_deck2 : v;
        int i = k / (L.ch + 1), j = k - i * (L.ch + 1);
        T u = T((float)(i - L.sp) / (float)L.sp);
        T xc = T((float)j / (float)L.ch);
        return top ? wing_point(g, u, xc, g[G_DECK_GAP], g[G_DECK_SCALE]) : wing_point(g, u, xc, T(0), T(1));
    }
    if (v < L.v_fin) {  // horizontal tail (common.h:215-236)
        int k = v - L.v_tail, i = k / 3, j = k - 3 * i;
        T hs = T(0.5f) * g[G_SPAN] * g[G_TAIL_SPAN];
        T hc = g[G_CHORD] * g[G_TAIL_CHORD];
        T x0 = T(-0.32f) * g[G_CHORD];
        T z0 = camber_line(g, T(0.82f)) * g[G_CHORD];
        T u = T((float)(i - 1)), au = tabs(u);
        T tc = hc * (T(1) + (g[G_TAIL_TAPER] - T(1)) * au);
        T xle = x0 - g[G_TAIL_SWEEP] * hc * au;
        T xc = T(0.5f * (float)j);
        T x = xle - xc * tc;
        T z = z0 + g[G_TAIL_ANGLE] * (x - xle) + au * hs * tsin(g[G_TAIL_DIHEDRAL]);
        return VT<T>(x, u * hs * tcos(g[G_TAIL_DIHEDRAL]), z);
    }
    if (v < L.v_deck2) {  // fins (common.h:238-254)
        int k = v - L.v_fin, f = k >> 2, c = k & 3;
        T hs = T(0.5f) * g[G_SPAN] * g[G_TAIL_SPAN];
        T z0 = camber_line(g, T(0.82f)) * g[G_CHORD];
        T y0, zr;
        int side;
        if (L.fin_kind == 3) {
            T yo = hs * T(0.94f) * tcos(g[G_TAIL_DIHEDRAL]);
            T zo = z0 + hs * T(0.94f) * tsin(g[G_TAIL_DIHEDRAL]);
            side = f == 0 ? -1 : 1; y0 = f == 0 ? -yo : yo; zr = zo;
        } else {
            int pf = L.fin_kind == 2 ? f - 1 : (L.fin_kind == 1 ? f : -1);  // -1: centre fin
            if (pf < 0) { side = 0; y0 = T(0); zr = z0; }
            else {
                T yp = hs * g[G_FIN_SPAN_POS] * tcos(g[G_TAIL_DIHEDRAL]);
                T zp = z0 + hs * g[G_FIN_SPAN_POS] * tsin(g[G_TAIL_DIHEDRAL]);
                side = pf == 0 ? -1 : 1; y0 = pf == 0 ? -yp : yp; zr = zp;
            }
        }
        return fin_vertex(g, y0, zr, side, c);
    }
    if (v < L.v_strut) {  // keel sheet (common.h:264-277)
        int k = v - L.v_keel, j = k >> 1;
        T xc = T((float)j / 3.0f);
        V3T<T> top = wing_point(g, T(0), xc, T(0), T(1));
        if (!(k & 1)) return top;
        return VT<T>(top.x, T(0), top.z - g[G_KEEL_H] * (T(0.35f) + T(0.65f) * xc));
    }
    // biplane struts at the kink stations (common.h:279-290)
    int k = v - L.v_strut, s = k >> 2, c = k & 3;
    T u = T(s == 0 ? -1 : 1) * g[G_KINK];
    T xc = (c & 1) ? T(0.55f) : T(0.25f);
    return c < 2 ? wing_point(g, u, xc, T(0), T(1)) : wing_point(g, u, xc, g[G_DECK_GAP], g[G_DECK_SCALE]);
}

// exact: first-pass terms of one triangle (common.h:313-319): mass and mass-weighted centroid
template <class T>
PR_HD void tri_mass_terms(V3T<T> a, V3T<T> b, V3T<T> d, T* mi, V3T<T>* mct) {
#pragma clang fp contract(off)
    V3T<T> cr = vcross(vsub(b, a), vsub(d, a));
    T A = T(0.5f) * tsqrt(vdot(cr, cr));
    T m = A * T(PR_AREAL);
    V3T<T> ct = vmul(vadd(vadd(a, b), d), T(1.0f / 3));
    *mi = m;
    *mct = vmul(ct, m);
}

// exact: second-pass inertia terms of one triangle about c (common.h:329-351).
// out[0..5]: lamina terms (added to xx,yy,zz / subtracted from xy,xz,yz),
// out[6..11]: through-thickness terms, same signs.
template <class T>
PR_HD void tri_inertia_terms(V3T<T> a, V3T<T> b, V3T<T> d, V3T<T> c, T* out) {
#pragma clang fp contract(off)
    V3T<T> cr = vcross(vsub(b, a), vsub(d, a));
    T A = T(0.5f) * tsqrt(vdot(cr, cr));
    T mi = A * T(PR_AREAL);
    V3T<T> ra = vsub(a, c), rb = vsub(b, c), rd = vsub(d, c);
    T exx = (ra.x * ra.x + rb.x * rb.x + rd.x * rd.x + ra.x * rb.x + ra.x * rd.x + rb.x * rd.x) / T(6);
    T eyy = (ra.y * ra.y + rb.y * rb.y + rd.y * rd.y + ra.y * rb.y + ra.y * rd.y + rb.y * rd.y) / T(6);
    T ezz = (ra.z * ra.z + rb.z * rb.z + rd.z * rd.z + ra.z * rb.z + ra.z * rd.z + rb.z * rd.z) / T(6);
    T exy = (T(2) * (ra.x * ra.y + rb.x * rb.y + rd.x * rd.y) + ra.x * rb.y + rb.x * ra.y + ra.x * rd.y +
             rd.x * ra.y + rb.x * rd.y + rd.x * rb.y) / T(12);
    T exz = (T(2) * (ra.x * ra.z + rb.x * rb.z + rd.x * rd.z) + ra.x * rb.z + rb.x * ra.z + ra.x * rd.z +
             rd.x * ra.z + rb.x * rd.z + rd.x * rb.z) / T(12);
    T eyz = (T(2) * (ra.y * ra.z + rb.y * rb.z + rd.y * rd.z) + ra.y * rb.z + rb.y * ra.z + ra.y * rd.z +
             rd.y * ra.z + rb.y * rd.z + rd.y * rb.z) / T(12);
    out[0] = mi * (eyy + ezz); out[1] = mi * (exx + ezz); out[2] = mi * (exx + eyy);
    out[3] = mi * exy; out[4] = mi * exz; out[5] = mi * eyz;
    T l = tsqrt(vdot(cr, cr));
    V3T<T> nn = l > T(1e-12f) ? vmul(cr, T(1) / l) : VT<T>(0, 0, 1);
    T th = mi * T(PR_WALL) * T(PR_WALL) / T(12);
    out[6] = th * (T(1) - nn.x * nn.x); out[7] = th * (T(1) - nn.y * nn.y); out[8] = th * (T(1) - nn.z * nn.z);
    out[9] = th * nn.x * nn.y; out[10] = th * nn.x * nn.z; out[11] = th * nn.y * nn.z;
}

// exact: panel record of one triangle relative to the shell COM (sim.cu:41-47)
template <class T>
PR_HD void panel_of(V3T<T> va, V3T<T> vb, V3T<T> vc, V3T<T> com, V3T<T>* cen, T* area, V3T<T>* nrm) {
#pragma clang fp contract(off)
    V3T<T> a = vsub(va, com), b = vsub(vb, com), c = vsub(vc, com);
    V3T<T> n = vcross(vsub(b, a), vsub(c, a));
    T A = T(0.5f) * tsqrt(vdot(n, n));
    *cen = vmul(vadd(vadd(a, b), c), T(1.0f / 3));
    *area = A;
    *nrm = A > T(1e-12f) ? vmul(n, T(0.5f) / A) : VT<T>(0, 0, 1);
}

// exact: symmetric 3x3 inverse in T (common.h:356-362 / 429-436)
template <class T>
PR_HD void invert_sym3_t(T xx, T yy, T zz, T xy, T xz, T yz, bool floor_det, T* inv) {
#pragma clang fp contract(off)
    T det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    T id = T(1) / (floor_det ? tmax(det, T(1e-24f)) : det);
    inv[0] = (yy * zz - yz * yz) * id; inv[1] = (xz * yz - xy * zz) * id; inv[2] = (xy * yz - xz * yy) * id;
    inv[3] = inv[1]; inv[4] = (xx * zz - xz * xz) * id; inv[5] = (xy * xz - xx * yz) * id;
    inv[6] = inv[2]; inv[7] = inv[5]; inv[8] = (xx * yy - xy * xy) * id;
}

}  // namespace pr
