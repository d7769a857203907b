"""ctypes bindings for the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only -- to check the GPU path's results or as the timed CPU
baseline, never as the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PLANES_ORACLE_SO: another build of the oracle (the planted-mistake tests load mutated builds)
ORACLE_SO = os.environ.get("PLANES_ORACLE_SO") or os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libplanes_ref.so")

N_GENES = 35
N_SF = 26
F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
IP = C.POINTER(C.c_int)

SCEN_FIELDS = ["speed", "pitch", "yaw", "roll", "height", "roll_rate", "pitch_rate", "yaw_rate",
               "air_rho", "skin_cf", "gust", "base_x", "base_y", "force_pitch", "force_yaw",
               "grip_x", "grip_y", "grip_z", "material_scale", "infill_ratio", "print_bias_x",
               "print_bias_y", "print_bias_z", "warp_pitch", "warp_roll", "base_z"]


def _load():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
    lib = C.CDLL(ORACLE_SO)
    for fx, P in (("f32", F32P), ("f64", F64P)):
        getattr(lib, f"or_mesh_{fx}").argtypes = [F32P, C.c_int, C.c_int, P, C.c_int, I32P, C.c_int, IP, IP]
        getattr(lib, f"or_mass_{fx}").argtypes = [F32P, C.c_int, C.c_int, P, IP, IP, IP]
        getattr(lib, f"or_panels_{fx}").argtypes = [F32P, C.c_int, C.c_int, P, P, P, C.c_int]
        getattr(lib, f"or_printed_{fx}").argtypes = [P, F32P, P, P]
        getattr(lib, f"or_launch_{fx}").argtypes = [F32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                                    C.c_float, C.c_int, C.c_float, P, P, P, IP, C.c_void_p]
        getattr(lib, f"or_traj_{fx}").argtypes = [F32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                                  C.c_float, C.c_int, C.c_int, P, C.c_void_p, C.c_int]
        getattr(lib, f"or_traj_{fx}").restype = C.c_int
        getattr(lib, f"or_fly_{fx}").argtypes = [F32P, I32P, I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                                 C.c_float, C.c_float, C.c_int, C.c_int, P, I32P]
        getattr(lib, f"or_fly_{fx}").restype = C.c_double
        getattr(lib, f"or_fly_res_{fx}").argtypes = [F32P, I32P, I32P, I32P, C.c_int, C.c_int, C.c_float, C.c_float,
                                                     C.c_float, C.c_int, C.c_int, P, I32P]
        getattr(lib, f"or_fly_res_{fx}").restype = C.c_double
    lib.or_aero_f64.argtypes = [C.c_int, F64P, F64P, F64P] + [F64P] * 5 + [C.c_double] * 4 + [C.c_int, F64P, F64P]
    lib.or_wind_f32.argtypes = [C.c_float] * 4 + [C.c_int, F32P, F32P]
    lib.or_wind_f64.argtypes = [C.c_double] * 4 + [C.c_int, F32P, F64P]
    lib.or_set_libm_float.argtypes = [C.c_int]
    for fn in ("or_sinf", "or_cosf", "or_tanf"):
        getattr(lib, fn).argtypes = [C.c_float]
        getattr(lib, fn).restype = C.c_float
    lib.or_hash32.argtypes = [C.c_uint32]
    lib.or_hash32.restype = C.c_uint32
    lib.or_lane_u01.argtypes = [C.c_int, C.c_int]
    lib.or_lane_u01.restype = C.c_float
    lib.or_lane_normal.argtypes = [C.c_int, C.c_int]
    lib.or_lane_normal.restype = C.c_float
    lib.or_phase.argtypes = [C.c_int]
    lib.or_scenario.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, F32P]
    lib.or_scenario_batch.argtypes = [I32P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, F32P]
    lib.or_bounds.argtypes = [F32P, F32P]
    lib.or_robust_scores.argtypes = [F32P, C.c_int, C.c_int, F32P]
    lib.or_ga_rank.argtypes = [F32P, C.c_int, I32P]
    lib.or_ga2_new.argtypes = [C.c_uint]
    lib.or_ga2_new.restype = C.c_void_p
    lib.or_ga2_free.argtypes = [C.c_void_p]
    lib.or_ga2_random.argtypes = [C.c_void_p, F32P, C.c_int]
    lib.or_ga2_breed.argtypes = [C.c_void_p, F32P, F32P, I32P, C.c_int, C.c_int, C.c_int, F32P]
    lib.or_ga_sanitize.argtypes = [F32P, C.c_int]
    lib.or_percentile.argtypes = [F32P, C.c_int, C.c_float]
    lib.or_percentile.restype = C.c_float
    lib.or_record_f32.argtypes = [F32P, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, F32P]
    lib.or_record_f32.restype = C.c_float
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


class libm_float:
    """Context: the oracle's float sin/cos/tan call libm's float functions (reference host build)."""

    def __enter__(self):
        lib().or_set_libm_float(1)

    def __exit__(self, *a):
        lib().or_set_libm_float(0)


def _dt(prec):
    return np.float32 if prec == 32 else np.float64


def bounds():
    lo = np.zeros(N_GENES, np.float32)
    hi = np.zeros(N_GENES, np.float32)
    lib().or_bounds(lo, hi)
    return lo, hi


def mesh(g, prec=32, ch=0, sp=0, cap_v=8192, cap_t=16384):
    g = np.ascontiguousarray(g, np.float32)
    V = np.zeros(cap_v * 3, _dt(prec))
    T = np.zeros(cap_t * 3, np.int32)
    nv, nt = C.c_int(), C.c_int()
    rc = getattr(lib(), f"or_mesh_f{prec}")(g, ch, sp, V, cap_v, T, cap_t, C.byref(nv), C.byref(nt))
    assert rc == 0
    return V[:3 * nv.value].reshape(-1, 3), T[:3 * nt.value].reshape(-1, 3)


def mass(g, prec=32, ch=0, sp=0):
    """-> (out20 = mass, invmass, com3, I6, invI9; printable; nt; nv)"""
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros(20, _dt(prec))
    pr, nt, nv = C.c_int(), C.c_int(), C.c_int()
    getattr(lib(), f"or_mass_f{prec}")(g, ch, sp, out, C.byref(pr), C.byref(nt), C.byref(nv))
    return out, pr.value, nt.value, nv.value


def panels(g, prec=32, ch=0, sp=0, cap=16384):
    g = np.ascontiguousarray(g, np.float32)
    c = np.zeros(cap * 3, _dt(prec))
    A = np.zeros(cap, _dt(prec))
    n = np.zeros(cap * 3, _dt(prec))
    nt = getattr(lib(), f"or_panels_f{prec}")(g, ch, sp, c, A, n, cap)
    assert nt >= 0
    return c[:3 * nt].reshape(-1, 3), A[:nt], n[:3 * nt].reshape(-1, 3)


def scenario(seed, preset=0, gust_rms_max=-1.0, warp_max=-1.0, wind_sd=-1.0):
    out = np.zeros(N_SF, np.float32)
    lib().or_scenario(int(seed), preset, gust_rms_max, warp_max, wind_sd, out)
    return out


def scenarios(seeds, preset=0, gust_rms_max=-1.0, warp_max=-1.0, wind_sd=-1.0):
    seeds = np.ascontiguousarray(seeds, np.int32)
    out = np.zeros(len(seeds) * N_SF, np.float32)
    lib().or_scenario_batch(seeds, len(seeds), preset, gust_rms_max, warp_max, wind_sd, out)
    return out.reshape(-1, N_SF)


def printed(mp20, sc26, prec=32):
    mp20 = np.ascontiguousarray(mp20, _dt(prec))
    sc26 = np.ascontiguousarray(sc26, np.float32)
    out = np.zeros(20, _dt(prec))
    sh = np.zeros(3, _dt(prec))
    getattr(lib(), f"or_printed_f{prec}")(mp20, sc26, out, sh)
    return out, sh


def wind(p, t, seed, sc26, prec=32):
    out = np.zeros(3, _dt(prec))
    sc26 = np.ascontiguousarray(sc26, np.float32)
    getattr(lib(), f"or_wind_f{prec}")(float(p[0]), float(p[1]), float(p[2]), float(t), int(seed), sc26, out)
    return out


def launch(g, seed, prec=32, preset=0, gust=-1.0, warp=-1.0, wsd=-1.0, plane_index=0, spread=0.0, ch=0, sp=0,
           scenario_override=None):
    """Launch state (pos3 q4 vel3 wb3), printed mass properties, CG shift and triangle count of
    one flight; scenario_override: a 26-field descriptor to launch under instead of seed's."""
    g = np.ascontiguousarray(g, np.float32)
    st = np.zeros(13, _dt(prec))
    mp = np.zeros(20, _dt(prec))
    sh = np.zeros(3, _dt(prec))
    nt = C.c_int()
    ov = None if scenario_override is None else np.ascontiguousarray(scenario_override, np.float32)
    getattr(lib(), f"or_launch_f{prec}")(g, ch, sp, int(seed), preset, gust, warp, wsd, plane_index, spread,
                                         st, mp, sh, C.byref(nt), None if ov is None else ov.ctypes.data)
    return st, mp, sh, nt.value


def mt19937(seed, n):
    """The first n raw outputs of std::mt19937(seed), the GA's engine."""
    out = np.zeros(n, np.uint32)
    lib().or_mt19937(C.c_uint(seed), n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def ga_breed(gs, acc, idx, gen, gens, seed=42, state=None):
    """One breeding pass (sim.cu:472-490) with the reference's persistent distributions; state:
    a handle from lib().or_ga2_new(seed) to continue a stream, else a fresh mt19937(seed)."""
    h = state if state is not None else lib().or_ga2_new(seed)
    gs = np.ascontiguousarray(gs, np.float32)
    pop = len(gs)
    nxt = np.zeros_like(gs)
    lib().or_ga2_breed(h, gs.reshape(-1), np.ascontiguousarray(acc, np.float32), np.ascontiguousarray(idx, np.int32),
                       pop, gen, gens, nxt.reshape(-1))
    if state is None:
        lib().or_ga2_free(h)
    return nxt


def trajectory(g, seed, nsteps, prec=32, preset=0, gust=-1.0, warp=-1.0, wsd=-1.0, max_steps=0, ch=0, sp=0,
               scenario_override=None, flip_normals=False):
    """States (pos3 q4 vel3 wb3) after each of the first nsteps steps."""
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros(nsteps * 13, _dt(prec))
    ov = None
    if scenario_override is not None:
        ov = np.ascontiguousarray(scenario_override, np.float32)
    n = getattr(lib(), f"or_traj_f{prec}")(g, ch, sp, int(seed), preset, gust, warp, wsd, max_steps, nsteps, out,
                                           None if ov is None else ov.ctypes.data, int(flip_normals))
    return out[:13 * n].reshape(-1, 13)


def fly(genomes, gi, seeds, prec=32, preset=0, gust=-1.0, warp=-1.0, wsd=-1.0, max_steps=0, nthreads=0,
        ch=0, sp=0):
    """Batched flights -> (fitness, integrated steps, wall seconds)."""
    genomes = np.ascontiguousarray(genomes, np.float32).reshape(-1)
    gi = np.ascontiguousarray(gi, np.int32)
    seeds = np.ascontiguousarray(seeds, np.int32)
    n = len(gi)
    fit = np.zeros(n, _dt(prec))
    steps = np.zeros(n, np.int32)
    if nthreads <= 0:
        nthreads = os.cpu_count() or 1
    wall = getattr(lib(), f"or_fly_f{prec}")(genomes, gi, seeds, n, ch, sp, preset, gust, warp, wsd, max_steps,
                                             nthreads, fit, steps)
    return fit, steps, wall


def fly_res(genomes, res, gi, seeds, prec=32, preset=0, gust=-1.0, warp=-1.0, wsd=-1.0, max_steps=0, nthreads=0):
    """Batched flights with a mesh resolution per genome (res [n_genomes, 2] = ch, sp)."""
    genomes = np.ascontiguousarray(genomes, np.float32).reshape(-1)
    res = np.ascontiguousarray(res, np.int32).reshape(-1)
    gi = np.ascontiguousarray(gi, np.int32)
    seeds = np.ascontiguousarray(seeds, np.int32)
    n = len(gi)
    fit = np.zeros(n, _dt(prec))
    steps = np.zeros(n, np.int32)
    if nthreads <= 0:
        nthreads = os.cpu_count() or 1
    wall = getattr(lib(), f"or_fly_res_f{prec}")(genomes, res, gi, seeds, n, preset, gust, warp, wsd, max_steps,
                                                 nthreads, fit, steps)
    return fit, steps, wall


def robust_scores(samples, pop, nseeds):
    samples = np.ascontiguousarray(samples, np.float32).reshape(-1)
    out = np.zeros(pop, np.float32)
    lib().or_robust_scores(samples, pop, nseeds, out)
    return out


def ga_rank(acc):
    acc = np.ascontiguousarray(acc, np.float32)
    idx = np.zeros(len(acc), np.int32)
    lib().or_ga_rank(acc, len(acc), idx)
    return idx


def random_genomes(n, rng_seed):
    """Genomes drawn like the reference's grand() from mt19937(rng_seed) (sim.cu:318-321)."""
    h = lib().or_ga2_new(rng_seed)
    out = np.zeros(n * N_GENES, np.float32)
    lib().or_ga2_random(h, out, n)
    lib().or_ga2_free(h)
    return out.reshape(n, N_GENES)


def record(g, seed, plane_index=0, spread=0.0, max_steps=0, rec_every=1, rec_cap=0):
    g = np.ascontiguousarray(g, np.float32)
    rec = np.zeros(max(rec_cap, 1) * 8, np.float32)
    fit = lib().or_record_f32(g, int(seed), plane_index, spread, max_steps, rec_every, rec_cap, rec)
    return fit, rec[:rec_cap * 8].reshape(-1, 8)


def aero(c, A, n, shift, q, vel, wb, wind, rho, cf, wp=0.0, wr=0.0, flip=False):
    """World-frame aero force and torque (FP64) of explicit panels."""
    c = np.ascontiguousarray(c, np.float64).reshape(-1)
    n = np.ascontiguousarray(n, np.float64).reshape(-1)
    A = np.ascontiguousarray(A, np.float64)
    F = np.zeros(3)
    T = np.zeros(3)
    a = [np.ascontiguousarray(x, np.float64) for x in (shift, q, vel, wb, wind)]
    lib().or_aero_f64(len(A), c, A, n, *a, rho, cf, wp, wr, int(flip), F, T)
    return F, T
