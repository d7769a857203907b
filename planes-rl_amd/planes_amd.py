"""Python binding of libplanes_hip.so (the MI355X fitness-evaluation library).

Mirrors the reference's GPU call surface (src/sim.h:9-13: panelbuf_alloc, gpu_prepare,
gpu_sim) plus the batched entry points the C ABI adds (include/planes_hip.h).  Every
call goes through the HIP library; there is no CPU path here.  Importing this module
without a built library raises immediately.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplanes_hip.so")
N_GENES = 35
N_SF = 26

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
VP = C.c_void_p


class SimParams(C.Structure):
    """pr_sim_params: scenario preset + flight settings (reference preset by default)."""
    _fields_ = [("preset", C.c_int), ("gust_rms_max", C.c_float), ("warp_max", C.c_float),
                ("wind_sd", C.c_float), ("max_steps", C.c_int), ("chunk", C.c_int), ("spread", C.c_float)]

    @classmethod
    def reference(cls, max_steps=7200, chunk=600, spread=0.0):
        return cls(0, 4.0, 0.0174532925, 2.0, max_steps, chunk, spread)

    @classmethod
    def config(cls, gust_rms_max=4.0, warp_max_deg=1.0, wind_sd=2.0, max_steps=7200, chunk=600):
        return cls(1, gust_rms_max, np.float32(np.deg2rad(warp_max_deg)), wind_sd, max_steps, chunk, 0.0)


NS_MAX = 16
NC_MAX = 64


class Stats(C.Structure):
    _fields_ = [("flight_steps", C.c_ulonglong), ("panel_steps", C.c_ulonglong), ("fly_ms", C.c_double),
                ("fly_launches", C.c_int), ("prep_ms", C.c_double), ("n_slots", C.c_int),
                ("s_nt_max", C.c_int * NS_MAX), ("s_L", C.c_int * NS_MAX), ("s_P", C.c_int * NS_MAX),
                ("s_launches", C.c_int * NS_MAX), ("s_flight_steps", C.c_ulonglong * NS_MAX),
                ("s_panel_steps", C.c_ulonglong * NS_MAX), ("s_fly_ms", C.c_double * NS_MAX),
                ("s_fused", C.c_int * NS_MAX), ("fused_ms", C.c_double), ("fused_launches", C.c_int),
                ("n_chunks", C.c_int), ("c_flights", C.c_int * NC_MAX), ("c_flight_steps", C.c_ulonglong * NC_MAX),
                ("c_panel_steps", C.c_ulonglong * NC_MAX), ("c_ms", C.c_double * NC_MAX),
                ("split_flight_steps", C.c_ulonglong), ("split_panel_steps", C.c_ulonglong)]

    def as_dict(self):
        d = {k: getattr(self, k) for k in ("flight_steps", "panel_steps", "fly_ms", "fly_launches", "prep_ms",
                                           "fused_ms", "fused_launches", "split_flight_steps",
                                           "split_panel_steps")}
        d["buckets"] = [{"nt_max": self.s_nt_max[i], "L": self.s_L[i], "P": self.s_P[i], "tail": i % 2 == 1,
                         "fused": bool(self.s_fused[i]),
                         "launches": self.s_launches[i], "flight_steps": self.s_flight_steps[i],
                         "panel_steps": self.s_panel_steps[i], "fly_ms": self.s_fly_ms[i]}
                        for i in range(self.n_slots) if self.s_L[i] > 0]
        d["chunks"] = [{"flights": self.c_flights[k], "flight_steps": self.c_flight_steps[k],
                        "panel_steps": self.c_panel_steps[k], "ms": self.c_ms[k]} for k in range(self.n_chunks)]
        return d


# pr_design_stats / pr_audit columns (evaluate_design's summary, sim.cu:291-309)
AUDIT_FIELDS = ("min", "p05", "p10", "p25", "median", "mean", "sd", "p75", "p90", "p95", "max", "cvar10", "failures")


def design_stats(samples):
    """evaluate_design's robustness statistics of one design's distances (host C++)."""
    x = np.ascontiguousarray(samples, np.float32)
    out = np.zeros(len(AUDIT_FIELDS), np.float32)
    _ck(LIB.pr_design_stats(x, len(x), out), "pr_design_stats")
    return dict(zip(AUDIT_FIELDS, out.tolist()))


class PlanesError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise PlanesError(f"libplanes_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    lib.pr_last_error.restype = C.c_char_p
    lib.pr_create.argtypes = [C.POINTER(VP), C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.pr_destroy.argtypes = [VP]
    lib.pr_destroy.restype = None
    lib.pr_prepare.argtypes = [VP, VP, VP, C.c_int]
    lib.pr_scenarios.argtypes = [VP, VP, C.c_int, C.POINTER(SimParams)]
    lib.pr_fly.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.POINTER(SimParams), VP, VP]
    lib.pr_init_flights.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.POINTER(SimParams), VP, VP]
    lib.pr_scores.argtypes = [VP, VP, C.c_int, C.c_int, VP]
    lib.pr_stats_get.argtypes = [VP, C.POINTER(Stats)]
    lib.pr_set_buckets.argtypes = [VP, C.c_int, I32P, I32P, I32P]
    lib.pr_fly_variants.argtypes = [C.POINTER(C.c_int), I32P, I32P, C.c_int]
    lib.pr_read_prepared.argtypes = [VP, C.c_int, VP, VP, VP, VP]
    lib.pr_read_panels.argtypes = [VP, C.c_longlong, C.c_int, F32P]
    lib.pr_read_scenarios.argtypes = [VP, C.c_int, F32P]
    lib.pr_read_genome_order.argtypes = [VP, C.c_int, I32P]
    lib.pr_read_states.argtypes = [VP, C.c_int, C.c_int, F32P]
    lib.pr_dev_alloc.argtypes = [C.POINTER(VP), C.c_size_t]
    lib.pr_dev_free.argtypes = [VP]
    lib.pr_h2d.argtypes = [VP, VP, C.c_size_t]
    lib.pr_d2h.argtypes = [VP, VP, C.c_size_t]
    lib.pr_device_name.argtypes = [C.c_char_p, C.c_int]
    lib.pr_fly_record.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.POINTER(SimParams), VP, VP, VP, C.c_int, C.c_int]
    lib.pr_read_extent.argtypes = [VP, C.c_int, F32P]
    lib.pr_set_graph.argtypes = [VP, C.c_int]
    lib.pr_set_tail_mapping.argtypes = [VP, C.c_int]
    lib.pr_set_fused.argtypes = [VP, C.c_int]
    lib.pr_set_list_order.argtypes = [VP, C.c_int]
    lib.pr_set_split.argtypes = [VP, C.c_int]
    lib.pr_sim.argtypes = [VP, C.c_int, C.c_int, VP, VP, C.c_int, C.c_int, C.c_float, C.c_int, C.POINTER(SimParams)]
    lib.pr_design_stats.argtypes = [F32P, C.c_int, F32P]
    lib.pr_audit.argtypes = [VP, VP, VP, C.c_int, C.c_int, C.c_int, C.POINTER(SimParams), VP, VP, F32P, C.c_int]
    lib.pr_gen_context.argtypes = [VP, C.c_int]
    lib.pr_gen_evaluate_async.argtypes = [VP, VP, VP, C.c_int, I32P, C.c_int, C.c_int, C.POINTER(SimParams), F32P,
                                          I32P, VP]
    lib.pr_gen_collect.argtypes = [VP, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    lib.pr_gen_read_samples.argtypes = [VP, C.c_int, C.c_int, VP]
    lib.pr_eval_noalloc.argtypes = [VP, C.c_int, VP, C.c_int, C.POINTER(SimParams), VP, VP, VP]
    lib.pr_gen_context.restype = VP
    lib.pr_set_prepare_precision.argtypes = [VP, C.c_int]
    lib.pr_mlp_last_error.restype = C.c_char_p
    lib.pr_pool_last_error.restype = C.c_char_p
    lib.pr_pool_create.argtypes = [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint]
    lib.pr_pool_destroy.argtypes = [VP]
    lib.pr_pool_destroy.restype = None
    lib.pr_pool_set_elite.argtypes = [VP, F32P, C.c_int]
    lib.pr_pool_get_elite.argtypes = [VP, F32P, C.c_int]
    lib.pr_pool_generation.argtypes = [VP, C.c_int, C.POINTER(SimParams), VP, C.c_int, C.c_float,
                                       np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
    lib.pr_pool_train.argtypes = [VP, VP, C.c_int, C.c_float, C.POINTER(C.c_float)]
    lib.pr_pool_set_collect.argtypes = [VP, C.c_int]
    lib.pr_pool_set_explore.argtypes = [VP, C.c_float]
    lib.pr_pool_set_parent_delta.argtypes = [VP, C.c_int]
    lib.pr_pool_set_key_weights.argtypes = [VP, C.c_float, C.c_float, C.c_float]
    lib.pr_pool_set_selection.argtypes = [VP, C.c_int]
    lib.pr_pool_set_target_norm.argtypes = [VP, C.c_int]
    lib.pr_pool_add_training.argtypes = [VP, VP, F32P, C.c_int]
    lib.pr_pool_champion.argtypes = [VP, F32P, C.POINTER(C.c_float)]
    lib.pr_pool_last.argtypes = [VP, VP, VP, VP, VP, C.POINTER(C.c_int)]
    lib.pr_mlp_param_count.restype = C.c_int
    lib.pr_mlp_create.argtypes = [C.POINTER(VP), C.c_int, C.c_uint]
    lib.pr_mlp_destroy.argtypes = [VP]
    lib.pr_mlp_destroy.restype = None
    lib.pr_mlp_get_params.argtypes = [VP, F32P, C.c_int]
    lib.pr_mlp_set_params.argtypes = [VP, F32P, C.c_int]
    lib.pr_mlp_predict.argtypes = [VP, VP, C.c_int, VP]
    lib.pr_mlp_loss_grad.argtypes = [VP, VP, VP, C.c_int, C.POINTER(C.c_float), F32P]
    lib.pr_mlp_train.argtypes = [VP, VP, VP, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]
    lib.pr_gene_bounds.argtypes = [F32P, F32P]
    lib.pr_gene_bounds.restype = None
    lib.pr_ga_new.argtypes = [C.c_uint]
    lib.pr_ga_new.restype = VP
    lib.pr_ga_free.argtypes = [VP]
    lib.pr_ga_free.restype = None
    lib.pr_ga_random.argtypes = [VP, F32P, C.c_int]
    lib.pr_ga_random.restype = None
    lib.pr_ga_sanitize.argtypes = [F32P, C.c_int]
    lib.pr_ga_sanitize.restype = None
    lib.pr_ga_rank.argtypes = [F32P, C.c_int, I32P]
    lib.pr_ga_rank.restype = None
    lib.pr_ga_breed.argtypes = [VP, F32P, F32P, I32P, C.c_int, C.c_int, C.c_int, F32P]
    lib.pr_ga_breed.restype = None
    lib.pr_robust_score_host.argtypes = [F32P, C.c_int]
    lib.pr_robust_score_host.restype = C.c_float
    lib.pr_gen_create.argtypes = [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int]
    lib.pr_gen_destroy.argtypes = [VP]
    lib.pr_gen_destroy.restype = None
    lib.pr_gen_evaluate.argtypes = [VP, VP, VP, C.c_int, I32P, C.c_int, C.c_int, C.POINTER(SimParams), F32P, I32P,
                                    VP, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    lib.pr_gen_stats.argtypes = [VP, C.POINTER(Stats), C.POINTER(Stats)]
    return lib


LIB = _load()


def _ck(rc, what):
    if rc != 0:
        raise PlanesError(f"{what} failed ({rc}): {LIB.pr_last_error().decode()}")


def device_name():
    buf = C.create_string_buffer(256)
    _ck(LIB.pr_device_name(buf, 256), "pr_device_name")
    return buf.value.decode()


def fly_variants():
    n = C.c_int()
    L = np.zeros(64, np.int32)
    P = np.zeros(64, np.int32)
    LIB.pr_fly_variants(C.byref(n), L, P, 64)
    return list(zip(L[:n.value].tolist(), P[:n.value].tolist()))


class DeviceBuffer:
    """Owned HBM allocation."""

    def __init__(self, nbytes):
        self.ptr = VP()
        self.nbytes = int(max(nbytes, 4))
        _ck(LIB.pr_dev_alloc(C.byref(self.ptr), self.nbytes), "pr_dev_alloc")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, (arr.nbytes, self.nbytes)
        _ck(LIB.pr_h2d(self.ptr, arr.ctypes.data, arr.nbytes), "pr_h2d")

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        _ck(LIB.pr_d2h(out.ctypes.data, self.ptr, out.nbytes), "pr_d2h")
        return out

    def free(self):
        if self.ptr:
            LIB.pr_dev_free(self.ptr)
            self.ptr = VP()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Planes:
    """One evaluation context (the reference's PanelBuf, sim.cu:15-21, plus the flight
    work lists).  Genomes, seeds and outputs live in buffers owned by the context."""

    def __init__(self, max_genomes, max_scenarios, max_flights, max_ch=0, max_sp=0, max_panels=0):
        self.h = VP()
        _ck(LIB.pr_create(C.byref(self.h), max_genomes, max_panels, max_scenarios, max_flights, max_ch, max_sp),
            "pr_create")
        self.max_genomes, self.max_scenarios, self.max_flights = max_genomes, max_scenarios, max_flights
        self.d_genomes = DeviceBuffer(4 * N_GENES * max_genomes)
        self.d_res = DeviceBuffer(8 * max_genomes)
        self.d_seeds = DeviceBuffer(4 * max_scenarios)
        self.d_fit = DeviceBuffer(4 * max_flights)
        self.d_steps = DeviceBuffer(4 * max_flights)
        self.d_scores = DeviceBuffer(4 * max_flights)
        self.n_genomes = 0
        self.n_scen = 0

    def close(self):
        if self.h:
            LIB.pr_destroy(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------------- upload + prepare
    def upload_genomes(self, genomes, res=None):
        g = np.ascontiguousarray(genomes, np.float32).reshape(-1, N_GENES)
        self.d_genomes.upload(g)
        if res is not None:
            self.d_res.upload(np.ascontiguousarray(res, np.int32).reshape(-1, 2))
        self.n_genomes = len(g)
        self._has_res = res is not None
        return len(g)

    def prepare(self, genomes=None, res=None, n=None):
        """gpu_prepare (sim.h:11): mesh + mass + panels for the uploaded genomes."""
        if genomes is not None:
            n = self.upload_genomes(genomes, res)
        n = self.n_genomes if n is None else n
        _ck(LIB.pr_prepare(self.h, self.d_genomes.ptr, self.d_res.ptr if self._has_res else None, n), "pr_prepare")
        self.n_prepared = n

    def upload_seeds(self, seeds):
        s = np.ascontiguousarray(seeds, np.int32)
        self.d_seeds.upload(s)
        self.n_scen = len(s)
        return len(s)

    def scenarios(self, seeds=None, params=None, n=None):
        if seeds is not None:
            n = self.upload_seeds(seeds)
        n = self.n_scen if n is None else n
        params = params or SimParams.reference()
        _ck(LIB.pr_scenarios(self.h, self.d_seeds.ptr, n, C.byref(params)), "pr_scenarios")

    # --------------------------------------------------------------- flights
    def fly_device(self, layout, n_genomes, n_scen, params):
        """Flies in place: results stay in d_fit / d_steps (HBM)."""
        _ck(LIB.pr_fly(self.h, layout, n_genomes, n_scen, C.byref(params), self.d_fit.ptr, self.d_steps.ptr),
            "pr_fly")
        return n_genomes * n_scen if layout == 0 else n_genomes

    def fly(self, layout, n_genomes, n_scen, params=None):
        params = params or SimParams.reference()
        nf = self.fly_device(layout, n_genomes, n_scen, params)
        return self.d_fit.download(np.float32, nf), self.d_steps.download(np.int32, nf)

    def sim(self, pop, seed, params=None, spread=0.0, vary=0, rec_every=1, rec_cap=0):
        """gpu_sim (sim.h:12) on the prepared genomes: fitness[pop] and, with rec_cap > 0,
        poses [pop, rec_cap, 8] (pos, q, alive)."""
        params = params or SimParams.reference()
        rec = DeviceBuffer(4 * 8 * pop * rec_cap) if rec_cap > 0 else None
        _ck(LIB.pr_sim(self.h, pop, seed, self.d_fit.ptr, rec.ptr if rec else None, rec_every, rec_cap, spread, vary,
                       C.byref(params)), "pr_sim")
        fit = self.d_fit.download(np.float32, pop)
        return (fit, rec.download(np.float32, 8 * pop * rec_cap).reshape(pop, rec_cap, 8)) if rec else fit

    def fly_record(self, layout, n_genomes, n_scen, params, rec_every, rec_cap):
        """gpu_sim with d_rec (sim.cu:117-128): returns fitness, steps, frames[n, rec_cap, 8]."""
        nf = n_genomes * n_scen if layout == 0 else n_genomes
        rec = DeviceBuffer(4 * 8 * max(1, nf * rec_cap))
        _ck(LIB.pr_fly_record(self.h, layout, n_genomes, n_scen, C.byref(params), self.d_fit.ptr, self.d_steps.ptr,
                              rec.ptr, rec_every, rec_cap), "pr_fly_record")
        frames = rec.download(np.float32, nf * rec_cap * 8).reshape(nf, rec_cap, 8)
        return self.d_fit.download(np.float32, nf), self.d_steps.download(np.int32, nf), frames

    def read_extent(self, n=None):
        n = self.n_prepared if n is None else n
        out = np.zeros(3 * n, np.float32)
        _ck(LIB.pr_read_extent(self.h, n, out), "pr_read_extent")
        return out.reshape(n, 3)

    def init_flights(self, layout, n_genomes, n_scen, params=None):
        params = params or SimParams.reference()
        _ck(LIB.pr_init_flights(self.h, layout, n_genomes, n_scen, C.byref(params), self.d_fit.ptr,
                                self.d_steps.ptr), "pr_init_flights")
        nf = n_genomes * n_scen if layout == 0 else n_genomes
        return self.read_states(0, nf)

    def evaluate(self, genomes, seeds, params=None, layout=0, res=None):  # noqa: D417
        """Batched evaluation: every genome under every seed (layout 0) or genome i under
        seed i (layout 1).  Returns fitness[n_genomes, n_seeds] (layout 0) or [n]."""
        params = params or SimParams.reference()
        self.prepare(genomes, res)
        self.scenarios(seeds, params)
        fit, steps = self.fly(layout, self.n_genomes, self.n_scen, params)
        if layout == 0:
            return fit.reshape(self.n_genomes, self.n_scen), steps.reshape(self.n_genomes, self.n_scen)
        return fit, steps

    def audit(self, genomes, n_scen, base_seed, params, res=None, nthreads=0, upload=True):
        """Batched robustness audit (evaluate_design for every genome at once): all genomes
        fly the same n_scen scenarios, seed base_seed + s*7919.  Returns (stats[n, 13] in
        AUDIT_FIELDS order, distances[n, n_scen]); distances stay in d_fit too."""
        if upload:
            self.upload_genomes(genomes, res)
        n = self.n_genomes
        st = np.zeros((n, len(AUDIT_FIELDS)), np.float32)
        _ck(LIB.pr_audit(self.h, self.d_genomes.ptr, self.d_res.ptr if self._has_res else None, n, n_scen,
                         base_seed, C.byref(params), self.d_fit.ptr, self.d_steps.ptr, st, nthreads), "pr_audit")
        self.n_scen = n_scen
        return st, self.d_fit.download(np.float32, n * n_scen).reshape(n, n_scen)

    def scores(self, samples, nseeds):
        s = np.ascontiguousarray(samples, np.float32).reshape(-1)
        pop = len(s) // nseeds
        buf = DeviceBuffer(s.nbytes)
        buf.upload(s)
        _ck(LIB.pr_scores(self.h, buf.ptr, pop, nseeds, self.d_scores.ptr), "pr_scores")
        return self.d_scores.download(np.float32, pop)

    def set_prepare_precision(self, bits):
        """32: bit-exact reference float meshing/mass; 64: FP64 prepare rounded once."""
        _ck(LIB.pr_set_prepare_precision(self.h, bits), "pr_set_prepare_precision")

    def set_tail_mapping(self, on=True):
        """False/0: one mapping per bucket; True/1: every bucket's last flights of a chunk fly
        one per wave; 2 (default): in the fused kernel only a bucket whose later-drained
        buckets hold fewer flights than there are waves does."""
        _ck(LIB.pr_set_tail_mapping(self.h, int(on)), "pr_set_tail_mapping")

    def set_fused(self, on=True):
        """One fused flight launch per chunk for all buckets (default) or one per bucket."""
        _ck(LIB.pr_set_fused(self.h, int(on)), "pr_set_fused")

    def set_list_order(self, sorted_by_size=True):
        """Chunk 0's work lists in a stable order of mesh size (default) or flight order."""
        _ck(LIB.pr_set_list_order(self.h, int(sorted_by_size)), "pr_set_list_order")

    def set_split(self, on=True):
        """Small chunks flown one flight per 2- or 4-wave workgroup (default) or like the others."""
        _ck(LIB.pr_set_split(self.h, int(on)), "pr_set_split")

    def set_graph(self, on=True):
        """Capture pr_fly's launch sequence as a hipGraph and replay it."""
        _ck(LIB.pr_set_graph(self.h, int(on)), "pr_set_graph")

    def set_buckets(self, nt_max, L, P):
        _ck(LIB.pr_set_buckets(self.h, len(nt_max), np.asarray(nt_max, np.int32), np.asarray(L, np.int32),
                               np.asarray(P, np.int32)), "pr_set_buckets")

    def stats(self):
        st = Stats()
        _ck(LIB.pr_stats_get(self.h, C.byref(st)), "pr_stats_get")
        return st.as_dict()

    # --------------------------------------------------------------- readback
    def read_prepared(self, n=None):
        n = self.n_prepared if n is None else n
        mass = np.zeros((n, 20), np.float32)
        nt = np.zeros(n, np.int32)
        nv = np.zeros(n, np.int32)
        off = np.zeros(n + 1, np.int32)
        _ck(LIB.pr_read_prepared(self.h, n, mass.ctypes.data, nt.ctypes.data, nv.ctypes.data, off.ctypes.data),
            "pr_read_prepared")
        return mass, nt, nv, off

    def read_panels(self, first, count):
        out = np.zeros(count * 7, np.float32)
        _ck(LIB.pr_read_panels(self.h, first, count, out), "pr_read_panels")
        return out.reshape(count, 7)

    def read_scenarios(self, n=None):
        n = self.n_scen if n is None else n
        out = np.zeros(n * N_SF, np.float32)
        _ck(LIB.pr_read_scenarios(self.h, n, out), "pr_read_scenarios")
        return out.reshape(n, N_SF)

    def read_genome_order(self, n):
        """Genome order of the last flight batch's chunk-0 work lists (sorted by mesh size)."""
        out = np.zeros(n, np.int32)
        _ck(LIB.pr_read_genome_order(self.h, n, out), "pr_read_genome_order")
        return out

    def read_states(self, first, n):
        """FlightState rows: pos3 q4 vel3 wb3 (float), step, done (int), fitness."""
        out = np.zeros(n * 16, np.float32)
        _ck(LIB.pr_read_states(self.h, first, n, out), "pr_read_states")
        out = out.reshape(n, 16)
        return {"state": out[:, :13].copy(), "step": out[:, 13].view(np.int32).copy(),
                "done": out[:, 14].view(np.int32).copy(), "fitness": out[:, 15].copy()}


# ------------------------------------------------------------------ GA host side (C ABI)
def gene_bounds():
    lo = np.zeros(N_GENES, np.float32)
    hi = np.zeros(N_GENES, np.float32)
    LIB.pr_gene_bounds(lo, hi)
    return lo, hi


class GA:
    """The GA's random stream (std::mt19937 + its distributions) and operators."""

    def __init__(self, seed=42):
        self.h = LIB.pr_ga_new(seed)

    def random(self, n):
        out = np.zeros(n * N_GENES, np.float32)
        LIB.pr_ga_random(self.h, out, n)
        return out.reshape(n, N_GENES)

    def breed(self, gs, acc, idx, gen, gens):
        gs = np.ascontiguousarray(gs, np.float32).reshape(-1)
        pop = len(acc)
        nxt = np.zeros(pop * N_GENES, np.float32)
        LIB.pr_ga_breed(self.h, gs, np.ascontiguousarray(acc, np.float32), np.ascontiguousarray(idx, np.int32), pop,
                        gen, gens, nxt)
        return nxt.reshape(pop, N_GENES)

    def __del__(self):
        try:
            LIB.pr_ga_free(self.h)
        except Exception:
            pass


def ga_rank(acc):
    acc = np.ascontiguousarray(acc, np.float32)
    idx = np.zeros(len(acc), np.int32)
    LIB.pr_ga_rank(acc, len(acc), idx)
    return idx


def ga_sanitize(genomes):
    g = np.ascontiguousarray(genomes, np.float32).copy()
    LIB.pr_ga_sanitize(g.reshape(-1), len(g))
    return g


class Generation:
    """pr_gen: one GA generation's batched evaluation (seeds x genomes, scores, ranking,
    top-8 holdout re-test on 128 scenarios)."""

    def __init__(self, pop, nseeds_max, max_ch=0, max_sp=0):
        self.h = VP()
        _ck(LIB.pr_gen_create(C.byref(self.h), pop, nseeds_max, max_ch, max_sp), "pr_gen_create")
        self.pop = pop
        self.d_genomes = DeviceBuffer(4 * N_GENES * pop)
        self.d_res = DeviceBuffer(8 * pop)
        self.has_res = False

    def upload(self, genomes, res=None):
        self.d_genomes.upload(np.ascontiguousarray(genomes, np.float32))
        self.has_res = res is not None
        if res is not None:
            self.d_res.upload(np.ascontiguousarray(res, np.int32))

    def evaluate(self, pop, seeds, params, holdout_k=8, want_samples=False):
        seeds = np.ascontiguousarray(seeds, np.int32)
        acc = np.zeros(pop, np.float32)
        idx = np.zeros(pop, np.int32)
        samples = np.zeros(pop * len(seeds), np.float32) if want_samples else None
        hold = C.c_float()
        champ = C.c_int()
        _ck(LIB.pr_gen_evaluate(self.h, self.d_genomes.ptr, self.d_res.ptr if self.has_res else None, pop, seeds,
                                len(seeds), holdout_k, C.byref(params), acc, idx,
                                None if samples is None else samples.ctypes.data, C.byref(hold), C.byref(champ)),
            "pr_gen_evaluate")
        out = {"acc": acc, "idx": idx, "holdout": hold.value, "champion": champ.value}
        if samples is not None:
            out["samples"] = samples.reshape(pop, len(seeds))
        return out

    def stats(self):
        a, b = Stats(), Stats()
        _ck(LIB.pr_gen_stats(self.h, C.byref(a), C.byref(b)), "pr_gen_stats")
        return a.as_dict(), b.as_dict()

    def set_fused(self, on=True):
        """Fused flight kernel (default) or one launch per bucket, for every flight context."""
        for which in (0, 1, 2):
            _ck(LIB.pr_set_fused(LIB.pr_gen_context(self.h, which), int(on)), "pr_set_fused")

    def set_list_order(self, sorted_by_size=True):
        """Work-list order (pr_set_list_order) for every flight context of the generation."""
        for which in (0, 1, 2):
            _ck(LIB.pr_set_list_order(LIB.pr_gen_context(self.h, which), int(sorted_by_size)), "pr_set_list_order")

    def set_split(self, on=True, holdout=False):
        """Small-chunk mapping (pr_set_split) of the population context (default on) and of
        the holdout contexts (default off: a pipelined holdout overlaps population flights)."""
        for which, flag in ((0, on), (1, holdout), (2, holdout)):
            _ck(LIB.pr_set_split(LIB.pr_gen_context(self.h, which), int(flag)), "pr_set_split")

    def set_graph(self, on=True):
        for which in (0, 1, 2):
            _ck(LIB.pr_set_graph(LIB.pr_gen_context(self.h, which), int(on)), "pr_set_graph")

    def set_tail_mapping(self, mode):
        """Tail-mapping mode (pr_set_tail_mapping) for every flight context of the generation."""
        for which in (0, 1, 2):
            _ck(LIB.pr_set_tail_mapping(LIB.pr_gen_context(self.h, which), int(mode)), "pr_set_tail_mapping")

    def set_buckets(self, nt_max, L, P):
        """Flight-kernel mapping per triangle-count bucket (e.g. from autotune) for every
        flight context of the generation."""
        for which in (0, 1, 2):
            _ck(LIB.pr_set_buckets(LIB.pr_gen_context(self.h, which), len(nt_max), np.asarray(nt_max, np.int32),
                                   np.asarray(L, np.int32), np.asarray(P, np.int32)), "pr_set_buckets")

    def evaluate_async(self, pop, seeds, params, holdout_k=8, want_samples=False):
        """Population part now (acc, idx[, samples]); the holdout batch keeps flying on its
        own stream until collect()."""
        seeds = np.ascontiguousarray(seeds, np.int32)
        acc = np.zeros(pop, np.float32)
        idx = np.zeros(pop, np.int32)
        samples = np.zeros(pop * len(seeds), np.float32) if want_samples else None
        _ck(LIB.pr_gen_evaluate_async(self.h, self.d_genomes.ptr, self.d_res.ptr if self.has_res else None, pop,
                                      seeds, len(seeds), holdout_k, C.byref(params), acc, idx,
                                      None if samples is None else samples.ctypes.data), "pr_gen_evaluate_async")
        out = {"acc": acc, "idx": idx}
        if samples is not None:
            out["samples"] = samples.reshape(pop, len(seeds))
        return out

    def read_samples(self, first, count, nseeds):
        """Flight distances [count x nseeds] of genomes first.. of the last population batch."""
        out = np.zeros(count * nseeds, np.float32)
        _ck(LIB.pr_gen_read_samples(self.h, int(first), int(count), out.ctypes.data), "pr_gen_read_samples")
        return out.reshape(count, nseeds)

    def collect(self):
        """Holdout score and champion of the oldest outstanding evaluate_async."""
        hold = C.c_float()
        champ = C.c_int()
        _ck(LIB.pr_gen_collect(self.h, C.byref(hold), C.byref(champ)), "pr_gen_collect")
        return {"holdout": hold.value, "champion": champ.value}

    def close(self):
        """Frees the generation's device buffers now (the destructor does it otherwise)."""
        if getattr(self, "h", None):
            LIB.pr_gen_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ geometry families (workload synthesis)
FAMILIES = ("delta", "elliptical", "plank", "generic")
FIN_LAYOUTS = ("centre", "twin", "triple", "outboard")


def mesh_triangles(genome, ch, sp):
    """Closed-form triangle count of a genome at mesh resolution (ch, sp) (common.h:177-292)."""
    lay = genome[26]
    nfin = 1 if lay < 0.25 else (2 if lay < 0.5 else (3 if lay < 0.75 else 2))
    bi = 1 if genome[13] > np.float32(0.02) else 0
    return 4 * sp * ch * (1 + bi) + 8 + 2 * nfin + 6 + 4 * bi


def family_population(n, seed=0, nt_min=200, nt_max=900, biplane_frac=0.25):
    """Mixed-family population for the 100k x 12 configuration: planform family (delta,
    elliptical, plank, generic), mono/biplane, centre/twin/triple/outboard fins, and a
    per-genome mesh resolution (ch, sp) whose triangle count is log-uniform in
    [nt_min, nt_max].  Returns (genomes [n,35] f32, res [n,2] int32, family ids)."""
    lo, hi = gene_bounds()
    rng = np.random.default_rng(seed)
    g = (lo + rng.random((n, N_GENES), dtype=np.float32) * (hi - lo)).astype(np.float32)
    fam = rng.integers(0, len(FAMILIES), n)

    def u(a, b, m):
        return rng.uniform(a, b, m).astype(np.float32)

    d = fam == 0  # delta: strong sweep, small taper, pointed tips, straight chord blend
    m = int(d.sum())
    g[d, 3], g[d, 2], g[d, 33] = u(0.6, 0.9, m), u(0.04, 0.25, m), u(0.0, 0.15, m)
    g[d, 24], g[d, 25] = u(-0.1, 0.2, m), u(-0.2, 0.1, m)
    e = fam == 1  # elliptical: rounded tips, moderate taper, little sweep
    m = int(e.sum())
    g[e, 33], g[e, 2], g[e, 3] = u(0.7, 1.0, m), u(0.4, 0.8, m), u(0.0, 0.2, m)
    p = fam == 2  # plank: straight, untapered, reflexed trailing edge
    m = int(p.sum())
    g[p, 3], g[p, 2], g[p, 10], g[p, 8] = u(0.0, 0.15, m), u(0.75, 1.0, m), u(0.05, 0.2, m), u(-0.06, 0.0, m)
    bi = rng.random(n) < biplane_frac
    g[bi, 13] = u(0.025, 0.06, int(bi.sum()))
    g[~bi, 13] = u(0.0, 0.019, int((~bi).sum()))
    fins = rng.integers(0, 4, n)
    g[:, 26] = (fins * 0.25 + rng.uniform(0.02, 0.23, n)).astype(np.float32)
    g = np.clip(g, lo, hi).astype(np.float32)
    # resolution: target triangle count log-uniform in [nt_min, nt_max]
    target = np.exp(rng.uniform(np.log(nt_min), np.log(nt_max), n))
    res = np.zeros((n, 2), np.int32)
    for i in range(n):
        b = 1 if g[i, 13] > np.float32(0.02) else 0
        base = mesh_triangles(g[i], 0, 0)
        tw = max(8.0, (target[i] - base) / (1 + b))  # triangles of one deck = 4 sp ch
        ch = int(np.clip(round(np.sqrt(tw / 8.0)), 2, 12))
        sp = int(np.clip(round(tw / (4 * ch)), 2, 24))
        while mesh_triangles(g[i], ch, sp) > nt_max and sp > 2:
            sp -= 1
        while mesh_triangles(g[i], ch, sp) < nt_min and sp < 24:
            sp += 1
        res[i] = (ch, sp)
    return g, res, fam


# ------------------------------------------------------------------ autotuner
NB_MAX = 8  # triangle-count buckets a context holds
def autotune(genomes, res, seeds, params, ranges=(192, 264, 320, 384, 528, 640, 768, 1024, 1536, 2048),
             sample_flights=98304,
             steps=600, out_json=None, verbose=False, merge=False, repeats=3):
    """Pick the flight-kernel mapping per triangle-count range: for each range (lo, hi],
    every compiled (L lanes per flight, P panels per lane) variant with L*P >= hi flies
    the same sample of real flights (genomes whose count falls in the range) for one
    chunk; the variant with the least kernel time per panel-step wins.  The ranges stay
    separate buckets by default (a wave runs as many panel slots as its fullest flight
    needs, so narrower ranges waste fewer); merge=True joins adjacent ranges with the same
    winner.  The default ranges end where the variants' capacities L*P do (8/16/32 x 33 =
    264/528/1056 with the LDS slots, x 24 = 192/384/768 in VGPRs), plus splits of the
    VGPR-only ranges.  More than 8 buckets (the context's limit) are merged down: adjacent
    ranges with the same winner first, then a range into the next one (whose mapping
    covers it) where the sample holds the fewest flights.  Returns (nt_max, L, P) lists;
    writes JSON if asked."""
    import json

    g = np.ascontiguousarray(genomes, np.float32).reshape(-1, N_GENES)
    r = None if res is None else np.ascontiguousarray(res, np.int32).reshape(-1, 2)
    nt = np.array([mesh_triangles(g[i], *(r[i] if r is not None else (5, 8))) for i in range(len(g))])
    nt_cap = int(nt.max())
    ranges = [x for x in ranges if x < nt_cap] + [min(y for y in ranges if y >= nt_cap)]
    variants = fly_variants()
    seeds = np.ascontiguousarray(seeds, np.int32)
    tune_params = SimParams(params.preset, params.gust_rms_max, params.warp_max, params.wind_sd,
                            min(params.max_steps, steps), params.chunk, 0.0)
    choice = []
    lo = 0
    max_ch = int(r[:, 0].max()) if r is not None else 0
    max_sp = int(r[:, 1].max()) if r is not None else 0
    for hi in ranges:
        sel = np.nonzero((nt > lo) & (nt <= hi))[0]
        cands = [(L, P) for (L, P) in variants if L * P >= hi]
        if len(sel) == 0 or not cands:
            lo = hi
            continue
        ng = max(1, min(len(sel), sample_flights // len(seeds)))
        gi = sel[:ng]
        pl = Planes(ng, len(seeds), ng * len(seeds), max_ch=max_ch, max_sp=max_sp)
        pl.set_tail_mapping(False)  # time the candidate mapping itself on every flight
        pl.prepare(g[gi], None if r is None else r[gi])
        pl.scenarios(seeds, tune_params)
        best = None
        for (L, P) in cands:
            big = max(variants, key=lambda v: v[0] * v[1])  # catch-all bucket; the sample never reaches it
            if hi >= big[0] * big[1]:
                pl.set_buckets([hi], [L], [P])
            else:
                pl.set_buckets([hi, big[0] * big[1]], [L, big[0]], [P, big[1]])
            pl.fly_device(0, ng, len(seeds), tune_params)  # warm
            cost, ms = None, 0.0
            for _ in range(repeats):  # the fastest of a few runs: timing noise only adds
                pl.fly_device(0, ng, len(seeds), tune_params)
                st = pl.stats()
                b0 = st["buckets"][0]
                # inside the fused kernel (at its occupancy) when the variant is compiled into
                # it -- the catch-all bucket holds no flights, so the launch time is bucket 0's
                t = st["fused_ms"] if b0["fused"] else b0["fly_ms"]
                c = t / max(1, b0["panel_steps"])
                if cost is None or c < cost:
                    cost, ms = c, t
            if verbose:
                print(f"  range ({lo},{hi}] L={L} P={P}: {ms:.3f} ms, {cost * 1e9:.3f} ps/panel-step")
            if best is None or cost < best[0]:
                best = (cost, L, P)
        pl.close()
        choice.append((hi, best[1], best[2], len(sel)))
        lo = hi
    merged = []
    for hi, L, P, n in choice:
        if merge and merged and merged[-1][1:3] == (L, P):
            merged[-1] = (hi, L, P, merged[-1][3] + n)
        else:
            merged.append((hi, L, P, n))
    merged[-1] = (max(merged[-1][0], nt_cap),) + merged[-1][1:]
    # a context sized for (max_ch, max_sp) needs its last bucket to cover that resolution's
    # triangle bound, not just the sample's largest count
    bound = 8 * max_ch * max_sp + 24 if max_ch and max_sp else 8 * 5 * 8 + 24
    if merged[-1][0] < bound:
        hi, L, P, n = merged[-1]
        if L * P >= bound:
            merged[-1] = (bound, L, P, n)
        else:
            cover = min((v for v in variants if v[0] * v[1] >= bound), key=lambda v: v[0] * v[1])
            merged.append((bound, cover[0], cover[1], 0))
    # a merged range takes the upper range's mapping, which covers both (L*P >= its nt_max)
    while len(merged) > NB_MAX:
        same = [i for i in range(len(merged) - 1) if merged[i][1:3] == merged[i + 1][1:3]]
        i = same[0] if same else min(range(len(merged) - 1), key=lambda k: merged[k][3])
        hi, L, P, n = merged[i + 1]
        merged[i:i + 2] = [(hi, L, P, n + merged[i][3])]
    nt_max = [m[0] for m in merged]
    Ls = [m[1] for m in merged]
    Ps = [m[2] for m in merged]
    if out_json:
        with open(out_json, "w") as f:
            json.dump({"nt_max": nt_max, "L": Ls, "P": Ps}, f)
    return nt_max, Ls, Ps


# ------------------------------------------------------------------ surrogate MLP
class Surrogate:
    """BF16 MLP 35 -> 256 -> 256 -> 256 -> 4 on the matrix cores (pr_mlp_*), trained on
    (genome, [mean, CVaR25, worst, oversize]) pairs produced by the flight simulation."""

    def __init__(self, max_batch=16384, seed=1):
        self.h = VP()
        rc = LIB.pr_mlp_create(C.byref(self.h), max_batch, seed)
        if rc:
            raise PlanesError(f"pr_mlp_create failed ({rc}): {LIB.pr_mlp_last_error().decode()}")
        self.max_batch = max_batch
        self.n_params = LIB.pr_mlp_param_count()
        self._bufs = {}

    def _ck(self, rc, what):
        if rc:
            raise PlanesError(f"{what} failed ({rc}): {LIB.pr_mlp_last_error().decode()}")

    def _dev(self, name, arr):
        arr = np.ascontiguousarray(arr, np.float32)
        b = self._bufs.get(name)
        if b is None or b.nbytes < arr.nbytes:
            b = DeviceBuffer(arr.nbytes)
            self._bufs[name] = b
        b.upload(arr)
        return b

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self._ck(LIB.pr_mlp_get_params(self.h, out, self.n_params), "pr_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._ck(LIB.pr_mlp_set_params(self.h, p, len(p)), "pr_mlp_set_params")

    def predict(self, genomes):
        """-> [n, 5]: mean, CVaR25, worst (m), oversize probability, predicted robust score."""
        g = np.ascontiguousarray(genomes, np.float32).reshape(-1, N_GENES)
        dg = self._dev("pg", g)
        out = self._bufs.get("pout")
        if out is None or out.nbytes < 20 * len(g):
            out = DeviceBuffer(20 * len(g))
            self._bufs["pout"] = out
        self._ck(LIB.pr_mlp_predict(self.h, dg.ptr, len(g), out.ptr), "pr_mlp_predict")
        return out.download(np.float32, 5 * len(g)).reshape(-1, 5)

    def predict_device(self, d_genomes, n, d_pred):
        self._ck(LIB.pr_mlp_predict(self.h, d_genomes, n, d_pred), "pr_mlp_predict")

    def loss_grad(self, genomes, targets):
        g = np.ascontiguousarray(genomes, np.float32).reshape(-1, N_GENES)
        dg = self._dev("lg", g)
        dt = self._dev("lt", np.ascontiguousarray(targets, np.float32).reshape(-1, 4))
        loss = C.c_float()
        grads = np.zeros(self.n_params, np.float32)
        self._ck(LIB.pr_mlp_loss_grad(self.h, dg.ptr, dt.ptr, len(g), C.byref(loss), grads), "pr_mlp_loss_grad")
        return loss.value, grads

    def train(self, genomes, targets, epochs=4, batch=4096, lr=1e-3, shuffle_seed=0):
        g = np.ascontiguousarray(genomes, np.float32).reshape(-1, N_GENES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
        if shuffle_seed is not None:
            perm = np.random.default_rng(shuffle_seed).permutation(len(g))
            g, t = g[perm], t[perm]
        dg = self._dev("tg", g)
        dt = self._dev("tt", t)
        loss = C.c_float()
        self._ck(LIB.pr_mlp_train(self.h, dg.ptr, dt.ptr, len(g), epochs, min(batch, self.max_batch), lr,
                                  C.byref(loss)), "pr_mlp_train")
        return loss.value

    def __del__(self):
        try:
            LIB.pr_mlp_destroy(self.h)
        except Exception:
            pass


class PoolGA:
    """Candidate-pool GA on the GPU (pr_pool_*): plain (fly the whole pool) or surrogate
    pre-screened (fly the elite + the best-predicted candidates)."""

    FIELDS = ("best", "elite_mean", "holdout", "wall_ms", "fly_ms", "mlp_ms", "flights", "pfs", "loss", "n_fly",
              "best_holdout", "rank_corr", "mae", "oversize_acc", "w_parent", "w_pred", "w_parent_pred")

    def __init__(self, n_elite, n_pool, n_fly, nseeds, buf_cap=2_000_000, sigma_frac=0.05, seed=1):
        self.h = VP()
        rc = LIB.pr_pool_create(C.byref(self.h), n_elite, n_pool, n_fly, nseeds, buf_cap, sigma_frac, seed)
        if rc:
            raise PlanesError(f"pr_pool_create failed ({rc}): {LIB.pr_pool_last_error().decode()}")
        self.n_elite = n_elite
        self.n_pool = n_pool

    def _ck(self, rc, what):
        if rc:
            raise PlanesError(f"{what} failed ({rc}): {LIB.pr_pool_last_error().decode()}")

    def set_elite(self, genomes):
        g = np.ascontiguousarray(genomes, np.float32).reshape(-1)
        self._ck(LIB.pr_pool_set_elite(self.h, g, len(g) // N_GENES), "pr_pool_set_elite")

    def elite(self):
        out = np.zeros(self.n_elite * N_GENES, np.float32)
        self._ck(LIB.pr_pool_get_elite(self.h, out, self.n_elite), "pr_pool_get_elite")
        return out.reshape(-1, N_GENES)

    def generation(self, gen, params, surrogate=None, train_epochs=2, lr=1e-3):
        out = np.zeros(len(self.FIELDS), np.float64)
        self._ck(LIB.pr_pool_generation(self.h, gen, C.byref(params), surrogate.h if surrogate else None,
                                        train_epochs, lr, out), "pr_pool_generation")
        return dict(zip(self.FIELDS, out.tolist()))

    def set_explore(self, frac):
        """Fraction of the surrogate fly set drawn at random from the pool."""
        self._ck(LIB.pr_pool_set_explore(self.h, float(frac)), "pr_pool_set_explore")

    def set_parent_delta(self, mode=True):
        """Mutants ranked by parent's simulated score + predicted change (True / 1, default), by
        a fitted combination ("fitted" / 2: weights of the parent's score, the mutant's and the
        parent's predicted score from the last exploration sample's least squares), or by the
        predicted score itself (False / 0)."""
        m = 2 if mode == "fitted" else int(bool(mode))
        self._ck(LIB.pr_pool_set_parent_delta(self.h, m), "pr_pool_set_parent_delta")

    def set_key_weights(self, w_parent=1.0, w_pred=1.0, w_parent_pred=-1.0):
        """Weights of the default key (set_parent_delta(True)): w_parent * parent's simulated
        score + w_pred * mutant's predicted score + w_parent_pred * parent's predicted score."""
        self._ck(LIB.pr_pool_set_key_weights(self.h, float(w_parent), float(w_pred), float(w_parent_pred)),
                 "pr_pool_set_key_weights")

    def set_selection(self, mode):
        """Fly-set choice of generations without a surrogate: "all" (the whole pool),
        "uniform" (elite + a uniform sample of the mutants, n_fly in total) or "parent"
        (elite + the mutants of the best-scoring parents) -- the no-network baselines."""
        m = {"all": 0, "uniform": 1, "parent": 2}[mode]
        self._ck(LIB.pr_pool_set_selection(self.h, m), "pr_pool_set_selection")

    def add_training(self, d_genomes, samples):
        """(genome, fitness) pairs for the replay buffer: device genomes [n x 35] and their
        host flight distances [n x nseeds] (targets computed, and normalised, as for flown
        generations)."""
        s = np.ascontiguousarray(samples, np.float32)
        self._ck(LIB.pr_pool_add_training(self.h, d_genomes, s.reshape(-1), s.shape[0]), "pr_pool_add_training")

    def set_target_norm(self, on=True):
        """Regression targets of each collected batch relative to the batch's medians over
        its flyable designs (removes the scenario set's common effect on the distances)."""
        self._ck(LIB.pr_pool_set_target_norm(self.h, int(bool(on))), "pr_pool_set_target_norm")

    def set_collect(self, on=True):
        """Plain generations also feed the surrogate's replay buffer (bootstrap)."""
        self._ck(LIB.pr_pool_set_collect(self.h, int(on)), "pr_pool_set_collect")

    def train(self, surrogate, epochs, lr=1e-3):
        loss = C.c_float()
        self._ck(LIB.pr_pool_train(self.h, surrogate.h, epochs, lr, C.byref(loss)), "pr_pool_train")
        return loss.value

    def last(self, targets=False):
        """The last generation: pool, fly set (flight order), simulated scores of the fly set
        and, with targets=True, the training targets of the fly set (collected generations)."""
        pool = np.zeros((self.n_pool, N_GENES), np.float32)
        nf = C.c_int()
        self._ck(LIB.pr_pool_last(self.h, None, None, None, None, C.byref(nf)), "pr_pool_last")
        fly = np.zeros((nf.value, N_GENES), np.float32)
        sc = np.zeros(nf.value, np.float32)
        tg = np.zeros((nf.value, 4), np.float32) if targets else None
        ptr = lambda a: None if a is None else a.ctypes.data
        self._ck(LIB.pr_pool_last(self.h, ptr(pool), ptr(fly), ptr(sc), ptr(tg), C.byref(nf)), "pr_pool_last")
        out = {"pool": pool, "fly": fly, "scores": sc}
        if targets:
            out["targets"] = tg
        return out

    def champion(self):
        g = np.zeros(N_GENES, np.float32)
        hs = C.c_float()
        self._ck(LIB.pr_pool_champion(self.h, g, C.byref(hs)), "pr_pool_champion")
        return g, hs.value

    def __del__(self):
        try:
            LIB.pr_pool_destroy(self.h)
        except Exception:
            pass
