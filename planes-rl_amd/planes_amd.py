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


class Stats(C.Structure):
    _fields_ = [("flight_steps", C.c_ulonglong), ("panel_steps", C.c_ulonglong), ("fly_ms", C.c_double),
                ("fly_launches", C.c_int), ("prep_ms", C.c_double)]


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
    lib.pr_read_states.argtypes = [VP, C.c_int, C.c_int, F32P]
    lib.pr_dev_alloc.argtypes = [C.POINTER(VP), C.c_size_t]
    lib.pr_dev_free.argtypes = [VP]
    lib.pr_h2d.argtypes = [VP, VP, C.c_size_t]
    lib.pr_d2h.argtypes = [VP, VP, C.c_size_t]
    lib.pr_device_name.argtypes = [C.c_char_p, C.c_int]
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

    def init_flights(self, layout, n_genomes, n_scen, params=None):
        params = params or SimParams.reference()
        _ck(LIB.pr_init_flights(self.h, layout, n_genomes, n_scen, C.byref(params), self.d_fit.ptr,
                                self.d_steps.ptr), "pr_init_flights")
        nf = n_genomes * n_scen if layout == 0 else n_genomes
        return self.read_states(0, nf)

    def evaluate(self, genomes, seeds, params=None, layout=0, res=None):
        """Batched evaluation: every genome under every seed (layout 0) or genome i under
        seed i (layout 1).  Returns fitness[n_genomes, n_seeds] (layout 0) or [n]."""
        params = params or SimParams.reference()
        self.prepare(genomes, res)
        self.scenarios(seeds, params)
        fit, steps = self.fly(layout, self.n_genomes, self.n_scen, params)
        if layout == 0:
            return fit.reshape(self.n_genomes, self.n_scen), steps.reshape(self.n_genomes, self.n_scen)
        return fit, steps

    def scores(self, samples, nseeds):
        s = np.ascontiguousarray(samples, np.float32).reshape(-1)
        pop = len(s) // nseeds
        buf = DeviceBuffer(s.nbytes)
        buf.upload(s)
        _ck(LIB.pr_scores(self.h, buf.ptr, pop, nseeds, self.d_scores.ptr), "pr_scores")
        return self.d_scores.download(np.float32, pop)

    def set_buckets(self, nt_max, L, P):
        _ck(LIB.pr_set_buckets(self.h, len(nt_max), np.asarray(nt_max, np.int32), np.asarray(L, np.int32),
                               np.asarray(P, np.int32)), "pr_set_buckets")

    def stats(self):
        st = Stats()
        _ck(LIB.pr_stats_get(self.h, C.byref(st)), "pr_stats_get")
        return {k: getattr(st, k) for k, _ in Stats._fields_}

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

    def read_states(self, first, n):
        """FlightState rows: pos3 q4 vel3 wb3 (float), step, done (int), fitness."""
        out = np.zeros(n * 16, np.float32)
        _ck(LIB.pr_read_states(self.h, first, n, out), "pr_read_states")
        out = out.reshape(n, 16)
        return {"state": out[:, :13].copy(), "step": out[:, 13].view(np.int32).copy(),
                "done": out[:, 14].view(np.int32).copy(), "fitness": out[:, 15].copy()}
