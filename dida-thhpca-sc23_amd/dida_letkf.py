"""Python binding of the LETKF engine's C ABI (include/dida_letkf.h).

ctypes over ``lib/libdida_letkf.so``.  Device buffers are torch tensors on a
ROCm device: torch is only the allocator and stream provider; every
computation runs in the library's HIP kernels.  Loading fails loudly when the
library has not been built -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdida_letkf.so")
SYNTH_PATH = os.path.join(HERE, "lib", "libdida_synth.so")

DIDA_OK = 0
DIDA_WARN_NOCONV = 1
DIDA_ERR_CAPACITY = -4
DIDA_ERR_NONFINITE = -6
GRAM_FP64, GRAM_BF16X3, GRAM_FP16X3, GRAM_BF16 = 0, 1, 2, 3
FORM_PLAIN, FORM_BATCHED = 0, 1
TIMING_STAGES, TIMING_PHASES = 1, 2

# Every entry point declared in include/dida_letkf.h.
EXPORTED = ["dida_letkf_params_default", "letkf_core_", "letkf_core_batch_",
            "letkf_core_r4_", "letkf_core_batch_r4_",
            "dida_letkf_core_batch_dev", "dida_letkf_core_batch_dev_noalloc",
            "dida_letkf_create", "dida_letkf_destroy",
            "dida_letkf_set_grid", "dida_letkf_sweep", "dida_letkf_sweep_f32",
            "dida_letkf_graph_capture",
            "dida_letkf_graph_launch", "dida_letkf_stats_get", "dida_letkf_debug_arrays",
            "dida_letkf_post", "dida_letkf_last_error", "dida_letkf_set_timing",
            "dida_letkf_stage_ms", "dida_letkf_phase_cycles", "dida_letkf_group_capture",
            "dida_letkf_group_launch", "dida_letkf_group_destroy", "dida_letkf_autotune",
            "dida_letkf_obs_operator", "dida_letkf_set_levels", "dida_struct_atm2endnn",
            "dida_struct_endnn2atm", "dida_struct_atm2endnn_a", "dida_struct_endnn2atm_a"]
PHASES = ["gram", "warm", "jacobi", "transform", "update"]
STAGES = ["bin", "select", "transform", "update"]


class Params(C.Structure):
    _fields_ = [("nanals", C.c_int32), ("nrow", C.c_int32), ("nlat", C.c_int32),
                ("nlon", C.c_int32), ("max_obs", C.c_int32), ("gc_filter", C.c_int32),
                ("local_dist", C.c_double), ("obs_polar", C.c_double),
                ("inflation_factor", C.c_double), ("rtps", C.c_int32), ("form", C.c_int32),
                ("rtps_alpha", C.c_double), ("gram", C.c_int32),
                ("jacobi_max_sweeps", C.c_int32), ("jacobi_tol", C.c_double),
                ("bin_deg", C.c_double), ("chain_len", C.c_int32), ("warm_start", C.c_int32),
                ("jacobi_predict", C.c_int32), ("vert_dist", C.c_double),
                ("cube_nz", C.c_int32), ("gc_filter_vertical", C.c_int32),
                ("cube_n", C.c_int32), ("overlap", C.c_int32), ("xcd_tiles", C.c_int32),
                ("bucket_chain", C.c_int32 * 4), ("pad_", C.c_int32)]


class TuningC(C.Structure):
    _fields_ = [("chain_len", C.c_int32), ("ncand", C.c_int32), ("bin_deg", C.c_double),
                ("ms_initial", C.c_double), ("ms_best", C.c_double), ("xcd_tiles", C.c_int32),
                ("bucketed", C.c_int32), ("bucket_lo", C.c_int32 * 4),
                ("bucket_chain", C.c_int32 * 4), ("bucket_count", C.c_int64 * 4),
                ("select_tile", C.c_int32), ("pad_", C.c_int32)]


class VertCoordC(C.Structure):
    _fields_ = [("type", C.c_int32), ("nz", C.c_int32), ("c1", C.c_void_p), ("c2", C.c_void_p),
                ("ptop", C.c_double), ("p0", C.c_double)]


class ObsRecordsC(C.Structure):
    _fields_ = [("nrec", C.c_int32), ("onz", C.c_int32), ("onv", C.c_int32), ("pad_", C.c_int32),
                ("vid", C.c_void_p), ("lat", C.c_void_p), ("lon", C.c_void_p),
                ("ph", C.c_void_p), ("yo", C.c_void_p)]


class ObsC(C.Structure):
    _fields_ = [("nobs", C.c_int32), ("pad_", C.c_int32), ("lat", C.c_void_p),
                ("lon", C.c_void_p), ("yb", C.c_void_p), ("dep", C.c_void_p),
                ("rinv", C.c_void_p), ("ph", C.c_void_p)]


class StatsC(C.Structure):
    _fields_ = [("nanal", C.c_int64), ("nobsl_sum", C.c_int64), ("nobsl_max", C.c_int32),
                ("nobsl_capped", C.c_int32), ("nonconverged", C.c_int32),
                ("max_sweeps_used", C.c_int32), ("capacity_overflow", C.c_int32),
                ("nonfinite", C.c_int32), ("sweeps_sum", C.c_int64)]


class SynthCfg(C.Structure):
    _fields_ = [("nlat", C.c_int32), ("nlon", C.c_int32), ("nz", C.c_int32), ("k", C.c_int32),
                ("nobs", C.c_int32), ("nclusters", C.c_int32), ("lmax_pert", C.c_int32),
                ("pad_", C.c_int32), ("cluster_frac", C.c_double),
                ("cluster_radius_km", C.c_double), ("sig_T", C.c_double),
                ("sig_uv", C.c_double), ("rscale_min", C.c_double), ("rscale_max", C.c_double),
                ("pert_T", C.c_double), ("pert_uv", C.c_double), ("pert_ps", C.c_double),
                ("pert_len_km", C.c_double), ("vert_corr", C.c_double), ("seed", C.c_uint64)]


_lib = None
_synth = None
_hip = None


class DidaError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DidaError(f"{LIB_PATH} is missing: build it with build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.dida_letkf_params_default.argtypes = [C.POINTER(Params)]
        L.dida_letkf_params_default.restype = None
        L.dida_letkf_last_error.restype = C.c_char_p
        L.dida_letkf_create.argtypes = [C.POINTER(vp), C.POINTER(Params), C.c_int32, C.c_int64]
        L.dida_letkf_destroy.argtypes = [vp]
        L.dida_letkf_set_grid.argtypes = [vp, vp, vp]
        L.dida_letkf_sweep.argtypes = [vp, C.POINTER(ObsC), vp, vp, vp, vp, vp]
        L.dida_letkf_sweep_f32.argtypes = [vp, C.POINTER(ObsC), vp, vp, vp, vp, vp]
        L.dida_letkf_graph_capture.argtypes = [vp, C.POINTER(ObsC), vp, vp, vp, vp, vp]
        L.dida_letkf_graph_launch.argtypes = [vp, vp]
        L.dida_letkf_stats_get.argtypes = [vp, C.POINTER(StatsC)]
        L.dida_letkf_debug_arrays.argtypes = [vp] + [C.POINTER(vp)] * 6
        L.dida_letkf_core_batch_dev.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp,
                                                vp, vp, vp, C.c_double, C.c_int32, vp, vp]
        L.dida_letkf_core_batch_dev_noalloc.argtypes = L.dida_letkf_core_batch_dev.argtypes
        L.dida_letkf_post.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, C.c_double,
                                      C.c_double, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp]
        L.dida_letkf_set_timing.argtypes = [vp, C.c_int32]
        L.dida_letkf_stage_ms.argtypes = [vp, C.POINTER(C.c_double)]
        L.dida_letkf_phase_cycles.argtypes = [vp, C.POINTER(C.c_double)]
        L.dida_letkf_group_capture.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp),
                                               C.POINTER(ObsC), C.POINTER(vp), C.POINTER(vp),
                                               C.POINTER(vp), C.POINTER(vp), vp]
        L.dida_letkf_group_launch.argtypes = [vp, vp]
        L.dida_letkf_obs_operator.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, vp,
                                              C.c_int32, vp, vp, C.c_int32,
                                              C.POINTER(VertCoordC), C.c_int32, vp, vp,
                                              C.POINTER(ObsRecordsC), vp, vp, vp, vp, vp, vp]
        L.dida_letkf_set_levels.argtypes = [vp, C.POINTER(VertCoordC), C.c_int32, C.c_int32,
                                            vp, vp]
        for nm in ("dida_struct_atm2endnn", "dida_struct_endnn2atm"):
            getattr(L, nm).argtypes = [C.c_int32] * 8 + [vp, vp, vp]
        for nm in ("dida_struct_atm2endnn_a", "dida_struct_endnn2atm_a"):
            getattr(L, nm).argtypes = [C.c_int32] * 5 + [vp, vp, vp]
        L.dida_letkf_autotune.argtypes = [vp, C.POINTER(ObsC), vp, vp, vp, vp, vp,
                                          C.POINTER(TuningC)]
        L.dida_letkf_group_destroy.argtypes = [vp]
        L.letkf_core_.argtypes = [vp] * 10
        L.letkf_core_.restype = None
        L.letkf_core_batch_.argtypes = [vp] * 13
        L.letkf_core_batch_.restype = None
        L.letkf_core_r4_.argtypes = [vp] * 10
        L.letkf_core_r4_.restype = None
        L.letkf_core_batch_r4_.argtypes = [vp] * 13
        L.letkf_core_batch_r4_.restype = None
        _lib = L
    return _lib


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise DidaError(f"{SYNTH_PATH} is missing: build it with build.py")
        S = C.CDLL(SYNTH_PATH)
        S.dida_synth_scratch_bytes.argtypes = [C.POINTER(SynthCfg)]
        S.dida_synth_scratch_bytes.restype = C.c_size_t
        S.dida_synth_generate.argtypes = [C.POINTER(SynthCfg)] + [C.c_void_p] * 12
        _synth = S
    return _synth


def hip():
    """The HIP runtime already loaded by torch (for raw device copies)."""
    global _hip
    if _hip is None:
        H = C.CDLL("libamdhip64.so")
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = lib().dida_letkf_last_error().decode()
        raise DidaError(f"{what} failed with status {rc}: {msg}")
    return rc


def d2h(ptr: int, n: int, dtype) -> np.ndarray:
    """Copy n elements of dtype from a device pointer to a new numpy array."""
    out = np.empty(n, dtype=dtype)
    if n:
        e = hip().hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2)
        if e != 0:
            raise DidaError(f"hipMemcpy D2H failed: {e}")
    return out


def default_params(**kw) -> Params:
    p = Params()
    lib().dida_letkf_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        if k == "bucket_chain":
            v = (C.c_int32 * 4)(*v)
        setattr(p, k, v)
    return p


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


@dataclass
class DeviceObs:
    """Observation arrays on the device (torch tensors)."""
    lat: object
    lon: object
    yb: object      # (nobs, k) float64
    dep: object
    rinv: object    # float32
    ph: object = None   # float64 obs pressure (vertical localisation only)

    def c(self) -> ObsC:
        return ObsC(int(self.lat.numel()), 0, self.lat.data_ptr(), self.lon.data_ptr(),
                    self.yb.data_ptr(), self.dep.data_ptr(), self.rinv.data_ptr(),
                    None if self.ph is None else self.ph.data_ptr())


class Group:
    """One HIP graph with a sweep of each of several engines (domains), run
    on parallel graph branches (dida_letkf_group_*)."""

    def __init__(self, engines, obs, xb, axb, xa, axa, stream):
        n = len(engines)
        vp = C.c_void_p
        arr = lambda xs: (vp * n)(*[x.data_ptr() for x in xs])
        self._keep = [engines, obs, xb, axb, xa, axa]
        obs_c = (ObsC * n)(*[o.c() for o in obs])
        h = vp()
        check(lib().dida_letkf_group_capture(C.byref(h), n, (vp * n)(*[e.h.value for e in engines]),
                                             obs_c, arr(xb), arr(axb), arr(xa), arr(axa),
                                             _stream_ptr(stream)), "group_capture")
        self.h = h

    def launch(self, stream=None) -> int:
        return check(lib().dida_letkf_group_launch(self.h, _stream_ptr(stream)), "group_launch")

    def close(self):
        if self.h:
            lib().dida_letkf_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """A dida_letkf_ctx: whole-sweep analyses on one device."""

    def __init__(self, params: Params, nobs_max: int, capacity: int = 0):
        self.params = params
        h = C.c_void_p()
        check(lib().dida_letkf_create(C.byref(h), C.byref(params), int(nobs_max),
                                      int(capacity)), "dida_letkf_create")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib().dida_letkf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, grid_lat, grid_lon):
        self._keep = [grid_lat, grid_lon]
        check(lib().dida_letkf_set_grid(self.h, _ptr(grid_lat), _ptr(grid_lon)), "set_grid")

    def set_levels(self, vcoord, ps_row, var_start, var_lev):
        """dida_letkf_set_levels (vertical localisation): vcoord dict(type, c1,
        c2, ptop, p0) with c1/c2 device float64 tensors of nz + 1 half levels."""
        c1 = vcoord["c1"]
        c2 = vcoord.get("c2", c1)
        self._keep_levels = (c1, c2)
        vc = VertCoordC(int(vcoord["type"]), int(c1.numel()) - 1, c1.data_ptr(), c2.data_ptr(),
                        float(vcoord["ptop"]), float(vcoord.get("p0", 0.0)))
        vs = np.ascontiguousarray(var_start, np.int32)
        vl = np.ascontiguousarray(var_lev, np.int32)
        check(lib().dida_letkf_set_levels(self.h, C.byref(vc), int(ps_row), len(vs),
                                          vs.ctypes.data, vl.ctypes.data), "set_levels")

    def sweep(self, obs: DeviceObs, xb, axb, xa, axa, stream=None) -> int:
        """dida_letkf_sweep, or dida_letkf_sweep_f32 when Xb' and Xa' are
        float32 tensors (means stay float64)."""
        oc = obs.c()
        f32 = getattr(xb, "dtype", None) is not None and str(xb.dtype).endswith("float32")
        if f32 != str(xa.dtype).endswith("float32"):
            raise TypeError("xb and xa must both be float32 or both float64")
        fn = lib().dida_letkf_sweep_f32 if f32 else lib().dida_letkf_sweep
        return check(fn(self.h, C.byref(oc), _ptr(xb), _ptr(axb), _ptr(xa), _ptr(axa),
                        _stream_ptr(stream)), "sweep")

    def graph_capture(self, obs: DeviceObs, xb, axb, xa, axa, stream) -> int:
        oc = obs.c()
        return check(lib().dida_letkf_graph_capture(self.h, C.byref(oc), _ptr(xb), _ptr(axb),
                                                    _ptr(xa), _ptr(axa), _stream_ptr(stream)),
                     "graph_capture")

    def autotune(self, obs: DeviceObs, xb, axb, xa, axa, stream=None) -> dict:
        """dida_letkf_autotune: keep the fastest layout for this workload."""
        oc, t = obs.c(), TuningC()
        check(lib().dida_letkf_autotune(self.h, C.byref(oc), _ptr(xb), _ptr(axb), _ptr(xa),
                                        _ptr(axa), _stream_ptr(stream), C.byref(t)), "autotune")
        out = {}
        for f, _ in TuningC._fields_:
            if f == "pad_":
                continue
            v = getattr(t, f)
            out[f] = list(v) if hasattr(v, "__len__") else v
        return out

    def graph_launch(self, stream=None) -> int:
        return check(lib().dida_letkf_graph_launch(self.h, _stream_ptr(stream)), "graph_launch")

    def set_timing(self, on: bool = True, phases: bool = False):
        """Stage events (on) and, for the batched form, in-kernel phase clocks."""
        mode = (TIMING_STAGES if on else 0) | (TIMING_PHASES if phases else 0)
        check(lib().dida_letkf_set_timing(self.h, mode), "set_timing")

    def stage_ms(self) -> dict:
        ms = (C.c_double * len(STAGES))()
        check(lib().dida_letkf_stage_ms(self.h, ms), "stage_ms")
        return {name: ms[i] for i, name in enumerate(STAGES)}

    def phase_cycles(self) -> dict:
        cyc = (C.c_double * len(PHASES))()
        check(lib().dida_letkf_phase_cycles(self.h, cyc), "phase_cycles")
        return {name: cyc[i] for i, name in enumerate(PHASES)}

    def stats(self, strict: bool = True) -> dict:
        """Diagnostics of the last sweep.  strict: raise on an error status
        (capacity overflow, non-finite analyses); otherwise return it in
        "status" with the counts."""
        st = StatsC()
        rc = lib().dida_letkf_stats_get(self.h, C.byref(st))
        if strict:
            check(rc, "stats")
        d = {f: getattr(st, f) for f, _ in StatsC._fields_ if f != "pad_"}
        d["status"] = rc
        return d

    def local_counts(self) -> np.ndarray:
        """Local obs count of every analysis of the last sweep (host copy)."""
        ptrs = [C.c_void_p() for _ in range(6)]
        check(lib().dida_letkf_debug_arrays(self.h, *[C.byref(p) for p in ptrs]), "debug")
        step = self.params.cube_n - self.params.overlap if self.params.overlap else \
            self.params.cube_n
        nsel = -(-self.params.nlat // step) * -(-self.params.nlon // step)
        return d2h(ptrs[0].value, nsel, np.int32)

    def debug(self, nobs: int) -> dict:
        """Copies of the last sweep's selection arrays (host numpy)."""
        ptrs = [C.c_void_p() for _ in range(6)]
        check(lib().dida_letkf_debug_arrays(self.h, *[C.byref(p) for p in ptrs]), "debug")
        ncol = self.params.nlat * self.params.nlon
        nobsl = d2h(ptrs[0].value, ncol, np.int32)
        offset = d2h(ptrs[1].value, ncol + 1, np.int64)
        tot = int(offset[-1])
        return {"nobsl": nobsl, "offset": offset,
                "csr_idx": d2h(ptrs[2].value, tot, np.int32),
                "csr_w": d2h(ptrs[3].value, tot, np.float64),
                "perm": d2h(ptrs[4].value, nobs, np.int32),
                "info": d2h(ptrs[5].value, ncol, np.int32)}


def core_batch_dev(nobsl, hxens, rdiaginv, dep, rloc, trans, wbar, inflation: float,
                   gram: int = GRAM_FP64, info=None, stream=None) -> int:
    """dida_letkf_core_batch_dev on torch tensors: hxens (batch, max_obs, k),
    rdiaginv/dep/rloc (batch, max_obs), trans (batch, k, k) with
    trans[b, j, i] = W_ij + wbar_i (Fortran order), wbar (batch, k)."""
    batch, max_obs, k = hxens.shape
    return check(lib().dida_letkf_core_batch_dev(
        batch, max_obs, k, _ptr(nobsl), _ptr(hxens), _ptr(rdiaginv), _ptr(dep), _ptr(rloc),
        _ptr(trans), _ptr(wbar), float(inflation), int(gram), _ptr(info),
        _stream_ptr(stream)), "core_batch_dev")


def letkf_core(hxens: np.ndarray, rdiaginv, dep, rloc, inflation: float, pflag: int = 0):
    """The Fortran-ABI drop-in letkf_core_ on host arrays.  hxens: (p, k).
    Returns (trans[i, j], wbar)."""
    hx = np.ascontiguousarray(hxens, dtype=np.float64)
    p, k = hx.shape
    ri = np.ascontiguousarray(rdiaginv, dtype=np.float32)
    dp = np.ascontiguousarray(dep, dtype=np.float64)
    rl = np.ascontiguousarray(rloc, dtype=np.float64)
    trans = np.zeros(k * k)
    wbar = np.zeros(k)
    ik, ip, pf = C.c_int32(k), C.c_int32(p), C.c_int32(pflag)
    rho = C.c_double(inflation)
    lib().letkf_core_(C.byref(ik), C.byref(ip), hx.ctypes.data, ri.ctypes.data, dp.ctypes.data,
                      rl.ctypes.data, trans.ctypes.data, wbar.ctypes.data, C.byref(rho),
                      C.byref(pf))
    return trans.reshape(k, k).T.copy(), wbar


def letkf_core_batch(nobsl, hxens, rdiaginv, dep, rloc, inflation: float):
    """letkf_core_batch_ on host arrays: hxens (batch, max_obs, k)."""
    hx = np.ascontiguousarray(hxens, dtype=np.float64)
    batch, max_obs, k = hx.shape
    nb = np.ascontiguousarray(nobsl, dtype=np.int32)
    ri = np.ascontiguousarray(rdiaginv, dtype=np.float32)
    dp = np.ascontiguousarray(dep, dtype=np.float64)
    rl = np.ascontiguousarray(rloc, dtype=np.float64)
    trans = np.zeros(batch * k * k)
    wbar = np.zeros(batch * k)
    vals = [C.c_int32(batch), C.c_int32(0), C.c_int32(max_obs), C.c_int32(k)]
    rho, pf = C.c_double(inflation), C.c_int32(0)
    lib().letkf_core_batch_(*[C.byref(v) for v in vals], nb.ctypes.data, hx.ctypes.data,
                            ri.ctypes.data, dp.ctypes.data, rl.ctypes.data, trans.ctypes.data,
                            wbar.ctypes.data, C.byref(rho), C.byref(pf))
    return trans.reshape(batch, k, k).transpose(0, 2, 1).copy(), wbar.reshape(batch, k)


def letkf_core_r4(hxens: np.ndarray, rdiaginv, dep, rloc, inflation: float, pflag: int = 0):
    """letkf_core_r4_ (REAL(4) arrays) on host arrays.  hxens: (p, k).
    Returns float32 (trans[i, j], wbar)."""
    hx = np.ascontiguousarray(hxens, dtype=np.float32)
    p, k = hx.shape
    ri = np.ascontiguousarray(rdiaginv, dtype=np.float32)
    dp = np.ascontiguousarray(dep, dtype=np.float32)
    rl = np.ascontiguousarray(rloc, dtype=np.float32)
    trans = np.zeros(k * k, np.float32)
    wbar = np.zeros(k, np.float32)
    ik, ip, pf = C.c_int32(k), C.c_int32(p), C.c_int32(pflag)
    rho = C.c_double(inflation)
    lib().letkf_core_r4_(C.byref(ik), C.byref(ip), hx.ctypes.data, ri.ctypes.data,
                         dp.ctypes.data, rl.ctypes.data, trans.ctypes.data, wbar.ctypes.data,
                         C.byref(rho), C.byref(pf))
    return trans.reshape(k, k).T.copy(), wbar


def letkf_core_batch_r4(nobsl, hxens, rdiaginv, dep, rloc, inflation: float):
    """letkf_core_batch_r4_ on host arrays: hxens (batch, max_obs, k), all
    REAL(4).  Returns float32 (trans[b, i, j], wbar[b])."""
    hx = np.ascontiguousarray(hxens, dtype=np.float32)
    batch, max_obs, k = hx.shape
    nb = np.ascontiguousarray(nobsl, dtype=np.int32)
    ri = np.ascontiguousarray(rdiaginv, dtype=np.float32)
    dp = np.ascontiguousarray(dep, dtype=np.float32)
    rl = np.ascontiguousarray(rloc, dtype=np.float32)
    trans = np.zeros(batch * k * k, np.float32)
    wbar = np.zeros(batch * k, np.float32)
    vals = [C.c_int32(batch), C.c_int32(0), C.c_int32(max_obs), C.c_int32(k)]
    rho, pf = C.c_double(inflation), C.c_int32(0)
    lib().letkf_core_batch_r4_(*[C.byref(v) for v in vals], nb.ctypes.data, hx.ctypes.data,
                               ri.ctypes.data, dp.ctypes.data, rl.ctypes.data,
                               trans.ctypes.data, wbar.ctypes.data, C.byref(rho), C.byref(pf))
    return trans.reshape(batch, k, k).transpose(0, 2, 1).copy(), wbar.reshape(batch, k)


def post(k, nrow, nlat, nlon, grid_lat, polar_lat, gamma1, gamma2, xb, axb, xa, axa, xa_out,
         spread_b, spread_a, stream=None) -> int:
    return check(lib().dida_letkf_post(k, nrow, nlat, nlon, _ptr(grid_lat), float(polar_lat),
                                       float(gamma1), float(gamma2), _ptr(xb), _ptr(axb),
                                       _ptr(xa), _ptr(axa), _ptr(xa_out), _ptr(spread_b),
                                       _ptr(spread_a), _stream_ptr(stream)), "post")


def struct_atm2endnn(ens_start, n_start, xb_atm, xb, stream=None) -> int:
    """dida_struct_atm2endnn on device tensors: xb_atm (n, nx, ny, ens) (the
    coupler's member-fastest block), xb (ny, nx, n_max, ens_max) updated in
    place at 1-based member / row starts."""
    n, nx, ny, ens = xb_atm.shape
    ny2, nx2, n_max, ens_max = xb.shape
    if (ny2, nx2) != (ny, nx):
        raise ValueError("grid shapes differ")
    return check(lib().dida_struct_atm2endnn(ens, n, nx, ny, ens_max, n_max, int(ens_start),
                                             int(n_start), _ptr(xb_atm), _ptr(xb),
                                             _stream_ptr(stream)), "struct_atm2endnn")


def struct_endnn2atm(ens_start, n_start, xb, xb_atm, stream=None) -> int:
    """dida_struct_endnn2atm: the block of xb (ny, nx, n_max, ens_max) into
    xb_atm (n, nx, ny, ens)."""
    n, nx, ny, ens = xb_atm.shape
    ny2, nx2, n_max, ens_max = xb.shape
    if (ny2, nx2) != (ny, nx):
        raise ValueError("grid shapes differ")
    return check(lib().dida_struct_endnn2atm(ens, n, nx, ny, ens_max, n_max, int(ens_start),
                                             int(n_start), _ptr(xb), _ptr(xb_atm),
                                             _stream_ptr(stream)), "struct_endnn2atm")


def struct_atm2endnn_a(n_start, axb_atm, axb, stream=None) -> int:
    """Means: axb_atm (n, nx, ny) into axb (ny, nx, n_max)."""
    n, nx, ny = axb_atm.shape
    return check(lib().dida_struct_atm2endnn_a(n, nx, ny, axb.shape[2], int(n_start),
                                               _ptr(axb_atm), _ptr(axb), _stream_ptr(stream)),
                 "struct_atm2endnn_a")


def struct_endnn2atm_a(n_start, axb, axb_atm, stream=None) -> int:
    n, nx, ny = axb_atm.shape
    return check(lib().dida_struct_endnn2atm_a(n, nx, ny, axb.shape[2], int(n_start),
                                               _ptr(axb), _ptr(axb_atm), _stream_ptr(stream)),
                 "struct_endnn2atm_a")


def obs_operator(grid_lat, grid_lon, xb, axb, ps_row, vcoord, var_start, var_lev, vid,
                 rec_lat, rec_lon, rec_ph, yo, periodic_lon=True, stream=None, with_ph=False):
    """dida_letkf_obs_operator on device tensors.  xb (ncol, nrow, k), axb
    (ncol, nrow); vcoord: dict(type, c1, c2, ptop, p0) with c1/c2 device
    float64 tensors; rec_ph (nrec, onz); yo (nrec, onz, onv).  Returns
    DeviceObs-ready arrays (lat, lon, yb (nobs, k), dep), and the scalar obs
    pressures after them with with_ph."""
    import torch
    ncol, nrow, k = xb.shape
    nrec, onz = rec_ph.shape
    onv = len(vid)
    ns = nrec * onz * onv
    dev = xb.device
    lat = torch.empty(ns, dtype=torch.float64, device=dev)
    lon = torch.empty(ns, dtype=torch.float64, device=dev)
    yb = torch.empty(ns, k, dtype=torch.float64, device=dev)
    dep = torch.empty(ns, dtype=torch.float64, device=dev)
    ph = torch.empty(ns, dtype=torch.float64, device=dev) if with_ph else None
    c1 = vcoord["c1"]
    c2 = vcoord.get("c2", c1)
    vc = VertCoordC(int(vcoord["type"]), int(c1.numel()) - 1, c1.data_ptr(), c2.data_ptr(),
                    float(vcoord["ptop"]), float(vcoord.get("p0", 0.0)))
    vid_h = np.ascontiguousarray(vid, np.int32)
    vs = np.ascontiguousarray(var_start, np.int32)
    vl = np.ascontiguousarray(var_lev, np.int32)
    rc = ObsRecordsC(nrec, onz, onv, 0, vid_h.ctypes.data, rec_lat.data_ptr(), rec_lon.data_ptr(),
                     rec_ph.data_ptr(), yo.data_ptr())
    check(lib().dida_letkf_obs_operator(k, nrow, grid_lat.numel(), grid_lon.numel(),
                                        _ptr(grid_lat), _ptr(grid_lon), int(bool(periodic_lon)),
                                        _ptr(xb), _ptr(axb), int(ps_row), C.byref(vc), len(vs),
                                        vs.ctypes.data, vl.ctypes.data, C.byref(rc), _ptr(lat),
                                        _ptr(lon), _ptr(yb), _ptr(dep),
                                        None if ph is None else _ptr(ph), _stream_ptr(stream)),
          "obs_operator")
    return (lat, lon, yb, dep, ph) if with_ph else (lat, lon, yb, dep)


# ----------------------------------------------------------------------------
# Synthetic workloads (benchmark / test inputs)

def synth_cfg(nlat, nlon, nz, k, nobs, seed, cluster_frac=0.0, nclusters=0,
              cluster_radius_km=800.0, sig_T=1.0, sig_uv=1.5, rscale=(1.0, 1.0),
              pert=(1.0, 1.5, 1.0), pert_len_km=500.0, vert_corr=0.8, lmax_pert=24) -> SynthCfg:
    return SynthCfg(nlat, nlon, nz, k, nobs, nclusters, lmax_pert, 0, cluster_frac,
                    cluster_radius_km, sig_T, sig_uv, rscale[0], rscale[1], pert[0], pert[1],
                    pert[2], pert_len_km, vert_corr, seed)


def synth_generate(cfg: SynthCfg, device="cuda", fp32_state=False, stream=None) -> dict:
    import torch
    n = 1 + 3 * cfg.nz
    ncol = cfg.nlat * cfg.nlon
    f64, f32 = torch.float64, torch.float32
    scratch = torch.empty(int(synth_lib().dida_synth_scratch_bytes(C.byref(cfg))),
                          dtype=torch.uint8, device=device)
    out = {"grid_lat": torch.empty(cfg.nlat, dtype=f32, device=device),
           "grid_lon": torch.empty(cfg.nlon, dtype=f32, device=device),
           "axb": torch.empty(ncol, n, dtype=f64, device=device),
           "lat": torch.empty(cfg.nobs, dtype=f64, device=device),
           "lon": torch.empty(cfg.nobs, dtype=f64, device=device),
           "yb": torch.empty(cfg.nobs, cfg.k, dtype=f64, device=device),
           "dep": torch.empty(cfg.nobs, dtype=f64, device=device),
           "rinv": torch.empty(cfg.nobs, dtype=f32, device=device)}
    if fp32_state:
        out["xb"] = torch.empty(ncol, n, cfg.k, dtype=f32, device=device)
    else:
        out["xb"] = torch.empty(ncol, n, cfg.k, dtype=f64, device=device)
    rc = synth_lib().dida_synth_generate(
        C.byref(cfg), scratch.data_ptr(), out["grid_lat"].data_ptr(), out["grid_lon"].data_ptr(),
        None if fp32_state else out["xb"].data_ptr(), out["xb"].data_ptr() if fp32_state else None,
        out["axb"].data_ptr(), out["lat"].data_ptr(), out["lon"].data_ptr(),
        out["yb"].data_ptr(), out["dep"].data_ptr(), out["rinv"].data_ptr(),
        _stream_ptr(stream).value)
    if rc != 0:
        raise DidaError(f"dida_synth_generate failed: {rc}")
    torch.cuda.synchronize()
    del scratch
    return out


# Workloads of the benchmark configurations.
def synth_region(data: dict, cfg: SynthCfg, j0: int, ny: int, i0: int, nx: int) -> dict:
    """A regional domain cut from a global synthetic workload: the ny x nx
    grid columns from latitude index j0 and longitude index i0 (no wrap), and
    the obs inside the box the columns span.  Columns are numbered
    jl + ny * il as in the global layout."""
    import torch
    dev = data["xb"].device
    if not (0 <= j0 and j0 + ny <= cfg.nlat and 0 <= i0 and i0 + nx <= cfg.nlon):
        raise ValueError("region outside the grid")
    il = torch.arange(i0, i0 + nx, device=dev)
    jl = torch.arange(j0, j0 + ny, device=dev)
    cols = (jl[None, :] + cfg.nlat * il[:, None]).reshape(-1)
    dlat, dlon = 180.0 / cfg.nlat, 360.0 / cfg.nlon
    lat_lo = -90.0 + j0 * dlat
    lat_hi = lat_lo + ny * dlat
    lon_lo = i0 * dlon
    lon_hi = lon_lo + nx * dlon
    inside = ((data["lat"] >= lat_lo) & (data["lat"] < lat_hi) & (data["lon"] >= lon_lo) &
              (data["lon"] < lon_hi))
    sel = torch.nonzero(inside).reshape(-1)
    return {"grid_lat": data["grid_lat"][j0:j0 + ny].contiguous(),
            "grid_lon": data["grid_lon"][i0:i0 + nx].contiguous(),
            "xb": data["xb"][cols].contiguous(), "axb": data["axb"][cols].contiguous(),
            "lat": data["lat"][sel].contiguous(), "lon": data["lon"][sel].contiguous(),
            "yb": data["yb"][sel].contiguous(), "dep": data["dep"][sel].contiguous(),
            "rinv": data["rinv"][sel].contiguous()}


def config_synth(i: int, scale: float = 1.0) -> tuple[SynthCfg, dict]:
    """Synthetic config i (0..4) and its analysis parameters.  `scale` < 1
    shrinks the grid and obs count proportionally (test sizes)."""
    seed = 0xD1DA0000 + i

    def g(nlat, nlon, nobs):
        s = max(scale, 1e-6) ** 0.5
        return max(4, int(round(nlat * s))), max(8, int(round(nlon * s))), max(1, int(round(nobs * scale)))

    if i == 0:
        nlat, nlon, nobs = g(72, 144, 10000)
        cfg = synth_cfg(nlat, nlon, 8, 20, nobs, seed)
        prm = dict(local_dist=1000.0, inflation_factor=1.1)
    elif i == 1:
        nlat, nlon, nobs = g(180, 360, 100000)
        cfg = synth_cfg(nlat, nlon, 16, 32, nobs, seed, cluster_frac=0.7, nclusters=6)
        # L is not given for this workload; 200 km makes the local obs counts
        # span 0 to about 1500 as its description requires
        prm = dict(local_dist=200.0, inflation_factor=1.1)
    elif i == 2:
        nlat, nlon, nobs = g(360, 720, 400000)
        cfg = synth_cfg(nlat, nlon, 32, 64, nobs, seed, cluster_frac=0.7, nclusters=6,
                        rscale=(0.1, 10.0))
        prm = dict(local_dist=600.0, inflation_factor=1.0, rtps=1, rtps_alpha=0.8)
    elif i == 4:
        nlat, nlon, nobs = g(720, 1440, 2000000)
        cfg = synth_cfg(nlat, nlon, 32, 128, nobs, seed, cluster_frac=0.5, nclusters=6,
                        rscale=(0.5, 3.0))
        prm = dict(local_dist=600.0, inflation_factor=1.0)
    else:
        raise ValueError(f"no single-domain synthetic config {i}")
    return cfg, prm
