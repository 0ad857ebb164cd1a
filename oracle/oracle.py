"""ctypes bindings of the C oracle (liboracle_letkf.so) -- TEST INFRASTRUCTURE.

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, as the checker.  The product library never
loads it.  Each function names the reference routine it restates; the C file
carries the file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_letkf.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcObs(C.Structure):
    _fields_ = [("nobs", C.c_int32), ("lat", C.c_void_p), ("lon", C.c_void_p),
                ("yb", C.c_void_p), ("dep", C.c_void_p), ("rinv", C.c_void_p)]


class OrcVcoord(C.Structure):
    _fields_ = [("type", C.c_int32), ("nz", C.c_int32), ("c1", C.c_void_p), ("c2", C.c_void_p),
                ("ptop", C.c_double), ("p0", C.c_double)]


class OrcParams(C.Structure):
    _fields_ = [("k", C.c_int32), ("nrow", C.c_int32), ("nlat", C.c_int32),
                ("nlon", C.c_int32), ("max_obs", C.c_int32), ("gc_filter", C.c_int32),
                ("local_dist", C.c_double), ("obs_polar", C.c_double),
                ("inflation", C.c_double)]


class OrcBins(C.Structure):
    _fields_ = [("bin_start", C.c_void_p), ("nband", C.c_int32), ("nlonbin", C.c_int32),
                ("bin_deg", C.c_double), ("perm", C.c_void_p)]


def _bins_c(bins):
    """OrcBins of bins = (bin_start, nband, nlonbin, bin_deg[, perm]) and the
    arrays it points into (keep them alive while it is used).  perm maps a
    binned position to the caller's obs index; without it the binned list's
    own order stands for the caller's (MAX_OBS truncation)."""
    if bins is None:
        return None, ()
    bs = np.ascontiguousarray(bins[0], dtype=np.int32)
    pm = None if len(bins) < 5 or bins[4] is None else np.ascontiguousarray(bins[4], np.int32)
    b = OrcBins(bs.ctypes.data, int(bins[1]), int(bins[2]), float(bins[3]),
                None if pm is None else pm.ctypes.data)
    return b, (bs, pm)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_latlon_dist.restype = C.c_double
        L.orc_latlon_dist.argtypes = [C.c_double] * 4
        L.orc_schr2.restype = C.c_double
        L.orc_schr2.argtypes = [C.c_double]
        L.orc_schr.restype = C.c_double
        L.orc_schr.argtypes = [C.c_double]
        L.orc_syev_jacobi.restype = C.c_int
        L.orc_syev_jacobi.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.orc_letkf_core.restype = C.c_int
        L.orc_letkf_core.argtypes = [C.c_int, C.c_int, _dp, _fp, _dp, _dp, _dp, _dp,
                                     C.c_double]
        L.orc_xb_2_xa.restype = None
        L.orc_xb_2_xa.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_select_column.restype = C.c_int
        L.orc_select_column.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcObs), C.c_double,
                                        C.c_double, _ip, _dp]
        L.orc_select_column_binned.restype = C.c_int
        L.orc_select_column_binned.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcObs), _ip,
                                               C.c_int, C.c_int, C.c_double, C.c_void_p,
                                               C.c_double, C.c_double, _ip, _dp]
        L.orc_bin_keys.restype = C.c_int
        L.orc_bin_keys.argtypes = [C.POINTER(OrcObs), C.c_double, C.c_double, _ip,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_bin_sort.restype = None
        L.orc_bin_sort.argtypes = [C.c_int, _ip, C.c_int, _ip, _ip]
        L.orc_candidate_ranges.restype = C.c_int
        L.orc_candidate_ranges.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_int, C.c_int, _ip]
        L.orc_sweep.restype = C.c_int
        L.orc_sweep.argtypes = [C.POINTER(OrcParams), _fp, _fp, C.POINTER(OrcObs), _dp, _dp,
                                _dp, _dp, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.POINTER(OrcBins), C.c_int]
        L.orc_sweep_compact.restype = C.c_int
        L.orc_sweep_compact.argtypes = [C.POINTER(OrcParams), _fp, _fp, C.POINTER(OrcObs),
                                        C.POINTER(OrcBins), _ip, C.c_int32, _dp, _dp, _dp, _dp,
                                        _ip, C.c_int]
        L.orc_sweep_cubes.restype = C.c_int
        L.orc_sweep_cubes.argtypes = [C.POINTER(OrcParams), _fp, _fp, C.POINTER(OrcObs), _dp,
                                      _dp, _dp, _dp, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                      C.c_int]
        for nm in ("orc_struct_atm2endnn", "orc_struct_endnn2atm"):
            getattr(L, nm).restype = None
            getattr(L, nm).argtypes = [C.c_int] * 8 + [_dp, _dp]
        for nm in ("orc_struct_atm2endnn_a", "orc_struct_endnn2atm_a"):
            getattr(L, nm).restype = None
            getattr(L, nm).argtypes = [C.c_int] * 5 + [_dp, _dp]
        L.orc_centre_obs.restype = None
        L.orc_centre_obs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_sweep_vertical.restype = C.c_int
        L.orc_sweep_vertical.argtypes = [C.POINTER(OrcParams), _fp, _fp, C.POINTER(OrcObs), _dp,
                                         _dp, _dp, _dp, _dp, C.c_void_p, C.POINTER(OrcVcoord),
                                         C.c_int, C.c_int, _ip, _ip, C.c_double, C.c_int,
                                         C.c_int, C.c_int, C.c_int]
        L.orc_obs_operator.restype = None
        L.orc_obs_operator.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, C.c_int,
                                       _dp, _dp, C.c_int, C.POINTER(OrcVcoord), _ip, _ip,
                                       C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp, _dp]
        L.orc_post.restype = None
        L.orc_post.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_double,
                               C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp,
                               _dp, _dp]
        _lib = L
    return _lib


# ----------------------------------------------------------------------------
def latlon_dist(x1, y1, x2, y2) -> float:
    """pro_info.F90:148-161 latlon_dist_r8."""
    return lib().orc_latlon_dist(float(x1), float(y1), float(x2), float(y2))


def schr2(r) -> float:
    """letkf_yb.F90:245-269."""
    return lib().orc_schr2(float(r))


def schr(r) -> float:
    """letkf_yb.F90:270-285."""
    return lib().orc_schr(float(r))


def syev(a: np.ndarray, max_sweeps: int = 100):
    """Symmetric eigensolve (lower triangle), ascending.  Returns (w, V, sweeps)."""
    n = a.shape[0]
    buf = np.array(np.asarray(a, dtype=np.float64).T, order="C").reshape(-1)  # col-major
    w = np.zeros(n)
    s = lib().orc_syev_jacobi(n, buf, w, max_sweeps)
    return w, buf.reshape(n, n).T.copy(), s


def letkf_core(hxens: np.ndarray, rdiaginv: np.ndarray, dep: np.ndarray, rloc: np.ndarray,
               inflation: float):
    """letkf_core.F90:14-233.  hxens: (p, k) array (row o = obs o's ensemble
    vector, i.e. the reference's column-major (k, p)).  Returns (trans, wbar,
    sweeps) with trans[i, j] = W_ij + wbar_i."""
    hx = np.ascontiguousarray(hxens, dtype=np.float64)
    p, k = hx.shape if hx.ndim == 2 and hx.size else (0, hx.shape[-1])
    trans = np.zeros(k * k)
    wbar = np.zeros(k)
    s = lib().orc_letkf_core(k, p, hx.reshape(-1) if p else np.zeros(1),
                             np.ascontiguousarray(rdiaginv, dtype=np.float32) if p else np.zeros(1, np.float32),
                             np.ascontiguousarray(dep, dtype=np.float64) if p else np.zeros(1),
                             np.ascontiguousarray(rloc, dtype=np.float64) if p else np.zeros(1),
                             trans, wbar, float(inflation))
    return trans.reshape(k, k).T.copy(), wbar, s


def letkf_core_r4(hxens: np.ndarray, rdiaginv: np.ndarray, dep: np.ndarray,
                  rloc: np.ndarray, inflation: float):
    """letkf_core.F90:14-233 built with yb_type = ayb_type = real(4)
    (da_config.inc:7-10): hxens, dep, rloc, trans, overline_w and every work
    array REAL(4); the products are sgemm (:78, :141, :168, :205), the
    eigensolve ssyevd (:101), here numpy's float32 matmul and LAPACK syevd
    through numpy.linalg.eigh.  Mixed expressions with REAL(8) operands
    (:111, :201) are evaluated in double and rounded on assignment, as
    Fortran does.  hxens: (p, k).  Returns float32 (trans[i, j], wbar)."""
    f4 = np.float32
    hx = np.asarray(hxens, f4).reshape(-1, np.asarray(hxens).shape[-1])
    p, k = hx.shape
    rr = np.sqrt(np.asarray(rdiaginv, f4) * np.asarray(rloc, f4))        # :64-66
    h = hx * rr[:, None]                                                  # :67
    work1 = (h.T @ h).astype(f4)                                          # :78
    eival, eivec = np.linalg.eigh(work1)                                  # :101
    eival = (eival.astype(np.float64) * 1.0 + (k - 1) / inflation).astype(f4)  # :110-113
    work1 = (eivec / eival[None, :]).astype(f4)                           # :129-135
    pa = (work1 @ eivec.T).astype(f4)                                     # :141
    h = h * rr[:, None]                                                   # :146-148
    work2 = (pa @ h.T).astype(f4)                                         # :168
    dep4 = np.asarray(dep, f4)
    wbar = np.zeros(k, f4)
    for o in range(p):                                                    # :190-195
        wbar = (wbar + work2[:, o] * dep4[o]).astype(f4)
    rho = np.sqrt((k - 1) / eival.astype(np.float64)).astype(f4)          # :199-200
    work1 = (eivec * rho[None, :]).astype(f4)
    trans = (work1 @ eivec.T).astype(f4)                                  # :205
    trans = (trans + wbar[:, None]).astype(f4)                            # :221-225
    return trans, wbar


def xb_2_xa(xb: np.ndarray, axb: np.ndarray, trans: np.ndarray):
    """letkf_compute.F90:231-301.  xb: (ncells, k), axb: (ncells,), trans: (k, k)
    with trans[i, j] as returned by letkf_core.  Returns (xa, axa)."""
    ncells, k = xb.shape
    xa = np.zeros(ncells * k)
    axa = np.zeros(ncells)
    tr = np.ascontiguousarray(np.asarray(trans, dtype=np.float64).T).reshape(-1)
    lib().orc_xb_2_xa(k, ncells, np.ascontiguousarray(xb, dtype=np.float64).reshape(-1),
                      np.ascontiguousarray(axb, dtype=np.float64), tr, xa, axa)
    return xa.reshape(ncells, k), axa


class Obs:
    """Scalar obs records in scan order; keeps the numpy buffers alive."""

    def __init__(self, lat, lon, yb, dep, rinv):
        self.lat = np.ascontiguousarray(lat, dtype=np.float64)
        self.lon = np.ascontiguousarray(lon, dtype=np.float64)
        self.yb = np.ascontiguousarray(yb, dtype=np.float64)  # (nobs, k)
        self.dep = np.ascontiguousarray(dep, dtype=np.float64)
        self.rinv = np.ascontiguousarray(rinv, dtype=np.float32)
        self.c = OrcObs(len(self.lat), self.lat.ctypes.data, self.lon.ctypes.data,
                        self.yb.ctypes.data, self.dep.ctypes.data, self.rinv.ctypes.data)

    def permuted(self, perm):
        return Obs(self.lat[perm], self.lon[perm], self.yb[perm], self.dep[perm],
                   self.rinv[perm])


def params(k, nrow, nlat, nlon, local_dist, inflation, max_obs=10000, gc_filter=True,
           obs_polar=90.0):
    return OrcParams(k, nrow, nlat, nlon, max_obs, 1 if gc_filter else 0, float(local_dist),
                     float(obs_polar), float(inflation))


def select_column(prm: OrcParams, obs: Obs, clat: float, clon: float):
    """gen_yb_dis over one global block (full scan in the given order)."""
    idx = np.zeros(max(prm.max_obs, 1), np.int32)
    rl = np.zeros(max(prm.max_obs, 1))
    p = lib().orc_select_column(C.byref(prm), C.byref(obs.c), float(clat), float(clon), idx, rl)
    return idx[:p].copy(), rl[:p].copy()


def bin_obs(obs: Obs, bin_deg: float, obs_polar: float):
    """Bin keys and the stable sort into binned order.  Returns
    (perm, bin_start, nband, nlonbin)."""
    key = np.zeros(max(len(obs.lat), 1), np.int32)
    nb, nl = C.c_int32(), C.c_int32()
    nbins = lib().orc_bin_keys(C.byref(obs.c), float(bin_deg), float(obs_polar), key,
                               C.byref(nb), C.byref(nl))
    perm = np.zeros(max(len(obs.lat), 1), np.int32)
    bin_start = np.zeros(nbins + 2, np.int32)
    lib().orc_bin_sort(len(obs.lat), key, nbins, perm, bin_start)
    return perm[:len(obs.lat)].copy(), bin_start, nb.value, nl.value


def select_column_binned(prm, obs_binned: Obs, bin_start, nband, nlonbin, bin_deg, clat, clon,
                         perm=None):
    """Selection over the binned list; perm (binned position -> caller index)
    decides which obs a column capped at MAX_OBS keeps (the smallest caller
    indices).  Returns binned indices and weights in bin order."""
    idx = np.zeros(max(prm.max_obs, 1), np.int32)
    rl = np.zeros(max(prm.max_obs, 1))
    pm = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
    p = lib().orc_select_column_binned(C.byref(prm), C.byref(obs_binned.c),
                                       np.ascontiguousarray(bin_start, dtype=np.int32),
                                       nband, nlonbin, float(bin_deg),
                                       None if pm is None else pm.ctypes.data, float(clat),
                                       float(clon), idx, rl)
    if p < 0:
        raise MemoryError("orc_select_column_binned")
    return idx[:p].copy(), rl[:p].copy()


def sweep(prm: OrcParams, grid_lat, grid_lon, obs: Obs, xb, axb, cols=None, bins=None,
          nthreads=0, want_trans=False):
    """letkf_compute_nooverlap (cube_n = 1).  xb: (ncol, nrow, k) flat-able,
    axb: (ncol, nrow).  Returns dict with xa, axa, nobsl (and trans, wbar)."""
    k, n = prm.k, prm.nrow
    ncol = prm.nlat * prm.nlon
    xb = np.ascontiguousarray(xb, dtype=np.float64).reshape(-1)
    axb = np.ascontiguousarray(axb, dtype=np.float64).reshape(-1)
    xa = np.zeros(ncol * n * k)
    axa = np.zeros(ncol * n)
    nobsl = np.full(ncol, -1, np.int32)
    trans = np.zeros(ncol * k * k) if want_trans else None
    wbar = np.zeros(ncol * k) if want_trans else None
    cols_a = None if cols is None else np.ascontiguousarray(cols, dtype=np.int32)
    bins_c, _keep = _bins_c(bins)
    nfail = lib().orc_sweep(C.byref(prm), np.ascontiguousarray(grid_lat, dtype=np.float32),
                            np.ascontiguousarray(grid_lon, dtype=np.float32), C.byref(obs.c),
                            xb, axb, xa, axa,
                            None if cols_a is None else cols_a.ctypes.data,
                            0 if cols_a is None else len(cols_a), nobsl.ctypes.data,
                            None if trans is None else trans.ctypes.data,
                            None if wbar is None else wbar.ctypes.data,
                            None if bins_c is None else C.byref(bins_c), int(nthreads))
    out = {"xa": xa.reshape(ncol, n, k), "axa": axa.reshape(ncol, n), "nobsl": nobsl,
           "nfail": nfail}
    if want_trans:
        out["trans"] = trans.reshape(ncol, k, k).transpose(0, 2, 1).copy()
        out["wbar"] = wbar.reshape(ncol, k)
    return out


def sweep_compact(prm: OrcParams, grid_lat, grid_lon, obs: Obs, bins, cols, xb_c, axb_c,
                  nthreads=0):
    """Sweep over the listed columns; xb_c (ncols, nrow, k), axb_c (ncols, nrow)
    hold those columns.  bins = (bin_start, nband, nlonbin, bin_deg[, perm]) or None.
    Returns (xa_c, axa_c, nobsl, nfail)."""
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    xb_c = np.ascontiguousarray(xb_c, dtype=np.float64)
    axb_c = np.ascontiguousarray(axb_c, dtype=np.float64)
    xa = np.zeros_like(xb_c)
    axa = np.zeros_like(axb_c)
    nobsl = np.zeros(len(cols), np.int32)
    bins_c, _keep = _bins_c(bins)
    nfail = lib().orc_sweep_compact(C.byref(prm), np.ascontiguousarray(grid_lat, np.float32),
                                    np.ascontiguousarray(grid_lon, np.float32), C.byref(obs.c),
                                    None if bins_c is None else C.byref(bins_c), cols, len(cols),
                                    xb_c.reshape(-1), axb_c.reshape(-1), xa.reshape(-1),
                                    axa.reshape(-1), nobsl, int(nthreads))
    return xa, axa, nobsl, nfail


def centre_obs(hxb: np.ndarray, yo: np.ndarray):
    """letkf_obs.F90:156-178.  hxb: (nobs, k).  Returns (yb, dep)."""
    nobs, k = hxb.shape
    yb = np.zeros(nobs * k)
    dep = np.zeros(nobs)
    lib().orc_centre_obs(k, nobs, np.ascontiguousarray(hxb, dtype=np.float64).reshape(-1),
                         np.ascontiguousarray(yo, dtype=np.float64), yb, dep)
    return yb.reshape(nobs, k), dep


def post(mode, grid_lat, nlat, nlon, polar_lat, gamma1, gamma2, alpha, xb, axb, xa, axa):
    """module_da_parallel.F90:216-259 (mode 0) and RTPS (mode 1).
    xb, xa: (ncol, nrow, k); axb, axa: (ncol, nrow)."""
    ncol, nrow, k = xb.shape
    out = np.zeros(xb.size)
    sb = np.zeros(ncol * nrow)
    sa = np.zeros(ncol * nrow)
    lib().orc_post(int(mode), k, nrow, nlat, nlon, np.ascontiguousarray(grid_lat, np.float32),
                   float(polar_lat), float(gamma1), float(gamma2), float(alpha),
                   np.ascontiguousarray(xb, np.float64).reshape(-1),
                   np.ascontiguousarray(axb, np.float64).reshape(-1),
                   np.ascontiguousarray(xa, np.float64).reshape(-1),
                   np.ascontiguousarray(axa, np.float64).reshape(-1), out, sb, sa)
    return out.reshape(xb.shape), sb.reshape(ncol, nrow), sa.reshape(ncol, nrow)


def obs_operator(grid_lat, grid_lon, xb, axb, ps_row, vcoord, var_start, var_lev, vid,
                 rec_lat, rec_lon, rec_ph, yo, periodic_lon=True):
    """letkf_obs.F90:120-178 / obs_make.F90:233-381 on host arrays.
    xb: (ncol, nrow, k) perturbations, axb: (ncol, nrow); vcoord: dict with
    type (0 sigma, 1 hybrid), c1, c2 (nz + 1 half levels), ptop, p0;
    rec_ph: (nrec, onz); yo: (nrec, onz, onv).  Returns (lat, lon, yb, dep)
    per scalar obs in (record, level, variable) order, yb (nobs, k)."""
    xb = np.ascontiguousarray(xb, np.float64)
    ncol, nrow, k = xb.shape
    glat = np.ascontiguousarray(grid_lat, np.float32)
    glon = np.ascontiguousarray(grid_lon, np.float32)
    c1 = np.ascontiguousarray(vcoord["c1"], np.float64)
    c2 = np.ascontiguousarray(vcoord.get("c2", vcoord["c1"]), np.float64)
    vc = OrcVcoord(int(vcoord["type"]), len(c1) - 1, c1.ctypes.data, c2.ctypes.data,
                   float(vcoord["ptop"]), float(vcoord.get("p0", 0.0)))
    rec_ph = np.ascontiguousarray(rec_ph, np.float64)
    nrec, onz = rec_ph.shape
    vid = np.ascontiguousarray(vid, np.int32)
    onv = len(vid)
    ns = nrec * onz * onv
    lat, lon, dep = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    yb = np.zeros(ns * k)
    lib().orc_obs_operator(k, nrow, len(glat), len(glon), glat, glon, int(bool(periodic_lon)),
                           xb.reshape(-1), np.ascontiguousarray(axb, np.float64).reshape(-1),
                           int(ps_row), C.byref(vc), np.ascontiguousarray(var_start, np.int32),
                           np.ascontiguousarray(var_lev, np.int32), nrec, onz, onv, vid,
                           np.ascontiguousarray(rec_lat, np.float64),
                           np.ascontiguousarray(rec_lon, np.float64), rec_ph.reshape(-1),
                           np.ascontiguousarray(yo, np.float64).reshape(-1), lat, lon, yb, dep)
    return lat, lon, yb.reshape(ns, k), dep


def sweep_cubes(prm: OrcParams, grid_lat, grid_lon, obs: Obs, xb, axb, cube_n, overlap,
                bins=None, nthreads=0, want_trans=False):
    """letkf_compute_nooverlap with cube_n x cube_n cubes (overlap = 0) or
    letkf_compute with overlapping cubes averaged per column (overlap > 0).
    Returns dict(xa, axa, nfail[, trans (nanal, k, k) as trans[a][i][e]])."""
    k, n = prm.k, prm.nrow
    ncol = prm.nlat * prm.nlon
    step = cube_n - overlap if overlap > 0 else cube_n
    if cube_n < 1 or overlap < 0 or step < 1:
        raise ValueError("cube_n / overlap out of range")
    nanal = -(-prm.nlat // step) * -(-prm.nlon // step)
    xa = np.zeros(ncol * n * k)
    axa = np.zeros(ncol * n)
    trans = np.zeros(nanal * k * k) if want_trans else None
    bins_c, _keep = _bins_c(bins)
    nfail = lib().orc_sweep_cubes(
        C.byref(prm), np.ascontiguousarray(grid_lat, np.float32),
        np.ascontiguousarray(grid_lon, np.float32), C.byref(obs.c),
        np.ascontiguousarray(xb, np.float64).reshape(-1),
        np.ascontiguousarray(axb, np.float64).reshape(-1), xa, axa,
        None if bins_c is None else C.cast(C.pointer(bins_c), C.c_void_p), int(cube_n),
        int(overlap), None if trans is None else trans.ctypes.data, int(nthreads))
    if nfail < 0:
        raise ValueError("cube_n / overlap out of range")
    out = {"xa": xa.reshape(ncol, n, k), "axa": axa.reshape(ncol, n), "nfail": nfail}
    if want_trans:
        out["trans"] = trans.reshape(nanal, k, k).transpose(0, 2, 1).copy()
    return out


def sweep_vertical(prm: OrcParams, grid_lat, grid_lon, obs: Obs, obs_ph, xb, axb, vcoord,
                   ps_row, var_start, var_lev, vert_dist, cube_nz, gc_filter_vertical=True,
                   bins=None, nthreads=0, cube_n=1):
    """letkf_compute_nooverlap_vertical with cube_n x cube_n horizontal cubes.
    obs_ph: obs pressure in the order of obs; vcoord as obs_operator.  Returns
    dict(xa, axa, nfail)."""
    k, n = prm.k, prm.nrow
    ncol = prm.nlat * prm.nlon
    xa = np.zeros(ncol * n * k)
    axa = np.zeros(ncol * n)
    c1 = np.ascontiguousarray(vcoord["c1"], np.float64)
    c2 = np.ascontiguousarray(vcoord.get("c2", vcoord["c1"]), np.float64)
    vc = OrcVcoord(int(vcoord["type"]), len(c1) - 1, c1.ctypes.data, c2.ctypes.data,
                   float(vcoord["ptop"]), float(vcoord.get("p0", 0.0)))
    bins_c, _keep = _bins_c(bins)
    vs = np.ascontiguousarray(var_start, np.int32)
    vl = np.ascontiguousarray(var_lev, np.int32)
    nfail = lib().orc_sweep_vertical(
        C.byref(prm), np.ascontiguousarray(grid_lat, np.float32),
        np.ascontiguousarray(grid_lon, np.float32), C.byref(obs.c),
        np.ascontiguousarray(obs_ph, np.float64), np.ascontiguousarray(xb, np.float64).reshape(-1),
        np.ascontiguousarray(axb, np.float64).reshape(-1), xa, axa,
        None if bins_c is None else C.cast(C.pointer(bins_c), C.c_void_p), C.byref(vc),
        int(ps_row), len(vs), vs, vl, float(vert_dist), int(cube_nz), int(bool(gc_filter_vertical)),
        int(cube_n), int(nthreads))
    return {"xa": xa.reshape(ncol, n, k), "axa": axa.reshape(ncol, n), "nfail": nfail}


def struct_atm2endnn(ens_start, n_start, xb_atm, xb):
    """data_trans.F90 struct_atm2endnn.  xb_atm: C array (n, nx, ny, ens)
    (= Fortran xb(ens, ny, nx, n)); xb: C array (ny, nx, n_max, ens_max),
    updated in place (1-based starts)."""
    n, nx, ny, ens = xb_atm.shape
    ny2, nx2, n_max, ens_max = xb.shape
    assert (ny2, nx2) == (ny, nx) and xb.flags.c_contiguous and xb.dtype == np.float64
    lib().orc_struct_atm2endnn(ens, n, nx, ny, ens_max, n_max, ens_start, n_start,
                               np.ascontiguousarray(xb_atm, np.float64).reshape(-1),
                               xb.reshape(-1))


def struct_endnn2atm(ens, n, ens_start, n_start, xb):
    """struct_endnn2atm: returns the (n, nx, ny, ens) block of xb (C array
    (ny, nx, n_max, ens_max))."""
    ny, nx, n_max, ens_max = xb.shape
    out = np.zeros(n * nx * ny * ens)
    lib().orc_struct_endnn2atm(ens, n, nx, ny, ens_max, n_max, ens_start, n_start,
                               np.ascontiguousarray(xb, np.float64).reshape(-1), out)
    return out.reshape(n, nx, ny, ens)


def struct_atm2endnn_a(n_start, axb_atm, axb):
    """struct_atm2endnn_a: axb_atm C (n, nx, ny) into axb C (ny, nx, n_max)."""
    n, nx, ny = axb_atm.shape
    assert axb.flags.c_contiguous and axb.dtype == np.float64
    lib().orc_struct_atm2endnn_a(n, nx, ny, axb.shape[2], n_start,
                                 np.ascontiguousarray(axb_atm, np.float64).reshape(-1),
                                 axb.reshape(-1))


def struct_endnn2atm_a(n, n_start, axb):
    ny, nx, n_max = axb.shape
    out = np.zeros(n * nx * ny)
    lib().orc_struct_endnn2atm_a(n, nx, ny, n_max, n_start,
                                 np.ascontiguousarray(axb, np.float64).reshape(-1), out)
    return out.reshape(n, nx, ny)
