"""GPU parity on the benchmark workloads themselves, at reduced grids.

Config 2 (k = 64, 70 % of the obs in six clusters, R x 0.1-10, L = 600 km,
RTPS 0.8, MAX_OBS 10 000, 32 levels) at half its obs count and columns, so
the cluster columns still reach MAX_OBS; config 4 (k = 128, half the obs in
clusters, R x 0.5-3, L = 600 km, 32 levels) at a tenth, in the chained form,
with the state in FP64 and in FP32.  Every column is analysed on the GPU;
the oracle (the reference's CPU path restated) analyses a sample of columns
-- every capped column up to a limit and a strided set -- and the
per-variable relative error of the analysis-mean increment and of Xa' on
the sample is held to the bar of the form: FP64 forms 1e-10, split
half-precision Grams 1e-4 (letkf_compute.F90:12-126 is the path restated).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _var_rows(nz):
    return {"ps": [0], "u": list(range(1, 1 + nz)), "v": list(range(1 + nz, 1 + 2 * nz)),
            "T": list(range(1 + 2 * nz, 1 + 3 * nz))}


def _sample(nobsl, max_obs, ncol, ncap=48, nstride=160):
    capped = np.nonzero(nobsl >= max_obs)[0]
    a = capped[np.linspace(0, len(capped) - 1, min(ncap, len(capped))).astype(int)] \
        if len(capped) else np.zeros(0, np.int64)
    b = np.arange(0, ncol, max(1, ncol // nstride))
    return np.unique(np.concatenate([a, b])).astype(np.int32)


def _oracle(oracle, cfg, prm, h, cols, xb_c, axb_c):
    n = 1 + 3 * cfg.nz
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bs, nb, nl = oracle.bin_obs(o, 3.0, 90.0)
    op = oracle.params(cfg.k, n, cfg.nlat, cfg.nlon, prm["local_dist"],
                       prm["inflation_factor"])
    xa, axa, nobsl, nfail = oracle.sweep_compact(op, h["grid_lat"], h["grid_lon"],
                                                 o.permuted(perm), (bs, nb, nl, 3.0, perm),
                                                 cols, xb_c, axb_c)
    assert nfail == 0
    if prm.get("rtps"):
        # relax to prior spread on the sampled columns (orc_post mode 1)
        ncs = len(cols)
        out, _, _ = oracle.post(1, np.zeros(ncs, np.float32), ncs, 1, 90.0, 0, 0,
                                prm["rtps_alpha"], xb_c, axb_c, xa, axa)
        xa = out - axa[..., None]
    return xa, axa, nobsl


def _errors(cfg, xa_g, axa_g, xa_o, axa_o, axb_c):
    out = {}
    for var, rows in _var_rows(cfg.nz).items():
        inc_o = axa_o[:, rows] - axb_c[:, rows]
        inc_g = axa_g[:, rows] - axb_c[:, rows]
        out[var] = np.linalg.norm(inc_g - inc_o) / np.linalg.norm(inc_o)
        out[var + "_xa"] = np.linalg.norm(xa_g[:, rows] - xa_o[:, rows]) / \
            np.linalg.norm(xa_o[:, rows])
    return out


def _run(dl, torch, cfg, prm, data, form, gram, f32=False):
    n = 1 + 3 * cfg.nz
    p = dl.default_params(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon, form=form,
                          gram=gram, **prm)
    eng = dl.Engine(p, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xb = data["xb"].float() if f32 else data["xb"]
    xa = torch.empty_like(xb)
    axa = torch.empty_like(data["axb"])
    eng.sweep(obs, xb, data["axb"], xa, axa)     # dida_letkf_sweep_f32 for float32 state
    st = eng.stats()
    nobsl = eng.debug(cfg.nobs)["nobsl"]
    eng.close()
    return xa, axa, st, nobsl


CONFIG2 = [("plain_fp64", 0, 0, 1e-10), ("batched_fp64", 1, 0, 1e-10),
           ("batched_bf16x3", 1, 1, 1e-4), ("batched_fp16x3", 1, 2, 1e-4)]


def test_config2_workload_against_oracle(cuda, dl, oracle):
    torch = cuda
    cfg, prm = dl.config_synth(2, 0.5)
    print(f"config 2 at half: {cfg.nlat}x{cfg.nlon}x{cfg.nz} k={cfg.k} nobs={cfg.nobs}",
          flush=True)
    data = dl.synth_generate(cfg)
    h = {k: data[k].cpu().numpy() for k in ("lat", "lon", "yb", "dep", "rinv", "grid_lat",
                                            "grid_lon")}
    ref = None
    for name, form, gram, tol in CONFIG2:
        print(f"  {name}", flush=True)
        xa, axa, st, nobsl = _run(dl, torch, cfg, prm, data, form, gram)
        assert st["status"] == 0 and st["nonconverged"] == 0
        if ref is None:
            assert st["nobsl_capped"] > 0       # MAX_OBS = 10000 binds in the clusters
            cols = _sample(nobsl, 10000, cfg.nlat * cfg.nlon)
            idx = torch.from_numpy(cols.astype(np.int64)).cuda()
            xb_c = data["xb"][idx].cpu().numpy()
            axb_c = data["axb"][idx].cpu().numpy()
            xa_o, axa_o, nobsl_o = _oracle(oracle, cfg, prm, h, cols, xb_c, axb_c)
            assert np.array_equal(nobsl_o, nobsl[cols])
            assert (nobsl_o >= 10000).sum() > 0
            ref = (xa_o, axa_o)
        idx = torch.from_numpy(cols.astype(np.int64)).cuda()
        err = _errors(cfg, xa[idx].cpu().numpy(), axa[idx].cpu().numpy(), ref[0], ref[1],
                      axb_c)
        print(f"    errors vs oracle: { {k: f'{v:.2e}' for k, v in err.items()} }", flush=True)
        assert max(err.values()) <= tol, (name, err)
        del xa, axa
        torch.cuda.empty_cache()


CONFIG4 = [("chained_fp64", 0, False, 1e-10), ("chained_bf16x3_f32", 1, True, 1e-4),
           ("chained_fp64_f32", 0, True, 1e-10)]


def test_config4_workload_against_oracle(cuda, dl, oracle):
    torch = cuda
    cfg, prm = dl.config_synth(4, 0.1)
    print(f"config 4 at a tenth: {cfg.nlat}x{cfg.nlon}x{cfg.nz} k={cfg.k} nobs={cfg.nobs}",
          flush=True)
    data = dl.synth_generate(cfg)
    h = {k: data[k].cpu().numpy() for k in ("lat", "lon", "yb", "dep", "rinv", "grid_lat",
                                            "grid_lon")}
    refs = {}
    cols = None
    for name, gram, f32, tol in CONFIG4:
        print(f"  {name}", flush=True)
        xa, axa, st, nobsl = _run(dl, torch, cfg, prm, data, 1, gram, f32=f32)
        assert st["status"] == 0 and st["nonconverged"] == 0
        if cols is None:
            assert st["nobsl_capped"] > 0
            cols = _sample(nobsl, 10000, cfg.nlat * cfg.nlon, ncap=24, nstride=96)
            idx = torch.from_numpy(cols.astype(np.int64)).cuda()
            axb_c = data["axb"][idx].cpu().numpy()
        if f32 not in refs:
            # the FP32 state's oracle runs on the same (rounded) values in FP64
            xb_c = (data["xb"][idx].float().double() if f32 else data["xb"][idx]).cpu().numpy()
            xa_o, axa_o, nobsl_o = _oracle(oracle, cfg, prm, h, cols, xb_c, axb_c)
            assert np.array_equal(nobsl_o, nobsl[cols])
            refs[f32] = (xa_o, axa_o)
        xa_o, axa_o = refs[f32]
        err = _errors(cfg, xa[idx].double().cpu().numpy(), axa[idx].cpu().numpy(), xa_o, axa_o,
                      axb_c)
        print(f"    errors vs oracle: { {k: f'{v:.2e}' for k, v in err.items()} }", flush=True)
        # Xa' of the FP32 state is rounded once to FP32 on store
        xtol = max(tol, 1e-6) if f32 else tol
        for k, v in err.items():
            assert v <= (xtol if k.endswith("_xa") else tol), (name, k, v)
        del xa, axa
        torch.cuda.empty_cache()
