"""GPU parity of the whole sweep (selection -> transform -> update) through the
C ABI, against the oracle on the same inputs.

Bars: bin permutation, local obs counts and selection index lists identical;
Gaspari-Cohn weights to 1e-12; per-variable analysis-increment and Xa'
relative error <= 1e-10 (FP64 plain form).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _setup(dl, torch, cfg, **over):
    data = dl.synth_generate(cfg)
    n = 1 + 3 * cfg.nz
    p = dict(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon)
    p.update(over)
    params = dl.default_params(**p)
    eng = dl.Engine(params, nobs_max=max(cfg.nobs, 1))
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa = torch.empty_like(data["xb"])
    axa = torch.empty_like(data["axb"])
    return data, params, eng, obs, xa, axa


def _var_rows(nz):
    return {"ps": [0], "u": list(range(1, 1 + nz)), "v": list(range(1 + nz, 1 + 2 * nz)),
            "T": list(range(1 + 2 * nz, 1 + 3 * nz))}


def _rel_err(a, b):
    d = np.linalg.norm(a - b)
    return d / max(np.linalg.norm(b), 1e-300)


def _oracle_sweep(oracle, data_h, params, bin_deg, cols=None):
    o = oracle.Obs(data_h["lat"], data_h["lon"], data_h["yb"], data_h["dep"], data_h["rinv"])
    perm, bin_start, nband, nlonbin = oracle.bin_obs(o, bin_deg, params.obs_polar)
    ob = o.permuted(perm)
    prm = oracle.params(params.nanals, params.nrow, params.nlat, params.nlon,
                        params.local_dist, params.inflation_factor, max_obs=params.max_obs,
                        gc_filter=bool(params.gc_filter), obs_polar=params.obs_polar)
    out = oracle.sweep(prm, data_h["grid_lat"], data_h["grid_lon"], ob, data_h["xb"],
                       data_h["axb"], cols=cols, bins=(bin_start, nband, nlonbin, bin_deg, perm))
    return out, perm, bin_start, nband, nlonbin, ob, prm


def _check_selection(oracle, dbg, perm, bin_start, nband, nlonbin, ob, prm, bin_deg, glat, glon,
                     cols):
    assert np.array_equal(dbg["perm"], perm)
    nlat = prm.nlat
    for col in cols:
        clat, clon = glat[col % nlat], glon[col // nlat]
        idx, w = oracle.select_column_binned(prm, ob, bin_start, nband, nlonbin, bin_deg,
                                             clat, clon, perm=perm)
        off, p = int(dbg["offset"][col]), int(dbg["nobsl"][col])
        assert p == len(idx), col
        assert np.array_equal(dbg["csr_idx"][off:off + p], idx), col
        assert np.abs(dbg["csr_w"][off:off + p] - w).max(initial=0.0) <= 1e-12, col


@pytest.mark.parametrize("case", ["uniform", "clustered_capped", "k128", "short_support"])
def test_sweep_matches_oracle(cuda, dl, oracle, case):
    torch = cuda
    bin_deg = 3.0
    if case == "uniform":
        cfg = dl.synth_cfg(36, 72, 4, 20, 2500, seed=0xD1DA0000)
        over = dict(local_dist=1000.0, inflation_factor=1.1, bin_deg=bin_deg)
    elif case == "short_support":
        # 2L/a below 0.197 rad: the weights' shorter arcsine series
        cfg = dl.synth_cfg(36, 72, 3, 16, 4000, seed=0xD1DA0005, cluster_frac=0.5,
                           nclusters=2, cluster_radius_km=500.0)
        over = dict(local_dist=600.0, inflation_factor=1.0, bin_deg=bin_deg)
    elif case == "k128":
        cfg = dl.synth_cfg(12, 24, 2, 128, 1500, seed=0xD1DA0004, cluster_frac=0.5,
                           nclusters=2, cluster_radius_km=800.0, rscale=(0.5, 3.0))
        over = dict(local_dist=1500.0, inflation_factor=1.05, bin_deg=bin_deg)
    else:
        cfg = dl.synth_cfg(45, 90, 3, 32, 6000, seed=0xD1DA0001, cluster_frac=0.7, nclusters=3,
                           cluster_radius_km=600.0, rscale=(0.1, 10.0))
        over = dict(local_dist=700.0, inflation_factor=1.0, max_obs=150, bin_deg=bin_deg,
                    obs_polar=80.0)
    print(f"sweep case {case}", flush=True)
    data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, **over)
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    assert st["status"] == 0 and st["nonconverged"] == 0 and st["capacity_overflow"] == 0
    if case == "clustered_capped":
        assert st["nobsl_capped"] > 0
    h = {k: v.cpu().numpy() for k, v in data.items()}
    out, perm, bin_start, nband, nlonbin, ob, prm = _oracle_sweep(oracle, h, params, bin_deg)
    dbg = eng.debug(cfg.nobs)
    assert np.array_equal(dbg["nobsl"], out["nobsl"])
    ncol = cfg.nlat * cfg.nlon
    _check_selection(oracle, dbg, perm, bin_start, nband, nlonbin, ob, prm, bin_deg,
                     h["grid_lat"], h["grid_lon"], range(0, ncol, 7))
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(cfg.nz).items():
        inc_g = axa_h[:, rows] - h["axb"][:, rows]
        inc_o = out["axa"][:, rows] - h["axb"][:, rows]
        assert _rel_err(inc_g, inc_o) <= TOL, var
        assert _rel_err(xa_h[:, rows], out["xa"][:, rows]) <= TOL, var


def test_capped_selection_follows_caller_order(cuda, dl, oracle):
    """MAX_OBS truncation keeps the obs of smallest caller index (gen_yb_dis
    scans in the caller's order and stops at MAX_OBS, letkf_yb.F90:98-115):
    for two bin sizes the device lists, mapped back through perm, equal the
    full-scan oracle on the caller's list, and the analyses agree."""
    torch = cuda
    cfg = dl.synth_cfg(45, 90, 3, 32, 6000, seed=0xD1DA0001, cluster_frac=0.7, nclusters=3,
                       cluster_radius_km=600.0, rscale=(0.1, 10.0))
    h = None
    res = []
    for bin_deg in (3.0, 1.3):
        print(f"capped selection bin_deg {bin_deg}", flush=True)
        data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, local_dist=700.0, max_obs=150,
                                                 bin_deg=bin_deg, obs_polar=80.0)
        eng.sweep(obs, data["xb"], data["axb"], xa, axa)
        st = eng.stats()
        assert st["status"] == 0 and st["nobsl_capped"] > 20
        if h is None:
            h = {k: v.cpu().numpy() for k, v in data.items()}
        dbg = eng.debug(cfg.nobs)
        o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
        prm = oracle.params(params.nanals, params.nrow, params.nlat, params.nlon,
                            params.local_dist, 1.0, max_obs=params.max_obs, obs_polar=80.0)
        capped = np.nonzero(dbg["nobsl"] >= params.max_obs)[0]
        cols = np.concatenate([capped[::3], np.arange(0, cfg.nlat * cfg.nlon, 97)])
        for col in cols:
            clat, clon = h["grid_lat"][col % cfg.nlat], h["grid_lon"][col // cfg.nlat]
            idx, w = oracle.select_column(prm, o, clat, clon)
            off, p = int(dbg["offset"][col]), int(dbg["nobsl"][col])
            got = dbg["perm"][dbg["csr_idx"][off:off + p]]
            order = np.argsort(got)
            assert np.array_equal(got[order], idx), (bin_deg, col)
            assert np.abs(dbg["csr_w"][off:off + p][order] - w).max(initial=0.0) <= 1e-12
        res.append(axa.cpu().numpy())
        eng.close()
    assert _rel_err(res[1] - h["axb"], res[0] - h["axb"]) <= 1e-12


def test_rtps_and_post(cuda, dl, oracle):
    torch = cuda
    cfg = dl.synth_cfg(24, 48, 2, 16, 1500, seed=7)
    data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, local_dist=900.0, rtps=1,
                                             rtps_alpha=0.8, bin_deg=4.0)
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    eng.stats()
    h = {k: v.cpu().numpy() for k, v in data.items()}
    out = _oracle_sweep(oracle, h, params, 4.0)[0]
    exp, _, _ = oracle.post(1, h["grid_lat"], cfg.nlat, cfg.nlon, 90.0, 0, 0, 0.8, h["xb"],
                            h["axb"], out["xa"], out["axa"])
    assert _rel_err(xa.cpu().numpy(), exp - out["axa"][..., None]) <= TOL
    # run_da_parallel's blend and spreads
    xo = torch.empty_like(data["xb"])
    sb = torch.empty_like(data["axb"])
    sa = torch.empty_like(data["axb"])
    dl.post(cfg.k, params.nrow, cfg.nlat, cfg.nlon, data["grid_lat"], 60.0, 0.6, 0.4,
            data["xb"], data["axb"], xa, axa, xo, sb, sa)
    torch.cuda.synchronize()
    e_out, e_sb, e_sa = oracle.post(0, h["grid_lat"], cfg.nlat, cfg.nlon, 60.0, 0.6, 0.4, 0.0,
                                    h["xb"], h["axb"], xa.cpu().numpy(), axa.cpu().numpy())
    assert np.allclose(xo.cpu().numpy(), e_out, rtol=1e-14, atol=1e-12)
    assert np.allclose(sb.cpu().numpy(), e_sb, rtol=1e-13, atol=1e-14)
    assert np.allclose(sa.cpu().numpy(), e_sa, rtol=1e-13, atol=1e-14)


def test_no_obs_and_all_polar(cuda, dl):
    """No usable obs: W = sqrt(rho) I, wbar = 0, so xbar_a = xbar_b and
    Xa' = sqrt(rho) Xb'."""
    torch = cuda
    cfg = dl.synth_cfg(12, 24, 2, 8, 300, seed=3)
    data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, local_dist=500.0,
                                             inflation_factor=1.21, obs_polar=-1.0)
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    assert st["nobsl_sum"] == 0
    assert torch.allclose(axa, data["axb"], rtol=0, atol=1e-12 * data["axb"].abs().max().item())
    # the update adds xbar_b and subtracts the new mean (xb_2_xa), so the
    # absolute rounding of Xa' scales with |xbar_b| (up to ~1000 for ps)
    atol = 8 * 2.2e-16 * data["axb"].abs().max().item()
    assert torch.allclose(xa, 1.1 * data["xb"], rtol=1e-13, atol=atol)


def test_graph_replays_match_eager(cuda, dl):
    torch = cuda
    cfg = dl.synth_cfg(24, 48, 2, 16, 2000, seed=11)
    data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, local_dist=800.0)
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    eng.stats()
    ref_xa, ref_axa = xa.clone(), axa.clone()
    s = torch.cuda.Stream()
    xa2, axa2 = torch.zeros_like(xa), torch.zeros_like(axa)
    with torch.cuda.stream(s):
        eng.graph_capture(obs, data["xb"], data["axb"], xa2, axa2, s)
    s.synchronize()
    for rep in range(3):
        xa2.zero_()
        axa2.zero_()
        torch.cuda.synchronize()
        eng.graph_launch(s)
        s.synchronize()
        _ = (data["xb"] * 2).sum().item()         # eager work between replays
        assert torch.equal(xa2, ref_xa) and torch.equal(axa2, ref_axa), rep
