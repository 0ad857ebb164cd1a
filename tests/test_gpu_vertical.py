"""GPU parity of the vertical-localisation sweep (one analysis per column and
cube of levels; letkf_compute_nooverlap_vertical, letkf_compute.F90:358-526)
through the C ABI, against the oracle's restatement on the same inputs.

Bars: per-variable analysis-increment and Xa' relative error <= 1e-10 with
the FP64 Gram, <= 1e-4 with the split BF16 Gram; with vert_dist so large that
every vertical weight is 1, the vertical sweep equals the engine's ordinary
sweep within 1e-10; a captured graph replays the eager result exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NZ = 6


def _var_rows(nz):
    return {"ps": [0], "u": list(range(1, 1 + nz)), "v": list(range(1 + nz, 1 + 2 * nz)),
            "T": list(range(1 + 2 * nz, 1 + 3 * nz))}


def _rel_err(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _setup(dl, torch, seed, k=20, nlat=36, nlon=72, nobs=2500):
    cfg = dl.synth_cfg(nlat, nlon, NZ, k, nobs, seed=seed)
    data = dl.synth_generate(cfg)
    rng = np.random.default_rng(seed)
    ph = rng.uniform(50.0, 1000.0, nobs)            # hPa, as the synthetic ps
    vc = {"type": 0, "c1": np.linspace(0.0, 1.0, NZ + 1), "ptop": 10.0}
    var_start = [0, 1, 1 + NZ, 1 + 2 * NZ]
    var_lev = [1, NZ, NZ, NZ]
    return cfg, data, ph, vc, var_start, var_lev


def _engine(dl, torch, cfg, data, ph, vc, vs, vl, **over):
    n = 1 + 3 * cfg.nz
    p = dict(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=1000.0,
             inflation_factor=1.1, bin_deg=4.0, form=dl.FORM_BATCHED)
    p.update(over)
    params = dl.default_params(**p)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    T = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")
    if params.vert_dist > 0:
        vcd = {"type": vc["type"], "c1": T(vc["c1"]), "ptop": vc["ptop"]}
        eng.set_levels(vcd, 0, vs, vl)
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"], T(ph))
    return params, eng, obs


def _oracle(oracle, h, ph, params, vc, vs, vl):
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bin_start, nband, nlonbin = oracle.bin_obs(o, params.bin_deg, params.obs_polar)
    prm = oracle.params(params.nanals, params.nrow, params.nlat, params.nlon,
                        params.local_dist, params.inflation_factor, max_obs=params.max_obs,
                        gc_filter=bool(params.gc_filter), obs_polar=params.obs_polar)
    return oracle.sweep_vertical(prm, h["grid_lat"], h["grid_lon"], o.permuted(perm), ph[perm],
                                 h["xb"], h["axb"], vc, 0, vs, vl, params.vert_dist,
                                 params.cube_nz, bool(params.gc_filter_vertical),
                                 bins=(bin_start, nband, nlonbin, params.bin_deg, perm),
                                 cube_n=params.cube_n)


@pytest.mark.parametrize("case", ["fp64_gc_cube1", "fp64_boxcar_cube4_capped", "bf16x3_cube2",
                                  "fp64_gc_hcube3", "fp64_gc_hcube2_cube3"])
def test_vertical_sweep_matches_oracle(cuda, dl, oracle, case):
    torch = cuda
    print(f"vertical sweep case {case}", flush=True)
    cfg, data, ph, vc, vs, vl = _setup(dl, torch, 0xD1DA0100 + len(case))
    if case == "fp64_gc_cube1":
        over, tol = dict(vert_dist=0.5, cube_nz=1, gram=dl.GRAM_FP64), 1e-10
    elif case == "fp64_boxcar_cube4_capped":
        over, tol = dict(vert_dist=0.7, cube_nz=4, gc_filter_vertical=0, gram=dl.GRAM_FP64,
                         max_obs=60), 1e-10
    elif case == "bf16x3_cube2":
        over, tol = dict(vert_dist=0.5, cube_nz=2, gram=dl.GRAM_BF16X3), 1e-4
    elif case == "fp64_gc_hcube3":
        over, tol = dict(vert_dist=0.5, cube_nz=1, cube_n=3, gram=dl.GRAM_FP64), 1e-10
    else:
        over, tol = dict(vert_dist=0.6, cube_nz=3, cube_n=2, gram=dl.GRAM_FP64), 1e-10
    params, eng, obs = _engine(dl, torch, cfg, data, ph, vc, vs, vl, **over)
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    ncube = -(-NZ // params.cube_nz)
    nh = -(-cfg.nlat // params.cube_n) * -(-cfg.nlon // params.cube_n)
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    assert st["nanal"] == nh * ncube
    if params.max_obs < 10000:
        assert st["nobsl_capped"] > 0
    h = {k: v.cpu().numpy() for k, v in data.items()}
    out = _oracle(oracle, h, ph, params, vc, vs, vl)
    assert out["nfail"] == 0
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(NZ).items():
        inc_g = axa_h[:, rows] - h["axb"][:, rows]
        inc_o = out["axa"][:, rows] - h["axb"][:, rows]
        assert _rel_err(inc_g, inc_o) <= tol, var
        assert _rel_err(xa_h[:, rows], out["xa"][:, rows]) <= tol, var
    # the localisation acts: the result differs from the ordinary sweep's
    if case == "fp64_gc_cube1":
        p2, eng2, _ = _engine(dl, torch, cfg, data, ph, vc, vs, vl, gram=dl.GRAM_FP64)
        xa2, axa2 = torch.empty_like(xa), torch.empty_like(axa)
        eng2.sweep(obs, data["xb"], data["axb"], xa2, axa2)
        eng2.stats()
        assert _rel_err(axa_h - h["axb"], axa2.cpu().numpy() - h["axb"]) > 1e-3
        # graph replays equal the eager sweep
        s = torch.cuda.Stream()
        xg, axg = torch.zeros_like(xa), torch.zeros_like(axa)
        with torch.cuda.stream(s):
            eng.graph_capture(obs, data["xb"], data["axb"], xg, axg, s)
        for _ in range(2):
            eng.graph_launch(s)
            s.synchronize()
            assert torch.equal(xg, xa) and torch.equal(axg, axa)
            xg.zero_()
            axg.zero_()
        eng2.close()
    eng.close()


def test_unit_vertical_weights_equal_the_ordinary_sweep(cuda, dl):
    torch = cuda
    print("vertical sweep with unit weights", flush=True)
    cfg, data, ph, vc, vs, vl = _setup(dl, torch, 0xD1DA0200, k=32)
    res = []
    for over in (dict(vert_dist=1e30, cube_nz=2), dict()):
        params, eng, obs = _engine(dl, torch, cfg, data, ph, vc, vs, vl, gram=dl.GRAM_FP64,
                                   **over)
        xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
        eng.sweep(obs, data["xb"], data["axb"], xa, axa)
        assert eng.stats()["status"] == 0
        res.append((xa.cpu().numpy(), axa.cpu().numpy()))
        eng.close()
    axb = data["axb"].cpu().numpy()
    for var, rows in _var_rows(NZ).items():
        assert _rel_err(res[0][1][:, rows] - axb[:, rows], res[1][1][:, rows] - axb[:, rows]) <= 1e-10
        assert _rel_err(res[0][0][:, rows], res[1][0][:, rows]) <= 1e-10


def test_set_levels_checks(cuda, dl):
    torch = cuda
    cfg = dl.synth_cfg(8, 16, NZ, 8, 50, seed=3)
    data = dl.synth_generate(cfg)
    params = dl.default_params(nanals=8, nrow=1 + 3 * NZ, nlat=8, nlon=16, vert_dist=1.0,
                               form=dl.FORM_BATCHED)
    eng = dl.Engine(params, nobs_max=50)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    c1 = torch.linspace(0, 1, NZ + 1, dtype=torch.float64, device="cuda")
    vcd = {"type": 0, "c1": c1, "ptop": 10.0}
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    with pytest.raises(dl.DidaError, match="levels not set"):
        eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    with pytest.raises(dl.DidaError, match="row"):
        eng.set_levels(vcd, 0, [0, 1, 1 + NZ], [1, NZ, NZ])          # T rows uncovered
    with pytest.raises(dl.DidaError, match="levels must be 1 or nz"):
        eng.set_levels(vcd, 0, [0, 1, 1 + NZ, 1 + 2 * NZ], [1, NZ, NZ, 2])
    eng.set_levels(vcd, 0, [0, 1, 1 + NZ, 1 + 2 * NZ], [1, NZ, NZ, NZ])
    with pytest.raises(dl.DidaError, match="ph is null"):
        eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    eng.close()
