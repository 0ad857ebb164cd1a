"""Plain-form sweeps over a selection workspace smaller than the sweep's
lists run in runs of analyses: results bit-identical to the one-run sweep
(the same analyses in the same order), within 1e-10 of the oracle, for
columns and horizontal cubes; a workspace smaller than one analysis's list
is an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(dl, torch, data, cfg, capacity, **over):
    p = dict(nanals=cfg.k, nrow=1 + 3 * cfg.nz, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=1000.0,
             inflation_factor=1.1, bin_deg=4.0)
    p.update(over)
    params = dl.default_params(**p)
    eng = dl.Engine(params, nobs_max=cfg.nobs, capacity=capacity)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    eng.close()
    return xa, axa, st, params


@pytest.mark.parametrize("cube_n,overlap", [(1, 0), (3, 1)])
def test_chunked_plain_sweep(cuda, dl, oracle, cube_n, overlap):
    torch = cuda
    print(f"chunked plain sweep cube {cube_n}/{overlap}", flush=True)
    cfg = dl.synth_cfg(30, 60, 2, 16, 2500, seed=5150 + cube_n)
    data = dl.synth_generate(cfg)
    xa1, axa1, st1, params = _run(dl, torch, data, cfg, 0, cube_n=cube_n, overlap=overlap)
    total = st1["nobsl_sum"]
    xa2, axa2, st2, _ = _run(dl, torch, data, cfg, total // 4 + 1, cube_n=cube_n,
                             overlap=overlap)
    assert st2["status"] == 0 and st2["nobsl_sum"] == total
    assert torch.equal(xa1, xa2) and torch.equal(axa1, axa2)
    h = {k: v.cpu().numpy() for k, v in data.items()}
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bs, nb, nl = oracle.bin_obs(o, params.bin_deg, params.obs_polar)
    prm = oracle.params(params.nanals, params.nrow, cfg.nlat, cfg.nlon, params.local_dist,
                        params.inflation_factor)
    out = oracle.sweep_cubes(prm, h["grid_lat"], h["grid_lon"], o.permuted(perm), h["xb"],
                             h["axb"], cube_n, overlap, bins=(bs, nb, nl, params.bin_deg, perm))
    inc_o = out["axa"] - h["axb"]
    assert np.linalg.norm(axa2.cpu().numpy() - h["axb"] - inc_o) <= 1e-10 * np.linalg.norm(inc_o)


def test_workspace_below_one_list_is_an_error(cuda, dl):
    torch = cuda
    cfg = dl.synth_cfg(12, 24, 2, 8, 800, seed=99)
    data = dl.synth_generate(cfg)
    p = dl.default_params(nanals=8, nrow=7, nlat=12, nlon=24, local_dist=2000.0, bin_deg=4.0)
    eng = dl.Engine(p, nobs_max=cfg.nobs, capacity=5)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    with pytest.raises(dl.DidaError, match="alone needs"):
        eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    eng.close()
