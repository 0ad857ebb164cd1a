"""The batched kernel's work queues: chains drained from one cost-sorted
queue or from per-die longitude-stripe groups (xcd_tiles) give bit-identical
analyses (a chain's columns and warm starts do not depend on which wave runs
it), including grids with fewer longitudes than groups and chains longer
than the grid; the grouped order matches the oracle to 1e-10 (FP64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(dl, torch, data, cfg, **over):
    n = 1 + 3 * cfg.nz
    p = dict(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=900.0,
             inflation_factor=1.1, bin_deg=4.0, form=dl.FORM_BATCHED, gram=dl.GRAM_FP64)
    p.update(over)
    params = dl.default_params(**p)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    eng.close()
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    return xa, axa, params


@pytest.mark.parametrize("nlat,nlon,chain_len", [(24, 48, 7), (30, 5, 64), (9, 17, 1)])
def test_grouped_queue_equals_sorted_queue(cuda, dl, nlat, nlon, chain_len):
    torch = cuda
    print(f"queue {nlat}x{nlon} chain {chain_len}", flush=True)
    cfg = dl.synth_cfg(nlat, nlon, 2, 20, 1500, seed=31 + nlon, cluster_frac=0.5, nclusters=2)
    data = dl.synth_generate(cfg)
    a = _run(dl, torch, data, cfg, chain_len=chain_len, xcd_tiles=0)
    b = _run(dl, torch, data, cfg, chain_len=chain_len, xcd_tiles=1)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_grouped_queue_matches_oracle(cuda, dl, oracle):
    torch = cuda
    cfg = dl.synth_cfg(36, 72, 2, 24, 2500, seed=77, cluster_frac=0.6, nclusters=3)
    data = dl.synth_generate(cfg)
    xa, axa, params = _run(dl, torch, data, cfg, chain_len=8, xcd_tiles=1)
    h = {k: v.cpu().numpy() for k, v in data.items()}
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bin_start, nband, nlonbin = oracle.bin_obs(o, params.bin_deg, params.obs_polar)
    prm = oracle.params(params.nanals, params.nrow, cfg.nlat, cfg.nlon, params.local_dist,
                        params.inflation_factor)
    out = oracle.sweep(prm, h["grid_lat"], h["grid_lon"], o.permuted(perm), h["xb"], h["axb"],
                       bins=(bin_start, nband, nlonbin, params.bin_deg, perm))
    inc_g = axa.cpu().numpy() - h["axb"]
    inc_o = out["axa"] - h["axb"]
    assert np.linalg.norm(inc_g - inc_o) <= 1e-10 * np.linalg.norm(inc_o)
    assert np.linalg.norm(xa.cpu().numpy() - out["xa"]) <= 1e-10 * np.linalg.norm(out["xa"])
