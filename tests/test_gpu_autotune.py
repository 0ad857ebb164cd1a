"""Autotuner of the batched form (dida_letkf_autotune) on the ragged
clustered workload of 1-degree config 1, scaled down: k = 32, 70 % of the obs
in six 800 km clusters, L = 200 km.

Bars: the kept layout is one of the candidates and is no slower than the
initial one; the sweep run with it matches the oracle (same bins) within
1e-10 per variable (FP64 batched form); a later sweep reproduces it bit for
bit.
"""
import numpy as np
import pytest

from test_gpu_sweep import _oracle_sweep, _rel_err, _var_rows

pytestmark = pytest.mark.gpu


def test_autotune_config1(cuda, dl, oracle):
    torch = cuda
    cfg, prm = dl.config_synth(1, 0.06)
    print(f"autotune on {cfg.nlat}x{cfg.nlon} k={cfg.k} nobs={cfg.nobs}", flush=True)
    data = dl.synth_generate(cfg)
    n = 1 + 3 * cfg.nz
    params = dl.default_params(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon,
                               form=dl.FORM_BATCHED, gram=dl.GRAM_FP64, **prm)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    t = eng.autotune(obs, data["xb"], data["axb"], xa, axa)
    print("tuning", t, flush=True)
    st = eng.stats()
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    assert t["chain_len"] in (2, 4, 8, 16, 32, 64) and t["ncand"] >= 6
    assert 0.0 < t["ms_best"] <= t["ms_initial"]
    # the layout is kept: a later sweep gives the same bits
    xa2, axa2 = torch.empty_like(xa), torch.empty_like(axa)
    eng.sweep(obs, data["xb"], data["axb"], xa2, axa2)
    assert torch.equal(xa, xa2) and torch.equal(axa, axa2)
    h = {key: v.cpu().numpy() for key, v in data.items()}
    out = _oracle_sweep(oracle, h, params, t["bin_deg"])[0]
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(cfg.nz).items():
        assert _rel_err(axa_h[:, rows] - h["axb"][:, rows],
                        out["axa"][:, rows] - h["axb"][:, rows]) <= 1e-10, var
        assert _rel_err(xa_h[:, rows], out["xa"][:, rows]) <= 1e-10, var
    eng.close()
