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
    # the per-bucket record: every analysis in one of the obs-count buckets,
    # and (when kept) a chain length for each bucket that holds analyses
    assert t["bucket_lo"] == [0, 64, 512, 2048]
    assert sum(t["bucket_count"]) == cfg.nlat * cfg.nlon
    assert sum(1 for c in t["bucket_count"] if c) >= 2        # a ragged workload
    if t["bucketed"]:
        assert all(c in (1, 2, 4, 8, 16, 32, 64) for c, m in
                   zip(t["bucket_chain"], t["bucket_count"]) if m)
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


@pytest.mark.parametrize("gram,tol", [(0, 1e-10), (1, 1e-4)])
def test_bucket_chains_match_oracle(cuda, dl, oracle, gram, tol):
    """A chain length per obs-count bucket (cold starts in the smallest
    bucket, long chains in the next, short ones above): the batched sweep
    matches the oracle within the form's bar, and its time is reported beside
    the single-length default."""
    torch = cuda
    cfg, prm = dl.config_synth(1, 0.06)
    data = dl.synth_generate(cfg)
    n = 1 + 3 * cfg.nz
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    h = {key: v.cpu().numpy() for key, v in data.items()}
    res = {}
    for name, bc in (("default", [0, 0, 0, 0]), ("buckets", [1, 64, 2, 8])):
        params = dl.default_params(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon,
                                   form=dl.FORM_BATCHED, gram=gram, bucket_chain=bc, **prm)
        eng = dl.Engine(params, nobs_max=cfg.nobs)
        eng.set_grid(data["grid_lat"], data["grid_lon"])
        xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
        eng.sweep(obs, data["xb"], data["axb"], xa, axa)
        st = eng.stats()
        assert st["status"] == 0 and st["nonconverged"] == 0, st
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            eng.sweep(obs, data["xb"], data["axb"], xa, axa)
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / 3
        out = _oracle_sweep(oracle, h, params, 3.0)[0]
        xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
        for var, rows in _var_rows(cfg.nz).items():
            assert _rel_err(axa_h[:, rows] - h["axb"][:, rows],
                            out["axa"][:, rows] - h["axb"][:, rows]) <= tol, (name, var)
            assert _rel_err(xa_h[:, rows], out["xa"][:, rows]) <= tol, (name, var)
        eng.close()
    print(f"gram {gram}: ms per sweep {res}", flush=True)
