"""Sweeps with the ensemble perturbations stored in FP32
(dida_letkf_sweep_f32), against the oracle run on the same values widened
to FP64.

Bars: with the FP64 Gram the means are computed in FP64 from exactly the
oracle's inputs, so the per-variable analysis increment stays within 1e-10
and Xa' within 1e-6 (one FP32 rounding of each output); with the BF16x3
Gram (k = 128, chained batched form) both within 1e-4.  The fused batched
form (k <= 64) reports DIDA_ERR_UNSUPPORTED.
"""
import numpy as np
import pytest

from test_gpu_sweep import _rel_err, _var_rows

pytestmark = pytest.mark.gpu


def _oracle(oracle, h, params, cube_n, overlap):
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bs, nb, nl = oracle.bin_obs(o, params.bin_deg, params.obs_polar)
    prm = oracle.params(params.nanals, params.nrow, params.nlat, params.nlon, params.local_dist,
                        params.inflation_factor, max_obs=params.max_obs)
    out = oracle.sweep_cubes(prm, h["grid_lat"], h["grid_lon"], o.permuted(perm), h["xb"],
                             h["axb"], cube_n, overlap, bins=(bs, nb, nl, params.bin_deg, perm))
    assert out["nfail"] == 0
    return out


@pytest.mark.parametrize("case", ["plain_k40", "plain_k100_cubes", "chained_k128_bf16x3",
                                  "plain_k128_chunks"])
def test_sweep_f32_state(cuda, dl, oracle, case):
    torch = cuda
    print(f"fp32 state {case}", flush=True)
    k = {"plain_k40": 40, "plain_k100_cubes": 100, "chained_k128_bf16x3": 128,
         "plain_k128_chunks": 128}[case]
    cfg = dl.synth_cfg(20, 40, 2, k, 1500, seed=0xD1DA0700 + len(case), cluster_frac=0.5,
                       nclusters=2, rscale=(0.5, 3.0))
    data = dl.synth_generate(cfg)
    p = dict(nanals=k, nrow=7, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=1200.0,
             inflation_factor=1.05, bin_deg=4.0, form=dl.FORM_PLAIN, gram=dl.GRAM_FP64)
    cube, tol_inc, tol_xa = (1, 0), 1e-10, 1e-6
    if case == "plain_k100_cubes":
        cube = (2, 1)
        p.update(cube_n=2, overlap=1)
    elif case == "chained_k128_bf16x3":
        p.update(form=dl.FORM_BATCHED, gram=dl.GRAM_BF16X3, chain_len=4)
        tol_inc = tol_xa = 1e-4
    params = dl.default_params(**p)
    xb32 = data["xb"].float()
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    cap = 0
    if case.endswith("chunks"):
        # a workspace of a third of the lists: the sweep runs in runs of analyses
        probe = dl.Engine(params, nobs_max=cfg.nobs)
        probe.set_grid(data["grid_lat"], data["grid_lon"])
        probe.sweep(obs, xb32, data["axb"], torch.empty_like(xb32), torch.empty_like(data["axb"]))
        cap = probe.stats()["nobsl_sum"] // 3 + 1
        probe.close()
    eng = dl.Engine(params, nobs_max=cfg.nobs, capacity=cap)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    xa32 = torch.empty_like(xb32)
    axa = torch.empty_like(data["axb"])
    eng.sweep(obs, xb32, data["axb"], xa32, axa)
    st = eng.stats()
    eng.close()
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    h = {key: v.cpu().numpy() for key, v in data.items()}
    h["xb"] = xb32.double().cpu().numpy()
    out = _oracle(oracle, h, params, *cube)
    xa_h, axa_h = xa32.double().cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(2).items():
        e_inc = _rel_err(axa_h[:, rows] - h["axb"][:, rows], out["axa"][:, rows] - h["axb"][:, rows])
        e_xa = _rel_err(xa_h[:, rows], out["xa"][:, rows])
        print(f"  {var}: increment {e_inc:.2e} Xa' {e_xa:.2e}", flush=True)
        assert e_inc <= tol_inc, var
        assert e_xa <= tol_xa, var


def test_sweep_f32_rejects_the_fused_form(cuda, dl):
    torch = cuda
    cfg = dl.synth_cfg(8, 16, 2, 32, 200, seed=5)
    data = dl.synth_generate(cfg)
    params = dl.default_params(nanals=32, nrow=7, nlat=cfg.nlat, nlon=cfg.nlon,
                               form=dl.FORM_BATCHED, gram=dl.GRAM_BF16X3)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xb32 = data["xb"].float()
    with pytest.raises(dl.DidaError, match="FP32 state"):
        eng.sweep(obs, xb32, data["axb"], torch.empty_like(xb32), torch.empty_like(data["axb"]))
    eng.close()
