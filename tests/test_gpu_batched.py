"""GPU parity of the batched form (cost-sorted chain queue, persistent fused
kernel, warm-started eigensolve) through the C ABI.

Bars: with the FP64 Gram the batched form meets the plain form's bar
(per-variable increment and Xa' relative error <= 1e-10 against the oracle);
the mixed-precision Grams (FP16 x3, BF16 x3) meet <= 1e-4 relative
increment error per variable, the precision form's stated tolerance.
"""
import numpy as np
import pytest

from test_gpu_sweep import _oracle_sweep, _rel_err, _setup, _var_rows

pytestmark = pytest.mark.gpu


def _run(dl, torch, cfg, **over):
    data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, **over)
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    return data, params, eng, obs, xa, axa, st


CASES = {
    "uniform_k20": (dict(nlat=36, nlon=72, nz=4, k=20, nobs=2500, seed=0xD1DA0000),
                    dict(local_dist=1000.0, inflation_factor=1.1)),
    "clustered_k64": (dict(nlat=30, nlon=60, nz=3, k=64, nobs=8000, seed=0xD1DA0002,
                           cluster_frac=0.7, nclusters=4, cluster_radius_km=800.0,
                           rscale=(0.1, 10.0)),
                      dict(local_dist=900.0, inflation_factor=1.0, rtps=1, rtps_alpha=0.8,
                           max_obs=1500)),
}


def _cfg(dl, spec):
    spec = dict(spec)
    return dl.synth_cfg(spec.pop("nlat"), spec.pop("nlon"), spec.pop("nz"), spec.pop("k"),
                        spec.pop("nobs"), **spec)


def _oracle_with_post(oracle, h, params, bin_deg, cfg):
    out = _oracle_sweep(oracle, h, params, bin_deg)[0]
    if params.rtps:
        exp, _, _ = oracle.post(1, h["grid_lat"], cfg.nlat, cfg.nlon, 90.0, 0, 0,
                                params.rtps_alpha, h["xb"], h["axb"], out["xa"], out["axa"])
        out["xa"] = exp - out["axa"][..., None]
    return out


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("chain_len", [1, 8])
def test_batched_fp64_matches_oracle(cuda, dl, oracle, case, chain_len):
    torch = cuda
    spec, over = CASES[case]
    cfg = _cfg(dl, spec)
    print(f"batched fp64 {case} chain {chain_len}", flush=True)
    data, params, eng, obs, xa, axa, st = _run(dl, torch, cfg, form=dl.FORM_BATCHED,
                                               gram=dl.GRAM_FP64, chain_len=chain_len,
                                               bin_deg=3.0, **over)
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    h = {k: v.cpu().numpy() for k, v in data.items()}
    out = _oracle_with_post(oracle, h, params, 3.0, cfg)
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(cfg.nz).items():
        assert _rel_err(axa_h[:, rows] - h["axb"][:, rows],
                        out["axa"][:, rows] - h["axb"][:, rows]) <= 1e-10, var
        assert _rel_err(xa_h[:, rows], out["xa"][:, rows]) <= 1e-10, var


@pytest.mark.parametrize("gram", ["FP16X3", "BF16X3"])
def test_batched_mixed_precision_error(cuda, dl, oracle, gram):
    torch = cuda
    spec, over = CASES["clustered_k64"]
    cfg = _cfg(dl, spec)
    g = getattr(dl, "GRAM_" + gram)
    print(f"batched {gram}", flush=True)
    data, params, eng, obs, xa, axa, st = _run(dl, torch, cfg, form=dl.FORM_BATCHED, gram=g,
                                               bin_deg=3.0, **over)
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    h = {k: v.cpu().numpy() for k, v in data.items()}
    out = _oracle_with_post(oracle, h, params, 3.0, cfg)
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    errs = {}
    for var, rows in _var_rows(cfg.nz).items():
        errs[var] = _rel_err(axa_h[:, rows] - h["axb"][:, rows],
                             out["axa"][:, rows] - h["axb"][:, rows])
        errs[var + "_xa"] = _rel_err(xa_h[:, rows], out["xa"][:, rows])
    print(gram, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for var in ("ps", "u", "v", "T"):
        assert errs[var] <= 1e-4, errs
        assert errs[var + "_xa"] <= 1e-4, errs


def test_batched_graph_replays(cuda, dl):
    torch = cuda
    cfg = dl.synth_cfg(24, 48, 2, 32, 3000, seed=21)
    data, params, eng, obs, xa, axa, st = _run(dl, torch, cfg, form=dl.FORM_BATCHED,
                                               gram=dl.GRAM_FP16X3, local_dist=800.0)
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
        _ = (data["xb"] * 3).sum().item()
        assert torch.equal(xa2, ref_xa) and torch.equal(axa2, ref_axa), rep
