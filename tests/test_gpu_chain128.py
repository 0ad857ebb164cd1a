"""The batched form for 64 < k <= 128: chains of analyses, each warm-starting
its eigensolve from the previous analysis of its chain.

Bars: per-variable increment and Xa' relative error against the oracle
<= 1e-10 on the FP64 Gram, for columns (MAX_OBS binding), horizontal cubes
with overlap, and a workspace smaller than the sweep's lists; <= 1e-4 on the
BF16x3 and FP16x3 Grams (clustered obs with R spanning 0.5-3, and MAX_OBS
binding);
fewer Jacobi sweeps than the plain form's cold starts; captured-graph
replays bit-identical to the eager sweep."""
import numpy as np
import pytest

from test_gpu_sweep import _rel_err, _var_rows

pytestmark = pytest.mark.gpu


def _check(oracle, data, params, xa, axa, cube_n=1, overlap=0, tol=1e-10):
    h = {k: v.cpu().numpy() for k, v in data.items()}
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bs, nb, nl = oracle.bin_obs(o, params.bin_deg, params.obs_polar)
    prm = oracle.params(params.nanals, params.nrow, params.nlat, params.nlon, params.local_dist,
                        params.inflation_factor, max_obs=params.max_obs)
    out = oracle.sweep_cubes(prm, h["grid_lat"], h["grid_lon"], o.permuted(perm), h["xb"],
                             h["axb"], cube_n, overlap, bins=(bs, nb, nl, params.bin_deg, perm))
    assert out["nfail"] == 0
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(2).items():
        inc_g = axa_h[:, rows] - h["axb"][:, rows]
        inc_o = out["axa"][:, rows] - h["axb"][:, rows]
        e_inc, e_xa = _rel_err(inc_g, inc_o), _rel_err(xa_h[:, rows], out["xa"][:, rows])
        print(f"  {var}: increment {e_inc:.2e} Xa' {e_xa:.2e}", flush=True)
        assert e_inc <= tol, var
        assert e_xa <= tol, var


@pytest.mark.parametrize("case", ["k100_chain5", "k128_capped", "k100_cubes", "k128_chunks",
                                  "k100_chunks_in_chain"])
def test_chained_batched_k128(cuda, dl, oracle, case):
    torch = cuda
    print(f"chained batched {case}", flush=True)
    k = 128 if case.startswith("k128") else 100
    cfg = dl.synth_cfg(20, 40, 2, k, 1500, seed=0xD1DA0500 + len(case), cluster_frac=0.5,
                       nclusters=2, rscale=(0.5, 3.0))
    data = dl.synth_generate(cfg)
    p = dict(nanals=k, nrow=7, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=1200.0,
             inflation_factor=1.05, bin_deg=4.0, form=dl.FORM_BATCHED, gram=dl.GRAM_FP64,
             chain_len=40 if case == "k100_chunks_in_chain" else 5)
    cap, cube = 0, (1, 0)
    if case == "k128_capped":
        p["max_obs"] = 200
    elif case == "k100_cubes":
        cube = (2, 1)
        p.update(cube_n=2, overlap=1)
    params = dl.default_params(**p)
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    if case.endswith("chunks") or case.endswith("in_chain"):
        probe = dl.Engine(params, nobs_max=cfg.nobs)
        probe.set_grid(data["grid_lat"], data["grid_lon"])
        x, ax = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
        probe.sweep(obs, data["xb"], data["axb"], x, ax)
        pst = probe.stats()
        # k128_chunks: a third of the lists (runs of whole chains);
        # in_chain: a third of one 40-analysis chain's lists (runs end inside chains)
        cap = pst["nobsl_sum"] // 3 + 1 if case == "k128_chunks" else \
            max(pst["nobsl_sum"] // 60, pst["nobsl_max"]) + 1
        probe.close()
    eng = dl.Engine(params, nobs_max=cfg.nobs, capacity=cap)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    eng.close()
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    if case == "k128_capped":
        assert st["nobsl_capped"] > 0
    if case == "k128_chunks":
        # runs start at chain boundaries: bit-identical to one whole-workspace sweep
        assert torch.equal(xa, x) and torch.equal(axa, ax)
    elif cap:
        # a chain split between runs keeps its analyses (chains are counted
        # from analysis 0) but its second part starts cold: the same within
        # the eigensolver's tolerance
        inc, inc1 = (axa - data["axb"]).cpu().numpy(), (ax - data["axb"]).cpu().numpy()
        assert _rel_err(inc, inc1) <= 1e-10
        assert _rel_err(xa.cpu().numpy(), x.cpu().numpy()) <= 1e-10
    _check(oracle, data, params, xa, axa, *cube)
    if case == "k100_chain5":
        # the plain form's cold starts take more sweeps
        pp = dl.default_params(**dict(p, form=dl.FORM_PLAIN))
        e2 = dl.Engine(pp, nobs_max=cfg.nobs)
        e2.set_grid(data["grid_lat"], data["grid_lon"])
        x2, a2 = torch.empty_like(xa), torch.empty_like(axa)
        e2.sweep(obs, data["xb"], data["axb"], x2, a2)
        s2 = e2.stats()
        e2.close()
        assert st["sweeps_sum"] < s2["sweeps_sum"], (st["sweeps_sum"], s2["sweeps_sum"])


@pytest.mark.parametrize("case", ["k128_bf16x3", "k100_bf16x3_capped", "k128_bf16x3_chunks",
                                  "k65_bf16x3", "k128_fp16x3", "k100_fp16x3_capped"])
def test_chained_batched_k128_split_gram(cuda, dl, oracle, case):
    """The BF16x3 and FP16x3 Grams (split hi/lo on half-precision MFMA, FP32
    accumulation; FP16 with a power-of-two scale per analysis) in the
    chained form, against the FP64 oracle; also with a workspace that holds a
    third of the sweep's selection lists (runs of analyses)."""
    torch = cuda
    print(f"chained batched {case}", flush=True)
    k = int(case[1:].split("_")[0])
    cfg = dl.synth_cfg(20, 40, 2, k, 2000, seed=0xD1DA0600 + len(case), cluster_frac=0.7,
                       nclusters=3, rscale=(0.5, 3.0))
    data = dl.synth_generate(cfg)
    gram = dl.GRAM_FP16X3 if "fp16x3" in case else dl.GRAM_BF16X3
    p = dict(nanals=k, nrow=7, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=1200.0,
             inflation_factor=1.05, bin_deg=4.0, form=dl.FORM_BATCHED, gram=gram,
             chain_len=4)
    if case.endswith("capped"):
        p["max_obs"] = 300
    params = dl.default_params(**p)
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    cap = 0
    if case.endswith("chunks"):
        probe = dl.Engine(params, nobs_max=cfg.nobs)
        probe.set_grid(data["grid_lat"], data["grid_lon"])
        x, ax = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
        probe.sweep(obs, data["xb"], data["axb"], x, ax)
        cap = probe.stats()["nobsl_sum"] // 3 + 1
        probe.close()
    eng = dl.Engine(params, nobs_max=cfg.nobs, capacity=cap)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    eng.close()
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    if case.endswith("capped"):
        assert st["nobsl_capped"] > 0
    if cap:
        # the runs reproduce one whole-workspace sweep bit for bit
        assert torch.equal(xa, x) and torch.equal(axa, ax)
    _check(oracle, data, params, xa, axa, tol=1e-4)


def test_chained_graph_replays(cuda, dl):
    """The chained form captured in a HIP graph: every replay reproduces the
    eager sweep bit for bit (the queue is dynamic, but a chain's analyses and
    warm starts do not depend on the workgroup that takes it)."""
    torch = cuda
    print("chained graph replays", flush=True)
    cfg = dl.synth_cfg(16, 32, 2, 100, 1200, seed=0xD1DA0800, cluster_frac=0.5, nclusters=2)
    data = dl.synth_generate(cfg)
    params = dl.default_params(nanals=100, nrow=7, nlat=cfg.nlat, nlon=cfg.nlon,
                               local_dist=1200.0, bin_deg=4.0, form=dl.FORM_BATCHED,
                               gram=dl.GRAM_BF16X3, chain_len=4)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    torch.cuda.synchronize()
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
        assert torch.equal(xa2, xa) and torch.equal(axa2, axa), rep
    eng.close()
