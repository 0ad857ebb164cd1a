"""The batching form's regional case: twelve independent domains of uneven
size (ensemble sizes 32, 48 and 64, 0.5-5 obs per column) captured as one
HIP graph with a parallel branch per domain (dida_letkf_group_*).

Bars: every replay of the group graph is bit-identical to eager sweeps of the
same contexts; the FP64 batched form matches the oracle within 1e-10
(per-variable increment and Xa') on every domain; BF16x3 within 1e-4.
"""
import numpy as np
import pytest

from test_gpu_sweep import _oracle_sweep, _rel_err, _var_rows

pytestmark = pytest.mark.gpu

# (ny, nx, k, obs per column) of the twelve domains, on 2-degree global fields
DOMAINS = [(8, 8, 32, 0.5), (12, 10, 48, 1.0), (16, 12, 64, 2.0), (10, 20, 32, 5.0),
           (24, 16, 48, 0.5), (20, 24, 64, 1.5), (30, 24, 32, 3.0), (8, 28, 48, 4.0),
           (16, 16, 64, 5.0), (28, 20, 32, 1.0), (12, 30, 48, 2.5), (24, 24, 64, 0.8)]
NZ = 3
BIN = 4.0


def _domains(dl, torch, gram):
    out = []
    for d, (ny, nx, k, dens) in enumerate(DOMAINS):
        nobs = int(dens * 90 * 180)
        cfg = dl.synth_cfg(90, 180, NZ, k, nobs, seed=0xD1DA0300 + d, rscale=(0.5, 3.0))
        g = dl.synth_generate(cfg)
        j0, i0 = 10 + (7 * d) % 40, 5 + (23 * d) % 140
        r = dl.synth_region(g, cfg, j0, ny, i0, nx)
        del g
        n = 1 + 3 * NZ
        params = dl.default_params(nanals=k, nrow=n, nlat=ny, nlon=nx, local_dist=500.0,
                                   inflation_factor=1.05, form=dl.FORM_BATCHED, gram=gram,
                                   bin_deg=BIN)
        eng = dl.Engine(params, nobs_max=max(1, r["lat"].numel()))
        eng.set_grid(r["grid_lat"], r["grid_lon"])
        obs = dl.DeviceObs(r["lat"], r["lon"], r["yb"], r["dep"], r["rinv"])
        out.append((r, params, eng, obs))
    return out


@pytest.mark.parametrize("gram,tuned", [("FP64", False), ("BF16X3", False), ("FP64", True)])
def test_domains_group_graph(cuda, dl, oracle, gram, tuned):
    """tuned: every domain's context autotuned first (its own bin size, chain
    lengths per obs-count bucket, queue order), so the graph records each
    domain with its own layout."""
    torch = cuda
    print(f"twelve domains, {gram}, tuned {tuned}", flush=True)
    doms = _domains(dl, torch, getattr(dl, "GRAM_" + gram))
    eager = []
    for r, params, eng, obs in doms:
        xa, axa = torch.empty_like(r["xb"]), torch.empty_like(r["axb"])
        if tuned:
            t = eng.autotune(obs, r["xb"], r["axb"], xa, axa)
            assert t["ms_best"] <= t["ms_initial"]
        eng.sweep(obs, r["xb"], r["axb"], xa, axa)
        st = eng.stats()
        assert st["status"] == 0 and st["nonconverged"] == 0, st
        eager.append((xa, axa))
    torch.cuda.synchronize()
    xa2 = [torch.zeros_like(e[0]) for e in eager]
    axa2 = [torch.zeros_like(e[1]) for e in eager]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        grp = dl.Group([d[2] for d in doms], [d[3] for d in doms], [d[0]["xb"] for d in doms],
                       [d[0]["axb"] for d in doms], xa2, axa2, s)
    s.synchronize()
    for rep in range(3):
        for t in xa2 + axa2:
            t.zero_()
        torch.cuda.synchronize()
        grp.launch(s)
        s.synchronize()
        _ = (doms[0][0]["xb"] * 2).sum().item()
        for i in range(len(doms)):
            assert torch.equal(xa2[i], eager[i][0]) and torch.equal(axa2[i], eager[i][1]), (rep, i)
    grp.close()
    bar = 1e-10 if gram == "FP64" else 1e-4
    for i, (r, params, eng, obs) in enumerate(doms):
        h = {key: v.cpu().numpy() for key, v in r.items()}
        out = _oracle_sweep(oracle, h, params, BIN)[0]
        xa_h, axa_h = eager[i][0].cpu().numpy(), eager[i][1].cpu().numpy()
        for var, rows in _var_rows(NZ).items():
            inc = _rel_err(axa_h[:, rows] - h["axb"][:, rows], out["axa"][:, rows] - h["axb"][:, rows])
            assert inc <= bar, (i, var, inc)
            assert _rel_err(xa_h[:, rows], out["xa"][:, rows]) <= bar, (i, var)
    for d in doms:
        d[2].close()
