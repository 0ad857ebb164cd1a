"""Failure reporting through the C ABI: an eigensolve that hits the sweep cap
and an analysis that meets a non-finite input.

The reference prints a failed eigensolve (letkf_core.F90:119-122) and its
Sunway caller stops on a NaN transform (letkf_compute.F90:168-187).  Here a
capped eigensolve sets info = -1 and the sweep status DIDA_WARN_NOCONV; an
analysis whose M or b is not finite sets info = -2, its transform (so its
columns) is NaN, the status is DIDA_ERR_NONFINITE, and every other analysis
is the one a clean run gives.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = {
    "plain": dict(form=0, gram=0, k=20),
    "batched_fp64": dict(form=1, gram=0, k=32),
    "batched_bf16x3": dict(form=1, gram=1, k=32),
    "chained_fp64": dict(form=1, gram=0, k=96),
}


def _engine(dl, torch, data, cfg, form, gram, **over):
    n = 1 + 3 * cfg.nz
    prm = dl.default_params(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon, form=form,
                            gram=gram, local_dist=900.0, inflation_factor=1.05, bin_deg=4.0,
                            **over)
    eng = dl.Engine(prm, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    return eng


@pytest.mark.parametrize("name", list(FORMS))
def test_sweep_cap_reports_nonconverged(cuda, dl, name):
    torch = cuda
    f = FORMS[name]
    cfg = dl.synth_cfg(16, 32, 2, f["k"], 1500, seed=21)
    data = dl.synth_generate(cfg)
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    print(f"sweep cap {name}", flush=True)
    # one sweep cannot converge a Jacobi that rotates (it ends after a sweep
    # that rotates nothing), and prediction off keeps it from stopping early
    eng = _engine(dl, torch, data, cfg, f["form"], f["gram"], jacobi_max_sweeps=1,
                  jacobi_predict=0)
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    dbg = eng.debug(cfg.nobs)
    assert st["status"] == dl.DIDA_WARN_NOCONV
    with_obs = dbg["nobsl"] > 0
    assert with_obs.sum() > 100
    assert st["nonconverged"] == int((dbg["info"] == -1).sum()) > 0
    assert np.all(dbg["info"][~with_obs] == 1)        # nothing to rotate: converged
    assert st["nonfinite"] == 0
    eng.close()


@pytest.mark.parametrize("name", list(FORMS))
def test_nonfinite_obs_flags_only_its_columns(cuda, dl, name):
    torch = cuda
    f = FORMS[name]
    cfg = dl.synth_cfg(16, 32, 2, f["k"], 1500, seed=22)
    data = dl.synth_generate(cfg)
    print(f"non-finite {name}", flush=True)
    xa0, axa0 = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng = _engine(dl, torch, data, cfg, f["form"], f["gram"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    eng.sweep(obs, data["xb"], data["axb"], xa0, axa0)
    assert eng.stats()["status"] == 0
    dbg = eng.debug(cfg.nobs)
    # poison one obs: a NaN in its Yb row (and an Inf innovation on another)
    j_nan, j_inf = 700, 1200
    yb = data["yb"].clone()
    yb[j_nan, 3] = float("nan")
    dep = data["dep"].clone()
    dep[j_inf] = float("inf")
    bad_obs = dl.DeviceObs(data["lat"], data["lon"], yb, dep, data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(bad_obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats(strict=False)
    dbg2 = eng.debug(cfg.nobs)
    ncol = cfg.nlat * cfg.nlon
    caller = dbg["perm"][dbg["csr_idx"]]
    col_of = np.repeat(np.arange(ncol), dbg["nobsl"])
    hit = np.zeros(ncol, bool)
    hit[col_of[(caller == j_nan) | (caller == j_inf)]] = True
    assert 0 < hit.sum() < ncol // 2
    assert st["status"] == dl.DIDA_ERR_NONFINITE
    assert st["nonfinite"] == int(hit.sum())
    assert np.array_equal(dbg2["info"] == -2, hit)
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    xa0_h, axa0_h = xa0.cpu().numpy(), axa0.cpu().numpy()
    assert np.all(np.isnan(axa_h[hit])) and np.all(np.isnan(xa_h[hit]))
    ok = ~hit
    assert np.all(np.isfinite(xa_h[ok])) and np.all(np.isfinite(axa_h[ok]))
    if f["form"] == 0:
        # no state between analyses: bit-equal to the clean run
        assert np.array_equal(xa_h[ok], xa0_h[ok]) and np.array_equal(axa_h[ok], axa0_h[ok])
    else:
        # a chain restarts cold after a flagged analysis, so the next ones
        # converge from another start: equal within the form's bar
        tol = 1e-10 if f["gram"] == 0 else 1e-4
        inc, inc0 = axa_h[ok] - data["axb"].cpu().numpy()[ok], axa0_h[ok] - \
            data["axb"].cpu().numpy()[ok]
        assert np.linalg.norm(inc - inc0) <= tol * np.linalg.norm(inc0)
        assert np.linalg.norm(xa_h[ok] - xa0_h[ok]) <= tol * np.linalg.norm(xa0_h[ok])
    # a clean sweep on the same context clears the flags
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    assert st["status"] == 0 and st["nonfinite"] == 0
    eng.close()


def test_nonfinite_in_letkf_core_batch(cuda, dl):
    """The batched core: info -2 and NaN trans / wbar for the analysis with a
    NaN Yb entry; the others equal a clean call."""
    torch = cuda
    rng = np.random.default_rng(5)
    k, counts, max_obs = 24, [10, 50, 50, 0], 50
    b = len(counts)
    hx = np.zeros((b, max_obs, k))
    for j, p in enumerate(counts):
        h = rng.normal(size=(p, k))
        hx[j, :p] = h - h.mean(axis=1, keepdims=True)
    ri = np.full((b, max_obs), 0.5, np.float32)
    dep = rng.normal(size=(b, max_obs))
    rl = np.ones((b, max_obs))
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    nb = dev(np.array(counts, np.int32))
    outs = []
    for poison in (False, True):
        h = hx.copy()
        if poison:
            h[1, 7, 2] = np.nan
        tr = torch.empty((b, k, k), dtype=torch.float64, device="cuda")
        wb = torch.empty((b, k), dtype=torch.float64, device="cuda")
        info = torch.empty(b, dtype=torch.int32, device="cuda")
        dl.core_batch_dev(nb, dev(h), dev(ri), dev(dep), dev(rl), tr, wb, 1.1, info=info)
        torch.cuda.synchronize()
        outs.append((tr.cpu().numpy(), wb.cpu().numpy(), info.cpu().numpy()))
    (t0, w0, i0), (t1, w1, i1) = outs
    assert np.all(i0 >= 1) and i1[1] == -2 and np.all(np.delete(i1, 1) >= 1)
    assert np.all(np.isnan(t1[1])) and np.all(np.isnan(w1[1]))
    keep = [0, 2, 3]
    assert np.array_equal(t1[keep], t0[keep]) and np.array_equal(w1[keep], w0[keep])


def test_workspace_grows_with_the_lists(cuda, dl, oracle):
    """A context sized by a small sweep grows its selection workspace when a
    later sweep keeps more obs (the result matches the oracle), drops the
    graph captured on the old buffers; a fixed workspace too small for the
    fused batched form is refused before any analysis runs."""
    torch = cuda
    cfg = dl.synth_cfg(18, 36, 2, 16, 2400, seed=23)
    data = dl.synth_generate(cfg)
    n = 1 + 3 * cfg.nz
    prm = dl.default_params(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon,
                            local_dist=900.0, bin_deg=4.0)
    eng = dl.Engine(prm, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    few = dl.DeviceObs(*(data[key][:200] for key in ("lat", "lon", "yb", "dep", "rinv")))
    full = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(few, data["xb"], data["axb"], xa, axa)
    small = eng.stats()["nobsl_sum"]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng.graph_capture(few, data["xb"], data["axb"], xa, axa, s)
    s.synchronize()
    eng.sweep(full, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    assert st["status"] == 0 and st["nobsl_sum"] > 4 * small
    with pytest.raises(dl.DidaError):
        eng.graph_launch(s)          # the graph referred to the replaced workspace
    h = {k: v.cpu().numpy() for k, v in data.items()}
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    op = oracle.params(cfg.k, n, cfg.nlat, cfg.nlon, 900.0, 1.0)
    ref = oracle.sweep(op, h["grid_lat"], h["grid_lon"], o, h["xb"], h["axb"])
    inc = axa.cpu().numpy() - h["axb"]
    inc_o = ref["axa"] - h["axb"]
    assert np.linalg.norm(inc - inc_o) <= 1e-10 * np.linalg.norm(inc_o)
    assert np.linalg.norm(xa.cpu().numpy() - ref["xa"]) <= 1e-10 * np.linalg.norm(ref["xa"])
    eng.close()
    # a fixed workspace below the sweep's lists: the fused form refuses the sweep
    prm_b = dl.default_params(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon,
                              local_dist=900.0, bin_deg=4.0, form=dl.FORM_BATCHED)
    eng = dl.Engine(prm_b, nobs_max=cfg.nobs, capacity=small + 10)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    with pytest.raises(dl.DidaError, match="status -4"):
        eng.sweep(full, data["xb"], data["axb"], xa, axa)
    eng.close()
