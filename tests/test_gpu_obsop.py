"""GPU observation operator H(Xb) + centring (dida_letkf_obs_operator)
against the oracle on the same inputs, and its outputs driving a sweep.

Bars: scalar obs positions identical; yb and dep within 1e-12 relative (the
ensemble mean is a wave reduction here and a sequential sum in the oracle;
log() may differ in the last bit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(dl, torch, k, vtype, seed):
    nz = 8
    cfg = dl.synth_cfg(45, 90, nz, k, 10, seed=seed)
    g = dl.synth_generate(cfg)
    rng = np.random.default_rng(seed)
    var_start = [0, 1, 1 + nz, 1 + 2 * nz]           # ps, u, v, T
    var_lev = [1, nz, nz, nz]
    if vtype == 0:
        vc = {"type": 0, "c1": np.linspace(0.0, 1.0, nz + 1), "ptop": 10.0}
    else:
        vc = {"type": 1, "c1": np.linspace(0.25, 0.0, nz + 1), "c2": np.linspace(0.0, 1.0, nz + 1),
              "ptop": 5.0, "p0": 1000.0}
    nrec, onz = 3000, 3
    lat = rng.uniform(-89.5, 89.5, nrec)
    lon = rng.uniform(0.0, 360.0, nrec)
    ph = np.sort(rng.uniform(5.0, 1100.0, (nrec, onz)), axis=1)   # hPa, some outside
    vid = [3, 1, 0] if seed % 2 else [2, 3]
    yo = rng.normal(size=(nrec, onz, len(vid)))
    return g, vc, var_start, var_lev, vid, lat, lon, ph, yo


@pytest.mark.parametrize("k,vtype", [(20, 0), (64, 1), (100, 0)])
def test_obs_operator_matches_oracle(cuda, dl, oracle, k, vtype):
    torch = cuda
    print(f"obs operator k={k} vcoord={vtype}", flush=True)
    g, vc, vs, vl, vid, lat, lon, ph, yo = _case(dl, torch, k, vtype, 11 + k)
    dev = "cuda"
    T = lambda x, dt=torch.float64: torch.tensor(np.ascontiguousarray(x), dtype=dt, device=dev)
    vcd = {key: (T(v) if key in ("c1", "c2") else v) for key, v in vc.items()}
    la, lo, yb, dep = dl.obs_operator(g["grid_lat"], g["grid_lon"], g["xb"], g["axb"], 0, vcd, vs,
                                      vl, vid, T(lat), T(lon), T(ph), T(yo))
    torch.cuda.synchronize()
    h = {key: g[key].cpu().numpy() for key in ("grid_lat", "grid_lon", "xb", "axb")}
    ola, olo, oyb, odep = oracle.obs_operator(h["grid_lat"], h["grid_lon"], h["xb"], h["axb"], 0,
                                              vc, vs, vl, vid, lat, lon, ph, yo)
    assert np.array_equal(la.cpu().numpy(), ola) and np.array_equal(lo.cpu().numpy(), olo)
    yb_h, dep_h = yb.cpu().numpy(), dep.cpu().numpy()
    assert np.abs(yb_h - oyb).max() <= 1e-12 * np.abs(oyb).max()
    assert np.abs(dep_h - odep).max() <= 1e-12 * max(np.abs(odep).max(), 1.0)
    assert np.allclose(yb_h.sum(axis=1), 0.0, atol=1e-9 * np.abs(oyb).max())


def test_obs_operator_feeds_a_sweep(cuda, dl):
    torch = cuda
    k = 32
    g, vc, vs, vl, vid, lat, lon, ph, yo = _case(dl, torch, k, 0, 5)
    dev = "cuda"
    T = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    vcd = {"type": 0, "c1": T(vc["c1"]), "ptop": vc["ptop"]}
    yo = 0.1 * yo
    la, lo, yb, dep = dl.obs_operator(g["grid_lat"], g["grid_lon"], g["xb"], g["axb"], 0, vcd, vs,
                                      vl, vid, T(lat), T(lon), T(ph), T(yo))
    rinv = torch.full((la.numel(),), 1.0, dtype=torch.float32, device=dev)
    nrow = g["axb"].shape[1]
    params = dl.default_params(nanals=k, nrow=nrow, nlat=45, nlon=90, local_dist=800.0,
                               form=dl.FORM_BATCHED, gram=dl.GRAM_FP64)
    eng = dl.Engine(params, nobs_max=la.numel())
    eng.set_grid(g["grid_lat"], g["grid_lon"])
    obs = dl.DeviceObs(la, lo, yb, dep, rinv)
    xa, axa = torch.empty_like(g["xb"]), torch.empty_like(g["axb"])
    eng.sweep(obs, g["xb"], g["axb"], xa, axa)
    st = eng.stats()
    assert st["status"] == 0 and st["nonconverged"] == 0 and st["nobsl_sum"] > 0, st
    assert torch.isfinite(xa).all() and torch.isfinite(axa).all()
    eng.close()
