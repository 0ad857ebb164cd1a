"""GPU parity of the horizontal-cube sweeps (cube_n x cube_n columns per
analysis, letkf_compute_nooverlap with cube_n > 1; overlapping cubes averaged
per column, letkf_compute + xb_add2_xa) through the C ABI, against the
oracle's restatement on the same inputs.

Bars: per-variable analysis-increment and Xa' relative error <= 1e-10 with
the FP64 forms (plain and batched), <= 1e-4 with the split BF16 Gram; RTPS
as the oracle's post step; a captured graph replays the eager result exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _var_rows(nz):
    return {"ps": [0], "u": list(range(1, 1 + nz)), "v": list(range(1 + nz, 1 + 2 * nz)),
            "T": list(range(1 + 2 * nz, 1 + 3 * nz))}


def _rel_err(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = {
    # name: (form, gram, cube_n, overlap, extra params, tolerance)
    "plain_fp64_cube3": ("plain", "fp64", 3, 0, {}, 1e-10),
    "batched_fp64_cube4_overlap2": ("batched", "fp64", 4, 2, {}, 1e-10),
    "batched_fp64_cube2_overlap1_capped": ("batched", "fp64", 2, 1, {"max_obs": 80}, 1e-10),
    "plain_fp64_cube4_overlap3_rtps": ("plain", "fp64", 4, 3, {"rtps": 1, "rtps_alpha": 0.7},
                                       1e-10),
    "batched_bf16x3_cube3_overlap1": ("batched", "bf16x3", 3, 1, {}, 1e-4),
    # k > 64: the plain form's update on the VALU kernel, trans in LDS
    "plain_fp64_k100_cube3_overlap1": ("plain", "fp64", 3, 1, {"k": 100}, 1e-10),
}


@pytest.mark.parametrize("case", list(CASES))
def test_cube_sweep_matches_oracle(cuda, dl, oracle, case):
    torch = cuda
    form, gram, cube_n, overlap, extra, tol = CASES[case]
    extra = dict(extra)
    k = extra.pop("k", 24)
    print(f"cube sweep case {case}", flush=True)
    cfg = dl.synth_cfg(37, 71, 3, k, 3000, seed=0xD1DA0300 + len(case), cluster_frac=0.5,
                       nclusters=3, cluster_radius_km=900.0, rscale=(0.5, 2.0))
    data = dl.synth_generate(cfg)
    n = 1 + 3 * cfg.nz
    bin_deg = 3.0
    p = dict(nanals=cfg.k, nrow=n, nlat=cfg.nlat, nlon=cfg.nlon, local_dist=900.0,
             inflation_factor=1.05, bin_deg=bin_deg, cube_n=cube_n, overlap=overlap,
             form=dl.FORM_PLAIN if form == "plain" else dl.FORM_BATCHED,
             gram=dl.GRAM_FP64 if gram == "fp64" else dl.GRAM_BF16X3)
    p.update(extra)
    params = dl.default_params(**p)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    step = cube_n - overlap if overlap else cube_n
    nanal = -(-cfg.nlat // step) * -(-cfg.nlon // step)
    assert st["status"] == 0 and st["nonconverged"] == 0 and st["nanal"] == nanal, st
    if "max_obs" in extra:
        assert st["nobsl_capped"] > 0
    h = {k: v.cpu().numpy() for k, v in data.items()}
    o = oracle.Obs(h["lat"], h["lon"], h["yb"], h["dep"], h["rinv"])
    perm, bin_start, nband, nlonbin = oracle.bin_obs(o, bin_deg, params.obs_polar)
    prm = oracle.params(params.nanals, n, cfg.nlat, cfg.nlon, params.local_dist,
                        params.inflation_factor, max_obs=params.max_obs)
    out = oracle.sweep_cubes(prm, h["grid_lat"], h["grid_lon"], o.permuted(perm), h["xb"],
                             h["axb"], cube_n, overlap, bins=(bin_start, nband, nlonbin, bin_deg, perm))
    assert out["nfail"] == 0
    exp_xa, exp_axa = out["xa"], out["axa"]
    if params.rtps:
        full, _, _ = oracle.post(1, h["grid_lat"], cfg.nlat, cfg.nlon, 90.0, 0, 0,
                                 params.rtps_alpha, h["xb"], h["axb"], exp_xa, exp_axa)
        exp_xa = full - exp_axa[..., None]
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(cfg.nz).items():
        inc_g = axa_h[:, rows] - h["axb"][:, rows]
        inc_o = exp_axa[:, rows] - h["axb"][:, rows]
        assert _rel_err(inc_g, inc_o) <= tol, var
        assert _rel_err(xa_h[:, rows], exp_xa[:, rows]) <= tol, var
    if case == "batched_fp64_cube4_overlap2":
        s = torch.cuda.Stream()
        xg, axg = torch.zeros_like(xa), torch.zeros_like(axa)
        with torch.cuda.stream(s):
            eng.graph_capture(obs, data["xb"], data["axb"], xg, axg, s)
        for _ in range(2):
            eng.graph_launch(s)
            s.synchronize()
            assert torch.equal(xg, xa) and torch.equal(axg, axa)
            xg.zero_()
            axg.zero_()
    eng.close()

