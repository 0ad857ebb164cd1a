"""GPU parity of the selection at the support edge.

Obs are placed on rings around chosen columns, at central angles within a
few parts in 1e5 of 2L/a in the reference's radians (REAL(4) pi), so that
about half of them fall in the thin shell that the device's FP32 unit-vector
test leaves open and that the FP64 unit-vector cosine decides against
cos(2L/a).  The device lists must equal the oracle's (the reference formula,
letkf_yb.F90:84-96 with latlon_dist_r8, pro_info.F90:157) entry for entry,
with weights within 1e-12, at supports inside the chord-angle range (L =
600 km, the shorter series, and 1000 km) and beyond it (L = 1500 km).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A_REF = float(np.float32(6371.393))        # pro_info.F90:148, REAL(4)
PI_REF = float(np.float32(np.arctan(1.0) * 4.0))  # pro_info.F90:156, REAL(4)


def _ring(clat, clon, theta, beta):
    """Points at central angles theta (reference radians) and bearings beta
    from (clat, clon): the destination formula on the sphere whose
    coordinates are the reference's radians."""
    s = PI_REF / 180.0
    x1, y1 = clat * s, clon * s
    x2 = np.arcsin(np.sin(x1) * np.cos(theta) + np.cos(x1) * np.sin(theta) * np.cos(beta))
    y2 = y1 + np.arctan2(np.sin(beta) * np.sin(theta) * np.cos(x1),
                         np.cos(theta) - np.sin(x1) * np.sin(x2))
    return x2 / s, np.mod(y2 / s, 360.0)


@pytest.mark.parametrize("local_dist", [600.0, 1000.0, 1500.0])
def test_selection_at_the_support_edge_matches_oracle(cuda, dl, oracle, local_dist):
    torch = cuda
    cfg = dl.synth_cfg(18, 36, 2, 16, 3200, seed=0xD1DA0030)
    print(f"support-edge selection L = {local_dist}", flush=True)
    data = dl.synth_generate(cfg)
    glat = data["grid_lat"].cpu().numpy().astype(np.float64)
    glon = data["grid_lon"].cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(0x5E11)
    centres = [(2, 3), (5, 10), (9, 17), (12, 25), (15, 31), (17, 0), (0, 20), (8, 8)]
    per = cfg.nobs // len(centres)
    th0 = 2.0 * local_dist / A_REF
    lat = np.empty(cfg.nobs)
    lon = np.empty(cfg.nobs)
    for c, (j, i) in enumerate(centres):
        sl = slice(c * per, cfg.nobs if c == len(centres) - 1 else (c + 1) * per)
        n = sl.stop - sl.start
        theta = th0 * (1.0 + rng.uniform(-3e-5, 3e-5, n))
        lat[sl], lon[sl] = _ring(glat[j], glon[i], theta, rng.uniform(0.0, 2.0 * np.pi, n))
    data["lat"].copy_(torch.from_numpy(lat))
    data["lon"].copy_(torch.from_numpy(lon))
    n_row = 1 + 3 * cfg.nz
    params = dl.default_params(nanals=cfg.k, nrow=n_row, nlat=cfg.nlat, nlon=cfg.nlon,
                               local_dist=local_dist, bin_deg=3.0, max_obs=100000)
    eng = dl.Engine(params, nobs_max=cfg.nobs)
    eng.set_grid(data["grid_lat"], data["grid_lon"])
    obs = dl.DeviceObs(data["lat"], data["lon"], data["yb"], data["dep"], data["rinv"])
    xa, axa = torch.empty_like(data["xb"]), torch.empty_like(data["axb"])
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    assert st["status"] == 0 and st["capacity_overflow"] == 0
    dbg = eng.debug(cfg.nobs)
    o = oracle.Obs(lat, lon, data["yb"].cpu().numpy(), data["dep"].cpu().numpy(),
                   data["rinv"].cpu().numpy())
    prm = oracle.params(cfg.k, n_row, cfg.nlat, cfg.nlon, local_dist, 1.0, max_obs=100000)
    for c, (j, i) in enumerate(centres):
        # the ring straddles the edge: its column keeps part of it and not all
        sl = slice(c * per, cfg.nobs if c == len(centres) - 1 else (c + 1) * per)
        idx, _ = oracle.select_column(prm, o, glat[j], glon[i])
        kept = np.count_nonzero((idx >= sl.start) & (idx < sl.stop))
        assert 0.2 * (sl.stop - sl.start) < kept < 0.8 * (sl.stop - sl.start), (j, i, kept)
    for col in range(cfg.nlat * cfg.nlon):
        idx, w = oracle.select_column(prm, o, glat[col % cfg.nlat], glon[col // cfg.nlat])
        off, p = int(dbg["offset"][col]), int(dbg["nobsl"][col])
        got = dbg["perm"][dbg["csr_idx"][off:off + p]]
        order = np.argsort(got)
        assert np.array_equal(got[order], idx), col
        assert np.abs(dbg["csr_w"][off:off + p][order] - w).max(initial=0.0) <= 1e-12, col
    eng.close()
