"""Pins of the oracle's vertical-localisation sweep
(letkf_compute_nooverlap_vertical, letkf_compute.F90:358-526): with the
vertical weights at 1 it is the ordinary sweep, level by level; obs beyond
2 vert_dist from a cube's centre do not act on that cube; each row is written
by exactly one cube."""
import numpy as np
import pytest

NZ = 6
VS, VL = [0, 1, 1 + NZ], [1, NZ, NZ]
SIGMA = {"type": 0, "c1": np.linspace(0.0, 1.0, NZ + 1), "ptop": 1000.0}


def _case(seed, k=8, nlat=6, nlon=10, nobs=120):
    rng = np.random.default_rng(seed)
    glat = (-75.0 + np.arange(nlat) * 30.0).astype(np.float32)
    glon = (np.arange(nlon) * 36.0).astype(np.float32)
    n = 1 + 2 * NZ
    ncol = nlat * nlon
    xb = rng.normal(size=(ncol, n, k))
    xb -= xb.mean(axis=2, keepdims=True)
    axb = rng.normal(size=(ncol, n))
    axb[:, 0] = 1.0e5 + 300.0 * rng.normal(size=ncol)
    lat = rng.uniform(-80, 80, nobs)
    lon = rng.uniform(0, 360, nobs)
    yb = rng.normal(size=(nobs, k))
    yb -= yb.mean(axis=1, keepdims=True)
    dep = rng.normal(size=nobs)
    rinv = (1.0 / rng.uniform(0.5, 2.0, nobs) ** 2).astype(np.float32)
    ph = rng.uniform(1.0e4, 1.0e5, nobs)
    return glat, glon, xb, axb, (lat, lon, yb, dep, rinv), ph


def _run(oracle, seed, vert_dist, cube_nz, gcv=True, drop=None):
    glat, glon, xb, axb, o, ph = _case(seed)
    k, n = xb.shape[2], xb.shape[1]
    obs = oracle.Obs(*o)
    if drop is not None:
        keep = ~drop(ph)
        obs = oracle.Obs(*(a[keep] for a in o))
        ph = ph[keep]
    prm = oracle.params(k, n, len(glat), len(glon), 3000.0, 1.1)
    return (oracle.sweep_vertical(prm, glat, glon, obs, ph, xb, axb, SIGMA, 0, VS, VL,
                                  vert_dist, cube_nz, gcv), oracle.sweep(prm, glat, glon, obs, xb, axb),
            ph, axb)


@pytest.mark.parametrize("cube_nz", [1, 2, 4, NZ])
def test_unit_vertical_weights_give_the_plain_sweep(oracle, cube_nz):
    v, s, _, _ = _run(oracle, 1, 1.0e30, cube_nz)
    assert v["nfail"] == 0
    assert np.allclose(v["xa"], s["xa"], atol=1e-11)
    assert np.allclose(v["axa"], s["axa"], atol=1e-11)


def test_distant_obs_do_not_act(oracle):
    # one cube of the top level only; obs more than 2 vert_dist below it in
    # ln p carry zero weight, so removing them changes nothing in those rows
    glat, glon, xb, axb, o, ph = _case(2)
    k, n = xb.shape[2], xb.shape[1]
    c1 = SIGMA["c1"]
    half = lambda ps: c1 * (ps - 1000.0) + 1000.0
    aps = axb[:, 0]
    ptop_full = np.array([0.5 * (half(p)[0] + half(p)[1]) for p in aps])
    vd = 0.2
    far = lambda p: np.log(p) - np.log(ptop_full.max()) > 2.0 * vd + 1e-9
    prm = oracle.params(k, n, len(glat), len(glon), 3000.0, 1.1)
    a = oracle.sweep_vertical(prm, glat, glon, oracle.Obs(*o), ph, xb, axb, SIGMA, 0, VS, VL,
                              vd, 1, True)
    keep = ~far(ph)
    assert 0 < keep.sum() < len(ph)
    b = oracle.sweep_vertical(prm, glat, glon, oracle.Obs(*(x[keep] for x in o)), ph[keep], xb,
                              axb, SIGMA, 0, VS, VL, vd, 1, True)
    top_rows = [VS[1], VS[2]]   # level 0 of the two 3-D variables
    assert np.allclose(a["xa"][:, top_rows], b["xa"][:, top_rows], atol=1e-12)
    assert np.allclose(a["axa"][:, top_rows], b["axa"][:, top_rows], atol=1e-12)


def test_every_row_written_once(oracle):
    v, s, _, axb = _run(oracle, 3, 0.3, 4)
    # re-centred analysis perturbations on every row, finite everywhere
    assert np.allclose(v["xa"].sum(axis=2), 0.0, atol=1e-14 * np.abs(v["axa"]).max() * 8)
    assert np.isfinite(v["axa"]).all() and (v["axa"] != 0).all()


@pytest.mark.parametrize("cube_n,cube_nz", [(2, 1), (3, 4), (4, NZ)])
def test_horizontal_cubes_with_unit_vertical_weights_are_the_cube_sweep(oracle, cube_n, cube_nz):
    # every vertical cube of a horizontal cube takes the same analysis, so the
    # result is the horizontal-cube sweep without vertical localisation
    glat, glon, xb, axb, o, ph = _case(3)
    k, n = xb.shape[2], xb.shape[1]
    obs = oracle.Obs(*o)
    prm = oracle.params(k, n, len(glat), len(glon), 3000.0, 1.1)
    v = oracle.sweep_vertical(prm, glat, glon, obs, ph, xb, axb, SIGMA, 0, VS, VL, 1.0e30,
                              cube_nz, True, cube_n=cube_n)
    c = oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, cube_n, 0)
    assert v["nfail"] == 0
    assert np.allclose(v["xa"], c["xa"], atol=1e-11)
    assert np.allclose(v["axa"], c["axa"], atol=1e-11)


def test_horizontal_cube_levels_use_the_centre_column_pressure(oracle):
    # cube_n = 3 on the 6 x 10 grid: the cube of columns (0..2) x (0..2) is
    # centred on column (1, 1); its analyses are those of the one-column sweep
    # at (1, 1), so column (1, 1) matches that sweep and any other column of
    # the cube takes the same transforms (its rows differ only through xb)
    glat, glon, xb, axb, o, ph = _case(4)
    k, n = xb.shape[2], xb.shape[1]
    obs = oracle.Obs(*o)
    prm = oracle.params(k, n, len(glat), len(glon), 3000.0, 1.1)
    v3 = oracle.sweep_vertical(prm, glat, glon, obs, ph, xb, axb, SIGMA, 0, VS, VL, 0.4, 2, True,
                               cube_n=3)
    v1 = oracle.sweep_vertical(prm, glat, glon, obs, ph, xb, axb, SIGMA, 0, VS, VL, 0.4, 2, True)
    c = 1 + len(glat) * 1
    assert np.allclose(v3["xa"][c], v1["xa"][c], atol=1e-12)
    assert np.allclose(v3["axa"][c], v1["axa"][c], atol=1e-12)
    # column (0, 0) with the centre's xb gives the centre's result
    xb2, axb2 = xb.copy(), axb.copy()
    xb2[0], axb2[0] = xb[c], axb[c]
    v3b = oracle.sweep_vertical(prm, glat, glon, obs, ph, xb2, axb2, SIGMA, 0, VS, VL, 0.4, 2,
                                True, cube_n=3)
    assert np.allclose(v3b["xa"][0], v3b["xa"][c], atol=1e-12)
    assert np.allclose(v3b["axa"][0], v3b["axa"][c], atol=1e-12)
