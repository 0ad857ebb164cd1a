"""Pins of the oracle's observation operator H(Xb) and centring
(letkf_obs.F90:120-178, 187-243; obs_make.F90:233-381): closed forms that the
bilinear / log-pressure interpolation reproduces exactly, the boundary
rules, and the centring identities."""
import numpy as np
import pytest

NLAT, NLON, NZ = 10, 20, 6
GLAT = (-90.0 + (np.arange(NLAT) + 0.5) * 18.0).astype(np.float32)
GLON = (np.arange(NLON) * 18.0).astype(np.float32)
# rows: ps, then T (NZ levels), then u (NZ levels)
VAR_START = np.array([0, 1, 1 + NZ], np.int32)
VAR_LEV = np.array([1, NZ, NZ], np.int32)
NROW = 1 + 2 * NZ
SIGMA = {"type": 0, "c1": np.linspace(0.0, 1.0, NZ + 1), "ptop": 1000.0}
HYBRID = {"type": 1, "c1": np.linspace(0.3, 0.0, NZ + 1), "c2": np.linspace(0.0, 1.0, NZ + 1),
          "ptop": 500.0, "p0": 1.0e5}


def _half(vc, ps):
    c1 = np.asarray(vc["c1"])
    if vc["type"] == 0:
        return c1 * (ps - vc["ptop"]) + vc["ptop"]
    return c1 * (vc["p0"] - vc["ptop"]) + np.asarray(vc["c2"]) * (ps - vc["ptop"]) + vc["ptop"]


def _state(k, f):
    """xb, axb with x(col, row, e) = f(ilat, ilon, row, e), split as mean +
    perturbation."""
    x = np.zeros((NLAT * NLON, NROW, k))
    for ilon in range(NLON):
        for ilat in range(NLAT):
            for row in range(NROW):
                x[ilat + NLAT * ilon, row] = [f(ilat, ilon, row, e) for e in range(k)]
    axb = x.mean(axis=2)
    return x - axb[..., None], axb


def _h(yb, yo, dep):
    return yb + (yo.reshape(-1) - dep)[:, None]


def test_bilinear_exact_for_linear_fields(oracle):
    k = 5
    rng = np.random.default_rng(1)
    a, b, c = rng.normal(size=k), rng.normal(size=k), rng.normal(size=k)
    xb, axb = _state(k, lambda i, j, r, e: a[e] + b[e] * float(GLAT[i]) + c[e] * float(GLON[j])
                     + (1.0e5 if r == 0 else 0.0))
    nrec = 40
    lat = rng.uniform(float(GLAT[0]), float(GLAT[-1]), nrec)
    lon = rng.uniform(float(GLON[0]), float(GLON[-1]), nrec)
    ph = np.full((nrec, 1), 5.0e4)
    yo = rng.normal(size=(nrec, 1, 3))
    la, lo, yb, dep = oracle.obs_operator(GLAT, GLON, xb, axb, 0, SIGMA, VAR_START, VAR_LEV,
                                          [0, 1, 2], lat, lon, ph, yo)
    h = _h(yb, yo, dep).reshape(nrec, 3, k)
    exp = a[None] + b[None] * lat[:, None] + c[None] * lon[:, None]
    assert np.allclose(h[:, 0] - 1.0e5, exp, atol=1e-9)
    assert np.allclose(h[:, 1], exp, atol=1e-10)   # 3-D variables constant in the vertical
    assert np.allclose(h[:, 2], exp, atol=1e-10)
    assert np.array_equal(la, np.repeat(lat, 3)) and np.array_equal(lo, np.repeat(lon, 3))


@pytest.mark.parametrize("vc", [SIGMA, HYBRID], ids=["sigma", "hybrid"])
def test_log_pressure_exact_and_nearest_outside(oracle, vc):
    k = 4
    rng = np.random.default_rng(2)
    ps_e = 1.0e5 + 500.0 * rng.normal(size=k)
    al, be = rng.normal(size=k), rng.normal(size=k)
    pf = {e: None for e in range(k)}
    for e in range(k):
        h = _half(vc, ps_e[e])
        pf[e] = 0.5 * (h[:-1] + h[1:])
    xb, axb = _state(k, lambda i, j, r, e: ps_e[e] if r == 0 else
                     (al[e] + be[e] * np.log(pf[e][(r - 1) % NZ])))
    # levels inside every member's column, above the top and below the bottom
    top = max(pf[e][0] for e in range(k))
    bot = min(pf[e][-1] for e in range(k))
    inside = np.exp(rng.uniform(np.log(top) + 1e-3, np.log(bot) - 1e-3, 8))
    # a record's levels in ascending pressure (the level search carries over
    # them): above the top, inside, below the bottom
    ph = np.concatenate([[0.5 * min(pf[e][0] for e in range(k))], np.sort(inside),
                         [1.5 * max(pf[e][-1] for e in range(k))]])[None, :]
    onz = ph.shape[1]
    yo = np.zeros((1, onz, 1))
    _, _, yb, dep = oracle.obs_operator(GLAT, GLON, xb, axb, 0, vc, VAR_START, VAR_LEV, [1],
                                        [float(GLAT[3])], [float(GLON[5])], ph, yo)
    h = _h(yb, yo, dep)
    assert np.allclose(h[0], [al[e] + be[e] * np.log(pf[e][0]) for e in range(k)], atol=1e-12)
    for l in range(1, 9):
        assert np.allclose(h[l], al + be * np.log(ph[0, l]), rtol=1e-12, atol=1e-12), l
    assert np.allclose(h[9], [al[e] + be[e] * np.log(pf[e][-1]) for e in range(k)], atol=1e-12)
    # the search is carried over the record's levels (obs_make.F90:253-258):
    # a level after a deeper one is not searched from the top again
    ph2 = np.array([[float(np.sort(inside)[-1]), 0.5 * min(pf[e][0] for e in range(k))]])
    _, _, yb2, dep2 = oracle.obs_operator(GLAT, GLON, xb, axb, 0, vc, VAR_START, VAR_LEV, [1],
                                          [float(GLAT[3])], [float(GLON[5])], ph2,
                                          np.zeros((1, 2, 1)))
    h2 = _h(yb2, np.zeros((1, 2, 1)), dep2)
    assert not np.allclose(h2[1], [al[e] + be[e] * np.log(pf[e][0]) for e in range(k)])


def test_grid_points_periodic_seam_and_edges(oracle):
    k = 3
    rng = np.random.default_rng(3)
    vals = rng.normal(size=(NLAT * NLON, NROW, k))
    xb, axb = vals - vals.mean(axis=2, keepdims=True), vals.mean(axis=2)
    ph = np.full((5, 1), 5.0e4)
    ilat = 4
    lat = np.full(5, float(GLAT[ilat]))
    # on grid point (4, 7); on the seam 3/4 of the way from the last column to
    # 360; below the first / above the last latitude (nearest row)
    lon = np.array([float(GLON[7]), float(GLON[-1]) + 13.5, 0.0, 100.0, 100.0])
    lat[3], lat[4] = -89.9, 89.9
    yo = np.zeros((5, 1, 1))
    _, _, yb, dep = oracle.obs_operator(GLAT, GLON, xb, axb, 0, SIGMA, VAR_START, VAR_LEV, [0],
                                        lat, lon, ph, yo)
    h = _h(yb, yo, dep)
    x = lambda i, j: vals[i + NLAT * j, 0]
    assert np.allclose(h[0], x(ilat, 7), atol=1e-13)
    t = (18.0 - 13.5) / 18.0  # weight of the last column
    assert np.allclose(h[1], t * x(ilat, NLON - 1) + (1 - t) * x(ilat, 0), atol=1e-13)
    assert np.allclose(h[2], x(ilat, 0), atol=1e-13)
    # lon 100 lies between columns 5 (90) and 6 (108)
    tl = (108.0 - 100.0) / 18.0
    assert np.allclose(h[3], tl * x(0, 5) + (1 - tl) * x(0, 6), atol=1e-13)
    assert np.allclose(h[4], tl * x(NLAT - 1, 5) + (1 - tl) * x(NLAT - 1, 6), atol=1e-13)


def test_centring(oracle):
    k = 6
    rng = np.random.default_rng(4)
    vals = rng.normal(size=(NLAT * NLON, NROW, k)) + 3.0
    vals[:, 0] += 1.0e5
    xb, axb = vals - vals.mean(axis=2, keepdims=True), vals.mean(axis=2)
    nrec = 12
    ph = np.sort(rng.uniform(2.0e4, 9.0e4, (nrec, 2)), axis=1)
    yo = rng.normal(size=(nrec, 2, 3))
    _, _, yb, dep = oracle.obs_operator(GLAT, GLON, xb, axb, 0, SIGMA, VAR_START, VAR_LEV,
                                        [0, 1, 2], rng.uniform(-80, 80, nrec),
                                        rng.uniform(0, 359, nrec), ph, yo)
    assert np.allclose(yb.sum(axis=1), 0.0, atol=1e-9)
    h = _h(yb, yo, dep)
    assert np.allclose(yo.reshape(-1) - dep, h.mean(axis=1), atol=1e-9)
