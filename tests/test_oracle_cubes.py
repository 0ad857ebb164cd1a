"""Pins of the oracle's horizontal-cube sweeps (letkf_compute_nooverlap with
cube_n > 1, letkf_compute.F90:12-126; letkf_compute with overlapping cubes,
:128-228, xb_add2_xa :304-354): cubes of one column are the ordinary sweep;
a cube's analysis is the one centred on its lower-middle column (a single obs
at a grid point with a tiny support is seen only by the cube centred there);
with no obs every transform is sqrt(rho) I, so the average over overlapping
cubes must return sqrt(rho) Xb' (a missing division by the cube count
fails); overlapping cubes average the covering cubes' analyses."""
import numpy as np
import pytest


def _grid(nlat=8, nlon=8):
    glat = (-70.0 + 20.0 * np.arange(nlat)).astype(np.float32)
    glon = (45.0 * np.arange(nlon)).astype(np.float32)
    return glat, glon


def _state(rng, ncol, n, k):
    xb = rng.normal(size=(ncol, n, k))
    xb -= xb.mean(axis=2, keepdims=True)
    axb = rng.normal(size=(ncol, n))
    return xb, axb


def _obs_at(oracle, glat, glon, pts, k, rng, dep=3.0):
    lat = np.array([float(glat[j]) for j, _ in pts])
    lon = np.array([float(glon[i]) for _, i in pts])
    yb = rng.normal(size=(len(pts), k))
    yb -= yb.mean(axis=1, keepdims=True)
    return oracle.Obs(lat, lon, yb, np.full(len(pts), dep), np.ones(len(pts), np.float32))


def _inc(out, axb):
    return np.abs(out["axa"] - axb).max(axis=1)


def test_cube_of_one_column_is_the_sweep(oracle):
    rng = np.random.default_rng(1)
    glat, glon = _grid()
    k, n = 6, 3
    xb, axb = _state(rng, 64, n, k)
    pts = [(int(j), int(i)) for j, i in rng.integers(0, 8, size=(12, 2))]
    obs = _obs_at(oracle, glat, glon, pts, k, rng)
    prm = oracle.params(k, n, 8, 8, 2500.0, 1.2)
    a = oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, 1, 0)
    b = oracle.sweep(prm, glat, glon, obs, xb, axb)
    assert np.array_equal(a["xa"], b["xa"]) and np.array_equal(a["axa"], b["axa"])


def test_cube_analysis_is_centred_on_the_lower_middle_column(oracle):
    rng = np.random.default_rng(2)
    glat, glon = _grid()
    k, n = 6, 3
    xb, axb = _state(rng, 64, n, k)
    obs = _obs_at(oracle, glat, glon, [(1, 1)], k, rng)   # at column (lat 1, lon 1)
    prm = oracle.params(k, n, 8, 8, 100.0, 1.0)           # support 200 km < spacing
    out = oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, 4, 0)
    inc = _inc(out, axb).reshape(8, 8)                    # [ilon, ilat]
    assert (inc[:4, :4] > 1e-3).all()                     # the cube centred on (1, 1)
    inc[:4, :4] = 0.0
    assert inc.max() < 1e-12
    ref = oracle.sweep(prm, glat, glon, obs, xb, axb, want_trans=True)
    c = 1 + 8 * 1
    assert np.allclose(out["xa"][c], ref["xa"][c], atol=1e-12, rtol=0)
    # every column of the cube takes the centre's transform
    for col in (0, 3 + 8 * 2, 2 + 8 * 3):
        xa_c, axa_c = oracle.xb_2_xa(xb[col], axb[col], ref["trans"][c])
        assert np.allclose(out["xa"][col], xa_c, atol=1e-12, rtol=0)
        assert np.allclose(out["axa"][col], axa_c, atol=1e-12, rtol=0)


def test_overlap_without_obs_averages_to_sqrt_rho(oracle):
    rng = np.random.default_rng(3)
    glat, glon = _grid(10, 9)
    k, n = 5, 4
    xb, axb = _state(rng, 90, n, k)
    obs = oracle.Obs(np.zeros(0), np.zeros(0), np.zeros((0, k)), np.zeros(0),
                     np.zeros(0, np.float32))
    rho = 1.3
    prm = oracle.params(k, n, 10, 9, 500.0, rho)
    for cube_n, overlap in ((4, 2), (3, 1), (5, 4), (2, 1)):
        out = oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, cube_n, overlap)
        assert np.allclose(out["xa"], np.sqrt(rho) * xb, atol=1e-13, rtol=0), (cube_n, overlap)
        assert np.allclose(out["axa"], axb, atol=1e-13, rtol=0)


def test_overlap_averages_the_covering_cubes(oracle):
    # cube_n 4, overlap 2: starts 0, 2, 4, 6; the cube starting at (2, 2) is
    # centred on (3, 3) and alone sees an obs there; column (3, 3) is covered
    # by 2 x 2 cubes, so its increment is a quarter of the one-column sweep's
    rng = np.random.default_rng(4)
    glat, glon = _grid()
    k, n = 6, 3
    xb, axb = _state(rng, 64, n, k)
    obs = _obs_at(oracle, glat, glon, [(3, 3)], k, rng)
    prm = oracle.params(k, n, 8, 8, 100.0, 1.0)
    out = oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, 4, 2)
    ref = oracle.sweep(prm, glat, glon, obs, xb, axb)
    c = 3 + 8 * 3
    assert np.allclose(out["axa"][c] - axb[c], 0.25 * (ref["axa"][c] - axb[c]), atol=1e-13,
                       rtol=0)
    assert np.allclose(out["xa"][c], 0.25 * ref["xa"][c] + 0.75 * xb[c], atol=1e-13, rtol=0)
    inc = _inc(out, axb).reshape(8, 8)
    assert (inc[2:6, 2:6] > 1e-4).all()
    inc[2:6, 2:6] = 0.0
    assert inc.max() < 1e-12


@pytest.mark.parametrize("cube_n,overlap", [(3, 0), (3, 1), (4, 2)])
def test_cubes_against_per_cube_analyses(oracle, cube_n, overlap):
    rng = np.random.default_rng(5 + cube_n + overlap)
    nlat, nlon, k, n = 11, 10, 6, 3
    glat = (-80.0 + 16.0 * np.arange(nlat)).astype(np.float32)
    glon = (36.0 * np.arange(nlon)).astype(np.float32)
    xb, axb = _state(rng, nlat * nlon, n, k)
    nobs = 150
    yb = rng.normal(size=(nobs, k))
    yb -= yb.mean(axis=1, keepdims=True)
    obs = oracle.Obs(rng.uniform(-85, 85, nobs), rng.uniform(0, 360, nobs), yb,
                     rng.normal(size=nobs), np.ones(nobs, np.float32))
    prm = oracle.params(k, n, nlat, nlon, 1500.0, 1.1)
    out = oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, cube_n, overlap, want_trans=True)
    ref = oracle.sweep(prm, glat, glon, obs, xb, axb, want_trans=True)
    step = cube_n - overlap if overlap else cube_n
    starts = lambda m: [(s, min(s + cube_n - 1, m - 1)) for s in range(0, m, step)]
    slat, slon = starts(nlat), starts(nlon)
    for ai, (si, ei) in enumerate(slon):
        for aj, (sj, ej) in enumerate(slat):
            a = aj + len(slat) * ai
            cen = (sj + ej) // 2 + nlat * ((si + ei) // 2)
            assert np.allclose(out["trans"][a], ref["trans"][cen], atol=1e-13, rtol=0)
    for col in range(nlat * nlon):
        ilat, ilon = col % nlat, col // nlat
        xs, ms = [], []
        for ai, (si, ei) in enumerate(slon):
            for aj, (sj, ej) in enumerate(slat):
                if si <= ilon <= ei and sj <= ilat <= ej:
                    x, m = oracle.xb_2_xa(xb[col], axb[col], out["trans"][aj + len(slat) * ai])
                    xs.append(x)
                    ms.append(m)
        assert np.allclose(out["xa"][col], np.mean(xs, axis=0), atol=1e-12, rtol=0), col
        assert np.allclose(out["axa"][col], np.mean(ms, axis=0), atol=1e-12, rtol=0), col


def test_bad_cube_arguments(oracle):
    glat, glon = _grid()
    obs = oracle.Obs(np.zeros(0), np.zeros(0), np.zeros((0, 4)), np.zeros(0),
                     np.zeros(0, np.float32))
    prm = oracle.params(4, 1, 8, 8, 500.0, 1.0)
    xb, axb = np.zeros((64, 1, 4)), np.zeros((64, 1))
    with pytest.raises(ValueError):
        oracle.sweep_cubes(prm, glat, glon, obs, xb, axb, 2, 2)
