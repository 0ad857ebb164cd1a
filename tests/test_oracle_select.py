"""Pins of the oracle's distance, Gaspari-Cohn and selection (gen_yb_dis).

Distance: against the haversine formula (a different formula) with the
reference's REAL(4) constants, and shown to differ from the exact-pi value, so
the float-rounded pi is pinned.  Gaspari-Cohn: against the textbook closed
form with rational constants (golden fixture) and the values the reference's
single-precision literals fix.  Selection: against a haversine brute force,
plus the in-order MAX_OBS truncation and the bin pruning.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_core.npz")
A_F32 = float(np.float32(6371.393))
PI_F32 = float(np.float32(np.pi))


def haversine(lat1, lon1, lat2, lon2, pi=PI_F32, a=A_F32):
    f = pi / 180.0
    p1, p2 = lat1 * f, lat2 * f
    dp, dl = p2 - p1, (lon2 - lon1) * f
    h = np.sin(dp / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(dl / 2) ** 2
    return 2 * a * np.arcsin(np.sqrt(np.minimum(h, 1.0)))


def test_float_constants():
    assert PI_F32 == 3.1415927410125732
    assert A_F32 == 6371.39306640625


def test_distance_matches_haversine(oracle):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        la1, la2 = rng.uniform(-90, 90, 2)
        lo1, lo2 = rng.uniform(0, 360, 2)
        d = oracle.latlon_dist(la1, lo1, la2, lo2)
        h = haversine(la1, lo1, la2, lo2)
        # acos loses ~sqrt(eps) at tiny angles; the sample avoids them mostly
        assert abs(d - h) < 2e-6 * max(h, 1.0) + 2e-4, (la1, lo1, la2, lo2)


def test_distance_uses_single_precision_pi(oracle):
    d = oracle.latlon_dist(0.0, 0.0, 0.0, 90.0)
    assert abs(d - haversine(0.0, 0.0, 0.0, 90.0)) < 1e-7
    assert abs(d - haversine(0.0, 0.0, 0.0, 90.0, pi=np.pi)) > 1e-4


def test_gc_closed_form(oracle):
    g = np.load(GOLD)
    for r, e in zip(g["gc_r"], g["gc_exact"]):
        v = oracle.schr2(r)
        if r > 2.0:
            assert v == 0.0
        else:
            # the REAL(4) literals 1.666667 and 0.6666667 differ from 5/3 and
            # 2/3 by 3.18e-7 and 1.99e-8; their terms carry r^2 and 1/r
            tol = 3.2e-7 * r * r + 2.0e-8 / max(r, 1.0) + 1e-14
            assert abs(v - e) <= tol, r


def test_gc_single_precision_literals(oracle):
    # values fixed by 1.666667f and 0.6666667f (letkf_yb.F90:256-261)
    assert oracle.schr2(0.0) == 1.0
    assert oracle.schr2(0.5) == 0.6848957538604736
    assert oracle.schr2(1.0) == 0.20833301544189453
    assert oracle.schr2(1.5) == 0.016493757565816225
    assert 1.26e-6 < oracle.schr2(2.0) < 1.27e-6     # exact GC gives 0 at r = 2
    assert oracle.schr2(np.nextafter(2.0, 3.0)) == 0.0
    assert oracle.schr(2.0) == 1.0 and oracle.schr(2.0000001) == 0.0


def _obs(rng, n, k, clustered=False):
    if clustered:
        lat = np.clip(rng.normal(30, 5, n), -89, 89)
        lon = np.mod(rng.normal(100, 8, n), 360)
    else:
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
        lon = rng.uniform(0, 360, n)
    yb = rng.normal(size=(n, k))
    return lat, lon, yb, rng.normal(size=n), rng.uniform(0.5, 2, n).astype(np.float32)


@pytest.mark.parametrize("polar", [90.0, 60.0])
def test_selection_bruteforce(oracle, polar):
    rng = np.random.default_rng(11)
    lat, lon, yb, dep, ri = _obs(rng, 3000, 4)
    obs = oracle.Obs(lat, lon, yb, dep, ri)
    L = 800.0
    prm = oracle.params(4, 1, 1, 1, L, 1.0, max_obs=100000, obs_polar=polar)
    for clat, clon in [(0.0, 0.0), (45.0, 359.0), (-89.0, 10.0), (88.75, 180.0)]:
        idx, w = oracle.select_column(prm, obs, np.float32(clat), np.float32(clon))
        h = haversine(lat, lon, float(np.float32(clat)), float(np.float32(clon)))
        inside = (h < 2 * L * (1 - 1e-7)) & (np.abs(lat) <= polar)
        border = np.abs(h - 2 * L) <= 2 * L * 1e-7
        got = np.zeros(len(lat), bool)
        got[idx] = True
        assert np.all(got[inside]) and not np.any(got[~inside & ~border])
        assert np.all(np.diff(idx) > 0)                       # scan order
        g = np.load(GOLD)
        assert np.allclose(w, np.interp(h[idx] / L, g["gc_r"], g["gc_exact"]), atol=2e-4)


def test_selection_max_obs_truncation(oracle):
    rng = np.random.default_rng(12)
    lat, lon, yb, dep, ri = _obs(rng, 4000, 4, clustered=True)
    obs = oracle.Obs(lat, lon, yb, dep, ri)
    full = oracle.params(4, 1, 1, 1, 1500.0, 1.0, max_obs=100000)
    capped = oracle.params(4, 1, 1, 1, 1500.0, 1.0, max_obs=500)
    i_full, w_full = oracle.select_column(full, obs, 30.0, 100.0)
    i_cap, w_cap = oracle.select_column(capped, obs, 30.0, 100.0)
    assert len(i_full) > 500 and len(i_cap) == 500
    assert np.array_equal(i_cap, i_full[:500]) and np.array_equal(w_cap, w_full[:500])


@pytest.mark.parametrize("L,bin_deg", [(600.0, 2.7), (1000.0, 4.5), (200.0, 0.9), (3000.0, 13.0)])
def test_binned_selection_equals_full_scan(oracle, L, bin_deg):
    rng = np.random.default_rng(int(L))
    lat, lon, yb, dep, ri = _obs(rng, 5000, 3)
    lat[:50] = rng.uniform(85, 90, 50)              # near the pole
    lon[50:100] = rng.uniform(359.5, 360.0, 50)     # near the seam
    obs = oracle.Obs(lat, lon, yb, dep, ri)
    perm, bin_start, nband, nlonbin = oracle.bin_obs(obs, bin_deg, 89.0)
    ob = obs.permuted(perm)
    prm = oracle.params(3, 1, 1, 1, L, 1.0, max_obs=2000, obs_polar=89.0)
    for clat in (-88.75, -40.0, 0.0, 33.3, 89.5):
        for clon in (0.0, 0.25, 179.0, 359.75):
            a = oracle.select_column(prm, ob, clat, clon)
            b = oracle.select_column_binned(prm, ob, bin_start, nband, nlonbin, bin_deg, clat, clon)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (clat, clon)


@pytest.mark.parametrize("bin_deg", [1.1, 2.7, 6.3])
def test_binned_truncation_follows_caller_order(oracle, bin_deg):
    """With MAX_OBS binding, the binned selection keeps the obs a full scan of
    the caller's list keeps (gen_yb_dis stops at MAX_OBS in its scan order,
    letkf_yb.F90:98-115), whatever the bin size; the bin order alone would
    keep a different set."""
    rng = np.random.default_rng(77)
    lat, lon, yb, dep, ri = _obs(rng, 6000, 3, clustered=True)
    obs = oracle.Obs(lat, lon, yb, dep, ri)
    perm, bin_start, nband, nlonbin = oracle.bin_obs(obs, bin_deg, 90.0)
    ob = obs.permuted(perm)
    prm = oracle.params(3, 1, 1, 1, 1200.0, 1.0, max_obs=300)
    ncapped = 0
    for clat in (-60.0, -10.0, 0.0, 25.0, 45.0, 70.0):
        for clon in (0.0, 90.0, 100.0, 200.0, 300.0):
            a_idx, a_w = oracle.select_column(prm, obs, clat, clon)
            b_idx, b_w = oracle.select_column_binned(prm, ob, bin_start, nband, nlonbin,
                                                     bin_deg, clat, clon, perm=perm)
            got = perm[b_idx]
            order = np.argsort(got)
            assert np.array_equal(got[order], a_idx), (clat, clon)
            assert np.array_equal(b_w[order], a_w), (clat, clon)
            if len(a_idx) == prm.max_obs:
                ncapped += 1
                c_idx, _ = oracle.select_column_binned(prm, ob, bin_start, nband, nlonbin,
                                                       bin_deg, clat, clon)
                assert not np.array_equal(np.sort(perm[c_idx]), a_idx), (clat, clon)
    assert ncapped >= 3


def test_bin_sort_is_stable(oracle):
    rng = np.random.default_rng(5)
    lat, lon, yb, dep, ri = _obs(rng, 2000, 2)
    obs = oracle.Obs(lat, lon, yb, dep, ri)
    perm, bin_start, nband, nlonbin = oracle.bin_obs(obs, 5.0, 90.0)
    band = np.clip(np.floor((lat + 90) / 5.0), 0, nband - 1)
    lb = np.clip(np.floor(np.mod(lon, 360) / 5.0), 0, nlonbin - 1)
    key = (band * nlonbin + lb).astype(np.int64)
    assert np.array_equal(perm, np.argsort(key, kind="stable"))
