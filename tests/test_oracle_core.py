"""Pins of the C oracle's letkf_core (letkf_core.F90:14-233).

Every check is independent of the oracle's own formula: closed forms (p = 0,
p = 1 Sherman-Morrison), invariants that fix W and wbar uniquely (W symmetric
positive definite with W (C + (k-1)/rho I) W = (k-1) I, and
(C + (k-1)/rho I) wbar = b), and the scipy-eigh golden fixtures.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_core.npz")


def _random_case(rng, k, p, centred=True):
    hx = rng.normal(size=(p, k))
    if centred:
        hx -= hx.mean(axis=1, keepdims=True)
    rinv = (1.0 / rng.uniform(0.5, 2.0, p) ** 2).astype(np.float32)
    dep = rng.normal(size=p)
    rloc = rng.uniform(0.0, 1.0, p)
    return hx, rinv, dep, rloc


def _c_and_b(hx, rinv, rloc, dep):
    r = rinv.astype(np.float64) * rloc
    yh = hx * np.sqrt(r)[:, None]
    return yh.T @ yh, hx.T @ (r * dep)


@pytest.mark.parametrize("k", [2, 5, 20, 64])
@pytest.mark.parametrize("rho", [1.0, 1.1])
def test_no_obs_gives_inflated_identity(oracle, k, rho):
    trans, wbar, _ = oracle.letkf_core(np.zeros((0, k)), np.zeros(0, np.float32), np.zeros(0),
                                       np.zeros(0), rho)
    assert np.allclose(trans, np.sqrt(rho) * np.eye(k), atol=1e-15, rtol=0)
    assert np.all(wbar == 0.0)


@pytest.mark.parametrize("k", [3, 8, 20, 40])
def test_single_obs_sherman_morrison(oracle, k):
    rng = np.random.default_rng(k)
    rho = 1.2
    y = rng.normal(size=k)
    y -= y.mean()
    ri = np.float32(1.7)
    rl, d = 0.6, 0.9
    trans, wbar, _ = oracle.letkf_core(y[None, :], np.array([ri]), np.array([d]),
                                       np.array([rl]), rho)
    r = float(ri) * rl
    a0 = (k - 1) / rho
    lam = r * (y @ y)
    yy = np.outer(y, y) / (y @ y)
    w_expect = np.sqrt(rho) * (np.eye(k) - yy) + np.sqrt((k - 1) / (lam + a0)) * yy
    wbar_expect = y * r * d / (lam + a0)
    assert np.allclose(wbar, wbar_expect, atol=1e-14, rtol=1e-12)
    assert np.allclose(trans, w_expect + wbar_expect[:, None], atol=1e-13, rtol=1e-12)


@pytest.mark.parametrize("k,p", [(4, 7), (20, 300), (32, 40), (64, 2000), (64, 10)])
def test_transform_invariants(oracle, k, p):
    rng = np.random.default_rng(1000 * k + p)
    rho = 1.1
    hx, rinv, dep, rloc = _random_case(rng, k, p)
    trans, wbar, sweeps = oracle.letkf_core(hx, rinv, dep, rloc, rho)
    assert sweeps > 0
    w = trans - wbar[:, None]
    c, b = _c_and_b(hx, rinv, rloc, dep)
    a = c + (k - 1) / rho * np.eye(k)
    scale = np.abs(a).max()
    assert np.allclose(w, w.T, atol=1e-13 * np.sqrt(rho), rtol=0)
    assert np.linalg.eigvalsh(0.5 * (w + w.T)).min() > 0
    assert np.abs(w @ a @ w - (k - 1) * np.eye(k)).max() < 1e-11 * scale
    assert np.abs(a @ wbar - b).max() < 1e-11 * max(1.0, np.abs(b).max())
    # zero-mean Yb: C 1 = 0, so W 1 = sqrt(rho) 1 and 1^T wbar = 0
    assert np.abs(w.sum(axis=1) - np.sqrt(rho)).max() < 1e-12
    assert abs(wbar.sum()) < 1e-12 * max(1.0, np.abs(wbar).max()) * k


def test_golden_fixtures(oracle):
    g = np.load(GOLD)
    for i in range(int(g["ncases"][0])):
        pre = f"c{i}_"
        trans, wbar, _ = oracle.letkf_core(g[pre + "hx"], g[pre + "rinv"], g[pre + "dep"],
                                           g[pre + "rloc"], float(g[pre + "rho"][0]))
        t_ref, w_ref = g[pre + "trans"], g[pre + "wbar"]
        assert np.abs(trans - t_ref).max() <= 1e-12 * max(1.0, np.abs(t_ref).max()), i
        assert np.abs(wbar - w_ref).max() <= 1e-12 * max(1.0, np.abs(w_ref).max()), i


def test_eigensolver_against_scipy(oracle):
    import scipy.linalg as sl
    rng = np.random.default_rng(7)
    for n in (2, 5, 17, 64):
        m = rng.normal(size=(n, n))
        a = m @ m.T
        a[:, 0] = a[0, :] = 0.0        # a null direction, as C 1 = 0
        w, v, s = oracle.syev(a)
        assert s >= 0
        w_ref = sl.eigvalsh(a)
        assert np.allclose(w, w_ref, atol=1e-13 * np.abs(w_ref).max(), rtol=0)
        assert np.allclose(v.T @ v, np.eye(n), atol=1e-13)
        assert np.abs(a @ v - v * w).max() < 1e-12 * np.abs(w_ref).max()


def test_planted_mistake_is_caught(oracle):
    """The invariant check must reject a transform with a dropped term."""
    rng = np.random.default_rng(3)
    k, p, rho = 10, 30, 1.1
    hx, rinv, dep, rloc = _random_case(rng, k, p)
    trans, wbar, _ = oracle.letkf_core(hx, rinv, dep, rloc, rho)
    c, _ = _c_and_b(hx, rinv, rloc, dep)
    a = c + (k - 1) / rho * np.eye(k)
    w = trans - wbar[:, None]
    wrong = w * 1.001  # e.g. a sqrt((k-1)) vs sqrt(k) slip
    assert np.abs(wrong @ a @ wrong - (k - 1) * np.eye(k)).max() > 1e-6 * np.abs(a).max()


# ---------------------------------------------------------------------------
# The REAL(4) build of letkf_core (yb_type = real(4), da_config.inc:7-10),
# restated in float32 numpy: pinned by the same closed forms and invariants at
# single-precision tolerance and against the FP64 analysis of the same values.

def _r4(x):
    return np.asarray(x, np.float32)


@pytest.mark.parametrize("k", [4, 20, 64])
def test_r4_no_obs_gives_inflated_identity(oracle, k):
    trans, wbar = oracle.letkf_core_r4(np.zeros((0, k)), np.zeros(0), np.zeros(0), np.zeros(0),
                                       1.1)
    assert trans.dtype == np.float32 and wbar.dtype == np.float32
    assert np.allclose(trans, np.sqrt(1.1) * np.eye(k), atol=2e-6, rtol=0)
    assert np.all(wbar == 0.0)


@pytest.mark.parametrize("k", [3, 20, 40])
def test_r4_single_obs_sherman_morrison(oracle, k):
    rng = np.random.default_rng(100 + k)
    rho = 1.2
    y = _r4(rng.normal(size=k))
    y -= y.mean()
    ri, rl, d = np.float32(1.7), np.float32(0.6), np.float32(0.9)
    trans, wbar = oracle.letkf_core_r4(y[None, :], [ri], [d], [rl], rho)
    yd = y.astype(np.float64)
    r = float(ri) * float(rl)
    a0 = (k - 1) / rho
    lam = r * (yd @ yd)
    yy = np.outer(yd, yd) / (yd @ yd)
    w_expect = np.sqrt(rho) * (np.eye(k) - yy) + np.sqrt((k - 1) / (lam + a0)) * yy
    wbar_expect = yd * r * float(d) / (lam + a0)
    assert np.abs(wbar - wbar_expect).max() <= 1e-5 * np.abs(wbar_expect).max()
    assert np.abs(trans - (w_expect + wbar_expect[:, None])).max() <= 2e-5


@pytest.mark.parametrize("k,p", [(8, 30), (20, 300), (64, 1000)])
def test_r4_close_to_the_fp64_analysis_of_the_same_values(oracle, k, p):
    """Single-precision arithmetic differs from the FP64 analysis of the same
    REAL(4) values by its own rounding only; a wrong step (a missing sqrt, a
    transposed product) is far outside this."""
    rng = np.random.default_rng(k * 1000 + p)
    hx, rinv, dep, rloc = _random_case(rng, k, p)
    hx4, dep4, rl4 = _r4(hx), _r4(dep), _r4(rloc)
    t4, w4 = oracle.letkf_core_r4(hx4, rinv, dep4, rl4, 1.05)
    t8, w8, _ = oracle.letkf_core(hx4.astype(np.float64), rinv, dep4.astype(np.float64),
                                  rl4.astype(np.float64), 1.05)
    assert np.abs(t4 - t8).max() <= 1e-4 * np.abs(t8).max()
    assert np.abs(w4 - w8).max() <= 1e-4 * np.abs(w8).max()
    # planted: rloc applied once instead of sqrt(R^-1 rloc) twice
    bad, _ = oracle.letkf_core_r4(hx4 * np.sqrt(rl4)[:, None], rinv, dep4, np.ones_like(rl4),
                                  1.05)
    assert np.abs(bad - t8).max() > 1e-2 * np.abs(t8).max()
