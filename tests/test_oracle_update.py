"""Pins of the oracle's update (xb_2_xa), centring and post-analysis blend."""
import numpy as np


def test_update_dense_product(oracle):
    rng = np.random.default_rng(1)
    k, ncells = 12, 30
    xb = rng.normal(size=(ncells, k))
    xb -= xb.mean(axis=1, keepdims=True)
    axb = rng.normal(size=ncells) * 100
    trans = rng.normal(size=(k, k))
    xa, axa = oracle.xb_2_xa(xb, axb, trans)
    full = xb @ trans + axb[:, None]              # xa(e) = sum_i trans(i, e) xb(i) + axb
    assert np.allclose(axa, full.mean(axis=1), atol=1e-12, rtol=1e-14)
    assert np.allclose(xa, full - full.mean(axis=1, keepdims=True), atol=1e-12, rtol=1e-13)


def test_update_mean_increment_identity(oracle):
    """With W 1 = sqrt(rho) 1 the analysis mean is xbar_b + Xb' wbar."""
    rng = np.random.default_rng(2)
    k, p, rho = 16, 50, 1.1
    hx = rng.normal(size=(p, k))
    hx -= hx.mean(axis=1, keepdims=True)
    trans, wbar, _ = oracle.letkf_core(hx, np.ones(p, np.float32), rng.normal(size=p),
                                       np.ones(p), rho)
    xb = rng.normal(size=(9, k))
    xb -= xb.mean(axis=1, keepdims=True)
    axb = rng.normal(size=9)
    xa, axa = oracle.xb_2_xa(xb, axb, trans)
    assert np.allclose(axa, axb + xb @ wbar, atol=1e-12)


def test_centring(oracle):
    rng = np.random.default_rng(3)
    hxb = rng.normal(size=(40, 7)) + 5
    yo = rng.normal(size=40)
    yb, dep = oracle.centre_obs(hxb, yo)
    assert np.allclose(yb.sum(axis=1), 0.0, atol=1e-12)
    assert np.allclose(dep, yo - hxb.mean(axis=1), atol=1e-14)


def test_post_blend_and_rtps(oracle):
    rng = np.random.default_rng(4)
    k, nrow, nlat, nlon = 8, 3, 4, 5
    glat = np.array([-80, -10, 10, 80], np.float32)
    ncol = nlat * nlon
    xb = rng.normal(size=(ncol, nrow, k))
    xb -= xb.mean(axis=2, keepdims=True)
    xa = 0.5 * rng.normal(size=(ncol, nrow, k))
    xa -= xa.mean(axis=2, keepdims=True)
    axb, axa = rng.normal(size=(ncol, nrow)), rng.normal(size=(ncol, nrow))
    out, sb, sa = oracle.post(0, glat, nlat, nlon, 60.0, 0.6, 0.4, 0.0, xb, axb, xa, axa)
    inside = np.abs(glat[np.arange(ncol) % nlat]) <= 60
    exp = 0.6 * (axb[..., None] + xb) + 0.4 * (axa[..., None] + xa)
    exp[~inside] = (axb[..., None] + xb)[~inside]
    assert np.allclose(out, exp, atol=1e-14)
    assert np.allclose(sb, np.sqrt((xb ** 2).mean(axis=2)), atol=1e-14)
    # RTPS: the analysis spread becomes alpha sb + (1 - alpha) sa
    out, sb, sa = oracle.post(1, glat, nlat, nlon, 90.0, 0, 0, 0.8, xb, axb, xa, axa)
    sa0 = np.sqrt((xa ** 2).mean(axis=2))
    assert np.allclose(sa, 0.8 * sb + 0.2 * sa0, atol=1e-13)
    assert np.allclose(out.mean(axis=2), axa, atol=1e-13)


def test_post_blend_special_cases(oracle):
    """gamma = (0, 1) returns the analysis, (1, 0) the background; outside the
    polar limit the background is kept whatever gamma is."""
    rng = np.random.default_rng(5)
    k, nrow, nlat, nlon = 6, 2, 3, 4
    glat = np.array([-70, 0, 70], np.float32)
    ncol = nlat * nlon
    xb = rng.normal(size=(ncol, nrow, k))
    xa = rng.normal(size=(ncol, nrow, k))
    axb, axa = rng.normal(size=(ncol, nrow)), rng.normal(size=(ncol, nrow))
    full_b, full_a = axb[..., None] + xb, axa[..., None] + xa
    out, _, _ = oracle.post(0, glat, nlat, nlon, 90.0, 0.0, 1.0, 0.0, xb, axb, xa, axa)
    assert np.array_equal(out, full_a)
    out, _, _ = oracle.post(0, glat, nlat, nlon, 90.0, 1.0, 0.0, 0.0, xb, axb, xa, axa)
    assert np.array_equal(out, full_b)
    out, _, _ = oracle.post(0, glat, nlat, nlon, 50.0, 0.0, 1.0, 0.0, xb, axb, xa, axa)
    polar = np.abs(glat[np.arange(ncol) % nlat]) > 50
    assert np.array_equal(out[polar], full_b[polar])
    assert np.array_equal(out[~polar], full_a[~polar])
