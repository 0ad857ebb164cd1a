"""GPU parity of the transform (letkf_core) through the C ABI.

Bar: FP64 plain form within 1e-10 relative of the oracle (the reference CPU
path restated) on identical inputs; the golden fixtures (scipy eigh) too.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_core.npz")
TOL = 1e-10


def _rel(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


def test_golden_through_letkf_core_(cuda, dl):
    g = np.load(GOLD)
    for i in range(int(g["ncases"][0])):
        pre = f"c{i}_"
        print(f"golden case {i}", flush=True)
        trans, wbar = dl.letkf_core(g[pre + "hx"], g[pre + "rinv"], g[pre + "dep"],
                                    g[pre + "rloc"], float(g[pre + "rho"][0]))
        assert _rel(trans, g[pre + "trans"]) < TOL, i
        assert _rel(wbar, g[pre + "wbar"]) < TOL, i


def _batch(rng, k, counts, max_obs):
    b = len(counts)
    hx = np.zeros((b, max_obs, k))
    ri = np.zeros((b, max_obs), np.float32)
    dep = np.zeros((b, max_obs))
    rl = np.zeros((b, max_obs))
    for j, p in enumerate(counts):
        h = rng.normal(size=(p, k)) * rng.uniform(0.1, 3.0)
        hx[j, :p] = h - h.mean(axis=1, keepdims=True)
        ri[j, :p] = (1.0 / rng.uniform(0.3, 3.0, p) ** 2).astype(np.float32)
        dep[j, :p] = rng.normal(size=p)
        rl[j, :p] = rng.uniform(0.0, 1.0, p)
    return hx, ri, dep, rl


@pytest.mark.parametrize("k", [2, 3, 16, 20, 32, 48, 64, 65, 96, 128])
def test_batch_matches_oracle(cuda, dl, oracle, k):
    rng = np.random.default_rng(k)
    counts = [0, 1, 2, 5, k - 1, k, k + 3, 300, 1000]
    max_obs = max(counts)
    hx, ri, dep, rl = _batch(rng, k, counts, max_obs)
    rho = 1.1
    print(f"batch k={k}", flush=True)
    trans, wbar = dl.letkf_core_batch(np.array(counts), hx, ri, dep, rl, rho)
    for j, p in enumerate(counts):
        t_o, w_o, s = oracle.letkf_core(hx[j, :p], ri[j, :p], dep[j, :p], rl[j, :p], rho)
        assert _rel(trans[j], t_o) < TOL, (k, p)
        assert _rel(wbar[j], w_o) < TOL, (k, p)


def test_device_batch_entry_and_info(cuda, dl, oracle):
    torch = cuda
    rng = np.random.default_rng(99)
    k, counts = 40, [0, 3, 64, 500]
    hx, ri, dep, rl = _batch(rng, k, counts, 500)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    trans = torch.zeros(len(counts), k, k, dtype=torch.float64, device="cuda")
    wbar = torch.zeros(len(counts), k, dtype=torch.float64, device="cuda")
    info = torch.zeros(len(counts), dtype=torch.int32, device="cuda")
    dl.core_batch_dev(dev(np.array(counts, np.int32)), dev(hx), dev(ri), dev(dep), dev(rl),
                      trans, wbar, 1.05, info=info)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    assert np.all(info > 0), info
    for j, p in enumerate(counts):
        t_o, w_o, _ = oracle.letkf_core(hx[j, :p], ri[j, :p], dep[j, :p], rl[j, :p], 1.05)
        assert _rel(trans[j].cpu().numpy().T, t_o) < TOL
        assert _rel(wbar[j].cpu().numpy(), w_o) < TOL


def test_ill_conditioned_obs_errors(cuda, dl, oracle):
    """R spanning 0.1-10x and a few near-duplicate obs (badly conditioned C)."""
    rng = np.random.default_rng(5)
    k, p = 64, 2000
    h = rng.normal(size=(p, k))
    h[1::2] = h[0::2] + 1e-9 * rng.normal(size=(p // 2, k))
    h -= h.mean(axis=1, keepdims=True)
    ri = (1.0 / (np.exp(rng.uniform(np.log(0.1), np.log(10.0), p)))).astype(np.float32)
    dep = rng.normal(size=p)
    rl = rng.uniform(0, 1, p)
    trans, wbar = dl.letkf_core(h, ri, dep, rl, 1.0)
    t_o, w_o, _ = oracle.letkf_core(h, ri, dep, rl, 1.0)
    assert _rel(trans, t_o) < TOL and _rel(wbar, w_o) < TOL


@pytest.mark.parametrize("k", [8, 20, 64, 128])
def test_r4_drop_ins(cuda, dl, oracle, k):
    """letkf_core_r4_ / letkf_core_batch_r4_ (yb_type = real(4)): the FP64
    analysis of the REAL(4) values rounded once (within one REAL(4) rounding
    of the oracle's FP64 result), and within 1e-4 of the reference's own
    single-precision arithmetic (the oracle's float32 restatement)."""
    rng = np.random.default_rng(500 + k)
    counts = [0, 1, 7, k + 5, 400]
    max_obs = max(counts)
    hx, ri, dep, rl = _batch(rng, k, counts, max_obs)
    hx4, dep4, rl4 = hx.astype(np.float32), dep.astype(np.float32), rl.astype(np.float32)
    rho = 1.1
    print(f"r4 drop-ins k={k}", flush=True)
    tb, wb = dl.letkf_core_batch_r4(np.array(counts), hx4, ri, dep4, rl4, rho)
    for j, p in enumerate(counts):
        t1, w1 = dl.letkf_core_r4(hx4[j, :p], ri[j, :p], dep4[j, :p], rl4[j, :p], rho)
        assert t1.dtype == np.float32 and np.array_equal(t1, tb[j]) and np.array_equal(w1, wb[j])
        t8, w8, _ = oracle.letkf_core(hx4[j, :p].astype(np.float64), ri[j, :p],
                                      dep4[j, :p].astype(np.float64),
                                      rl4[j, :p].astype(np.float64), rho)
        # one REAL(4) rounding of a value within the FP64 bar (1e-10 of the largest)
        ulp = 2.0 ** -24
        assert np.all(np.abs(t1 - t8) <= ulp * np.abs(t8) + 1e-10 * np.abs(t8).max()), (k, p)
        assert np.all(np.abs(w1 - w8) <= ulp * np.abs(w8) + 1e-10 * np.abs(w8).max() + 1e-300), \
            (k, p)
        t4, w4 = oracle.letkf_core_r4(hx4[j, :p], ri[j, :p], dep4[j, :p], rl4[j, :p], rho)
        assert np.abs(t1 - t4).max() <= 1e-4 * np.abs(t4).max(), (k, p)
        if p:
            assert np.abs(w1 - w4).max() <= 1e-4 * np.abs(w4).max(), (k, p)
