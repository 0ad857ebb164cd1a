"""Write the golden fixtures of the LETKF transform (letkf_core) and the
selection weights.

The expected values come from an implementation independent of the C
oracle: numpy/scipy, with scipy.linalg.eigh as the symmetric eigensolver (the
reference calls MKL dsyevd at letkf_core.F90:98) and the reference's formulas
(letkf_core.F90:58-225) written with dense numpy algebra.  The Gaspari-Cohn
fixtures use the textbook rational constants 5/3 and 2/3 (Gaspari & Cohn
1999, eq. 4.10); the reference's single-precision literals move them by at
most ~1e-7.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def core_expected(hx, rinv, dep, rloc, rho):
    """hx: (p, k).  Returns trans[i, j] = W_ij + wbar_i and wbar."""
    p, k = hx.shape
    r = rinv.astype(np.float64) * rloc
    yh = hx * np.sqrt(r)[:, None]
    c = yh.T @ yh
    lam, v = sl.eigh(c)
    lt = lam + (k - 1) / rho
    pa = (v / lt) @ v.T
    w = (v * np.sqrt((k - 1) / lt)) @ v.T
    wbar = pa @ (hx.T @ (r * dep)) if p else np.zeros(k)
    return w + wbar[:, None], wbar


def gc_exact(r):
    r = np.asarray(r, dtype=np.float64)
    out = np.zeros_like(r)
    a = (r <= 1.0)
    b = (r > 1.0) & (r <= 2.0)
    ra = r[a]
    out[a] = -0.25 * ra**5 + 0.5 * ra**4 + 0.625 * ra**3 - 5.0 / 3.0 * ra**2 + 1.0
    rb = r[b]
    out[b] = (rb**5 / 12.0 - 0.5 * rb**4 + 0.625 * rb**3 + 5.0 / 3.0 * rb**2 - 5.0 * rb + 4.0
              - 2.0 / (3.0 * rb))
    return out


def main():
    rng = np.random.default_rng(20230917)
    cases = {}
    idx = 0
    for k in (4, 20, 32, 64):
        for p in (0, 1, 7, 300):
            hx = rng.normal(size=(p, k))
            hx -= hx.mean(axis=1, keepdims=True)          # centred like letkf_obs.F90:167
            rinv = (1.0 / rng.uniform(0.5, 2.0, p) ** 2).astype(np.float32)
            dep = rng.normal(size=p)
            rloc = gc_exact(rng.uniform(0.0, 2.0, p))
            rho = float(rng.choice([1.0, 1.1, 1.3]))
            trans, wbar = core_expected(hx, rinv, dep, rloc, rho)
            pre = f"c{idx}_"
            cases[pre + "hx"] = hx
            cases[pre + "rinv"] = rinv
            cases[pre + "dep"] = dep
            cases[pre + "rloc"] = rloc
            cases[pre + "rho"] = np.array([rho])
            cases[pre + "trans"] = trans
            cases[pre + "wbar"] = wbar
            idx += 1
    cases["ncases"] = np.array([idx])
    r = np.linspace(0.0, 2.2, 221)
    cases["gc_r"] = r
    cases["gc_exact"] = gc_exact(r)
    np.savez_compressed(os.path.join(HERE, "golden_core.npz"), **cases)
    print(f"wrote {idx} core cases and {len(r)} Gaspari-Cohn points")


if __name__ == "__main__":
    main()
