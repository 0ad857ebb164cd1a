"""The reference's own CPU-sized workload at its full size: 2.5 degrees
(72 x 144 columns), 8 levels (n = 25 rows: ps, u, v, T), k = 20, 10 000
point obs of T/u/v, Gaspari-Cohn L = 1000 km, multiplicative rho = 1.1.

Bars: the plain FP64 form against the oracle, selection lists identical and
per-variable increment and Xa' relative error <= 1e-10; the batched BF16x3
form <= 1e-4."""
import numpy as np
import pytest

from test_gpu_sweep import _oracle_sweep, _rel_err, _setup, _var_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["plain_fp64", "batched_bf16x3"])
def test_reference_workload_full_size(cuda, dl, oracle, form):
    torch = cuda
    print(f"reference workload {form}", flush=True)
    cfg, prm = dl.config_synth(0)
    assert (cfg.nlat, cfg.nlon, cfg.nz, cfg.k, cfg.nobs) == (72, 144, 8, 20, 10000)
    over = dict(prm, bin_deg=3.0)
    tol = 1e-10
    if form == "batched_bf16x3":
        over.update(form=dl.FORM_BATCHED, gram=dl.GRAM_BF16X3)
        tol = 1e-4
    data, params, eng, obs, xa, axa = _setup(dl, torch, cfg, **over)
    eng.sweep(obs, data["xb"], data["axb"], xa, axa)
    st = eng.stats()
    dbg = eng.debug(cfg.nobs)
    eng.close()
    assert st["status"] == 0 and st["nonconverged"] == 0, st
    h = {key: v.cpu().numpy() for key, v in data.items()}
    out = _oracle_sweep(oracle, h, params, 3.0)[0]
    assert np.array_equal(dbg["nobsl"], out["nobsl"])
    xa_h, axa_h = xa.cpu().numpy(), axa.cpu().numpy()
    for var, rows in _var_rows(cfg.nz).items():
        e_inc = _rel_err(axa_h[:, rows] - h["axb"][:, rows],
                         out["axa"][:, rows] - h["axb"][:, rows])
        e_xa = _rel_err(xa_h[:, rows], out["xa"][:, rows])
        print(f"  {var}: increment {e_inc:.2e} Xa' {e_xa:.2e}", flush=True)
        assert e_inc <= tol, var
        assert e_xa <= tol, var
