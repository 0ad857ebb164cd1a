"""GPU coupler layout conversions (dida_struct_*) against the oracle's
restatement of data_trans.F90: bit-identical copies, state entries outside
the block untouched, both directions, ragged tile edges."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ens,n,nx,ny,ens_max,n_max,es,ns", [
    (3, 5, 7, 11, 8, 12, 4, 6), (1, 1, 9, 40, 1, 1, 1, 1), (16, 33, 45, 90, 64, 97, 17, 2),
    (64, 1, 20, 37, 64, 97, 1, 1)])
def test_struct_ens_matches_oracle(cuda, dl, oracle, ens, n, nx, ny, ens_max, n_max, es, ns):
    torch = cuda
    print(f"struct ens={ens} n={n} nx={nx} ny={ny}", flush=True)
    rng = np.random.default_rng(ens + 10 * n + nx)
    atm = rng.normal(size=(n, nx, ny, ens))
    state = rng.normal(size=(ny, nx, n_max, ens_max))
    exp = state.copy()
    oracle.struct_atm2endnn(es, ns, atm, exp)
    d_state = torch.tensor(state, device="cuda")
    dl.struct_atm2endnn(es, ns, torch.tensor(atm, device="cuda"), d_state)
    assert np.array_equal(d_state.cpu().numpy(), exp)
    d_atm = torch.zeros(n, nx, ny, ens, dtype=torch.float64, device="cuda")
    dl.struct_endnn2atm(es, ns, d_state, d_atm)
    assert np.array_equal(d_atm.cpu().numpy(), oracle.struct_endnn2atm(ens, n, es, ns, exp))


@pytest.mark.parametrize("n,nx,ny,n_max,ns", [(1, 9, 40, 97, 1), (32, 45, 90, 97, 2),
                                              (33, 7, 65, 40, 8)])
def test_struct_mean_matches_oracle(cuda, dl, oracle, n, nx, ny, n_max, ns):
    torch = cuda
    print(f"struct mean n={n} nx={nx} ny={ny}", flush=True)
    rng = np.random.default_rng(n + ny)
    atm = rng.normal(size=(n, nx, ny))
    state = rng.normal(size=(ny, nx, n_max))
    exp = state.copy()
    oracle.struct_atm2endnn_a(ns, atm, exp)
    d_state = torch.tensor(state, device="cuda")
    dl.struct_atm2endnn_a(ns, torch.tensor(atm, device="cuda"), d_state)
    assert np.array_equal(d_state.cpu().numpy(), exp)
    d_atm = torch.zeros(n, nx, ny, dtype=torch.float64, device="cuda")
    dl.struct_endnn2atm_a(ns, d_state, d_atm)
    assert np.array_equal(d_atm.cpu().numpy(), oracle.struct_endnn2atm_a(n, ns, exp))


def test_struct_bounds_are_checked(cuda, dl):
    torch = cuda
    atm = torch.zeros(2, 3, 4, 5, dtype=torch.float64, device="cuda")
    st = torch.zeros(4, 3, 6, 5, dtype=torch.float64, device="cuda")
    with pytest.raises(dl.DidaError, match="outside"):
        dl.struct_atm2endnn(2, 1, atm, st)          # members 2..6 of 5
    with pytest.raises(dl.DidaError, match="outside"):
        dl.struct_atm2endnn(1, 6, atm, st)          # rows 6..7 of 6
