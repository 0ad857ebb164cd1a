"""Pins of the oracle's coupler layout conversions (data_trans.F90:2-57): the
Fortran loops equal a permutation of array axes -- xb_atm(ens, ny, nx, n)
into the (ens_max, n_max, nx, ny) state block at the given 1-based starts --
with every other state entry untouched, and the two directions invert each
other."""
import numpy as np
import pytest


@pytest.mark.parametrize("ens,n,nx,ny,ens_max,n_max,es,ns", [
    (3, 5, 4, 6, 8, 12, 4, 6), (1, 1, 3, 2, 1, 1, 1, 1), (5, 2, 2, 3, 5, 7, 1, 6)])
def test_struct_is_an_axis_permutation(oracle, ens, n, nx, ny, ens_max, n_max, es, ns):
    rng = np.random.default_rng(ens * 100 + n)
    atm = rng.normal(size=(n, nx, ny, ens))              # Fortran xb(ens, ny, nx, n)
    state = rng.normal(size=(ny, nx, n_max, ens_max))    # Fortran xb2(ens_max, n_max, nx, ny)
    before = state.copy()
    oracle.struct_atm2endnn(es, ns, atm, state)
    expect = before.copy()
    expect[:, :, ns - 1:ns - 1 + n, es - 1:es - 1 + ens] = np.transpose(atm, (2, 1, 0, 3))
    assert np.array_equal(state, expect)
    back = oracle.struct_endnn2atm(ens, n, es, ns, state)
    assert np.array_equal(back, atm)


@pytest.mark.parametrize("n,nx,ny,n_max,ns", [(4, 3, 5, 9, 3), (1, 2, 7, 1, 1)])
def test_struct_means(oracle, n, nx, ny, n_max, ns):
    rng = np.random.default_rng(n * 7 + ny)
    atm = rng.normal(size=(n, nx, ny))                   # Fortran axb(ny, nx, n)
    state = rng.normal(size=(ny, nx, n_max))             # Fortran axb2(n_max, nx, ny)
    before = state.copy()
    oracle.struct_atm2endnn_a(ns, atm, state)
    expect = before.copy()
    expect[:, :, ns - 1:ns - 1 + n] = np.transpose(atm, (2, 1, 0))
    assert np.array_equal(state, expect)
    assert np.array_equal(oracle.struct_endnn2atm_a(n, ns, state), atm)
