"""The allocation-free batched core (dida_letkf_core_batch_dev_noalloc)
through the C ABI, against dida_letkf_core_batch_dev on the same inputs: two
calls in a row, the second smaller, each bit-equal to the allocating entry
point (the same transform kernel reads the same offsets); batches above the
fixed offset array are refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(torch, rng, batch, k, max_obs):
    counts = rng.integers(0, max_obs + 1, batch).astype(np.int32)
    counts[:3] = [0, 1, max_obs]
    hx = np.zeros((batch, max_obs, k))
    for j, p in enumerate(counts):
        h = rng.normal(size=(p, k)) * rng.uniform(0.2, 3.0)
        hx[j, :p] = h - h.mean(axis=1, keepdims=True)
    ri = (1.0 / rng.uniform(0.3, 3.0, (batch, max_obs)) ** 2).astype(np.float32)
    dep = rng.normal(size=(batch, max_obs))
    rl = rng.uniform(0.0, 1.0, (batch, max_obs))
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    return [dev(counts), dev(hx), dev(ri), dev(dep), dev(rl)]


def _call(dl, torch, fn, args, batch, k, max_obs, rho):
    trans = torch.full((batch, k, k), -1.0, dtype=torch.float64, device="cuda")
    wbar = torch.full((batch, k), -1.0, dtype=torch.float64, device="cuda")
    info = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(batch, max_obs, k, *[ptr(t) for t in args], ptr(trans), ptr(wbar), float(rho),
            dl.GRAM_FP64, ptr(info), stream)
    assert rc == 0, dl.lib().dida_letkf_last_error()
    torch.cuda.synchronize()
    return trans, wbar, info


def test_noalloc_core_equals_the_allocating_one(cuda, dl):
    torch = cuda
    lib = dl.lib()
    rng = np.random.default_rng(31)
    for batch, k, max_obs, rho in ((300, 48, 400, 1.1), (37, 20, 120, 1.0)):
        print(f"noalloc core: batch {batch}, k {k}, max_obs {max_obs}", flush=True)
        args = _inputs(torch, rng, batch, k, max_obs)
        t1, w1, i1 = _call(dl, torch, lib.dida_letkf_core_batch_dev_noalloc, args, batch, k,
                           max_obs, rho)
        t0, w0, i0 = _call(dl, torch, lib.dida_letkf_core_batch_dev, args, batch, k, max_obs,
                           rho)
        assert torch.equal(i1, i0) and bool((i0 >= 1).all())
        assert torch.equal(t1, t0) and torch.equal(w1, w0)


def test_noalloc_core_refuses_batches_above_its_array(dl):
    lib = dl.lib()
    rc = lib.dida_letkf_core_batch_dev_noalloc(1048577, 1, 8, None, None, None, None, None,
                                               None, None, 1.0, 0, None, None)
    assert rc == -1 and b"1048576" in lib.dida_letkf_last_error()
