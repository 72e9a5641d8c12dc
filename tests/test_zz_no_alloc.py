"""izem_psr_batch_noalloc (the chemistry substep that allocates nothing)
against izem_psr_batch through the C ABI: the same end states, step counts
and work counters, bit for bit, on calls in a row whose sizes shrink, in
index order and with the stiffness order (whose buffers are the static
scratch); sizes above the maximum are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
RU = 8.314510e7
MAX_CELLS = 1 << 20


def _cells(thd, n, seed):
    from phd_amd import ics
    T, p, Y = ics.config3_cells(thd, n, seed=seed)
    Wmix = 1.0 / (Y / (np.array(thd.W) * 1000.0)).sum(axis=1)
    return T, (p * 10.0) * Wmix / (RU * T), Y


def _run(s, fn, T, rho, Y):
    import torch
    n = len(T)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    nst = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    st = np.zeros(4, dtype=np.uint64)
    assert s.lib.izem_chem_stats(s.ctx, st.ctypes.data) == 0
    assert fn(s.ctx, n, C.c_void_p(dT.data_ptr()), C.c_void_p(drho.data_ptr()),
              C.c_void_p(dY.data_ptr()), 1e-6, C.c_void_p(nst.data_ptr())) == 0
    s.sync()
    assert s.status() == 0
    assert s.lib.izem_chem_stats(s.ctx, st.ctypes.data) == 0
    return dT.cpu().numpy(), dY.cpu().numpy(), nst.cpu().numpy(), st


def test_psr_batch_noalloc_matches_psr_batch(mech, thd):
    import torch
    from phd_amd.ics import Box
    from phd_amd.solver import Options, Solver
    s = Solver(Box(8, 8, 1, 3, 2, (1, 1, 1)), Options(), mech=mech)
    for order in (1, 0):
        assert s.lib.izem_chem_order(s.ctx, order) == 0
        # shrinking sizes: above the 65536 resident lanes (the order sorts),
        # then below them
        for n, seed in ((150_001, 21), (70_003, 22), (4_133, 23)):
            T, rho, Y = _cells(thd, n, seed)
            a = _run(s, s.lib.izem_psr_batch, T, rho, Y)
            b = _run(s, s.lib.izem_psr_batch_noalloc, T, rho, Y)
            for x, y in zip(a, b):
                assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
            assert np.all(a[2] > 0)
            print(f"order {order}, {n} cells: identical, {a[2].sum()} steps")
    z = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    p = C.c_void_p(z.data_ptr())
    assert s.lib.izem_psr_batch_noalloc(s.ctx, MAX_CELLS + 1, p, p, p, 1e-6, None) == -1
    assert s.lib.izem_psr_batch_noalloc(s.ctx, 0, p, p, p, 1e-6, None) == 0
