"""Accuracy bars of the chemistry substep referenced to DVODE itself.

The reference's DVODE (oracle/_ref, compiled from its sources) is run twice
per cell: at the tolerances under test, and at rtol 1e-11 / atol 1e-17 as the
exact solution.  Errors are measured in DVODE's own norm: each component z_i
of a cell's state (T, Y_1..Y_K) scaled by rtol |z_i| + atol (DVEWSET, ITOL =
1), and per cell the root mean square over the components (DVNORM).  Bars:

  * per-cell error norm: the largest and the mean over the cells within
    twice DVODE's at the same tolerances;
  * each component (T and every Y_k): the mean scaled error over the cells
    within three times DVODE's (the means of single components over a few
    dozen cells scatter by about 2x between two equally accurate
    integrators; the per-cell norm above is the tight bar).
"""
from __future__ import annotations

import numpy as np

TIGHT = (1e-11, 1e-17)


def compare(zg, zd, zt, rtol, atol):
    """zg, zd, zt: (cells, components) end states of the code under test,
    of DVODE at (rtol, atol) and of DVODE at TIGHT."""
    w = rtol * np.abs(zt) + atol
    sg, sd = np.abs(zg - zt) / w, np.abs(zd - zt) / w
    ng, nd = np.sqrt(np.mean(sg ** 2, axis=1)), np.sqrt(np.mean(sd ** 2, axis=1))
    return dict(norm_max=(ng.max(), nd.max()), norm_mean=(ng.mean(), nd.mean()),
                comp_mean=(sg.mean(axis=0), sd.mean(axis=0)))


def dvode_errors(rc, T, rho, Y, dt, Tg, Yg, rtol=1e-6, atol=1e-12):
    n = len(T)
    zt, zd = np.zeros((n, 1 + Y.shape[1])), np.zeros((n, 1 + Y.shape[1]))
    for c in range(n):
        a = rc.psr(T[c], rho[c], Y[c], dt, *TIGHT)
        b = rc.psr(T[c], rho[c], Y[c], dt, rtol, atol)
        assert a[2] == 2 and b[2] == 2, c
        zt[c, 0], zt[c, 1:], zd[c, 0], zd[c, 1:] = a[0], a[1], b[0], b[1]
    zg = np.concatenate([np.asarray(Tg)[:, None], np.asarray(Yg)], axis=1)
    return compare(zg, zd, zt, rtol, atol)


def assert_like_dvode(e, factor=2.0, comp_factor=3.0, floor=0.0):
    """floor: component means below this many tolerance units pass (after
    one substep of a barely reacting field the temperature error is
    round-off, ~1e-7 units, and the ratio of two such means says nothing)."""
    (gm, dm), (ga, da), (gc, dc) = e["norm_max"], e["norm_mean"], e["comp_mean"]
    assert gm <= factor * dm, ("per-cell norm, max", gm, dm)
    assert ga <= factor * da, ("per-cell norm, mean", ga, da)
    # a component DVODE integrates exactly (an inert species) must stay
    # within one tolerance unit
    bad = (gc > np.where(dc > 0, comp_factor * dc, 1.0)) & (gc > floor)
    assert not bad.any(), ("component means", np.nonzero(bad)[0], gc[bad], dc[bad])


def summary(e) -> str:
    (gm, dm), (ga, da), (gc, dc) = e["norm_max"], e["norm_mean"], e["comp_mean"]
    r = np.max(np.where(dc > 0, gc / np.where(dc > 0, dc, 1.0), 0.0))
    return (f"error norm max {gm:.2f} (DVODE {dm:.2f}), mean {ga:.3f} (DVODE {da:.3f}), "
            f"worst component mean ratio {r:.2f}")
