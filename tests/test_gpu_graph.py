"""izem_step replayed from its captured HIP graph against the same call run
eagerly: after every step the conserved state, temperature, primitive
arrays, sensor and time state are bit-identical (each replay follows eager
work and a synchronisation of the program's own); a change of arrays
recaptures; the stiffness order is refused in graph mode."""
from __future__ import annotations

import numpy as np
import pytest

from tests.flowcase import OracleRun, small_state

pytestmark = pytest.mark.gpu


def _pair(oracle, adi, thd, mech, kind, n, reaction):
    import torch
    from phd_amd.solver import Options, Solver
    box, v, T = small_state(adi, thd, kind, n)
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.prim_all()
    out = []
    for _ in range(2):
        s = Solver(box, Options(reaction=reaction), adi=adi, thd=thd,
                   mech=mech if reaction else None)
        for name in ("v", "T", "W_i", "cp", "ha", "sensor"):
            getattr(s, name).copy_(torch.from_numpy(getattr(run, name)))
        s.set_time_state([1.0, 1.0, 1.0, 1.0, 1.0, 0.0])
        out.append(s)
    torch.cuda.synchronize()
    return out


def _same(a, b):
    for name in ("v", "T", "W_i", "cp", "ha", "sensor"):
        x, y = getattr(a, name).cpu().numpy(), getattr(b, name).cpu().numpy()
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name
    assert np.array_equal(a.time_state().view(np.uint64), b.time_state().view(np.uint64))


@pytest.mark.parametrize("kind,n,reaction", [("2d", 24, 1), ("3d", 12, 0), ("3d", 12, 1)])
def test_graph_step_equals_eager(oracle, adi, thd, mech, kind, n, reaction):
    eager, graph = _pair(oracle, adi, thd, mech, kind, n, reaction)
    for it in range(6):
        eager.step(graph=False)
        graph.step(graph=True)
        eager.sync()
        graph.sync()
        _same(eager, graph)
        if it == 2:
            # new arrays with the same contents: the graph is captured again
            for name in ("v", "T", "W_i", "cp", "ha", "sensor"):
                setattr(graph, name, getattr(graph, name).clone())
    assert eager.status() == 0 and graph.status() == 0


def test_graph_step_refuses_stiffness_order(oracle, adi, thd, mech):
    from phd_amd import _lib
    eager, graph = _pair(oracle, adi, thd, mech, "2d", 8, 1)
    assert graph.lib.izem_chem_order(graph.ctx, 1) == 0
    with pytest.raises(_lib.IzemError):
        graph.step(graph=True)
    assert graph.lib.izem_chem_order(graph.ctx, 0) == 0
    graph.step(graph=True)
    eager.step(graph=False)
    _same(eager, graph)
