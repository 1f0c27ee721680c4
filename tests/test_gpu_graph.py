"""The step sequence replayed from a captured hipGraph (erfx_sim_set_graph)
against the same steps launched directly: bit-identical after every replay,
with eager work and host synchronisation between the replays, for dry, moist
(Kessler) and terrain cases; a change of dt re-captures the graph."""
import numpy as np
import pytest

from erf_amd import config as C, problems as P

pytestmark = pytest.mark.gpu

FIELDS = (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL)


def _pair(cfg, f):
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    a, b = Sim(cfg), Sim(cfg)
    a.load(f)
    b.load(f)
    a.set_graph(True)
    return a, b


def _same(a, b):
    bad = []
    for fid in FIELDS:
        x, y = a.download(fid), b.download(fid)
        if not np.array_equal(x, y, equal_nan=True):
            bad.append(fid)
    return bad


@pytest.mark.parametrize("case", ["thermal", "moist", "hill"])
def test_graph_replay_matches_direct(case):
    if case == "thermal":
        cfg, f = P.rising_thermal(32, 24, 16, seed=5)
    elif case == "moist":
        cfg, f = P.moist_bubble(24, 20, 16, seed=7)
    else:
        cfg, f = P.hill_flow(24, 20, 16, seed=9)
    a, b = _pair(cfg, f)
    t = 0.0
    for rep in range(3):                      # three replays of the same graph
        a.step(t, 1.0, 2)
        b.step(t, 1.0, 2)
        t += 2.0
        assert _same(a, b) == [], (case, rep)
        b.download(C.F_CONS)                  # eager work between replays
    a.step(t, 1.0, 3)                         # one replay plus one direct step
    b.step(t, 1.0, 3)
    t += 3.0
    assert _same(a, b) == []
    a.step(t, 0.5, 2)                         # new dt: captured again
    b.step(t, 0.5, 2)
    assert _same(a, b) == []
    a.close()
    b.close()
