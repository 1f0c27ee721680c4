"""AMReX plotfile writer / reader and the fcompare-style comparison
(erf_amd/plotfile.py), CPU only.

  * a write / read round trip is exact, and the FAB and Cell_H headers carry
    the box, component count and per-component min / max;
  * the writer's Header has the structure of the reference's own gold Header
    (tests/golden/ScalarAdvDiff_order2.Header, copied from
    Tests/ERFGoldFiles/ScalarAdvDiff_order2): same line kinds in the same
    order, same values where the case fixes them;
  * the oracle's end state of that gold case, written as a plotfile with the
    derived plot variables and read back, passes fcompare against the gold
    data at the regression suite's tolerance (2e-10), and fcompare fails a
    field with a planted 1e-8 error.
"""
import os

import numpy as np

from erf_amd import config as C, plotfile
from helpers import valid
import test_gold_regression_cases as G

HERE = os.path.dirname(os.path.abspath(__file__))


def test_round_trip_and_headers(tmp_path):
    rng = np.random.default_rng(3)
    data = rng.normal(size=(3, 5, 6, 7))
    p = str(tmp_path / "plt00007")
    plotfile.write(p, ["a", "b", "c"], data, 1.25, 7, [0.0, -1.0, 0.0], [7.0, 5.0, 2.5])
    r = plotfile.read(p)
    assert r["names"] == ["a", "b", "c"]
    assert r["time"] == 1.25 and r["nstep"] == 7 and r["n"] == [7, 6, 5]
    assert r["lo"] == [0.0, -1.0, 0.0] and r["hi"] == [7.0, 5.0, 2.5]
    assert np.array_equal(r["data"], data)
    head = open(os.path.join(p, "Level_0", "Cell_D_00000"), "rb").readline().decode()
    assert head.endswith("((0,0,0) (6,5,4) (0,0,0)) 3\n")
    mins = [float(x) for x in r["cell_h"].split("\n")[11].rstrip(",").split(",")]
    assert np.array_equal(mins, data.reshape(3, -1).min(axis=1))


def test_header_matches_reference_layout(tmp_path):
    gold = open(os.path.join(HERE, "golden", "ScalarAdvDiff_order2.Header")).read().split("\n")
    z = np.load(os.path.join(HERE, "golden", "ScalarAdvDiff_order2.npz"))
    p = str(tmp_path / "plt00020")
    plotfile.write(p, list(z["names"]), z["data"], float(z["time"]), int(z["nstep"]), list(z["lo"]), list(z["hi"]))
    mine = open(os.path.join(p, "Header")).read().split("\n")
    nvar = int(gold[1])
    # magic, variable count and names, dimension, time, finest level, the
    # domain corners, the blank line, the domain box, the step, the cell size,
    # coordinate system, boundary width: line for line
    head = 2 + nvar + 11
    for g, m in zip(gold[:head], mine[:head]):
        gt, mt = g.split(), m.split()
        assert len(gt) == len(mt), (g, m)
        for a, b in zip(gt, mt):
            try:
                assert float(a) == float(b), (g, m)
            except ValueError:
                assert a == b, (g, m)
    # then "level nboxes time", the level's step, one extent line per box and
    # direction, and the name of the level's data
    lvl, nb, t = gold[head].split()
    assert mine[head].split()[0] == lvl and float(mine[head].split()[2]) == float(t)
    assert gold[head + 1] == mine[head + 1]
    assert gold[head + 2 + 3 * int(nb)] == mine[head + 2 + 3] == "Level_0/Cell"


def test_oracle_end_state_passes_fcompare(tmp_path):
    name = "ScalarAdvDiff_order2"
    z = np.load(os.path.join(HERE, "golden", name + ".npz"))
    cfg, orc, t_end = G.run_case(name)
    L = orc.L
    cons = valid(orc.get(C.F_CONS), cfg, L, C.F_CONS)
    u = valid(orc.get(C.F_XVEL), cfg, L, C.F_XVEL)[0]
    v = valid(orc.get(C.F_YVEL), cfg, L, C.F_YVEL)[0]
    w = valid(orc.get(C.F_ZVEL), cfg, L, C.F_ZVEL)[0]
    d = plotfile.derive(cfg, cons, u, v, w)
    names = list(z["names"])
    p = str(tmp_path / "plt00020")
    plotfile.write(p, names, np.stack([d[n] for n in names]), t_end, 20, list(z["lo"]), list(z["hi"]))
    mine = plotfile.read(p)
    gold = dict(names=names, data=z["data"])
    rep = plotfile.fcompare(mine, gold)
    assert all(ok for _, _, ok in rep.values()), rep
    # a planted error beyond the tolerance in one variable fails it
    bad = dict(mine)
    bad["data"] = mine["data"].copy()
    bad["data"][4, 3, 3, 3] += 1e-8
    rep = plotfile.fcompare(bad, gold)
    assert not rep["scalar"][2] and all(ok for n, (_, _, ok) in rep.items() if n != "scalar")
