"""Launch-level options of the sim against the plain path:

* the step sequence replayed from a captured hipGraph (erfx_sim_set_graph):
  bit-identical after every replay, with eager work and host synchronisation
  between the replays, for dry, moist (Kessler) and terrain cases; a change of
  dt re-captures the graph;
* every variant of the acoustic column kernel (erfx_sim_set_column_variant)
  and the autotuner that picks one (erfx_sim_autotune): bit-identical."""
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


@pytest.mark.parametrize("variant", list(range(C.COL_VARIANTS)))
def test_column_variants_identical(variant):
    cfg, f = P.rising_thermal(40, 24, 16, seed=11, adv=C.ADV_UPWIND3)
    a, b = _pair(cfg, f)
    a.set_graph(False)
    a.set_column_variant(variant)
    a.step(0.0, 1.0, 3)
    b.step(0.0, 1.0, 3)
    assert _same(a, b) == []
    a.close()
    b.close()


def test_autotune_picks_a_variant_and_keeps_results():
    cfg, f = P.moist_bubble(24, 20, 16, seed=13)
    a, b = _pair(cfg, f)
    a.set_graph(False)
    a.autotune(True)
    a.step(0.0, 1.0, 3)          # 3 steps x 10 substeps: each of the 10 variants timed 3 times
    b.step(0.0, 1.0, 3)
    v, ms = a.column_variant()
    assert 0 <= v < C.COL_VARIANTS and all(m > 0 for m in ms) and ms[v] == min(ms)
    a.step(3.0, 1.0, 2)          # with the chosen variant
    b.step(3.0, 1.0, 2)
    assert _same(a, b) == []
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["flat37", "flat_moist_50", "hill21", "hill_moist_30", "flat_deep128",
                                  "flat_tc32_40", "flat_tc32_80", "flat_tc32_deep100"])
def test_tile_form_matches_column_form(case):
    """The tile form of the acoustic column kernel (the default) against the
    one-lane-per-column form on level counts that leave the last wave's chunk
    partial and trailing waves empty (37 levels: chunks of 3 over 16 waves;
    21 levels on terrain: chunks of 3 over 8 waves), flat and terrain, dry
    and moist: bit-identical after several steps.  Boxes of fewer than
    16,384 columns take 16-column tiles (32 chunks of levels) and those of
    8,192 or more 32-column tiles (16 chunks), also with 97..128 levels: 40 x
    24 x 128 (16 columns, chunks of 4), 96 x 96 x 40 and x 80 (32 columns,
    chunks of 3 and 5), 96 x 96 x 100 (32 columns, chunks of 7)."""
    if case == "flat37":
        cfg, f = P.rising_thermal(40, 24, 37, seed=21, adv=C.ADV_UPWIND3)
    elif case == "flat_deep128":
        cfg, f = P.rising_thermal(40, 24, 128, seed=25, adv=C.ADV_UPWIND3)
    elif case.startswith("flat_tc32"):
        nz = {"flat_tc32_40": 40, "flat_tc32_80": 80, "flat_tc32_deep100": 100}[case]
        cfg, f = P.rising_thermal(96, 96, nz, seed=26, adv=C.ADV_UPWIND3)
    elif case == "flat_moist_50":
        cfg, f = P.moist_bubble(24, 20, 50, seed=22)
    elif case == "hill21":
        cfg, f = P.hill_flow(24, 20, 21, seed=23)
    else:
        cfg, f = P.hill_flow(24, 20, 30, seed=24, moist=True)
    a, b = _pair(cfg, f)
    a.set_graph(False)
    a.set_column_variant(7)          # one lane per column
    a.step(0.0, 1.0, 3)
    b.step(0.0, 1.0, 3)
    assert _same(a, b) == []
    a.close()
    b.close()
