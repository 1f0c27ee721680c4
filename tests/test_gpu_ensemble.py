"""Batched ensembles (erfx_ensemble): members of different box sizes advanced
with one launch per kernel for all of them.  Bar: every member bit-identical
to the same sim stepped alone (itself bit-identical to the oracle, which one
member also checks directly), with and without graph replay, across the
changes of the stage time steps as t grows, for dry CBL, moist (Kessler) and
terrain members; members whose kernel sequences differ are refused."""
import numpy as np
import pytest

from erf_amd import config as C, problems as P
from helpers import valid, rel_err
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

FIELDS = (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL)


def _members(kind):
    if kind == "cbl":
        return list(P.cbl_ensemble(3, seed=4, sizes=((16, 16, 16), (24, 20, 16), (20, 28, 24))))
    if kind == "moist":
        return [P.moist_bubble(nx, ny, nz, seed=s) for (nx, ny, nz), s in
                (((24, 20, 16), 1), ((16, 16, 20), 2), ((28, 16, 16), 3))]
    return [P.hill_flow(nx, ny, nz, seed=s) for (nx, ny, nz), s in (((24, 20, 16), 5), ((20, 16, 16), 6))]


def _pairs(cases):
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    a, b = [], []
    for cfg, f in cases:
        x, y = Sim(cfg), Sim(cfg)
        x.load(f)
        y.load(f)
        a.append(x)
        b.append(y)
    return a, b


def _diff(a, b):
    bad = []
    for m, (x, y) in enumerate(zip(a, b)):
        for fid in FIELDS:
            if not np.array_equal(x.download(fid), y.download(fid), equal_nan=True):
                bad.append((m, fid))
    return bad


@pytest.mark.parametrize("kind", ["cbl", "moist", "hill"])
@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_matches_members_alone(kind, graph):
    from erf_amd.lib import Ensemble
    cases = _members(kind)
    dt = 0.5 if kind == "cbl" else 1.0
    a, b = _pairs(cases)
    ens = Ensemble(a)
    ens.set_graph(graph)
    t = 0.0
    for n in (2, 1, 3):       # t crosses 1 and 2 (dt = 0.5) or 2 and 4 (dt = 1)
        ens.step(t, dt, n)
        for s in b:
            s.step(t, dt, n)
        t += n * dt
        assert _diff(a, b) == [], (kind, graph, t)
    launches, cells = ens.info()
    assert cells == sum(c.nx * c.ny * c.nz for c, _ in cases)
    assert launches > 0
    ens.close()
    for s in a + b:
        s.close()


# A batched member's scalar update is the per-cell form; a sim stepped alone
# uses the k-marching form, whose tiling depends on the box: at these sizes
# (at least 256 work-groups) its default tiles of 15 rows x 12 levels, while
# the small members above take the tiling for small boxes.
@pytest.mark.parametrize("kind", ["cbl_u5", "moist_u3_weno3", "hill_weno5", "cbl_u5_mixed", "moist_mixed"])
def test_ensemble_matches_member_alone_default_march_tiling(kind):
    """also for the mixed-precision form (FP32 face values), whose march and
    per-cell forms evaluate each face by the same expressions"""
    from erf_amd.lib import Ensemble
    if kind.endswith("_mixed"):
        cfg, f = P.cbl(256, 256, 64, seed=2) if kind.startswith("cbl") else P.moist_bubble(256, 256, 64, seed=3)
        cfg.precision = C.PREC_MIXED
        case, dt = (cfg, f), 0.5 if kind.startswith("cbl") else 1.0
    elif kind == "cbl_u5":
        case, dt = P.cbl(256, 256, 64, seed=2), 0.5
    elif kind == "moist_u3_weno3":   # dry scalars U3, moist scalars WENO3
        case, dt = P.moist_bubble(256, 256, 64, seed=3), 1.0
    else:
        case, dt = P.hill_flow(256, 256, 64, seed=4, adv=C.ADV_UPWIND3, scal_adv=C.ADV_WENO5), 1.0
    a, b = _pairs([case])
    ens = Ensemble(a)
    ens.step(0.0, dt, 2)
    b[0].step(0.0, dt, 2)
    assert _diff(a, b) == [], kind
    ens.close()
    for s in a + b:
        s.close()


@pytest.mark.parametrize("streams", [1, 3])
def test_ensemble_stream_groups(streams):
    """members split over 1 or 3 groups (streams): the same results"""
    from erf_amd.lib import Ensemble
    cases = _members("cbl")
    a, b = _pairs(cases)
    ens = Ensemble(a)
    ens.set_streams(streams)
    ens.step(0.0, 0.5, 3)
    for s in b:
        s.step(0.0, 0.5, 3)
    assert _diff(a, b) == []
    ens.close()
    for s in a + b:
        s.close()


def test_ensemble_member_against_oracle():
    from erf_amd.lib import Ensemble
    cases = _members("cbl")
    a, _ = _pairs(cases)
    ens = Ensemble(a)
    ens.step(0.0, 0.5, 2)
    cfg, f = cases[1]
    orc = Oracle(cfg)
    orc.load(f)
    orc.step(0.0, 0.5, 2)
    for fid in FIELDS:
        x = valid(a[1].download(fid), cfg, a[1].L, fid)
        y = valid(orc.get(fid), cfg, orc.L, fid)
        assert np.array_equal(x, y), (fid, rel_err(x, y))
    ens.close()
    for s in a:
        s.close()


def test_ensemble_autotune_keeps_results():
    from erf_amd.lib import Ensemble
    cases = _members("cbl")
    a, b = _pairs(cases)
    ens = Ensemble(a)
    ens.autotune(True)
    ens.step(0.0, 0.5, C.COL_VARIANTS)     # one step per column variant
    v, ms = ens.column_variant()
    assert 0 <= v < C.COL_VARIANTS and all(m > 0 for m in ms) and ms[v] == min(ms)
    ens.step(0.5 * C.COL_VARIANTS, 0.5, 2)
    for s in b:
        s.step(0.0, 0.5, C.COL_VARIANTS + 2)
    assert _diff(a, b) == []
    ens.close()
    for s in a + b:
        s.close()


def test_ensemble_refuses_different_sequences():
    from erf_amd.lib import Ensemble, ErfxError
    c1, f1 = P.rising_thermal(16, 16, 16, seed=1)
    c2, f2 = P.rising_thermal(20, 16, 16, seed=2)
    c2.mri_ratio = 4                   # a different substep count: other kernels
    a, _ = _pairs([(c1, f1), (c2, f2)])
    ens = Ensemble(a)
    ens.set_streams(1)                 # both in one group (one batch)
    with pytest.raises(ErfxError):
        ens.step(0.0, 1.0, 1)
    ens.close()
    for s in a:
        s.close()
