"""The reference's regression cases on the GPU: the same initial states as the
oracle's gold tests, ERF's time loop with its dt control (erfx_sim_run:
ComputeDt / estTimeStep before every step), bit-identical to the oracle, which
in turn matches the reference's gold files (test_gold_*.py)."""
import numpy as np
import pytest

from erf_amd import config as C
from helpers import valid
from oracle_lib import Oracle
import test_gold_regression_cases as G

pytestmark = pytest.mark.gpu

FIELDS = (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL)


@pytest.mark.parametrize("name", G.CASES)
def test_gold_case_on_gpu(name):
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    cfg, init, run, nsteps = G._case(name)
    # initial state from the oracle's restatement of the case
    print("gold case %s on the GPU" % name, flush=True)
    o0 = Oracle(cfg)
    G.init_oracle(o0, init)
    state = {fid: o0.get(fid) for fid in (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL, C.F_R0, C.F_P0, C.F_PI0)}
    cfl, fdt, ffdt, fratio, cmax, shrink = G.run_args(run)
    dtc = C.dt_control(cfl=cfl, fixed_dt=fdt, fixed_fast_dt=ffdt, fixed_mri_ratio=fratio, change_max=cmax,
                       init_shrink=shrink)
    sim = Sim(cfg)
    sim.load(state)
    ray = G.rayleigh_profiles(name, cfg)
    if ray is not None:
        sim.set_rayleigh(*ray)
    t_gpu, n_gpu = sim.run(0.0, nsteps, dtc)
    _, orc, t_orc = G.run_case(name)
    assert n_gpu == nsteps
    assert t_gpu == t_orc, (t_gpu, t_orc)
    diff = {}
    for fid in FIELDS:
        a = valid(sim.download(fid), cfg, sim.L, fid)
        b = valid(orc.get(fid), cfg, orc.L, fid)
        nd = int((a != b).sum())
        if nd:
            diff[fid] = (nd, float(np.abs(a - b).max()))
    sim.close()
    assert not diff, "not bit-identical: %s" % diff
