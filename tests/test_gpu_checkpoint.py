"""Checkpoint and restart of the sim (erfx_sim_write_checkpoint /
erfx_sim_read_checkpoint, the role of ERF's checkpoint files,
Source/IO/ERF_Checkpoint.cpp:25-590): a run stopped after k steps, written,
read into a fresh sim and stepped on gives the uninterrupted run bit for bit,
for a dry box, a moist box with rain (surface rain accumulated) and a terrain
grid with Rayleigh damping; a checkpoint of another grid is refused."""
import os
import tempfile

import numpy as np
import pytest

from erf_amd import config as C, problems as P

pytestmark = pytest.mark.gpu

FIELDS = (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL)


def _case(kind):
    if kind == "dry":
        cfg, f = P.cbl(20, 16, 16, seed=5)
        return cfg, f, None, 0.5
    if kind == "moist":
        cfg, f = P.moist_bubble(20, 16, 16, seed=13)
        return cfg, f, None, 1.0
    cfg, f = P.moist_terrain(24, 20, 24, seed=3)
    return cfg, f, P.rayleigh_profiles_uniform(cfg, f=f), 1.0


@pytest.mark.parametrize("kind", ["dry", "moist", "terrain"])
def test_restart_is_bit_identical(kind):
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    print("checkpoint / restart: %s" % kind, flush=True)
    cfg, f, ray, dt = _case(kind)
    n1, n2 = 3, 4

    def fresh():
        s = Sim(cfg)
        s.load(f)
        if ray is not None:
            s.set_rayleigh(*ray)
        return s

    whole = fresh()
    whole.step(0.0, dt, n1 + n2)
    first = fresh()
    first.step(0.0, dt, n1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "chk%05d" % n1)
        first.write_checkpoint(path, n1 * dt, n1)
        first.close()
        second = Sim(cfg)                       # nothing loaded: everything comes from the files
        t, n = second.read_checkpoint(path)
    assert (t, n) == (n1 * dt, n1)
    second.step(t, dt, n2)
    ids = list(FIELDS)
    if cfg.moisture_type != C.MOIST_NONE:
        ids.append(C.D_RAIN_ACCUM)
    for fid in ids:
        a, b = whole.download(fid), second.download(fid)
        assert np.array_equal(a, b), (fid, int((a != b).sum()))
    whole.close()
    second.close()


def test_checkpoint_of_another_grid_is_refused():
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim, ErfxError
    cfg, f = P.cbl(20, 16, 16, seed=5)
    s = Sim(cfg)
    s.load(f)
    cfg2, _ = P.cbl(24, 16, 16, seed=5)
    s2 = Sim(cfg2)
    with tempfile.TemporaryDirectory() as d:
        s.write_checkpoint(os.path.join(d, "chk"), 0.0, 0)
        with pytest.raises(ErfxError):
            s2.read_checkpoint(os.path.join(d, "chk"))
    s.close()
    s2.close()
