"""The no-allocation variant of erf_fast_rhs_N (erfx_fast_rhs_N_noalloc: its
work fields in a static device array, no context and no device allocation)
against the context hook erfx_fast_rhs_N on the same inputs: bit-identical
outputs, over several calls in a row on one box and then on a smaller box
(the array keeps the previous call's contents)."""
import ctypes

import numpy as np
import pytest

from erf_amd import config as C, problems as P
from helpers import ALL_IDS
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _noalloc():
    from erf_amd.lib import lib
    return lib().erfx_fast_rhs_N_noalloc      # argument types from erf_amd.lib.SIGNATURES


def _pair(nx, ny, nz, seed):
    """Two sims holding the same state: every buffer of the integrator after
    one oracle step of the rising thermal."""
    from erf_amd.lib import Sim
    cfg, f = P.rising_thermal(nx, ny, nz, seed=seed, adv=C.ADV_UPWIND3)
    orc = Oracle(cfg)
    orc.load(f)
    orc.step(0.0, 1.0, 1)
    sims = []
    for _ in range(2):
        s = Sim(cfg)
        for fid in ALL_IDS:
            s.upload(fid, orc.get(fid))
        sims.append(s)
    return cfg, sims


def test_noalloc_fast_rhs_N_matches_context_hook():
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import lib, check
    fn = _noalloc()
    # the second box is smaller than the first: the static array then holds the
    # first box's data beyond the second's extent
    for nx, ny, nz, seed in [(40, 32, 24, 3), (20, 16, 12, 5)]:
        print("no-alloc fast_rhs_N on %dx%dx%d" % (nx, ny, nz), flush=True)
        cfg, (a, b) = _pair(nx, ny, nz, seed)
        nsub = cfg.mri_ratio
        dtau = 1.0 / nsub
        beta_s = 0.1

        def st(sim, s):
            return ctypes.byref(sim.state(s))

        for step, nrk in ((0, 0), (0, 2), (1, 2), (2, 2)):
            prev = C.S_OLD if step == 0 else C.S_SUM
            check(lib().erfx_fast_rhs_N(a.ctx, step, nrk, st(a, C.F_SLOW), st(a, prev), st(a, C.S_NEW),
                                        a.ptr(C.D_PRIM), a.ptr(C.D_PI_STAGE), a.ptr(C.D_FAST_COEFFS),
                                        st(a, C.S_SUM), st(a, C.S_SCRATCH), dtau, beta_s, 1.0 / nsub))
            check(fn(ctypes.byref(cfg), b.stream, step, nrk, st(b, C.F_SLOW), st(b, prev), st(b, C.S_NEW),
                     b.ptr(C.D_PRIM), b.ptr(C.D_PI_STAGE), b.ptr(C.D_FAST_COEFFS),
                     st(b, C.S_SUM), st(b, C.S_SCRATCH), dtau, beta_s, 1.0 / nsub))
            a.sync()
            b.sync()
        diff = {}
        for fid in ALL_IDS:
            x, y = a.download(fid), b.download(fid)
            # cells neither call writes keep the oracle's NaN fill on both sides
            d = (x != y) & ~(np.isnan(x) & np.isnan(y))
            nd = int(d.sum())
            if nd:
                diff[fid] = (nd, float(np.abs(x - y)[d].max()))
        a.close()
        b.close()
        assert not diff, "not bit-identical: %s" % diff


def test_noalloc_rejects_a_box_beyond_its_scratch():
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import lib, State
    fn = _noalloc()
    cfg = C.default_config(128, 128, 128, 100.0, 100.0, 100.0)
    dummy = State(1, 1, 1, 1)           # never dereferenced: the size check comes first
    rc = fn(ctypes.byref(cfg), None, 0, 0, ctypes.byref(dummy), ctypes.byref(dummy), ctypes.byref(dummy),
            1, 1, 1, ctypes.byref(dummy), ctypes.byref(dummy), 0.1, 0.1, 1.0)
    assert rc == -1                     # ERFX_ERR_ARG
    assert b"static buffer" in lib().erfx_last_error()
