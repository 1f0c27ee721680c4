"""The mixed-precision form (ERFX_PREC_MIXED: FP32 face values of the advected
quantities on perturbations about an FP64 face reference, everything else
FP64) against the FP64 form on the GPU: relative L2 and Linf differences of
rho, rho*theta, u, v, w (and qv, qc, qr with moisture) after a run, within the
bounds stated here; the numbers are printed."""
import numpy as np
import pytest

from erf_amd import config as C, problems as P
from helpers import valid

pytestmark = pytest.mark.gpu


def _errors(cfg, f, dt, nsteps, rayleigh=None):
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    cm = C.Config.from_buffer_copy(cfg)
    cm.precision = C.PREC_MIXED
    sims = []
    for c in (cfg, cm):
        s = Sim(c)
        s.load(f)
        if rayleigh is not None:
            s.set_rayleigh(*rayleigh)
        s.step(0.0, dt, nsteps)
        sims.append(s)
    L = sims[0].L
    out = {}
    for fid, comps in ((C.F_CONS, [(C.RHO, "rho"), (C.RHOTHETA, "rhotheta")] +
                        ([(C.RHOQ1, "qv"), (C.RHOQ2, "qc"), (C.RHOQ3, "qr")] if cfg.moisture_type else [])),
                       (C.F_XVEL, [(0, "u")]), (C.F_YVEL, [(0, "v")]), (C.F_ZVEL, [(0, "w")])):
        a = valid(sims[0].download(fid), cfg, L, fid)
        b = valid(sims[1].download(fid), cfg, L, fid)
        for c, name in comps:
            d = b[c] - a[c]
            ref = a[c]
            if fid == C.F_CONS and c in (C.RHO, C.RHOTHETA):   # perturbation scale of the field
                ref = ref - ref.mean()
            nrm = np.sqrt((ref ** 2).mean())
            out[name] = (float(np.sqrt((d ** 2).mean()) / nrm) if nrm > 0 else 0.0,
                         float(np.abs(d).max() / np.abs(ref).max()) if np.abs(ref).max() > 0 else 0.0)
    for s in sims:
        s.close()
    return out


def test_mixed_cbl_u5():
    cfg, f = P.cbl(32, 32, 32, seed=3)
    e = _errors(cfg, f, 0.5, 20)
    print("CBL 32^3 U5, 20 steps, mixed vs FP64 (rel L2, rel Linf):", e)
    for k, (l2, linf) in e.items():
        assert l2 < 1e-4 and linf < 1e-3, (k, l2, linf)


def test_mixed_moist_terrain_weno5():
    cfg, f = P.moist_terrain(24, 20, 16, seed=5)
    e = _errors(cfg, f, 1.0, 10, rayleigh=P.rayleigh_profiles_uniform(cfg, f=f))
    print("moist terrain 24x20x16 WENO5, 10 steps, mixed vs FP64 (rel L2, rel Linf):", e)
    for k, (l2, linf) in e.items():
        assert l2 < 1e-4 and linf < 1e-3, (k, l2, linf)
