"""GPU checks of the microphysics step with the learned stand-in
(erfx_kessler_mlp_advance) against plain Kessler (erfx_kessler_advance):

* with an empty training range every active cell falls back to Kessler, and
  the step is bit-identical to Kessler's;
* with any network (here an untrained one with large outputs) the projection
  conserves qv + qc + qr in every cell (no rain, so no sedimentation), keeps
  every species >= 0 and ties theta to the vapour change by Kessler's latent
  heat term, theta/T * L/c_p;
* training on Kessler's own tendencies (harvested from moist runs, seeds 0-2)
  cuts the held-out error (seed 3) by at least 5x against the untrained
  network, and a 10-step run with the trained network stays within stated
  bounds of plain Kessler (relative L2: rho 1e-4, rho theta 1e-3, qv 1e-2,
  qc and qr 0.1); the numbers are printed;
* the cost per microphysics call of both, and the gain of the network over
  Kessler, are printed (recorded, not required: there is none on MI355X).
"""
import ctypes
import time

import numpy as np
import pytest

from erf_amd import config as C, problems as P
from helpers import valid, rel_err

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _sims(cfg, f):
    from erf_amd.lib import Sim
    a, b = Sim(cfg), Sim(cfg)
    a.load(f)
    b.load(f)
    return a, b


def test_fallback_everywhere_is_kessler_bit_for_bit():
    from erf_amd.lib import MLP
    from erf_amd import emulator as E
    cfg, f = P.moist_bubble(24, 20, 16, seed=31)
    ka, em = _sims(cfg, f)
    net = MLP(64)
    net.set_params(E.init_params(1))
    inf = 1e300
    net.set_scaling(np.zeros(8), np.ones(8), np.full(8, inf), np.full(8, -inf), np.ones(4))
    em.set_mlp(net)
    ka.step(0.0, 1.0, 2)
    em.step(0.0, 1.0, 2)
    n_mlp, n_fb = em.mlp_counts()
    assert n_mlp == 0 and n_fb > 0
    for fid in (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL, C.D_RAIN_ACCUM):
        a, b = ka.download(fid), em.download(fid)
        assert np.array_equal(a, b), fid
    ka.close()
    em.close()


def test_projection_conserves_water_keeps_species_nonnegative():
    torch = _torch()
    from erf_amd.lib import MLP, lib, check
    from erf_amd import emulator as E
    cfg, f = P.moist_bubble(24, 20, 16, seed=33)
    f[C.F_CONS][C.RHOQ3][:] = 0.0                       # no rain: no sedimentation
    rng = np.random.default_rng(33)                     # cloud everywhere: every cell active
    f[C.F_CONS][C.RHOQ2][:] = f[C.F_CONS][C.RHO] * rng.uniform(0.0, 1e-3, f[C.F_CONS][C.RHO].shape)
    from erf_amd.lib import Sim
    sim = Sim(cfg)
    sim.load(f)
    net = MLP(64)
    p = E.init_params(5)
    p[-4:] = [3.0, -4.0, 5.0, 1.0]                      # large, unbalanced outputs
    net.set_params(p)
    big = 1e300
    net.set_scaling(np.zeros(8), np.ones(8), np.full(8, -big), np.full(8, big), np.array([2e-3, 2e-3, 2e-3, 1.0]))
    before = valid(sim.download(C.F_CONS), cfg, sim.L, C.F_CONS)
    n_mlp, n_fb = ctypes.c_long(), ctypes.c_long()
    check(lib().erfx_kessler_mlp_advance(sim.ctx, net.h, sim.ptr(C.F_CONS), sim.ptr(C.F_ZVEL), 1.0, None,
                                         ctypes.byref(n_mlp), ctypes.byref(n_fb)))
    after = valid(sim.download(C.F_CONS), cfg, sim.L, C.F_CONS)
    assert n_mlp.value == cfg.nx * cfg.ny * cfg.nz and n_fb.value == 0
    rho = before[C.RHO]
    q0 = [np.maximum(before[c] / rho, 0.0) for c in (C.RHOQ1, C.RHOQ2, C.RHOQ3)]
    q1 = [after[c] / rho for c in (C.RHOQ1, C.RHOQ2, C.RHOQ3)]
    for q in q1:
        assert np.all(q >= 0.0)
    tw0, tw1 = sum(q0), sum(q1)
    assert np.max(np.abs(tw1 - tw0)) <= 1e-15
    moved = np.abs(q1[0] - q0[0]) > 1e-6
    assert moved.sum() > 100                            # the network did change the cells
    # theta follows the vapour change with Kessler's latent heat factor
    from oracle_lib import temperature
    th0, th1 = before[C.RHOTHETA] / rho, after[C.RHOTHETA] / rho
    T0 = temperature(rho, before[C.RHOTHETA], before[C.RHOQ1] / rho)
    expect = -(th0 / T0) * (2.501e6 / cfg.c_p) * (q1[0] - q0[0])
    assert np.max(np.abs((th1 - th0) - expect)) <= 1e-9 * np.max(np.abs(expect))
    sim.close()
    net.close()


def test_training_gain_and_emulated_run():
    torch = _torch()
    from erf_amd.lib import MLP, Sim
    from erf_amd import emulator as E
    train = E.harvest(seeds=(0, 1, 2), nsteps=20, shape=(64, 64, 32))
    test = E.harvest(seeds=(3,), nsteps=20, shape=(64, 64, 32))
    assert train.shape[0] > 20000 and test.shape[0] > 5000
    sc = E.Scaling(train)
    untrained = MLP(64)
    untrained.set_params(E.init_params(0))
    sc.apply(untrained)
    mse0 = E.mse(untrained, sc, test)
    t0 = time.time()
    net, sc, losses = E.fit(train, steps=1500, batch=4096, lr=2e-3, scaling=sc, log_every=500)
    t_fit = time.time() - t0
    mse1 = E.mse(net, sc, test)
    print("training rows %d, held-out rows %d, fit %.1f s, losses %s" % (train.shape[0], test.shape[0], t_fit, losses))
    print("held-out MSE (scaled): untrained %.4g, trained %.4g" % (mse0, mse1))
    assert mse1 < mse0 / 5.0
    # 10 steps of the moist bubble (seed 3) with the network vs plain Kessler
    cfg, f = P.moist_bubble(64, 64, 32, seed=3)
    ka, em = _sims(cfg, f)
    em.set_mlp(net)
    ka.step(0.0, 1.0, 10)
    em.step(0.0, 1.0, 10)
    n_mlp, n_fb = em.mlp_counts()
    a = valid(ka.download(C.F_CONS), cfg, ka.L, C.F_CONS)
    b = valid(em.download(C.F_CONS), cfg, em.L, C.F_CONS)
    errs = {name: float(np.linalg.norm(b[c] - a[c]) / np.linalg.norm(a[c]))
            for name, c in (("rho", C.RHO), ("rhotheta", C.RHOTHETA), ("qv", C.RHOQ1), ("qc", C.RHOQ2),
                            ("qr", C.RHOQ3))}
    print("10 steps, network cells %d, fallback cells %d, relative L2 vs Kessler %s" % (n_mlp, n_fb, errs))
    assert n_mlp > 0
    # rho differs only through the dynamics' response to the heating
    assert errs["rho"] < 1e-4 and errs["rhotheta"] < 1e-3 and errs["qv"] < 1e-2
    assert errs["qc"] < 0.1 and errs["qr"] < 0.1
    # microphysics cost per call on the last state, both ways, and the gain of
    # the network over Kessler it gives: recorded, not required (on this GPU
    # plain Kessler is a memory-bound sweep and the network adds work to it,
    # so the ratio is below 1)
    from erf_amd.lib import lib, check
    cost = {}
    for name, call in (("kessler", lambda s: lib().erfx_kessler_advance(s.ctx, s.ptr(C.F_CONS), 1.0, None)),
                       ("network", lambda s: lib().erfx_kessler_mlp_advance(s.ctx, net.h, s.ptr(C.F_CONS),
                                                                            s.ptr(C.F_ZVEL), 1.0, None, None, None))):
        check(call(ka))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            check(call(ka))
        torch.cuda.synchronize()
        cost[name] = (time.perf_counter() - t0) / 20 * 1e3
        print("%s microphysics on %dx%dx%d: %.3f ms per call" % (name, cfg.nx, cfg.ny, cfg.nz, cost[name]))
    print("gain of the network over Kessler (cost per call): %.2fx (%s)"
          % (cost["kessler"] / cost["network"], "none" if cost["network"] >= cost["kessler"] else "faster"))
    ka.close()
    em.close()
