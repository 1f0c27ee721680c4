"""Training driver of the learned Kessler stand-in (erf_amd/csrc/erfx_mlp.hip).

The network is trained on the GPU by the library's own forward, backward and
Adam kernels (`MLP.train_step`) on rows of Kessler's own local tendencies,
harvested from moist runs of the library's dycore (`Sim.set_sampling`).
Feature scaling: inputs are centred and divided by their standard deviation;
each output is divided by its standard deviation.  The raw training range of
every input is kept: outside it the step falls back to Kessler.
"""
import time

import numpy as np

from . import config as C, problems as P
from .lib import MLP, Sim

NI, NO = 8, 4


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the emulator needs a GPU")
    return torch


def harvest(seeds=(0, 1), nsteps=20, shape=(32, 32, 24), dt=1.0, max_rows=2_000_000, case=None, every=1,
            per_step=None, log=None):
    """Rows (n, 12) of (T, p, rho, qv, qc, qr, w, dt, dqv, dqc, dqr, dtheta) from
    runs with Kessler, one run per seed: case(seed) -> (cfg, fields), the moist
    bubble of `shape` by default; rows from every `every`-th step, at most
    max_rows per run.  With per_step, each harvested step contributes at most
    that many of its rows, drawn uniformly at random on the GPU (seeded by the
    run's seed), so a long run's rows cover its whole time span.  log(msg), if
    given, is called after each run."""
    torch = _torch()
    out = []
    for seed in seeds:
        t0 = time.perf_counter()
        cfg, f = case(seed) if case else P.moist_bubble(*shape, seed=seed)
        sim = Sim(cfg)
        sim.load(f)
        del f
        buf = torch.zeros((max_rows, 12), dtype=torch.float32, device="cuda")
        if per_step is not None:
            cells = cfg.nx * cfg.ny * cfg.nz
            tmp = torch.zeros((cells, 12), dtype=torch.float32, device="cuda")
            g = torch.Generator(device="cuda")
            g.manual_seed(1000 + seed)
        torch.cuda.synchronize()
        off = 0
        for n in range(nsteps):
            take = (n + 1) % every == 0 and off < max_rows
            if take:
                if per_step is None:
                    sim.set_sampling(buf.data_ptr() + off * 12 * 4, max_rows - off)
                else:
                    sim.set_sampling(tmp.data_ptr(), cells)
            sim.step(n * dt, dt, 1)
            if take:
                cnt = sim.sample_count()
                sim.set_sampling(None, 0)
                if per_step is None:
                    off += min(cnt, max_rows - off)
                else:
                    k = min(cnt, per_step, max_rows - off)
                    if k < cnt:
                        idx = torch.randperm(cnt, device="cuda", generator=g)[:k]
                        buf[off:off + k] = tmp[idx]
                    else:
                        buf[off:off + k] = tmp[:k]
                    off += k
                    torch.cuda.synchronize()   # the gather reads tmp before the sim's stream writes it again
        out.append(buf[:off].cpu().numpy().astype(np.float64))
        sim.close()
        del buf
        if per_step is not None:
            del tmp
        if log:
            log("seed %d: %d rows from %d steps (%.0f s)" % (seed, off, nsteps, time.perf_counter() - t0))
    return np.concatenate(out, axis=0)


class Scaling:
    def __init__(self, rows):
        x, y = rows[:, :NI], rows[:, NI:]
        self.x_mean = x.mean(axis=0)
        self.x_scale = np.maximum(x.std(axis=0), 1e-12 * np.maximum(np.abs(self.x_mean), 1e-300))
        self.x_scale = np.where(self.x_scale > 0, self.x_scale, 1.0)
        self.x_lo = x.min(axis=0)
        self.x_hi = x.max(axis=0)
        s = y.std(axis=0)
        self.y_scale = np.where(s > 0, s, 1.0)

    def x(self, rows):
        return ((rows[:, :NI] - self.x_mean) / self.x_scale).astype(np.float32)

    def y(self, rows):
        return (rows[:, NI:] / self.y_scale).astype(np.float32)

    def apply(self, net):
        net.set_scaling(self.x_mean, self.x_scale, self.x_lo, self.x_hi, self.y_scale)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (parameter layout of include/erfx.h:
    W1 [8][128], b1, W2 [128][128], b2, W3 [128][4], b3)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP.NPARAM)
    off = 0
    for fi, fo in ((NI, 128), (128, 128), (128, NO)):
        lim = np.sqrt(6.0 / (fi + fo))
        p[off:off + fi * fo] = rng.uniform(-lim, lim, fi * fo)
        off += fi * fo + fo          # weights, then the biases (left at 0)
    assert off == MLP.NPARAM
    return p


def fit(rows, steps=2000, batch=4096, lr=2e-3, seed=0, scaling=None, log_every=0):
    """Train a fresh network on `rows` with Adam minibatches drawn on the host;
    returns (net, scaling, losses)."""
    torch = _torch()
    sc = scaling or Scaling(rows)
    X = torch.from_numpy(sc.x(rows)).cuda()
    Y = torch.from_numpy(sc.y(rows)).cuda()
    n = X.shape[0]
    batch = min(batch, n)
    net = MLP(batch)
    net.set_params(init_params(seed))
    sc.apply(net)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    losses = []
    xb = torch.empty((batch, NI), dtype=torch.float32, device="cuda")
    yb = torch.empty((batch, NO), dtype=torch.float32, device="cuda")
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), device="cuda", generator=g)
        torch.index_select(X, 0, idx, out=xb)
        torch.index_select(Y, 0, idx, out=yb)
        torch.cuda.synchronize()
        want = log_every and (it % log_every == 0 or it == steps - 1)
        l = net.train_step(xb.data_ptr(), yb.data_ptr(), batch, lr * (0.1 ** (it / steps)), want_loss=bool(want))
        if want:
            losses.append((it, l))
    torch.cuda.synchronize()
    return net, sc, losses


def mse(net, sc, rows, chunk=100_000):
    """Mean squared error of the scaled outputs over `rows`."""
    torch = _torch()
    tot, cnt = 0.0, 0
    for a in range(0, rows.shape[0], chunk):
        r = rows[a:a + chunk]
        x = torch.from_numpy(sc.x(r)).cuda()
        y = torch.zeros((r.shape[0], NO), dtype=torch.float32, device="cuda")
        net.forward(x.data_ptr(), y.data_ptr(), r.shape[0])
        torch.cuda.synchronize()
        d = y.cpu().numpy().astype(np.float64) - sc.y(r)
        tot += float((d * d).sum())
        cnt += d.size
    return tot / max(cnt, 1)
