"""Shared helpers for the parity tests (test infrastructure)."""
import ctypes

import numpy as np

from erf_amd import config as C

# every buffer id that the oracle and the sim share
ALL_IDS = ([C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL, C.F_R0, C.F_P0, C.F_PI0,
            C.D_PRIM, C.D_PI_STAGE, C.D_FAST_COEFFS, C.D_XVEL_OLD, C.D_YVEL_OLD, C.D_ZVEL_OLD] +
           [C.state_id(s, c) for s in range(5) for c in range(4)])


def valid(a, cfg, L, fid, grow=0):
    """Slice of the valid region (faces included for staggered fields)."""
    ng = L.ng
    ex = ey = ez = 0
    if fid in (C.F_XVEL, C.D_XVEL_OLD, C.F_Z_ND) or (fid >= C.D_STATE_BASE and (fid - C.D_STATE_BASE) % 4 == 1):
        ex = 1
    if fid in (C.F_YVEL, C.D_YVEL_OLD, C.F_Z_ND) or (fid >= C.D_STATE_BASE and (fid - C.D_STATE_BASE) % 4 == 2):
        ey = 1
    if fid in (C.F_ZVEL, C.D_ZVEL_OLD, C.F_Z_ND) or (fid >= C.D_STATE_BASE and (fid - C.D_STATE_BASE) % 4 == 3):
        ez = 1
    g = grow
    return a[:, ng - g:ng + cfg.nz + ez + g, ng - g:ng + cfg.ny + ey + g, ng - g:ng + cfg.nx + ex + g]


def rel_err(a, b):
    """max |a-b| / max |b| (0 when both vanish)."""
    d = np.abs(a - b)
    if not np.all(np.isfinite(d)):
        return np.inf
    s = np.abs(b).max()
    return float(d.max() / s) if s > 0 else float(d.max())


def comp_rel_err(a, b):
    """relative error per component (list)."""
    return [rel_err(a[c], b[c]) for c in range(a.shape[0])]
