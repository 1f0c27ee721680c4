"""The hooks on Array4-shaped views (the _v entry points) against the pointer
hooks: every field held as ERF allocates it -- cons and prim with ng ghost
cells, momenta and velocities with ng - 1 and the extra face along their
direction, pi_stage and the eddy diffusivities with 1, the fast coefficients
with none -- each with its own padded row stride, so no view is in the
library's common layout.  The outputs over each view's extent are
bit-identical to the pointer hooks' on the same state; S_prev passed as the
same views as S_data (the reference's aliasing after the first substep)."""
import ctypes

import numpy as np
import pytest

from erf_amd import config as C
import test_gpu_hooks as H

pytestmark = pytest.mark.gpu


class ErfField:
    """One field of the box in an ERF-shaped allocation (a torch tensor on the
    GPU, rows padded by 3 elements), filled from the sim's common layout."""

    def __init__(self, sim, fid, ghost, nodal, ncomp):
        import torch
        from erf_amd.lib import View
        L = sim.L
        n = (L.nx, L.ny, L.nz)
        self.lo = [-ghost] * 3
        self.hi = [n[d] - 1 + (1 if nodal == d else 0) + ghost for d in range(3)]
        ext = [self.hi[d] - self.lo[d] + 1 for d in range(3)]
        self.ext = ext
        self.fid, self.ncomp, self.ng = fid, ncomp, L.ng
        full = torch.zeros((ncomp, ext[2], ext[1], ext[0] + 3), dtype=torch.float64, device="cuda")
        a = sim.download(fid)[:ncomp]
        full[:, :, :, :ext[0]] = torch.from_numpy(np.ascontiguousarray(self._region(a))).cuda()
        self.t = full
        v = View()
        v.p = full.data_ptr()
        for d in range(3):
            v.lo[d], v.hi[d] = self.lo[d], self.hi[d]
            v.ixtype[d] = 1 if nodal == d else 0
        v.js = ext[0] + 3
        v.ks = v.js * ext[1]
        v.ns = v.ks * ext[2]
        v.ncomp = ncomp
        self.v = v

    def _region(self, a):
        g = self.ng
        return a[:, self.lo[2] + g:self.hi[2] + g + 1, self.lo[1] + g:self.hi[1] + g + 1,
                 self.lo[0] + g:self.hi[0] + g + 1]

    def differs(self, sim_ref):
        """points of this view that differ from the same field of sim_ref:
        (on the valid region, on the ghosts, largest difference, first point)"""
        got = self.t[:, :, :, :self.ext[0]].cpu().numpy()
        want = self._region(sim_ref.download(self.fid)[:self.ncomp])
        # cells neither side ever writes hold NaN on both (the oracle's fill
        # value, copied into both sims): those agree
        d = (got != want) & ~(np.isnan(got) & np.isnan(want))
        if not d.any():
            return None
        g = [-self.lo[e] for e in range(3)]
        nv = [self.ext[e] - 2 * g[e] for e in range(3)]
        val = d[:, g[2]:g[2] + nv[2], g[1]:g[1] + nv[1], g[0]:g[0] + nv[0]]
        first = np.argwhere(d)[0]
        return (int(val.sum()), int(d.sum() - val.sum()), float(np.abs(got - want)[d].max()),
                [int(first[0]), int(first[1]) - g[2], int(first[2]) - g[1], int(first[3]) - g[0]])


def _fields(sim, cfg):
    ng = sim.L.ng
    f = {}

    def add(key, fid, ghost, nodal, ncomp):
        f[key] = ErfField(sim, fid, ghost, nodal, ncomp)

    for s in range(5):
        add((s, 0), C.state_id(s, 0), ng, -1, cfg.ncons)
        for c, d in ((1, 0), (2, 1), (3, 2)):
            add((s, c), C.state_id(s, c), ng - 1, d, 1)
    for key, fid, d in (("u", C.F_XVEL, 0), ("v", C.F_YVEL, 1), ("w", C.F_ZVEL, 2),
                        ("u_old", C.D_XVEL_OLD, 0), ("v_old", C.D_YVEL_OLD, 1), ("w_old", C.D_ZVEL_OLD, 2)):
        add(key, fid, ng - 1, d, 1)
    add("prim", C.D_PRIM, ng, -1, cfg.ncons - 1)
    add("pi", C.D_PI_STAGE, 1, -1, 1)
    add("fc", C.D_FAST_COEFFS, 0, 2, 5)
    if cfg.les_type != C.LES_NONE:
        add("eddy", C.D_EDDY, 1, -1, 13)
    import torch
    torch.cuda.synchronize()        # the copies ran on torch's stream, the hooks run on the sim's
    return f


def _sv(f, s):
    from erf_amd.lib import StateView
    return StateView(f[(s, 0)].v, f[(s, 1)].v, f[(s, 2)].v, f[(s, 3)].v)


PROBLEMS = []


def _check(f, ref, keys, what=""):
    """record the views that differ from the pointer hook's fields (points on
    the valid region, on the ghosts, largest difference, first point (n, k, j, i))"""
    bad = {k: f[k].differs(ref) for k in keys}
    bad = {k: v for k, v in bad.items() if v}
    if bad:
        PROBLEMS.append((what, bad))


def _state_keys(*states):
    return [(s, c) for s in states for c in range(4)]


@pytest.mark.parametrize("kind", ["thermal", "cbl", "moist1", "hill", "inflow_diff"])
def test_view_hooks_match_pointer_hooks(kind):
    from erf_amd.lib import lib, check, Sim
    print("view hooks: %s" % kind, flush=True)
    cfg, orc, ref, ids, dt = H._setup(kind)
    # the view side starts from the same state
    sim = Sim(cfg)
    for fid in ids:
        sim.upload(fid, orc.get(fid))
    f = _fields(sim, cfg)
    st = lambda s: ctypes.byref(ref.state(s))
    B = ctypes.byref
    del PROBLEMS[:]
    _check(f, ref, list(f), "initial state")

    # pre_update: cons_to_prim on S_new
    ng = ref.L.ng
    check(lib().erfx_cons_to_prim(ref.ctx, ref.ptr(C.state_id(C.S_NEW, 0)), ref.ptr(C.D_PRIM),
                                  ref.ptr(C.D_PI_STAGE), ng))
    check(lib().erfx_cons_to_prim_v(sim.ctx, B(f[(C.S_NEW, 0)].v), B(f["prim"].v), B(f["pi"].v), ng))
    ref.sync(); sim.sync()
    _check(f, ref, ["prim", "pi"], "cons_to_prim")

    # turbulent viscosity (LES)
    if cfg.les_type != C.LES_NONE:
        check(lib().erfx_turbulent_viscosity(ref.ctx, ref.ptr(C.state_id(C.S_OLD, 0)), ref.ptr(C.D_XVEL_OLD),
                                             ref.ptr(C.D_YVEL_OLD), ref.ptr(C.D_ZVEL_OLD), ref.ptr(C.D_EDDY), None))
        check(lib().erfx_turbulent_viscosity_v(sim.ctx, B(f[(C.S_OLD, 0)].v), B(f["u_old"].v), B(f["v_old"].v),
                                               B(f["w_old"].v), B(f["eddy"].v)))
        ref.sync(); sim.sync()
        _check(f, ref, ["eddy"], "turbulent_viscosity")
    eddy_p = ref.ptr(C.D_EDDY) if cfg.les_type != C.LES_NONE else None
    eddy_v = B(f["eddy"].v) if cfg.les_type != C.LES_NONE else None

    # slow_rhs_pre: F_slow from S_new, averaged momenta seeded in S_scratch
    check(lib().erfx_slow_rhs_pre(ref.ctx, 0, dt / 3, st(C.F_SLOW), st(C.S_NEW), ref.ptr(C.D_PRIM),
                                  st(C.S_SCRATCH), ref.ptr(C.F_XVEL), ref.ptr(C.F_YVEL), ref.ptr(C.F_ZVEL),
                                  eddy_p, None))
    fs, sn, sc = _sv(f, C.F_SLOW), _sv(f, C.S_NEW), _sv(f, C.S_SCRATCH)
    check(lib().erfx_slow_rhs_pre_v(sim.ctx, 0, dt / 3, B(fs), B(sn), B(f["prim"].v), B(sc), B(f["u"].v),
                                    B(f["v"].v), B(f["w"].v), eddy_v))
    ref.sync(); sim.sync()
    _check(f, ref, _state_keys(C.F_SLOW, C.S_SCRATCH), "slow_rhs_pre")

    # the fast coefficients and two substeps (the second with S_prev == S_data)
    dtau = dt / cfg.mri_ratio
    beta_s = 0.1 if cfg.implicit_substepping else -1.0
    check(lib().erfx_make_fast_coeffs(ref.ctx, ref.ptr(C.D_FAST_COEFFS), st(C.S_NEW), ref.ptr(C.D_PRIM),
                                      ref.ptr(C.D_PI_STAGE), dtau, beta_s))
    check(lib().erfx_make_fast_coeffs_v(sim.ctx, B(f["fc"].v), B(sn), B(f["prim"].v), B(f["pi"].v), dtau, beta_s))
    ref.sync(); sim.sync()
    _check(f, ref, ["fc"], "make_fast_coeffs")
    T = "T" if cfg.use_terrain else "N"
    so, ss = _sv(f, C.S_OLD), _sv(f, C.S_SUM)
    for step in (0, 1):
        nrk = 2
        prev_p = st(C.S_OLD) if step == 0 else st(C.S_SUM)
        prev_v = so if step == 0 else ss
        check(getattr(lib(), "erfx_fast_rhs_" + T)(ref.ctx, step, nrk, st(C.F_SLOW), prev_p, st(C.S_NEW),
                                                 ref.ptr(C.D_PRIM), ref.ptr(C.D_PI_STAGE), ref.ptr(C.D_FAST_COEFFS),
                                                 st(C.S_SUM), st(C.S_SCRATCH), dtau, beta_s, 1.0 / cfg.mri_ratio))
        check(getattr(lib(), "erfx_fast_rhs_%s_v" % T)(sim.ctx, step, nrk, B(fs), B(prev_v), B(sn), B(f["prim"].v),
                                                     B(f["pi"].v), B(f["fc"].v), B(ss), B(sc), dtau, beta_s,
                                                     1.0 / cfg.mri_ratio))
        ref.sync(); sim.sync()
        _check(f, ref, _state_keys(C.S_SUM, C.S_SCRATCH), "fast_rhs step %d" % step)
        # the apply_bcs between substeps
        check(lib().erfx_apply_bcs(ref.ctx, st(C.S_SUM), ref.ptr(C.F_XVEL), ref.ptr(C.F_YVEL), ref.ptr(C.F_ZVEL),
                                   1, 1, 1, 0))
        check(lib().erfx_apply_bcs_v(sim.ctx, B(ss), B(f["u"].v), B(f["v"].v), B(f["w"].v), 1, 1, 1, 0))
        ref.sync(); sim.sync()
        _check(f, ref, _state_keys(C.S_SUM) + ["u", "v", "w"], "apply_bcs after step %d" % step)

    # slow_rhs_post: the slow scalars, S_new <- S_sum
    check(lib().erfx_slow_rhs_post(ref.ctx, 2, dt, st(C.F_SLOW), st(C.S_OLD), st(C.S_NEW), st(C.S_SUM),
                                   ref.ptr(C.D_PRIM), st(C.S_SCRATCH), None, None, eddy_p))
    check(lib().erfx_slow_rhs_post_v(sim.ctx, 2, dt, B(fs), B(so), B(sn), B(ss), B(f["prim"].v), B(sc), eddy_v))
    ref.sync(); sim.sync()
    _check(f, ref, _state_keys(C.F_SLOW, C.S_NEW, C.S_SUM), "slow_rhs_post")

    # post_update on S_new with the full ghost widths
    check(lib().erfx_apply_bcs(ref.ctx, st(C.S_NEW), ref.ptr(C.F_XVEL), ref.ptr(C.F_YVEL), ref.ptr(C.F_ZVEL),
                               ng, ng - 1, 0, 0))
    check(lib().erfx_apply_bcs_v(sim.ctx, B(sn), B(f["u"].v), B(f["v"].v), B(f["w"].v), ng, ng - 1, 0, 0))
    ref.sync(); sim.sync()
    _check(f, ref, _state_keys(C.S_NEW) + ["u", "v", "w"], "apply_bcs post_update")

    # microphysics on the new cons
    if cfg.moisture_type != C.MOIST_NONE:
        check(lib().erfx_kessler_advance(ref.ctx, ref.ptr(C.state_id(C.S_NEW, 0)), dt, None))
        check(lib().erfx_kessler_advance_v(sim.ctx, B(f[(C.S_NEW, 0)].v), dt))
        ref.sync(); sim.sync()
        _check(f, ref, [(C.S_NEW, 0)], "kessler")
    sim.close()
    ref.close()
    assert not PROBLEMS, "\n".join("%s: %s" % p for p in PROBLEMS)


def test_view_hook_rejects_a_short_view_before_device_work():
    from erf_amd.lib import lib, Sim
    cfg, orc, ref, ids, dt = H._setup("thermal")
    f = _fields(ref, cfg)
    v = f["prim"].v
    v.lo[0] += 1                       # one ghost short on the x-lo side
    rc = lib().erfx_cons_to_prim_v(ref.ctx, ctypes.byref(f[(C.S_NEW, 0)].v), ctypes.byref(v),
                                   ctypes.byref(f["pi"].v), ref.L.ng)
    assert rc == -1 and b"does not cover" in lib().erfx_last_error()
    ref.close()
