"""ctypes binding of the HIP dycore library (erf_amd/liberfx.so, include/erfx.h).

The library is the product: every call runs hand-written gfx950 kernels.  There
is no CPU fallback; loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

from . import config as C

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liberfx.so")
_lib = None

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int


class State(ctypes.Structure):
    """erfx_state: device pointers {cons, xmom, ymom, zmom}."""
    _fields_ = [("cons", _P), ("xmom", _P), ("ymom", _P), ("zmom", _P)]


class View(ctypes.Structure):
    """erfx_view: one field of one box as ERF allocates it (an Array4)."""
    _fields_ = [("p", _P), ("lo", _I * 3), ("hi", _I * 3), ("js", ctypes.c_long), ("ks", ctypes.c_long),
                ("ns", ctypes.c_long), ("ncomp", _I), ("ixtype", _I * 3)]


class StateView(ctypes.Structure):
    """erfx_state_view: views of {cons, xmom, ymom, zmom}."""
    _fields_ = [("cons", View), ("xmom", View), ("ymom", View), ("zmom", View)]


_V = ctypes.POINTER(View)
_SV = ctypes.POINTER(StateView)

# every symbol include/erfx.h declares, with its ctypes signature
SIGNATURES = {
    "erfx_check_view": (_I, [ctypes.POINTER(C.Config), _I, _V]),
    "erfx_cons_to_prim_v": (_I, [_P, _V, _V, _V, _I]),
    "erfx_apply_bcs_v": (_I, [_P, _SV, _V, _V, _V, _I, _I, _I, _I]),
    "erfx_make_fast_coeffs_v": (_I, [_P, _V, _SV, _V, _V, _D, _D]),
    "erfx_fast_rhs_N_v": (_I, [_P, _I, _I, _SV, _SV, _SV, _V, _V, _V, _SV, _SV, _D, _D, _D]),
    "erfx_fast_rhs_T_v": (_I, [_P, _I, _I, _SV, _SV, _SV, _V, _V, _V, _SV, _SV, _D, _D, _D]),
    "erfx_turbulent_viscosity_v": (_I, [_P, _V, _V, _V, _V, _V]),
    "erfx_slow_rhs_pre_v": (_I, [_P, _I, _D, _SV, _SV, _V, _SV, _V, _V, _V, _V]),
    "erfx_slow_rhs_post_v": (_I, [_P, _I, _D, _SV, _SV, _SV, _SV, _V, _SV, _V]),
    "erfx_kessler_advance_v": (_I, [_P, _V, _D]),
    "erfx_last_error": (ctypes.c_char_p, []),
    "erfx_config_defaults": (None, [ctypes.POINTER(C.Config)]),
    "erfx_ghost_cells": (_I, [ctypes.POINTER(C.Config)]),
    "erfx_layout_for": (_I, [ctypes.POINTER(C.Config), ctypes.POINTER(C.Layout)]),
    "erfx_ctx_create": (_I, [ctypes.POINTER(C.Config), _P, ctypes.POINTER(_P)]),
    "erfx_ctx_destroy": (_I, [_P]),
    "erfx_ctx_layout": (_I, [_P, ctypes.POINTER(C.Layout)]),
    "erfx_set_base_state": (_I, [_P, _P, _P, _P]),
    "erfx_set_rayleigh": (_I, [_P, _P, _P, _P, _P]),
    "erfx_set_terrain": (_I, [_P, _P]),
    "erfx_cons_to_prim": (_I, [_P, _P, _P, _P, _I]),
    "erfx_apply_bcs": (_I, [_P, ctypes.POINTER(State), _P, _P, _P, _I, _I, _I, _I]),
    "erfx_make_fast_coeffs": (_I, [_P, _P, ctypes.POINTER(State), _P, _P, _D, _D]),
    "erfx_fast_rhs_N": (_I, [_P, _I, _I, ctypes.POINTER(State), ctypes.POINTER(State), ctypes.POINTER(State),
                             _P, _P, _P, ctypes.POINTER(State), ctypes.POINTER(State), _D, _D, _D]),
    "erfx_fast_rhs_N_noalloc": (_I, [ctypes.POINTER(C.Config), _P, _I, _I, ctypes.POINTER(State),
                                     ctypes.POINTER(State), ctypes.POINTER(State), _P, _P, _P,
                                     ctypes.POINTER(State), ctypes.POINTER(State), _D, _D, _D]),
    "erfx_fast_rhs_T": (_I, [_P, _I, _I, ctypes.POINTER(State), ctypes.POINTER(State), ctypes.POINTER(State),
                             _P, _P, _P, ctypes.POINTER(State), ctypes.POINTER(State), _D, _D, _D]),
    "erfx_turbulent_viscosity": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "erfx_slow_rhs_pre": (_I, [_P, _I, _D, ctypes.POINTER(State), ctypes.POINTER(State), _P,
                               ctypes.POINTER(State), _P, _P, _P, _P, _P]),
    "erfx_slow_rhs_post": (_I, [_P, _I, _D, ctypes.POINTER(State), ctypes.POINTER(State), ctypes.POINTER(State),
                                ctypes.POINTER(State), _P, ctypes.POINTER(State), _P, _P, _P]),
    "erfx_kessler_advance": (_I, [_P, _P, _D, _P]),
    "erfx_mlp_create": (_I, [_I, ctypes.POINTER(_P)]),
    "erfx_mlp_destroy": (_I, [_P]),
    "erfx_mlp_set_params": (_I, [_P, _P]),
    "erfx_mlp_get_params": (_I, [_P, _P]),
    "erfx_mlp_set_scaling": (_I, [_P, _P, _P, _P, _P, _P]),
    "erfx_mlp_forward": (_I, [_P, _P, _P, _I]),
    "erfx_mlp_gradient": (_I, [_P, _P, _P, _I, _P, ctypes.POINTER(_D)]),
    "erfx_mlp_train_step": (_I, [_P, _P, _P, _I, _D, ctypes.POINTER(_D)]),
    "erfx_kessler_samples": (_I, [_P, _P, _P, _D, _P, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
    "erfx_kessler_mlp_advance": (_I, [_P, _P, _P, _P, _D, _P, ctypes.POINTER(ctypes.c_long),
                                      ctypes.POINTER(ctypes.c_long)]),
    "erfx_sim_set_mlp": (_I, [_P, _P]),
    "erfx_sim_mlp_counts": (_I, [_P, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
    "erfx_sim_set_sampling": (_I, [_P, _P, ctypes.c_long]),
    "erfx_sim_sample_count": (_I, [_P, ctypes.POINTER(ctypes.c_long)]),
    "erfx_sim_create": (_I, [ctypes.POINTER(C.Config), ctypes.POINTER(_P)]),
    "erfx_sim_destroy": (_I, [_P]),
    "erfx_sim_layout": (_I, [_P, ctypes.POINTER(C.Layout)]),
    "erfx_sim_field_size": (ctypes.c_long, [_P, _I]),
    "erfx_sim_field_ptr": (_P, [_P, _I]),
    "erfx_sim_ctx": (_P, [_P]),
    "erfx_sim_upload": (_I, [_P, _I, _P]),
    "erfx_sim_download": (_I, [_P, _I, _P]),
    "erfx_sim_step": (_I, [_P, _D, _D, _I]),
    "erfx_sim_step_async": (_I, [_P, _D, _D, _I]),
    "erfx_sim_sync": (_I, [_P]),
    "erfx_sim_set_graph": (_I, [_P, _I]),
    "erfx_sim_autotune": (_I, [_P, _I]),
    "erfx_sim_column_variant": (_I, [_P, ctypes.POINTER(_I), _P]),
    "erfx_sim_set_column_variant": (_I, [_P, _I]),
    "erfx_sim_stream": (_P, [_P]),
    "erfx_sim_est_dt": (_I, [_P, ctypes.POINTER(C.DtControl), ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "erfx_sim_run": (_I, [_P, _D, _I, ctypes.POINTER(C.DtControl), ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "erfx_sim_integrals": (_I, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "erfx_sim_timer_reset": (_I, [_P]),
    "erfx_sim_timer_stop": (_I, [_P]),
    "erfx_sim_write_checkpoint": (_I, [_P, ctypes.c_char_p, _D, _I]),
    "erfx_sim_read_checkpoint": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "erfx_sim_timer_read": (_I, [_P, ctypes.c_char_p, _I, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_long),
                                 ctypes.POINTER(_D)]),
    "erfx_ensemble_create": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "erfx_ensemble_destroy": (_I, [_P]),
    "erfx_ensemble_step": (_I, [_P, _D, _D, _I]),
    "erfx_ensemble_set_graph": (_I, [_P, _I]),
    "erfx_ensemble_set_streams": (_I, [_P, _I]),
    "erfx_ensemble_autotune": (_I, [_P, _I]),
    "erfx_ensemble_set_column_variant": (_I, [_P, _I]),
    "erfx_ensemble_column_variant": (_I, [_P, ctypes.POINTER(_I), _P]),
    "erfx_ensemble_info": (_I, [_P, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
}


def lib():
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch ships its own libamdhip64, and a
        # process that loads the system one first (through this library) leaves
        # torch without a device.  Importing torch first makes the library bind
        # to the runtime torch loaded, so both share buffers and streams.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("erf_amd/liberfx.so is not built: run __graft_entry__.build() "
                               "(make -C erf_amd/csrc)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class ErfxError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise ErfxError("erfx error %d: %s" % (rc, lib().erfx_last_error().decode()))


class Sim:
    """Whole-step driver owning every buffer in HBM (erfx_sim)."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = _P()
        check(lib().erfx_sim_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.L = C.Layout()
        check(lib().erfx_sim_layout(self.h, ctypes.byref(self.L)))

    def close(self):
        if getattr(self, "h", None):
            lib().erfx_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, fid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        n = lib().erfx_sim_field_size(self.h, fid)
        if a.size != n:
            raise ValueError("field %d: %d values given, %d expected" % (fid, a.size, n))
        check(lib().erfx_sim_upload(self.h, fid, a.ctypes.data))

    def download(self, fid):
        n = lib().erfx_sim_field_size(self.h, fid)
        if n < 0:
            raise ValueError("unknown field %d" % fid)
        if fid == C.D_RAIN_ACCUM:      # one k = 0 plane: (j, i) with ghosts
            a = np.empty((self.L.sz // self.L.sy, self.L.sy))
        else:
            a = np.empty((n // self.L.nelem,) + self.L.shape3())
        check(lib().erfx_sim_download(self.h, fid, a.ctypes.data))
        return a

    def load(self, fields):
        for fid, a in fields.items():
            self.upload(fid, a)

    def step(self, t0, dt, n):
        check(lib().erfx_sim_step(self.h, t0, dt, n))

    def run(self, t0, max_steps, dtc):
        """ERF's time loop with its dt control; returns (t_end, steps_done)."""
        t = _D()
        n = _I()
        check(lib().erfx_sim_run(self.h, t0, max_steps, ctypes.byref(dtc), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def integrals(self):
        """(mass, rho*theta, scalar) volume sums, exactly rounded."""
        m, r, q = _D(), _D(), _D()
        check(lib().erfx_sim_integrals(self.h, ctypes.byref(m), ctypes.byref(r), ctypes.byref(q)))
        return m.value, r.value, q.value

    def est_dt(self, dtc):
        dt = _D()
        r = _I()
        check(lib().erfx_sim_est_dt(self.h, ctypes.byref(dtc), ctypes.byref(dt), ctypes.byref(r)))
        return dt.value, r.value

    def write_checkpoint(self, path, time, step):
        """The persistent state into directory `path` (erfx_sim_write_checkpoint)."""
        check(lib().erfx_sim_write_checkpoint(self.h, path.encode(), time, step))

    def read_checkpoint(self, path):
        """Restore the state written by write_checkpoint; returns (time, step)."""
        t, n = _D(), _I()
        check(lib().erfx_sim_read_checkpoint(self.h, path.encode(), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def step_async(self, t0, dt, n):
        check(lib().erfx_sim_step_async(self.h, t0, dt, n))

    def sync(self):
        check(lib().erfx_sim_sync(self.h))

    def autotune(self, on=True):
        """Time the acoustic column kernel's variants on the next launches and keep the fastest."""
        check(lib().erfx_sim_autotune(self.h, 1 if on else 0))

    def column_variant(self):
        """(variant in use or -1 while timing, best ms per variant)."""
        v = _I()
        ms = (ctypes.c_float * C.COL_VARIANTS)()
        check(lib().erfx_sim_column_variant(self.h, ctypes.byref(v), ms))
        return v.value, list(ms)

    def set_column_variant(self, v):
        check(lib().erfx_sim_set_column_variant(self.h, int(v)))

    def set_graph(self, on=True):
        """Replay the step sequence from a captured hipGraph (two steps per graph)."""
        check(lib().erfx_sim_set_graph(self.h, 1 if on else 0))

    def ptr(self, fid):
        return lib().erfx_sim_field_ptr(self.h, fid)

    def state(self, s):
        """erfx_state of integrator buffer s (config.S_OLD ... config.F_SLOW)."""
        return State(*(self.ptr(C.state_id(s, c)) for c in range(4)))

    def set_rayleigh(self, ubar, vbar, wbar, thetabar):
        """Rayleigh reference profiles (nz+1 values each) on the sim's context."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (ubar, vbar, wbar, thetabar)]
        for a in arrs:
            if a.size != self.cfg.nz + 1:
                raise ValueError("Rayleigh profiles need nz+1 values")
        check(lib().erfx_set_rayleigh(self.ctx, *(a.ctypes.data for a in arrs)))

    def set_mlp(self, net):
        """Microphysics by the learned stand-in (an MLP), or Kessler for None."""
        self._mlp = net
        check(lib().erfx_sim_set_mlp(self.h, net.h if net is not None else None))

    def mlp_counts(self):
        a, b = ctypes.c_long(), ctypes.c_long()
        check(lib().erfx_sim_mlp_counts(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_sampling(self, rows_dev, max_rows):
        check(lib().erfx_sim_set_sampling(self.h, rows_dev, int(max_rows)))

    def sample_count(self):
        n = ctypes.c_long()
        check(lib().erfx_sim_sample_count(self.h, ctypes.byref(n)))
        return n.value

    @property
    def ctx(self):
        return lib().erfx_sim_ctx(self.h)

    @property
    def stream(self):
        return lib().erfx_sim_stream(self.h)


class Ensemble:
    """Sims advanced together, every kernel of the step launched once for all
    members (erfx_ensemble).  The members stay owned by the caller."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (_P * len(self.sims))(*[s.h for s in self.sims])
        h = _P()
        check(lib().erfx_ensemble_create(arr, len(self.sims), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().erfx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, t0, dt, n):
        check(lib().erfx_ensemble_step(self.h, t0, dt, n))

    def set_graph(self, on=True):
        check(lib().erfx_ensemble_set_graph(self.h, 1 if on else 0))

    def set_streams(self, n):
        check(lib().erfx_ensemble_set_streams(self.h, int(n)))

    def autotune(self, on=True):
        check(lib().erfx_ensemble_autotune(self.h, 1 if on else 0))

    def set_column_variant(self, v):
        check(lib().erfx_ensemble_set_column_variant(self.h, int(v)))

    def column_variant(self):
        v = _I()
        ms = (ctypes.c_float * C.COL_VARIANTS)()
        check(lib().erfx_ensemble_column_variant(self.h, ctypes.byref(v), ms))
        return v.value, list(ms)

    def info(self):
        """(kernel launches per step, total cells of the members)"""
        a, b = ctypes.c_long(), ctypes.c_long()
        check(lib().erfx_ensemble_info(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class MLP:
    """The learned Kessler stand-in (erfx_mlp): 8 -> 128 -> 128 -> 4, GELU,
    BF16 weights and activations with FP32 accumulation, trained on the GPU."""
    NPARAM = 18180

    def __init__(self, capacity):
        h = _P()
        check(lib().erfx_mlp_create(int(capacity), ctypes.byref(h)))
        self.h = h
        self.capacity = int(capacity)

    def close(self):
        if getattr(self, "h", None):
            lib().erfx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.NPARAM
        check(lib().erfx_mlp_set_params(self.h, a.ctypes.data))

    def get_params(self):
        a = np.empty(self.NPARAM, dtype=np.float32)
        check(lib().erfx_mlp_get_params(self.h, a.ctypes.data))
        return a

    def set_scaling(self, x_mean, x_scale, x_lo, x_hi, y_scale):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x_mean, x_scale, x_lo, x_hi, y_scale)]
        self._scaling = arrs
        check(lib().erfx_mlp_set_scaling(self.h, *(a.ctypes.data for a in arrs)))

    def forward(self, x_dev, y_dev, n):
        """x_dev, y_dev: device pointers (e.g. torch tensor data_ptr())."""
        check(lib().erfx_mlp_forward(self.h, x_dev, y_dev, int(n)))

    def gradient(self, x_dev, t_dev, n):
        g = np.empty(self.NPARAM, dtype=np.float32)
        loss = _D()
        check(lib().erfx_mlp_gradient(self.h, x_dev, t_dev, int(n), g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, x_dev, t_dev, n, lr, want_loss=True):
        loss = _D()
        check(lib().erfx_mlp_train_step(self.h, x_dev, t_dev, int(n), float(lr),
                                        ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None
