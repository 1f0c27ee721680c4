"""CPU checks of the C ABI (no GPU calls): the library loads, exports every
function include/erfx.h declares, the ctypes mirrors match the C structs byte
for byte, and the device-free entry points (defaults, ghost width, layout)
give ERF's values."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from erf_amd import config as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "erfx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(erfx_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from erf_amd.lib import lib, SIGNATURES
    names = declared_functions()
    assert len(names) >= 25
    l = lib()
    missing = [n for n in names if not hasattr(l, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared set
    assert sorted(SIGNATURES) == names


def _c_layout():
    """sizeof/offsetof of the ABI structs as a C compiler sees them."""
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "erfx.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
    printf("erfx_config %zu\n", sizeof(erfx_config));
    printf("erfx_layout %zu\n", sizeof(erfx_layout));
    printf("erfx_state %zu\n", sizeof(erfx_state));
    F(erfx_config, dx) F(erfx_config, prob_lo) F(erfx_config, bc) F(erfx_config, ncons)
    F(erfx_config, dycore_horiz_upw_frac) F(erfx_config, molec_diff_type) F(erfx_config, les_type)
    F(erfx_config, Cs) F(erfx_config, surface_heat_flux) F(erfx_config, c_p)
    F(erfx_config, abl_geo_forcing) F(erfx_config, mri_ratio) F(erfx_config, implicit_substepping)
    F(erfx_config, inflow_qc) F(erfx_config, prob_hi) F(erfx_config, no_substepping)
    F(erfx_config, wall_velocity) F(erfx_config, abl_pressure_grad)
    F(erfx_layout, sy) F(erfx_layout, nelem)
    printf("erfx_view %zu\n", sizeof(erfx_view));
    F(erfx_view, lo) F(erfx_view, hi) F(erfx_view, js) F(erfx_view, ns) F(erfx_view, ncomp) F(erfx_view, ixtype)
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).decode()
    return dict((a, int(b)) for a, b in (line.split() for line in out.strip().splitlines()))


def test_ctypes_structs_match_c_header():
    lay = _c_layout()
    assert lay["erfx_config"] == ctypes.sizeof(C.Config)
    assert lay["erfx_layout"] == ctypes.sizeof(C.Layout)
    from erf_amd.lib import View
    assert lay["erfx_view"] == ctypes.sizeof(View)
    for key, off in lay.items():
        if "." not in key:
            continue
        T, m = key.split(".")
        cls = {"erfx_config": C.Config, "erfx_layout": C.Layout, "erfx_view": View}[T]
        assert getattr(cls, m).offset == off, key


def test_defaults_ghosts_and_layout_without_gpu():
    from erf_amd.lib import lib
    cfg = C.Config()
    lib().erfx_config_defaults(ctypes.byref(cfg))
    # ERF defaults (ERF_AdvStruct.H:281-284, ERF_DataStruct.H, ERF_TurbStruct.H:175)
    assert cfg.dycore_horiz_adv == C.ADV_UPWIND3 and cfg.dryscal_vert_adv == C.ADV_UPWIND3
    assert cfg.implicit_substepping == 1 and cfg.force_stage1_single_substep == 1
    assert cfg.c_p == 1004.5 and abs(cfg.Pr_t - 1.0 / 3.0) < 1e-16
    assert lib().erfx_ghost_cells(ctypes.byref(cfg)) == 3
    cfg.dycore_horiz_adv = C.ADV_UPWIND5
    assert lib().erfx_ghost_cells(ctypes.byref(cfg)) == 4
    cfg.nx, cfg.ny, cfg.nz = 256, 256, 128
    L = C.Layout()
    assert lib().erfx_layout_for(ctypes.byref(cfg), ctypes.byref(L)) == 0
    assert (L.nx, L.ny, L.nz, L.ng) == (256, 256, 128, 4)
    assert L.sy % 16 == 0 and L.sy >= 256 + 1 + 8
    assert L.sz == L.sy * (256 + 1 + 8)
    assert L.nelem == L.sz * (128 + 1 + 8)
    # the Python mirror agrees
    L2 = C.layout_for(cfg)
    assert (L2.sy, L2.sz, L2.nelem) == (L.sy, L.sz, L.nelem)


def test_null_arguments_are_rejected_without_touching_the_device():
    from erf_amd.lib import lib
    assert lib().erfx_layout_for(None, None) == -1
    assert lib().erfx_ctx_create(None, None, None) == -1
    assert b"null" in lib().erfx_last_error()
    assert lib().erfx_sim_create(None, None) == -1
    assert lib().erfx_ghost_cells(None) == -1


def _view(lo, hi, ncomp, ixtype, p=1):
    from erf_amd.lib import View
    v = View()
    v.p = p
    for d in range(3):
        v.lo[d], v.hi[d], v.ixtype[d] = lo[d], hi[d], ixtype[d]
    nx, ny, nz = (hi[d] - lo[d] + 1 for d in range(3))
    v.js, v.ks, v.ns = nx, nx * ny, nx * ny * nz
    v.ncomp = ncomp
    return v


def test_view_checks_without_gpu():
    """erfx_check_view: ERF's own allocation of each field passes (cons with
    ng ghosts, momenta with ng - 1 and one more face along their direction);
    too few ghosts, a wrong index type, too few components, a missing pointer
    or overlapping strides give ERFX_ERR_ARG with the reason."""
    from erf_amd.lib import lib
    cfg = C.default_config(16, 12, 10, 100.0, 100.0, 100.0)
    ng = lib().erfx_ghost_cells(ctypes.byref(cfg))
    n = (16, 12, 10)
    cons = _view([-ng] * 3, [n[d] - 1 + ng for d in range(3)], cfg.ncons, [0, 0, 0])
    assert lib().erfx_check_view(ctypes.byref(cfg), 0, ctypes.byref(cons)) == 0
    g = ng - 1
    for role, d in ((1, 0), (2, 1), (3, 2), (4, 0), (5, 1), (6, 2)):
        ix = [1 if e == d else 0 for e in range(3)]
        mom = _view([-g] * 3, [n[e] - 1 + ix[e] + g for e in range(3)], 1, ix)
        assert lib().erfx_check_view(ctypes.byref(cfg), role, ctypes.byref(mom)) == 0, role
        short = _view([-g] * 3, [n[e] - 1 + g for e in range(3)], 1, ix)       # the extra face missing
        assert lib().erfx_check_view(ctypes.byref(cfg), role, ctypes.byref(short)) == -1
        assert b"does not cover" in lib().erfx_last_error()
    thin = _view([-(ng - 1)] * 3, [n[d] - 1 + ng - 1 for d in range(3)], cfg.ncons, [0, 0, 0])
    assert lib().erfx_check_view(ctypes.byref(cfg), 0, ctypes.byref(thin)) == -1
    wrong = _view([-ng] * 3, [n[d] + ng for d in range(3)], cfg.ncons, [1, 0, 0])
    assert lib().erfx_check_view(ctypes.byref(cfg), 0, ctypes.byref(wrong)) == -1
    assert b"index type" in lib().erfx_last_error()
    few = _view([-ng] * 3, [n[d] - 1 + ng for d in range(3)], cfg.ncons - 1, [0, 0, 0])
    assert lib().erfx_check_view(ctypes.byref(cfg), 0, ctypes.byref(few)) == -1
    assert b"components" in lib().erfx_last_error()
    cons.js = 3
    assert lib().erfx_check_view(ctypes.byref(cfg), 0, ctypes.byref(cons)) == -1
    assert b"strides" in lib().erfx_last_error()
    cons.p = None
    assert lib().erfx_check_view(ctypes.byref(cfg), 0, ctypes.byref(cons)) == -1
