"""AMReX plotfiles: writer, reader, the derived variables ERF plots and an
fcompare-style comparison.

The on-disk format is the one ERF writes through amrex::WriteMultiLevelPlotfile
(Source/IO/ERF_Plotfile.cpp:186-...) and that the reference's regression gold
files hold (Tests/ERFGoldFiles/<case>): a text `Header`, and for level 0 a text
`Level_0/Cell_H` naming the FAB files `Level_0/Cell_D_<n>`, each a one-line
FAB header followed by little-endian FP64 values, i fastest, one component
after another.  One level, one box per file here.

`derive` makes ERF's plot variables from the state (ERF_Plotfile.cpp:234-400,
ERF_Derive.cpp): face velocities averaged to cell centres, pressure from the
EOS, theta and scalars per unit mass, the base state; `fcompare` reports, per
variable, the largest absolute difference and that difference relative to the
largest magnitude of the reference, and fails a variable whose two errors both
exceed the tolerances, as AMReX's fcompare does in the regression suite
(Tests/CMakeLists.txt:3-4, rtol = atol = 2e-10).
"""
import os
import re

import numpy as np

from . import config as C

_FAB_REAL = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"


def _g17(x):
    return "%.17g" % float(x)


def write(path, names, data, time, step, prob_lo, prob_hi):
    """Write a single-level plotfile.  data: (nvar, nz, ny, nx) float64."""
    data = np.ascontiguousarray(data, dtype="<f8")
    nvar, nz, ny, nx = data.shape
    if len(names) != nvar:
        raise ValueError("one name per variable")
    lvl = os.path.join(path, "Level_0")
    os.makedirs(lvl, exist_ok=True)
    dx = [(prob_hi[d] - prob_lo[d]) / n for d, n in enumerate((nx, ny, nz))]
    box = "((0,0,0) (%d,%d,%d) (0,0,0))" % (nx - 1, ny - 1, nz - 1)
    h = ["HyperCLaw-V1.1", str(nvar)] + list(names)
    h += ["3", _g17(time), "0", " ".join(_g17(v) for v in prob_lo) + " ",
          " ".join(_g17(v) for v in prob_hi) + " ", "", box + " ", "%d " % step,
          " ".join(_g17(v) for v in dx) + " ", "0", "0", "0 1 %s" % _g17(time), str(step)]
    for d in range(3):
        h.append("%s %s" % (_g17(prob_lo[d]), _g17(prob_hi[d])))
    h.append("Level_0/Cell")
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("\n".join(h) + "\n")
    with open(os.path.join(lvl, "Cell_D_00000"), "wb") as f:
        f.write(("%s(%s %d\n" % (_FAB_REAL, box[1:], nvar)).encode())
        f.write(data.tobytes())
    mins = ",".join("%.16e" % v for v in data.reshape(nvar, -1).min(axis=1)) + ","
    maxs = ",".join("%.16e" % v for v in data.reshape(nvar, -1).max(axis=1)) + ","
    ch = ["1", "1", str(nvar), "0", "(1 0", box, ")", "1", "FabOnDisk: Cell_D_00000 0", "",
          "1,%d" % nvar, mins, "", "1,%d" % nvar, maxs, ""]
    with open(os.path.join(lvl, "Cell_H"), "w") as f:
        f.write("\n".join(ch))


def read(path):
    """Read a single-level plotfile: dict(names, time, lo, hi, n, nstep, data,
    cell_h) with data (nvar, nz, ny, nx) float64 (any number of boxes and FAB
    files; byte order as each FAB header states)."""
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    if not lines[0].startswith("HyperCLaw"):
        raise ValueError("not an AMReX plotfile: %s" % path)
    nvar = int(lines[1])
    names = lines[2:2 + nvar]
    pos = 2 + nvar
    time = float(lines[pos + 1])
    lo = [float(x) for x in lines[pos + 3].split()]
    hi = [float(x) for x in lines[pos + 4].split()]
    dom = re.findall(r"-?\d+", lines[pos + 6])
    n = [int(dom[3]) - int(dom[0]) + 1, int(dom[4]) - int(dom[1]) + 1, int(dom[5]) - int(dom[2]) + 1]
    nstep = int(lines[pos + 7])
    data = np.full((nvar, n[2], n[1], n[0]), np.nan)
    with open(os.path.join(path, "Level_0", "Cell_H")) as f:
        ch = f.read()
    boxes = re.findall(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\)", ch)
    fabs = re.findall(r"FabOnDisk: (\S+) (\d+)", ch)
    for (b, (fname, off)) in zip(boxes, fabs):
        b = [int(x) for x in b]
        with open(os.path.join(path, "Level_0", fname), "rb") as f:
            f.seek(int(off))
            head = f.readline().decode()
            m = re.match(r"FAB \(\(8, \(([\d ]+)\)\),\(8, \(([\d ]+)\)\)\)\(\((-?\d+),(-?\d+),(-?\d+)\) "
                         r"\((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\) (\d+)", head)
            if not m:
                raise ValueError("bad FAB header in %s: %r" % (fname, head))
            order = m.group(2).split()
            dt = np.dtype("<f8") if order == list("87654321") else np.dtype(">f8")
            nc = int(m.group(12))
            sh = (b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)
            arr = np.frombuffer(f.read(8 * nc * sh[0] * sh[1] * sh[2]), dtype=dt).reshape((nc,) + sh)
        data[:, b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] = arr
    return dict(names=names, time=time, lo=lo, hi=hi, n=n, nstep=nstep, data=data, cell_h=ch)


def derive(cfg, cons, u, v, w, r0=None, p0=None):
    """ERF's plot variables from valid-region arrays: cons (ncons, nz, ny, nx),
    u (nz, ny, nx+1), v (nz, ny+1, nx), w (nz+1, ny, nx); r0, p0 the base state
    on cells.  Returns {name: (nz, ny, nx) array} (ERF_Plotfile.cpp:234-400)."""
    rho, rt = cons[C.RHO], cons[C.RHOTHETA]
    moist = cfg.moisture_type != C.MOIST_NONE
    qv = cons[C.RHOQ1] / rho if moist else np.zeros_like(rho)
    p = C.p_0 * np.power(C.R_d * rt * (1.0 + C.R_v / C.R_d * qv) / C.p_0, C.Gamma)
    out = {
        "density": rho, "rhotheta": rt, "rhoadv_0": cons[C.RHOSCALAR], "scalar": cons[C.RHOSCALAR] / rho,
        "x_velocity": 0.5 * (u[:, :, :-1] + u[:, :, 1:]),
        "y_velocity": 0.5 * (v[:, :-1, :] + v[:, 1:, :]),
        "z_velocity": 0.5 * (w[:-1] + w[1:]),
        "theta": rt / rho, "pressure": p, "temp": p / (C.R_d * rho * (1.0 + C.R_v / C.R_d * qv)),
    }
    if moist:
        out.update(qv=qv, qc=cons[C.RHOQ2] / rho, qrain=cons[C.RHOQ3] / rho)
    if r0 is not None:
        out["dens_hse"] = r0
    if p0 is not None:
        out["pres_hse"] = p0
    return out


def fcompare(a, b, rtol=2e-10, atol=2e-10):
    """Compare plotfile dicts a (tested) and b (reference) variable by variable:
    {name: (abs_err, rel_err, ok)}; a variable fails when both errors exceed
    their tolerances.  Variables of b missing from a fail."""
    out = {}
    av = dict(zip(a["names"], a["data"]))
    for name, ref in zip(b["names"], b["data"]):
        if name not in av:
            out[name] = (np.inf, np.inf, False)
            continue
        d = float(np.abs(av[name] - ref).max())
        scale = float(np.abs(ref).max())
        r = d / scale if scale > 0 else d
        out[name] = (d, r, not (d > atol and r > rtol))
    return out
