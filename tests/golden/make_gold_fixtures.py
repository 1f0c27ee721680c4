"""Extract gold plotfiles of the reference's regression tests
(Tests/ERFGoldFiles/<case>, AMReX plotfiles) into small fixtures:
  density_current_step10.npz  the y-invariant j = 0 plane of DensityCurrent;
  <case>.npz                  the full arrays of the small cases listed in SMALL.

Run from the repository root in a container that has the reference mounted:
    python tests/golden/make_gold_fixtures.py /root/reference
The plotfiles are read as data (text headers, raw FP64 FABs); nothing of the
reference is imported or executed.
"""
import os
import re
import sys

import numpy as np


def read_plotfile(path):
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    assert lines[0].startswith("HyperCLaw")
    nvar = int(lines[1])
    names = lines[2:2 + nvar]
    pos = 2 + nvar
    dim = int(lines[pos]); time = float(lines[pos + 1])
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
            assert m, head
            order = m.group(2).split()
            dt = np.dtype("<f8") if order == list("87654321") else np.dtype(">f8")
            nc = int(m.group(12))
            sh = (b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)
            arr = np.frombuffer(f.read(8 * nc * sh[0] * sh[1] * sh[2]), dtype=dt).reshape((nc,) + sh)
        data[:, b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] = arr
    return dict(names=names, time=time, lo=lo, hi=hi, n=n, nstep=nstep, data=data, cell_h=ch)


SMALL = ["ScalarDiffusionSine", "ScalarDiffusionGaussian", "ScalarAdvectionUniformU", "ScalarAdvectionShearedU",
         "TaylorGreenAdvecting", "TaylorGreenAdvectingDiffusing"] + \
    ["ScalarAdvDiff_" + v for v in ("order2", "order3", "order4", "order5", "order6", "weno3", "weno3z", "weno5",
                                    "weno5z", "wenomzq3")] + \
    ["CouetteFlow_x", "CouetteFlow_y", "PoiseuilleFlow_x", "PoiseuilleFlow_y", "EkmanSpiral", "RayleighDamping"]


def save_small(ref, case):
    pf = read_plotfile(os.path.join(ref, "Tests", "ERFGoldFiles", case))
    assert np.all(np.isfinite(pf["data"]))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), case + ".npz"),
                        names=np.array(pf["names"]), time=pf["time"], nstep=pf["nstep"],
                        lo=np.array(pf["lo"]), hi=np.array(pf["hi"]), n=np.array(pf["n"]), data=pf["data"])
    print("wrote fixture", case, pf["names"], pf["data"].shape, "time", pf["time"], "steps", pf["nstep"])


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    only = sys.argv[2:]
    for case in SMALL:
        if not only or case in only:
            save_small(ref, case)
    if only:
        return
    pf = read_plotfile(os.path.join(ref, "Tests", "ERFGoldFiles", "DensityCurrent"))
    d = pf["data"]
    assert np.all(np.isfinite(d))
    # min/max per FAB and component recorded in Cell_H must match the data read
    mins = re.findall(r"^([-\d.e+,]+),\s*$", pf["cell_h"], flags=re.M)
    assert len(mins) >= 4
    # the case is two-dimensional: every j plane is identical
    assert np.all(d == d[:, :, :1, :])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "density_current_step10.npz"),
                        names=np.array(pf["names"]), time=pf["time"], nstep=pf["nstep"],
                        lo=np.array(pf["lo"]), hi=np.array(pf["hi"]), n=np.array(pf["n"]),
                        plane=d[:, :, 0, :])
    print("wrote fixture:", pf["names"], d.shape, "time", pf["time"], "steps", pf["nstep"])


if __name__ == "__main__":
    main()
