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
import shutil
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from erf_amd import plotfile  # noqa: E402


def read_plotfile(path):
    return plotfile.read(path)


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
    # the text Header of one gold plotfile, to check the writer's format against
    shutil.copyfile(os.path.join(ref, "Tests", "ERFGoldFiles", "ScalarAdvDiff_order2", "Header"),
                    os.path.join(os.path.dirname(os.path.abspath(__file__)), "ScalarAdvDiff_order2.Header"))
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
    # the same case on a terrain-following grid of flat 100 m levels over a
    # 50 m computational dz (detJ = 2): its y-invariant j = 0 plane
    pf = read_plotfile(os.path.join(ref, "Tests", "ERFGoldFiles", "DensityCurrent_detJ2"))
    d = pf["data"]
    assert np.all(np.isfinite(d)) and np.all(d == d[:, :, :1, :])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "density_current_detJ2_step10.npz"),
                        names=np.array(pf["names"]), time=pf["time"], nstep=pf["nstep"],
                        lo=np.array(pf["lo"]), hi=np.array(pf["hi"]), n=np.array(pf["n"]),
                        plane=d[:, :, 0, :])
    print("wrote fixture:", pf["names"], d.shape, "time", pf["time"], "steps", pf["nstep"])


if __name__ == "__main__":
    main()
