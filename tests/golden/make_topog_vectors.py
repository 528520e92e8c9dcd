"""Writes tests/golden/topog_funcs.npz: the reference's own TopogFunc
subroutines (src/TopogFuncs.f90) evaluated on a vertex grid for each function
with fixed parameters, and its Bicubic (src/Interp2d.f90:214-282) on seeded
data and points -- both compiled into oracle/_ref by `make -C oracle ref`."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import REF_TOPOG, ref_bicubic, ref_topog  # noqa: E402

XV = np.linspace(-60.0, 60.0, 61)
YV = np.linspace(-8.0, 8.0, 17)
PARAMS = {"x2slopes": [0.5, 0.1, 20.0], "usgs": [0.3, 20.0], "flume": [31.0, 2.4, 8.5, 2.0, 0.3, 20.0],
          "channel power law": [0.2, 4.0, 1.7], "channel trapezium": [0.2, 4.0, 0.5],
          "xbislope": [30.0, 5.0, 10.0], "xtrislope": [30.0, 15.0, 3.0, 8.0, -20.0, 25.0],
          "xhump": [2.0, 25.0], "xtanh": [5.0, 3.0, 12.0], "xyparab": [0.01, 0.02]}


def main():
    out = {"xv": XV, "yv": YV}
    for n in REF_TOPOG:
        key = n.replace(" ", "_")
        out[f"{key}__params"] = np.array(PARAMS[n])
        out[f"{key}__b0"] = ref_topog(n, PARAMS[n], XV, YV)
    rng = np.random.default_rng(11)
    z = rng.uniform(0.0, 500.0, (24, 18))
    bx = np.concatenate([rng.uniform(2.0, 23.0, 400), np.arange(2.0, 23.0)])
    by = np.concatenate([rng.uniform(2.0, 17.0, 400), np.full(21, 5.0)])
    out.update(bicubic_z=z, bicubic_x=bx, bicubic_y=by, bicubic_out=ref_bicubic(z, bx, by))
    np.savez(os.path.join(HERE, "topog_funcs.npz"), **out)
    print("wrote", len(REF_TOPOG), "functions")


if __name__ == "__main__":
    main()
