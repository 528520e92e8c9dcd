"""Write the reference's hydraulic RHS on small decks as fixtures.

Runs in the build container only (it needs the reference's Fortran sources,
built by `make -C oracle ref`).  For each deck of
kestrel_amd.workloads.FIXTURE_CASES it takes the state after 12 oracle steps
from the deck's cap, rebuilds the ghost-tile w reconstructions as
LoadSourceConditions does, calls the reference's CalculateHydraulicRHS
(src/HydraulicRHS.f90:64-136) through oracle/_ref, and stores inputs and
outputs in rhs_<case>.npz.

    python tests/golden/make_rhs_vectors.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from kestrel_amd import workloads  # noqa: E402
from oracle.oracle import Oracle, ref_rhs  # noqa: E402


def main():
    for case in workloads.FIXTURE_CASES:
        wl = workloads.fixture_case(case)
        o = Oracle(wl.cfg)
        o.set_topography(wl.b0v)
        o.set_state(wl.init)
        o.step_n(12)
        st = o.state()
        o2 = Oracle(wl.cfg)
        o2.set_topography(wl.b0v)
        from kestrel_amd.scenario import InitialState
        init = InitialState(st["w"], st["rhoHnu"], st["rhoHnv"], st["Hnpsi"], st["tile_state"],
                            np.zeros_like(st["w"]))
        o2.set_state(init, t=st["t"])
        f13 = o2.fields13()
        sx, sy = o2.ghost_slopes()
        hs = np.zeros(wl.cfg.ntiles, np.int8)
        e, i, adv = ref_rhs(wl.cfg, st["tile_state"], hs, wl.b0v, f13, sx, sy,
                            recompute_ghost=True, substep=1, t=st["t"], grid_t=st["t"])
        out = os.path.join(HERE, f"rhs_{case}.npz")
        np.savez_compressed(out, w=st["w"], rhoHnu=st["rhoHnu"], rhoHnv=st["rhoHnv"],
                            Hnpsi=st["Hnpsi"], tile_state=st["tile_state"], b0v=wl.b0v, t=st["t"],
                            ddte=e, ddti=i, advised=adv,
                            config=json.dumps(wl.cfg.as_dict()))
        print(case, os.path.getsize(out), "bytes", "advised", adv)


if __name__ == "__main__":
    main()
