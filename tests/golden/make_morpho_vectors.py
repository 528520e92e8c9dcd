"""Write the reference's morphodynamic RHS and redistribution on small decks
as fixtures.

Runs in the build container only (it needs the reference's Fortran sources,
built by `make -C oracle ref`).  For each deck of
kestrel_amd.workloads.MORPHO_FIXTURE_CASES it runs the oracle 30 steps from
the deck's cap with the redistribution probe on, then

* calls the reference's CalculateMorphodynamicRHS
  (src/MorphodynamicRHS.f90:72-149) through oracle/_ref on the final state
  (fields and per-tile vertex bt as the oracle holds them), and
* if a redistribution happened, replays the reference's AddToRedistList and
  RedistributeGrid (src/Redistribute.f90:69-244) on the recorded intermed0 /
  intermed3 and candidate sequence,
* calls the reference's CalculateHydraulicRHS (src/HydraulicRHS.f90:64-136) on
  the final conserved fields over the per-tile vertex bt views (ghost w
  reconstructions rebuilt as LoadSourceConditions does),

and stores inputs and the reference's outputs in morpho_<case>.npz.  Only the
fields these routines read or write are kept (see FIELDS_RHS, FIELDS_REDIST).

    python tests/golden/make_morpho_vectors.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from kestrel_amd import workloads  # noqa: E402
from oracle.oracle import Oracle, ref_morpho_rhs, ref_redistribute, ref_rhs  # noqa: E402
from kestrel_amd.scenario import InitialState  # noqa: E402

# 0-based indices into the 13 fields: w rhoHnu rhoHnv Hnpsi Hn u v psi rho b0 bt dbdx dbdy
FIELDS_RHS = [4, 5, 6, 7, 10, 11, 12]
FIELDS_REDIST = [0, 3, 4, 9, 10, 11, 12]
STEPS = 30


def main():
    for case in workloads.MORPHO_FIXTURE_CASES:
        wl = workloads.morpho_fixture_case(case)
        o = Oracle(wl.cfg)
        o.set_topography(wl.b0v)
        o.set_state(wl.init)
        o.probe_enable(True)
        o.step_n(STEPS)
        st = o.state()
        f13 = o.fields13()
        btt = o.tile_bt()
        emd, dtbt = ref_morpho_rhs(wl.cfg, st["tile_state"], wl.b0v, f13, btt)
        rec = dict(tile_state=st["tile_state"], b0v=wl.b0v, f_rhs=f13[FIELDS_RHS], btt=btt,
                   emd=emd, dtbt=dtbt, config=json.dumps(wl.cfg.as_dict()))
        # hydraulic RHS over the eroded/deposited bed
        o2 = Oracle(wl.cfg)
        o2.set_topography(wl.b0v)
        o2.set_state(InitialState(st["w"], st["rhoHnu"], st["rhoHnv"], st["Hnpsi"], st["tile_state"],
                                  np.zeros_like(st["w"])), t=st["t"])
        o2.set_tile_bt(btt)
        sx, sy = o2.ghost_slopes()
        hs = np.zeros(wl.cfg.ntiles, np.int8)
        e, i, adv = ref_rhs(wl.cfg, st["tile_state"], hs, wl.b0v, o2.fields13(), sx, sy,
                            recompute_ghost=True, substep=1, t=st["t"], grid_t=st["t"], bt_tiles=btt)
        rec.update(h_w=st["w"], h_rhu=st["rhoHnu"], h_rhv=st["rhoHnv"], h_hnpsi=st["Hnpsi"],
                   h_t=st["t"], h_ddte=e, h_ddti=i, h_adv=adv)
        p = o.probe()
        if p is not None:
            fout, btout, refine = ref_redistribute(wl.cfg, p["tile_state"], wl.b0v, p["i0"],
                                                   p["i3_before"], p["cand"])
            k, i, j, ex = p["cand"]
            rec.update(r_tile_state=p["tile_state"],
                       r_f0=p["i0"][0][FIELDS_REDIST], r_bt0=p["i0"][1],
                       r_f3=p["i3_before"][0][FIELDS_REDIST], r_bt3=p["i3_before"][1],
                       r_k=k, r_i=i, r_j=j, r_ex=ex,
                       r_fout=fout[FIELDS_REDIST], r_btout=btout, r_refine=refine)
        out = os.path.join(HERE, f"morpho_{case}.npz")
        np.savez_compressed(out, **rec)
        print(case, os.path.getsize(out), "bytes", "candidates", 0 if p is None else len(p["cand"][0]))


if __name__ == "__main__":
    main()
