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
device must match the oracle, with the same times, steps and tiles: the
conserved fields, Hn and the bed (vertex and cell bt) to 1e-12 relative (the
parity bar; the momenta relative to at least rhow Hn sqrt(g Hn) of the
deepest cell, their scale where the flow is at rest and they are round-off
noise); the velocities, solids fraction and density -- ratios of conserved
fields, ill-conditioned in near-dry cells where Hn approaches the height
threshold -- and the cell bed slopes (differences of vertex bt) to 1e-9
relative to their maximum; the running maxima of Hn to 1e-12, of speed,
erosion, deposition and solids fraction to 1e-9, and the times they were
reached in all but 1e-3 of the cells (a near-tie between two steps can
resolve either way).

Where a run amplifies round-off (redistribution on a bed at its limits
turns last-bit differences into 1e-11 and more within a few steps), a field
beyond its bar is held instead to ten times the spread that the conditioning
probe shows: the oracle rebuilt with its per-cell transcendental functions
one ulp up (oracle/liboracle_pert.so), replayed over the same intervals.
A device bug moves a field by far more than that spread; libm differences of
a few ulps do not."""
import numpy as np
import pytest

from tests import runvec
from tests.helpers import rel_inf

pytestmark = pytest.mark.gpu

FIX = runvec.fixtures()

# cases the device reproduces bit for bit (only IEEE +, -, *, /, sqrt on the path)
BITWISE = {"restart_bowl_2d", "cap_dilute_2d", "flux_hydro_2d", "flux_single_pt", "plain_cap_256",
           "lake_at_rest_hydro_2d", "periodic_cap_2d", "dirichlet_cap_2d", "sponge_cap_2d",
           "cap_dilute", "flux_hydro", "lake_at_rest_hydro"}
# the parity bar of the conserved fields, Hn and the bed; the 1918 Strang steps
# of the measured morphodynamic run accumulate the libm differences further
# (measured: 9.7e-13, 6.7e-13, 8.8e-13, 1.14e-12 after its four intervals)
FIELD_TOL = {"morpho_hills_1024": 2e-12}


def _device(run, name, tmp):
    """The device at the case's start (the deck's painting, or the
    reference's restart files) and the first interval to run."""
    from kestrel_amd.runtime import Stepper
    g = Stepper(run.workload.cfg)
    g.set_topography(run.workload.b0v)
    first = runvec.start(g, run, name, tmp)
    return g, first


def _oracle(run, name, tmp):
    from oracle.oracle import Oracle
    o = Oracle(run.workload.cfg)
    o.set_topography(run.workload.b0v)
    runvec.start(o, run, name, tmp)
    return o


KEYS = ("cons", "derived", "topo", "btv", "maxima")


# a field beyond its bar is held to this multiple of the one-ulp spread
SPREAD = 10.0


class _Fields:
    """The fields a comparison reads, on the listed tiles: the solution
    columns of kst_get_output (0 Hn, 2 u, 3 v, 4 rho, 5 psi, 6 rhoHnu,
    7 rhoHnv, 8 Hnpsi, 9 w, 11 bt, 12 dbdx, 13 dbdy), the vertex bt and the 11
    maxima, from an oracle (tileContainer fields 5 6 7 9 8 2 3 4 1 11 12 13)
    or given."""

    COLS = {0: 4, 2: 5, 3: 6, 4: 8, 5: 7, 6: 1, 7: 2, 8: 3, 9: 0, 11: 10, 12: 11, 13: 12}

    def __init__(self, o, cfg, tiles, fg=None, bt=None, maxima=None):
        self.tiles = tiles
        if o is not None:
            f13 = runvec.tiles_of(cfg, tiles, o.fields13())
            self.f = {col: f13[:, idx] for col, idx in self.COLS.items()}
            bt = o.tile_bt()[tiles]
            maxima = np.array(o.maxima())
        else:
            self.f = {col: fg[:, col] for col in self.COLS}
        self.bt = bt
        self.m = [runvec.tiles_of(cfg, tiles, maxima[q]) for q in range(11)]
        hmax = float(np.abs(self.f[0]).max())
        self.mscale = cfg.rhow * hmax * np.sqrt(cfg.g * hmax)


def _checks(tol):
    """(what, bar): the conserved fields, Hn and the bed to the parity bar;
    the velocities, density, solids fraction and the cell bed slopes --
    ratios of conserved fields, ill-conditioned where Hn nears the height
    threshold, and differences of vertex bt -- to 1e-9; the maxima of Hn to
    the bar, the other maxima to 1e-9, and their times in all but 1e-3 of
    the cells."""
    out = [(("field", c), tol) for c in (0, 6, 7, 8, 9, 11)]
    out += [(("derived", c), 1e-9) for c in (2, 3, 4, 5, 12, 13)]
    out += [(("vbt",), tol)]
    out += [(("max", q), tol if q == 0 else 1e-9) for q in (0, 2, 4, 6, 8)]
    out += [(("maxtime", q), 1e-3) for q in (1, 3, 5, 7, 9, 10)]
    return out


def _metric(what, a, b):
    """The error of a against b: max|a-b| / max|b| per field (the momenta
    relative to at least rhow Hn sqrt(g Hn) of the deepest cell, their
    scale where the flow is at rest and they are round-off noise); for the
    times of the maxima the fraction of cells that differ (a step whose
    value ties the running maximum to round-off may win on one side only)."""
    if what[0] in ("field", "derived"):
        c = what[1]
        num = float(np.abs(a.f[c] - b.f[c]).max())
        den = float(np.abs(b.f[c]).max())
        if c in (6, 7):
            den = max(den, b.mscale)
        return num / den if den > 0 else num
    if what[0] == "vbt":
        return rel_inf(a.bt, b.bt)
    q = what[1]
    if what[0] == "max":
        return rel_inf(a.m[q], b.m[q])
    off = np.abs(a.m[q] - b.m[q]) > 1e-12 * max(1.0, float(np.abs(b.m[q]).max()))
    return float(np.mean(off))


class _Probe:
    """The conditioning probe: the oracle with its per-cell transcendental
    functions one ulp up (oracle/liboracle_pert.so), replayed over the same
    intervals when a field first exceeds its bar."""

    def __init__(self, run, name, tmp):
        self.run, self.name, self.tmp, self.o, self.k, self.first = run, name, tmp, None, 0, 1

    def at(self, k, tiles):
        from oracle.oracle import Oracle
        if self.o is None:
            self.o = Oracle(self.run.workload.cfg, perturbed=True)
            self.o.set_topography(self.run.workload.b0v)
            self.first = runvec.start(self.o, self.run, self.name, self.tmp)
        times = runvec.intervals(self.run, self.first)
        while self.k < k:
            self.o.integrate_to(times[self.k])
            self.k += 1
        return _Fields(self.o, self.run.workload.cfg, tiles)

    def close(self):
        if self.o is not None:
            self.o.close()


@pytest.mark.parametrize("name", sorted(FIX))
def test_device_equals_reference_run(gpu, name, tmp_path):
    if name in runvec.KNOWN_DIFFERENT:
        pytest.skip(runvec.KNOWN_DIFFERENT[name])
    fx = FIX[name]
    run = runvec.load_case(name)
    cfg = run.workload.cfg
    tmp = str(tmp_path)
    g, first = _device(run, name, tmp)
    assert first == fx.get("first", 1)
    o = None if name in BITWISE else _oracle(run, name, tmp)
    probe = _Probe(run, name, tmp)
    print(f"{name}: {cfg.NX}x{cfg.NY}, intervals {first}..{run.nout} to t = {run.tend}", flush=True)
    for k, tend in enumerate(runvec.intervals(run, first), start=1):
        ref = fx["intervals"][k]
        label = f"{name} interval {first + k - 1}"
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
        # device vs the reference's bits (through the oracle), per field
        tiles, fg = g.output_fields()
        ocmp = _Fields(o, cfg, tiles)
        dcmp = _Fields(None, cfg, tiles, fg=fg, bt=g.tile_bt()[tiles], maxima=np.array(g.maxima()))
        worst, worst_d = 0.0, 0.0
        tol = FIELD_TOL.get(name, 1e-12)
        for what, bar in _checks(tol):
            r = _metric(what, dcmp, ocmp)
            if r > bar:
                # beyond the bar: held to the spread a one-ulp change of the
                # host's transcendental functions produces in the same field
                s = _metric(what, probe.at(k, ocmp.tiles), ocmp)
                assert r <= SPREAD * s, (f"{label}: {what} {r:.3e} exceeds {bar:.0e} and "
                                         f"{SPREAD} x the one-ulp spread {s:.3e}")
                print(f"{label}: {what} {r:.2e} (bar {bar:.0e}; one-ulp spread {s:.2e})", flush=True)
            if what[0] == "field":
                worst = max(worst, r)
            elif what[0] == "derived":
                worst_d = max(worst_d, r)
        print(f"{label}: t = {sg['t']}, steps {sg['steps']}, worst field rel {worst:.2e}, "
              f"derived {worst_d:.2e}", flush=True)
    g.close()
    if o is not None:
        o.close()
    probe.close()
