"""Batches (kst_batch_*): scenarios advanced together give the same bits as
each scenario advanced alone, in FP64 and in mixed precision, whatever the
mix, the targets and the moment a scenario joins; a scenario that stops on an
error leaves the others untouched."""
import numpy as np
import pytest

from kestrel_amd import config as C
from kestrel_amd import workloads
from kestrel_amd.runtime import SLOT_DONE, SLOT_ERROR, SLOT_FREE, Batch, KestrelError
from tests.helpers import FIELDS, oracle_for, stepper_for

pytestmark = pytest.mark.gpu

STATS = ("steps", "rejected", "rhs_evals", "cell_updates", "active_tiles", "t")


def members(prec, n=6):
    """The n smallest scenarios of the ensemble (sides 128-192)."""
    picks = []
    for k in range(64):
        w = workloads.ensemble_member(k, precision=prec)
        if w.cfg.NX <= 192:
            picks.append(w)
        if len(picks) == n:
            break
    return picks


def same(a, b, label):
    assert np.array_equal(a["tile_state"], b["tile_state"]), f"{label}: tile states"
    assert a["t"] == b["t"], f"{label}: t {a['t']} vs {b['t']}"
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f"{label}: {f} differs"


@pytest.mark.parametrize("prec", [C.PREC_FP64, C.PREC_MIXED])
def test_batch_equals_alone(gpu, prec):
    wls = members(prec)
    alone = []
    for w in wls:
        g = stepper_for(w)
        st = g.step_n(60)
        alone.append((st, g.state()))
        g.close()
    b = Batch(8)
    gs = [stepper_for(w) for w in wls]
    slots = [b.attach(g) for g in gs]
    for s in slots:
        b.start(s, nsteps=60)
    status = b.run()
    for k, (s, g) in enumerate(zip(slots, gs)):
        assert status[s] == SLOT_DONE
        st = b.stats(s)
        for key in STATS:
            assert st[key] == alone[k][0][key], f"{wls[k].name}: {key} {st[key]} vs {alone[k][0][key]}"
        same(g.state(), alone[k][1], wls[k].name)
    assert all(status[q] == SLOT_FREE for q in range(len(slots), 8))
    b.close()


def test_batch_fp64_is_the_oracle(gpu):
    """A batched FP64 scenario against the oracle, bit for bit."""
    w = members(C.PREC_FP64, 1)[0]
    o = oracle_for(w)
    o.step_n(40)
    b = Batch(2)
    g = stepper_for(w)
    s = b.attach(g)
    b.start(s, nsteps=40)
    b.run()
    same(g.state(), o.state(), w.name)


def test_batch_targets_join_and_leave(gpu):
    """Scenarios with different targets (steps or a time), one joining after
    the others have run part of the way and a finished one leaving: each
    result is its result alone."""
    wls = members(C.PREC_MIXED, 4)
    ref = []
    for k, w in enumerate(wls):
        g = stepper_for(w)
        st = g.step_n(30 + 10 * k) if k % 2 == 0 else g.integrate_to(2.0 + k)
        ref.append(g.state())
        g.close()
    b = Batch(3)
    gs = [stepper_for(w) for w in wls]
    s0, s1 = b.attach(gs[0]), b.attach(gs[1])
    b.start(s0, nsteps=30)
    b.start(s1, tend=3.0)
    b.run(max_iters=16)                 # part of the way
    s2 = b.attach(gs[2])
    b.start(s2, nsteps=50)
    status = b.run()
    assert status[s0] == status[s1] == status[s2] == SLOT_DONE
    b.detach(s0)
    s3 = b.attach(gs[3])                # takes the freed slot
    assert s3 == s0
    b.start(s3, tend=5.0)
    status = b.run()
    assert status[s3] == SLOT_DONE
    for k in range(4):
        same(gs[k].state(), ref[k], wls[k].name)
    b.close()


def test_batch_error_isolated(gpu):
    """A cap that flows into a domain-edge tile halts its own scenario (as
    AddTile does, src/UpdateTiles.f90:61-65); the others run to the end."""
    bad = workloads.plain_cap(128, 0, tile=16, cap_x=-30.0, cap_radius=10.0)
    good = members(C.PREC_FP64, 2)
    ref = []
    for w in good:
        g = stepper_for(w)
        g.step_n(80)
        ref.append(g.state())
        g.close()
    b = Batch(4)
    gb = stepper_for(bad)
    gs = [stepper_for(w) for w in good]
    sb = b.attach(gb)
    sg = [b.attach(g) for g in gs]
    b.start(sb, nsteps=2000)
    for s in sg:
        b.start(s, nsteps=80)
    status = b.run()
    assert status[sb] == SLOT_ERROR
    with pytest.raises(KestrelError):
        b.stats(sb)
    for k, s in enumerate(sg):
        assert status[s] == SLOT_DONE
        same(gs[k].state(), ref[k], good[k].name)
    b.close()


def test_batch_rejects_mixed_settings(gpu):
    b = Batch(2)
    b.attach(stepper_for(members(C.PREC_FP64, 1)[0]))
    with pytest.raises(KestrelError):
        b.attach(stepper_for(members(C.PREC_MIXED, 1)[0]))
    b.close()


def test_full_ensemble_batch_equals_alone(gpu):
    """The whole batched ensemble at its stated size: 64 scenarios, sides
    log-uniform 128-1024, mixed precision, in one batch, 100 steps each;
    every member's counters and fields are the same bits as its run alone."""
    n, steps = 64, 100
    wls = workloads.ensemble(n, steps, precision=C.PREC_MIXED)
    assert len(wls) == 64 and max(w.cfg.NX for w in wls) >= 512
    alone = []
    for w in wls:
        g = stepper_for(w)
        st = g.step_n(steps)
        alone.append((st, g.state()))
        g.close()
    b = Batch(n)
    gs = [stepper_for(w) for w in wls]
    slots = [b.attach(g) for g in gs]
    for s in slots:
        b.start(s, nsteps=steps)
    status = b.run()
    for k, (s, g) in enumerate(zip(slots, gs)):
        assert status[s] == SLOT_DONE, wls[k].name
        st = b.stats(s)
        for key in STATS:
            assert st[key] == alone[k][0][key], f"{wls[k].name}: {key} {st[key]} vs {alone[k][0][key]}"
        same(g.state(), alone[k][1], wls[k].name)
    b.close()
    for g in gs:
        g.close()


def test_batch_mixes_limiters(gpu):
    """Scenarios with each of the reference's limiters in one batch (the
    kernels take the limiter per scenario then): every one the same bits as
    alone, in FP64 and in mixed precision."""
    for prec in (C.PREC_FP64, C.PREC_MIXED):
        wls = members(prec, 5)
        for w, lim in zip(wls, ("minmod1", "minmod2", "vanalbada", "weno", "none")):
            w.cfg.limiter = C.LIMITERS[lim]
        alone = []
        for w in wls:
            g = stepper_for(w)
            st = g.step_n(40)
            alone.append((st, g.state()))
            g.close()
        b = Batch(8)
        gs = [stepper_for(w) for w in wls]
        slots = [b.attach(g) for g in gs]
        for s in slots:
            b.start(s, nsteps=40)
        status = b.run()
        for k, (s, g) in enumerate(zip(slots, gs)):
            assert status[s] == SLOT_DONE
            st = b.stats(s)
            for key in STATS:
                assert st[key] == alone[k][0][key], f"{wls[k].name} ({prec}): {key}"
            same(g.state(), alone[k][1], f"{wls[k].name} ({prec})")
        b.close()


def test_batch_with_morphodynamics_equals_alone(gpu):
    """Morphodynamic scenarios (the Strang step: three morphodynamic stages,
    the redistribution walk, the face topography of the new bed, the second
    hydraulic operator, the bed commit) of different sizes and closures in
    one batch with hydraulic ones: every one the same bits as alone,
    counters, fields and the bed."""
    wls = [workloads.morpho_cap(64, 0, tile=16, radius=10.0, erosion="simple", deposition="simple",
                                transition="step", damping="rat3"),
           workloads.morpho_cap(96, 0, tile=16, radius=14.0, psi=0.2),
           workloads.morpho_cap(64, 0, tile=16, radius=8.0, drag="pouliquen", erosion="fluid",
                                deposition="simple", transition="none", damping="none")]
    wls += members(C.PREC_FP64, 2)
    steps = 40
    alone = []
    for w in wls:
        g = stepper_for(w)
        st = g.step_n(steps)
        alone.append((st, g.state(), g.tile_bt()))
        g.close()
    assert any(a[0]["redist_cells"] > 0 for a in alone[:3]), "no redistribution exercised"
    b = Batch(8)
    gs = [stepper_for(w) for w in wls]
    slots = [b.attach(g) for g in gs]
    for s in slots:
        b.start(s, nsteps=steps)
    status = b.run()
    for k, (s, g) in enumerate(zip(slots, gs)):
        assert status[s] == SLOT_DONE, wls[k].name
        st = b.stats(s)
        for key in STATS:
            assert st[key] == alone[k][0][key], f"{wls[k].name}: {key} {st[key]} vs {alone[k][0][key]}"
        same(g.state(), alone[k][1], wls[k].name)
        assert np.array_equal(g.tile_bt(), alone[k][2]), f"{wls[k].name}: bed"
    b.close()
