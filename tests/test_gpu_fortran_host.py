"""The Fortran side of the boundary: kestrel_amd/fortran/kst_host (built from
kst_host.f90 over the iso_c_binding module kestrel_step_binding) drives the
library as a Kestrel IntegrateTo would -- kst_create, kst_set_topography,
kst_set_state once, then kst_integrate_to and kst_get_state per output time --
and its results are the same bits as the ctypes path's."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from kestrel_amd import workloads
from kestrel_amd.build import HOST
from kestrel_amd.runtime import KstStats
from tests.helpers import stepper_for

pytestmark = pytest.mark.gpu

CASES = {
    "plain_cap": lambda: workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0),
    "morpho_cap": lambda: workloads.morpho_cap(64, 0, tile=16, erosion="fluid", deposition="simple",
                                               transition="none", damping="none"),
}
TOUTS = [0.1, 0.25, 0.5]


@pytest.mark.parametrize("case", list(CASES))
def test_fortran_host_matches_ctypes(gpu, tmp_path, case):
    if not os.path.exists(HOST):
        pytest.fail(f"{HOST} is missing: build it with python -m kestrel_amd.build (needs AMD flang)")
    wl = CASES[case]()
    cfg = wl.cfg
    assert cfg.n_sources == 0
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(bytes(cfg))
        f.write(np.int64(len(TOUTS)).tobytes())
        f.write(np.asarray(TOUTS, dtype=np.float64).tobytes())
        f.write(np.float64(0.0).tobytes())
        for a in (wl.b0v, wl.init.w, wl.init.rhoHnu, wl.init.rhoHnv, wl.init.Hnpsi):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(wl.init.tile_state, dtype=np.int8).tobytes())
    r = subprocess.run([HOST, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = cfg.NX * cfg.NY
    rec = ctypes.sizeof(KstStats) + 8 + 5 * 8 * n
    raw = open(fout, "rb").read()
    assert len(raw) == rec * len(TOUTS)
    g = stepper_for(wl)
    for k, tk in enumerate(TOUTS):
        blk = raw[k * rec:(k + 1) * rec]
        st = KstStats.from_buffer_copy(blk[:ctypes.sizeof(KstStats)]).as_dict()
        off = ctypes.sizeof(KstStats)
        t = np.frombuffer(blk[off:off + 8], dtype=np.float64)[0]
        fields = np.frombuffer(blk[off + 8:], dtype=np.float64).reshape(5, cfg.NY, cfg.NX)
        ref_st = g.integrate_to(tk)
        ref = g.state()
        assert t == ref["t"] == tk
        for key in ("steps", "rejected", "rhs_evals", "cell_updates", "active_tiles"):
            assert st[key] == ref_st[key], (case, tk, key)
        for q, f in enumerate(("w", "rhoHnu", "rhoHnv", "Hnpsi", "Hn")):
            assert np.array_equal(fields[q], ref[f]), (case, tk, f)
