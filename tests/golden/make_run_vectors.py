"""Write tests/golden/run_vectors.json from the reference program's own Run.
(--params re-reads only the RunSet scalars of the cases, keeping their runs.)

    make -C oracle refrun
    python tests/golden/make_run_vectors.py [-j N] [case ...]
    python tests/golden/make_run_vectors.py --params [case ...]

Runs oracle/_ref/run/ref_run (every Fortran file of the reference compiled
where it lies, with oracle/ref_run.f90 as its main program) on each case of
tests/runvec.py and records, for the initial state and after every output
interval: grid%t, grid%dt, the accepted steps and revised-dt messages of the
interval (counted from the program's 't = ...' and 'revised nextT' lines),
the active list, the ghost set and the digests of the active tiles' state.
Cases with sampled topography get it as a vertex file (oracle/refrun.py
write_b0).  A restart case (tests/runvec.py RESTART) runs its first deck, keeps
the files its restart reads (RunInfo.txt, the last solution file and its _topo
companion, Maximums.txt) in tests/golden/restart/<case>/ (gzip), then runs the
case's deck, which resumes from them.  Existing entries of other cases are kept.
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from kestrel_amd import scenario as S  # noqa: E402
from oracle import refrun  # noqa: E402
from tests import runvec  # noqa: E402


def make(name: str, work: str) -> dict:
    path, src = runvec.CASES[name]
    run = runvec.load_case(name)
    b0file = None
    if src is not None:
        c = run.workload.cfg
        xv, yv = S.vertex_coords(c)
        b0file = os.path.join(work, "b0.bin")
        refrun.write_b0(b0file, run.workload.b0v, float(xv[0]), float(yv[0]), c.dx, c.dy)
    if name in runvec.RESTART:
        refrun.run(runvec.RESTART[name], work, b0file, keep=False, name="first")
        keep_restart_files(name, os.path.join(work, "out"))
    prm = refrun.params(path, os.path.join(work, "params"))
    R = refrun.run(path, work, b0file, keep=False)
    recs = []
    for k, d in enumerate(R["dumps"]):
        rec = runvec.digests_from_tiles(d["u"], d["bt"], d["maxima"])
        rec.update(t=d["t"], dt=d["dt"], active=[int(q) for q in d["active"]],
                   ghost=sorted(int(q) for q in d["ghost"]))
        if k > 0:
            rec.update(steps=R["steps"][k - 1], revised=R["revised"][k - 1])
        recs.append(rec)
    print(f"{name}: {len(recs) - 1} intervals, steps {R['steps']}, revised {R['revised']}, "
          f"reference wall {R['wall']:.1f} s", flush=True)
    return {"deck": os.path.relpath(path, runvec.HERE), "topography": src, "params": prm,
            "nout": run.nout, "tend": run.tend, "first": (R["intervals"] or [1])[0], "intervals": recs}


def keep_restart_files(name: str, out: str) -> None:
    """The files a restart from out reads, as the reference wrote them."""
    import gzip
    import shutil
    from kestrel_amd import restart
    dest = os.path.join(runvec.RESTART_DIR, name)
    if os.path.isdir(dest):
        shutil.rmtree(dest)
    os.makedirs(dest)
    info = restart.read_run_info(os.path.join(out, "RunInfo.txt"))
    last = f"{info.last_file:06d}.txt"
    shutil.copy(os.path.join(out, "RunInfo.txt"), os.path.join(dest, "RunInfo.txt"))
    for f in (last, last + "_topo", "Maximums.txt"):
        with open(os.path.join(out, f), "rb") as a, gzip.GzipFile(os.path.join(dest, f + ".gz"), "wb",
                                                                  compresslevel=9, mtime=0) as b:
            b.write(a.read())


def _one(name: str):
    with tempfile.TemporaryDirectory() as work:
        return name, make(name, work)


def main(names, jobs: int = 1):
    from concurrent.futures import ProcessPoolExecutor
    out = runvec.fixtures() if os.path.exists(runvec.FIXTURE) else {}
    names = names or list(runvec.CASES)
    with ProcessPoolExecutor(max_workers=max(1, jobs)) as ex:
        for name, rec in ex.map(_one, names):
            out[name] = rec
            with open(runvec.FIXTURE, "w") as f:
                json.dump(out, f, indent=1, sort_keys=True)


def refresh_params(names):
    out = runvec.fixtures()
    for name in names or list(out):
        with tempfile.TemporaryDirectory() as work:
            out[name]["params"] = refrun.params(runvec.CASES[name][0], os.path.join(work, "params"))
    with open(runvec.FIXTURE, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    if args[:1] == ["--params"]:
        refresh_params(args[1:])
        sys.exit(0)
    jobs = 1
    if args[:1] == ["-j"]:
        jobs, args = int(args[1]), args[2:]
    main(args, jobs)
