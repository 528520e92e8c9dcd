"""Output diagnostics on the CPU: the oracle's CalculateVolume
(src/Output.f90:617-735) pinned by an independent restatement over the
downloaded state, by the exact sum, by a closed form and by the reference's
own conservation check (tests/testlib.jl:104-158); and the Volume.txt writer
pinned by the Fortran runtime's own formatting of the reference's format
strings (src/Output.f90:664-673, 720-727)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from kestrel_amd import config as C, output, scenario as S, workloads
from oracle.oracle import volume_from_state
from tests.helpers import oracle_for

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def flat_lake(level=0.5, n_tiles=8, tile=16):
    """A level lake of depth `level` over a flat bed (b0 = 0), disc of radius 40 m."""
    cfg = C.default_config(n_tiles, n_tiles, tile, tile, dx=1.0)
    C.hydraulic_only(cfg)
    C.derive(cfg)
    b0v = np.zeros((cfg.NY + 1, cfg.NX + 1))
    caps = [S.Cap(0.0, 0.0, 40.0, level, shape="level")]
    init = S.initial_state(cfg, b0v, caps)
    return workloads.Workload("flat_lake", cfg, b0v, init, 0, caps, [])


def test_volume_matches_restatement_over_state():
    """ko_volume reads tileContainer's stored Hn, rho, bt and slopes; the
    restatement re-derives them from the downloaded state: same bits."""
    wl = workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0)
    o = oracle_for(wl)
    for _ in range(2):
        o.step_n(40)
        got = o.volume()
        ref = volume_from_state(wl.cfg, wl.b0v, o.state())
        assert got[0] == ref["volume"]
        assert got[2] == ref["mass"]
        assert got[1] == ref["bed"]
        assert got[4] == ref["solids"] * wl.cfg.rhos


@pytest.mark.parametrize("morpho", [False, True])
def test_kahan_sums_are_within_two_ulps_of_the_exact_sums(morpho):
    wl = (workloads.morpho_cap(64, 0, tile=16, radius=10.0) if morpho
          else workloads.plain_cap(128, 0, tile=16, cap_x=0.0, cap_radius=10.0))
    o = oracle_for(wl)
    o.step_n(30)
    v = o.volume()
    terms = o.volume_terms()
    a = wl.cfg.dx * wl.cfg.dy
    for col, q in ((0, 0), (2, 1), (4, 3)):
        exact = math.fsum(terms[:, q])
        scale = wl.cfg.rhos if col == 4 else 1.0
        assert abs(v[col] / a / scale - exact) <= 2 * math.ulp(exact), (col, v[col], exact)
    bed = math.fsum(terms[:, 2])
    assert abs(v[1] / a - bed) <= 2 * math.ulp(np.abs(terms[:, 2]).sum())


def test_flat_lake_volume_is_the_closed_form():
    wl = flat_lake()
    o = oracle_for(wl)
    v = o.volume()
    st = o.state()
    n_wet = int((st["w"] > 0.0).sum())
    assert n_wet > 100
    assert v[0] == 0.5 * n_wet
    assert v[2] == wl.cfg.rhow * 0.5 * n_wet
    assert v[1] == 0.0 and v[4] == 0.0


def test_conservation_as_the_reference_checks_it():
    """tests/testlib.jl:104-158: flow + bed volume and solids mass (flow +
    bed) are conserved to 1e-10 relative with no flux sources."""
    wl = workloads.morpho_cap(128, 0, tile=16, radius=10.0)
    o = oracle_for(wl)
    v0 = o.volume()
    o.step_n(60)
    v1 = o.volume()
    assert v1[1] != 0.0          # the bed has changed
    assert abs((v0[0] + v0[1]) - (v1[0] + v1[1])) / (v0[0] + v0[1]) < 1e-10
    rhos = wl.cfg.rhos
    s0 = (v0[4] + v0[5]) / rhos
    s1 = (v1[4] + v1[5]) / rhos
    assert abs(s0 - s1) / s0 < 1e-10


def test_volume_line_rejects_wrong_width():
    with pytest.raises(ValueError):
        output.volume_line(0.0, [1.0] * 5)


VALUES = [1.5, -2.25e-300, 0.0, 123456.789, 1e100, -7e-5, 2.0 / 3.0, 1e-5, 9.9999999999999995e-1,
          5e-324, 1.7976931348623157e308, -1234.5678e-99]
TIMES = [0.0, 0.125, 0.375, 1.005, 12345.675, 99999999.99, -0.5]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not present")
def test_volume_format_matches_the_fortran_runtime(tmp_path):
    """The reference's own format strings, written by flang's runtime, against
    output.volume_header / volume_line / fortran_es."""
    src = f"""program fmt
  implicit none
  real(8) :: x({len(VALUES)}), t({len(TIMES)})
  integer :: i
  read(*,*) x
  read(*,*) t
  write(*,fmt="(a12,6(a2,a24))") '        time', ', ', '                  volume', ', ', &
     '        total_bed_volume', ', ', '              total_mass', ', ', '                bed_mass', ', ', &
     '       total_solids_mass', ', ', '         bed_solids_mass'
  do i = 1, {len(TIMES)}
     write(*,fmt="(f12.2,6(a2, es24.15))") t(i), ', ', x(mod(i, {len(VALUES)}) + 1), ', ', &
        x(mod(i + 1, {len(VALUES)}) + 1), ', ', x(mod(i + 2, {len(VALUES)}) + 1), ', ', &
        x(mod(i + 3, {len(VALUES)}) + 1), ', ', x(mod(i + 4, {len(VALUES)}) + 1), ', ', &
        x(mod(i + 5, {len(VALUES)}) + 1)
  end do
  do i = 1, {len(VALUES)}
     write(*,fmt="(ES18.10E3)") x(i)
  end do
end program
"""
    f90 = tmp_path / "fmt.f90"
    f90.write_text(src)
    exe = tmp_path / "fmt"
    subprocess.run([FLANG, str(f90), "-o", str(exe)], check=True, capture_output=True)
    # shortest round-trip decimals, read back exactly by list-directed input
    stdin = " ".join(repr(v) for v in VALUES) + "\n" + " ".join(repr(t) for t in TIMES) + "\n"
    got = subprocess.run([str(exe)], input=stdin, check=True, capture_output=True,
                         text=True).stdout.splitlines()
    want = [output.volume_header()]
    n = len(VALUES)
    for i in range(1, len(TIMES) + 1):
        want.append(output.volume_line(TIMES[i - 1], [VALUES[(i + k) % n] for k in range(6)]))
    want += [output.fortran_es(v, 18, 10, 3) for v in VALUES]
    assert got == want


def test_write_volume_appends_after_the_header(tmp_path):
    p = str(tmp_path / "Volume.txt")
    output.write_volume(p, 0.0, [1.0, 0.0, 1000.0, 0.0, 0.0, 0.0], create=True)
    output.write_volume(p, 1.0, [1.0, 0.0, 1000.0, 0.0, 0.0, 0.0])
    lines = open(p).read().splitlines()
    assert lines[0] == output.volume_header() and len(lines) == 3
    # the reference's own reader (tests/testlib.jl:122-133) takes awk field 2
    # of line 2 and the last line, with the comma stripped
    assert float(lines[1].split()[1].replace(",", "")) == 1.0
    output.write_volume(p, 2.0, [2.0] * 6, create=True)
    assert len(open(p).read().splitlines()) == 2


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not present")
def test_solution_format_matches_the_fortran_runtime(tmp_path):
    """OutputSolutionData_txt's 2D header and row formats and the _topo
    formats (src/Output.f90:812-838, 876-893, 908-918) written by flang's
    runtime, against the writer's lines for the same values."""
    rng = np.random.default_rng(7)
    rows = []
    for r in range(6):
        v = rng.standard_normal(19) * 10.0 ** rng.integers(-120, 120, 19)
        v[r] = 0.0
        v[(r + 3) % 19] = 5e-324 if r % 2 else -1e-310
        rows.append((int(rng.integers(1, 99999999)), v))
    names = [f"'{c.rjust(18)}'" for c in output.SOLUTION_COLUMNS[1:]]
    hdr = ", ', ', ".join(names)
    src = f"""program sol
  implicit none
  real(8) :: v(19), b(2)
  integer :: k, r
  write (*, fmt="(a8,a2,18(a18,a2),a18)", advance='yes') '    tile', ', ', &
     {hdr}
  do r = 1, {len(rows)}
     read (*, *) k, v
     write (*, fmt="(i8, a2, 18(ES18.10E3, a2), ES18.10E3)", advance='yes') k, ', ', &
        v(1), ', ', v(2), ', ', v(3), ', ', v(4), ', ', v(5), ', ', v(6), ', ', v(7), ', ', &
        v(8), ', ', v(9), ', ', v(10), ', ', v(11), ', ', v(12), ', ', v(13), ', ', v(14), ', ', &
        v(15), ', ', v(16), ', ', v(17), ', ', v(18), ', ', v(19)
     write (*, *)
  end do
  write (*, fmt="(a8, a2, a18, a2, a18)", advance='yes') '    tile', ', ', '         B0_vertex', ', ', &
     '         Bt_vertex'
  read (*, *) k, b
  write (*, fmt="(i8, a2, ES18.10E3, a2, ES18.10E3)", advance='yes') k, ', ', b(1), ', ', b(2)
end program
"""
    f90 = tmp_path / "sol.f90"
    f90.write_text(src)
    exe = tmp_path / "sol"
    subprocess.run([FLANG, str(f90), "-o", str(exe)], check=True, capture_output=True)
    stdin = "".join(f"{k} " + " ".join(repr(float(x)) for x in v) + "\n" for k, v in rows)
    stdin += "17 -2.5e-3 1.25e101\n"
    got = subprocess.run([str(exe)], input=stdin, check=True, capture_output=True,
                         text=True).stdout.splitlines()
    want = [output.solution_header()]
    for k, v in rows:
        want += [output._row(k, v), ""]
    want += [output.topo_header(), output._row(17, [-2.5e-3, 1.25e101])]
    assert got == want


def test_solution_lines_layout():
    """Tiles in list order, rows j outer and cells i inner, an empty record
    after each row, tile number id + 1, GridToPhysical coordinates."""
    cfg = C.default_config(3, 3, 4, 4, dx=2.0)
    C.derive(cfg)
    tiles = np.array([4, 1])
    fields = np.arange(2 * 14 * 16, dtype=float).reshape(2, 14, 4, 4)
    lines = output.solution_lines(cfg, tiles, fields)
    assert len(lines) == 1 + 2 * 4 * 5
    first = [float(x) for x in lines[1].split(",")]
    xc, yc = S.cell_centres(cfg)
    assert first[0] == 5 and first[1] == pytest.approx(xc[4]) and first[2] == pytest.approx(yc[4])
    assert first[5:] == pytest.approx(list(fields[0, :, 0, 0]))
    assert lines[5] == "" and lines[21].startswith("       2,")


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not present")
def test_maxima_format_matches_the_fortran_runtime(tmp_path):
    """OutputMaximums_txt's header, row format and empty records
    (src/Output.f90:1299-1352) written by flang's runtime for a 1-tile grid
    of 2 x 2 cells, against maxima_lines."""
    cfg = C.default_config(3, 3, 2, 2, dx=1.0)
    C.derive(cfg)
    rng = np.random.default_rng(3)
    maxima = [rng.standard_normal((cfg.NY, cfg.NX)) * 10.0 ** rng.integers(-50, 50) for _ in range(11)]
    maxima[10][0, 0] = -1.0
    xc, yc = S.cell_centres(cfg)
    names = ", ', ', ".join(f"'{c.rjust(18)}'" for c in output.MAXIMA_COLUMNS[1:])
    src = f"""program mx
  implicit none
  real(8) :: v(15)
  integer :: i, j
  write (*, fmt="(a8,a2,14(a18,a2), a18)", advance='yes') '    tile', ', ', &
     {names}
  do j = 1, 2
     do i = 1, 2
        read (*, *) v
        write (*, fmt="(i8,a2,14(ES18.10E3, a2), ES18.10E3)", advance='yes') 5, ', ', &
           v(1), ', ', v(2), ', ', v(3), ', ', v(4), ', ', v(5), ', ', v(6), ', ', v(7), ', ', &
           v(8), ', ', v(9), ', ', v(10), ', ', v(11), ', ', v(12), ', ', v(13), ', ', v(14), ', ', v(15)
     end do
     write (*, *)
  end do
  write (*, *)
end program
"""
    f90 = tmp_path / "mx.f90"
    f90.write_text(src)
    exe = tmp_path / "mx"
    subprocess.run([FLANG, str(f90), "-o", str(exe)], check=True, capture_output=True)
    stdin = ""
    for j in range(2):
        for i in range(2):
            J, I = 1 * 2 + j, 1 * 2 + i
            vals = [xc[I], yc[J], 0.0, 0.0] + [m[J, I] for m in maxima]
            stdin += " ".join(repr(float(v)) for v in vals) + "\n"
    got = subprocess.run([str(exe)], input=stdin, check=True, capture_output=True,
                         text=True).stdout.splitlines()
    assert got == output.maxima_lines(cfg, [4], maxima)
