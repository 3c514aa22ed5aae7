"""The edit descriptors the library writes tea.out with (tl_run_deck_out:
field_summary.f90:41-81, tea_solve.f90:201-203 and 356-371, timestep.f90:58)
against a Fortran compiler's own output for the same values: a small program
compiled with amdflang (the compiler the reference's driver is built with in
this image) writes each value with each descriptor the reference uses (E14.7,
E15.7, E14.6, E26.17, E16.7, F14.6, F11.7, 1PE9.2), and every record must
equal the library's text, exponents beyond 99 included. The list-directed
items of diffuse.f90:66-69 are compiler-defined and not compared. CPU only;
skipped where amdflang is absent."""
import os
import shutil
import subprocess

import numpy as np
import pytest

FLANG = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

# (kind, w, d): 0 Ew.d, 1 1PEw.d, 2 Fw.d
DESCRIPTORS = [(0, 14, 7), (0, 15, 7), (0, 14, 6), (0, 26, 17), (0, 16, 7), (2, 14, 6), (2, 11, 7), (1, 9, 2)]


def values():
    rng = np.random.default_rng(11)
    v = [0.0, 1.0, -1.0, 0.5, 100.0, 9436.127524375916, 124.53439159789798, 15.131590925116196, 0.004,
         1.0e-15, 9.999999999999999e-1, 9.9999999e-6, 0.999999995, 123456.789, -3.25e-7, 1.0e-100, 2.5e120,
         float("nan"), float("inf"), float("-inf")]
    v += list(rng.uniform(-1, 1, 20) * 10.0 ** rng.integers(-20, 20, 20))
    return v


def fortran_records(tmp_path, vals):
    lines = ["program fmt", "  use, intrinsic :: ieee_arithmetic", "  implicit none",
             "  real(8) :: v(%d)" % len(vals), "  integer :: i"]
    # 17 significant digits with a d exponent: each REAL(8) literal is the double itself
    for i, x in enumerate(vals, 1):
        if np.isnan(x):
            lit = "ieee_value(1d0, ieee_quiet_nan)"
        elif np.isinf(x):
            lit = "ieee_value(1d0, ieee_%s_inf)" % ("positive" if x > 0 else "negative")
        else:
            lit = ("%.17e" % x).replace("e", "d")
        lines.append("  v(%d) = %s" % (i, lit))
    lines += ["  do i = 1, %d" % len(vals)]
    for kind, w, d in DESCRIPTORS:
        desc = {0: "E%d.%d", 1: "1PE%d.%d", 2: "F%d.%d"}[kind] % (w, d)
        lines.append("    write(*,'(%s)') v(i)" % desc)
    lines += ["  end do", "end program fmt"]
    src = tmp_path / "fmt.f90"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "fmt"
    subprocess.run([FLANG, "-O0", str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return out.split("\n")


@pytest.mark.skipif(FLANG is None, reason="amdflang not available")
def test_descriptors_match_fortran(hiplib, tmp_path):
    vals = values()
    recs = fortran_records(tmp_path, vals)
    k = 0
    bad = []
    for x in vals:
        for kind, w, d in DESCRIPTORS:
            mine = hiplib.format_fortran(kind, x, w, d)
            if recs[k] != mine:
                bad.append((x, (kind, w, d), recs[k], mine))
            k += 1
    assert not bad, bad[:8]


def test_rejects_out_of_range_digits(hiplib):
    """d must stay below w (the text of every edit then fits its buffers),
    and Ew.d needs d >= 1."""
    for kind, w, d in [(0, 10, 10), (0, 60, 59 + 1), (1, 9, 9), (2, 5, 7), (0, 10, 0), (0, 61, 3), (0, 0, 0)]:
        with pytest.raises(hiplib.TealeafError):
            hiplib.format_fortran(kind, 1.25, w, d)
    # the widest accepted edits still format
    assert len(hiplib.format_fortran(0, -1.0e-300, 60, 59)) == 60
    assert len(hiplib.format_fortran(1, 1.0e300, 60, 59)) == 60
