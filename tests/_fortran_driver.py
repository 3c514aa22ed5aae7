"""Runs the reference's own Fortran driver (tea.f90 and its modules, linked
against libtealeaf_mi355x.so by oracle/fortran_driver.mk) on a deck in a
scratch directory, as the reference is run: ./tea.in in, ./tea.out out."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "oracle", "_ref", "tea_leaf_mi355x")


def available():
    return os.path.exists(BINARY)


def run(deck, workdir, timeout=300):
    """Returns (returncode, stdout+stderr, tea.out text or None)."""
    os.makedirs(workdir, exist_ok=True)
    shutil.copy(deck, os.path.join(workdir, "tea.in"))
    proc = subprocess.run([BINARY], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=timeout)
    out = os.path.join(workdir, "tea.out")
    text = open(out).read() if os.path.exists(out) else None
    return proc.returncode, proc.stdout, text


def step_lines(tea_out):
    """(step, [vol, mass, density, energy, u]) of every field-summary block."""
    rows = []
    for line in tea_out.splitlines():
        if line.startswith(" step:"):
            rows.append((int(line[6:13]), [float(line[13 + 26 * i:13 + 26 * (i + 1)]) for i in range(5)]))
    return rows
