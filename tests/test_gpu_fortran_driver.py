"""Drop-in proof: the reference's own, unchanged Fortran driver (tea.f90,
tea_solve.f90, ... compiled from the reference sources and linked against
libtealeaf_mi355x.so by oracle/fortran_driver.mk) runs decks on the GPU
through the *_kernel_ocl_ entry points. Its tea.out must carry the
reference's known-answer verdict and the same field summaries as the
library's own restatement of that driver in drop-in mode
(tl_run_deck_out, TL_MODE_DROPIN), which issues the same entry-point calls
in the same order."""
import os

import pytest

import _fortran_driver as fd
from tealeaf_opencl_amd import decks

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not fd.available(), reason="Fortran driver not built (oracle/fortran_driver.mk)")]

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _check(hiplib, deck, tmp_path, kat):
    rc, out, f_out = fd.run(deck, str(tmp_path / "fortran"))
    assert rc == 0, out[-2000:]
    ours = tmp_path / "ours.out"
    hiplib.run_deck(deck, tmp_path / "ours.log", mode=0, tea_out=str(ours))
    a, b = fd.step_lines(f_out), fd.step_lines(ours.read_text())
    assert a and [s for s, _ in a] == [s for s, _ in b]
    worst = max(abs(x - y) / abs(y) for (_, ra), (_, rb) in zip(a, b) for x, y in zip(ra, rb))
    assert worst <= 1e-14, worst
    its_f = [int(l[16:]) for l in f_out.splitlines() if l.startswith("Iteration count ")]
    its_o = [int(l[16:]) for l in ours.read_text().splitlines() if l.startswith("Iteration count ")]
    assert its_f == its_o
    if kat:
        assert " This test is considered PASSED" in f_out.splitlines(), f_out[-1500:]


@pytest.mark.parametrize("kat", [1, 2, 6])
def test_reference_driver_kat(hiplib, tmp_path, kat):
    """KAT 1 (Jacobi, 10x10), 2 (CG, 960^2, 87 steps), 6 (PPCG with
    halo_depth 4, 500^2)."""
    _check(hiplib, os.path.join(DECKS, "kat%d.in" % kat), tmp_path, True)


@pytest.mark.parametrize("solver,prec", [("tl_use_chebyshev", "jac_block"), ("tl_use_cg", "jac_diag"),
                                         ("tl_use_ppcg", "none")])
def test_reference_driver_solvers(hiplib, tmp_path, solver, prec):
    """Chebyshev, CG and PPCG decks with each preconditioner on a grid whose
    tiles are partial in both directions."""
    extra = ["tl_ch_cg_presteps=20", "tl_ch_cg_epslim=1.0e300"] if solver == "tl_use_chebyshev" else []
    deck = decks.write_deck(str(tmp_path / "d.in"), 200, 148, decks.seeded_states(2, 2, True), solver=solver,
                            preconditioner=prec, end_step=3, eps=1e-12, max_iters=4000, extra=extra)
    _check(hiplib, deck, tmp_path, False)
