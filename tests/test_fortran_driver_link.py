"""The reference's own Fortran driver links against libtealeaf_mi355x.so
(oracle/fortran_driver.mk, built by build() where the reference sources are
present) and binds every entry point it calls. Without a GPU it must get as
far as reading tea.in and calling initialise_ocl_, which stops the run with
the reference-style fatal error (ocl_errors.cpp:175-194) and MPI_Abort."""
import os

import pytest

import _fortran_driver as fd

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _gpu_present():
    import tealeaf_opencl_amd as tl
    try:
        return tl.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(not fd.available(), reason="Fortran driver not built (no reference sources or toolchain)")
def test_fortran_driver_reaches_initialise_ocl(tmp_path):
    if _gpu_present():
        pytest.skip("a GPU is present: the GPU test runs the driver to completion")
    rc, out, tea_out = fd.run(os.path.join(DECKS, "kat1.in"), str(tmp_path), timeout=120)
    assert rc != 0
    assert "initialise_ocl_" in out and "no HIP device" in out, out
    # the Fortran side read the deck and wrote its echo before the first
    # device call
    assert tea_out is not None and "test_problem" in tea_out
