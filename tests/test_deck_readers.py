"""The library's tea.in reader (tealeaf_opencl_amd/csrc/tl_deck.cpp, used by
tl_run_deck) against the oracle's (oracle/tl_oracle.c), on the CPU: both
restate read_input.f90:38-327 (keywords, defaults with their single-precision
literals, the dx/100 state shrink, the PPCG default inner-step count),
parse.f90:143-354 (100-character records, the continuation word, '!' lines)
and the C++ side channel of ocl_init.cpp:131-220. Every setting and every
state must come out identical, for the reference's known-answer decks and for
synthetic decks that exercise the keywords."""
import glob
import os

import numpy as np
import pytest

from tealeaf_opencl_amd import decks

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "decks")


def both(hiplib, oracle, path):
    got = hiplib.parse_deck(path)
    want = oracle.parse_deck_values(path)
    assert want is not None, path
    return got, want


def assert_same(got, want):
    (gi, gr, gs), (wi, wr, ws) = got, want
    assert gi == wi
    assert np.array_equal(gr, wr)
    assert np.array_equal(gs, ws)


@pytest.mark.parametrize("name", sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "*.in"))))
def test_known_answer_decks(hiplib, oracle, name):
    assert_same(*both(hiplib, oracle, os.path.join(GOLDEN, name)))


SYNTHETIC = {
    "config1": lambda p: decks.config1_deck(p, seed=1),
    "config2_cg": lambda p: decks.config2_deck(p, n=4096, steps=20),
    "config2_cheby": lambda p: decks.config2_deck(p, solver="tl_use_chebyshev", n=4096, steps=20),
    "config2_ppcg_fresh": lambda p: decks.config2_deck(p, solver="tl_use_ppcg", n=4096, steps=2,
                                                       extra=["tl_ppcg_fresh_z"]),
    "config3_mixed": lambda p: decks.config3_deck(p, precision="mixed"),
    "config5_fp64": lambda p: decks.config5_deck(p, precision="fp64", n=8192, steps=20),
    "ppcg_default_inner": lambda p: decks.write_deck(p, 300, 200, decks.seeded_states(3, 4, True),
                                                     solver="tl_use_ppcg", preconditioner="jac_diag", halo_depth=2),
    "jacobi_reflective": lambda p: decks.write_deck(p, 64, 48, decks.seeded_states(2, 2, False),
                                                    solver="tl_use_jacobi", preconditioner="none",
                                                    extra=["reflective_boundary", "profiler_on", "verbose_on",
                                                           "tl_check_result"]),
    "inverse_density": lambda p: decks.write_deck(p, 128, 96, decks.seeded_states(4, 3, True),
                                                  coefficient="tl_coefficient_inverse_density",
                                                  extra=["tl_ch_cg_presteps=7", "tl_ch_cg_epslim=0.5"]),
}


@pytest.mark.parametrize("name", sorted(SYNTHETIC))
def test_synthetic_decks(hiplib, oracle, tmp_path, name):
    path = SYNTHETIC[name](str(tmp_path / (name + ".in")))
    got, want = both(hiplib, oracle, path)
    assert_same(got, want)
    assert got[0][18] >= 2  # the background and at least one painted state


def test_comments_and_continuations(hiplib, oracle, tmp_path):
    """'!' lines, a state spread over three records with the continuation
    word, a record longer than 100 characters (cut there) and the
    side-channel keywords."""
    path = tmp_path / "c.in"
    path.write_text("\n".join([
        "*tea",
        "! a comment line with tl_use_ppcg in it",
        "state 1 density=100.0 energy=0.0001",
        "state 2 density=0.1 energy=25.0 geometry=rectangle \\",
        "  xmin=1.0 xmax=3.5 \\",
        "  ymin=2.0 ymax=4.25",
        "x_cells=40 y_cells=30 xmin=0.0 ymin=0.0 xmax=10.0 ymax=10.0" + " " * 60 + "end_step=99",
        "initial_timestep=0.004 end_step=3",
        "tl_use_cg tl_eps=1.0e-12 tl_max_iters=321",
        "halo_depth=1 profiler_on",
        "tl_preconditioner_type=jac_block",
        "*endtea",
    ]) + "\n")
    got, want = both(hiplib, oracle, str(path))
    assert_same(got, want)
    ints = got[0]
    assert ints[0] == 40 and ints[2] == 3 and ints[5] == 321 and ints[18] == 2


@pytest.mark.parametrize("case", ["ppcg_block_halo2", "words_after_preconditioner"])
def test_rejected_deck(hiplib, oracle, tmp_path, case):
    """Decks both readers refuse: PPCG + block Jacobi with halo_depth > 1
    (read_input.f90:299-301), and more words after tl_preconditioner_type on
    its record (read_input.f90:175-190 takes every remaining word of the record
    as a preconditioner name)."""
    if case == "ppcg_block_halo2":
        path = decks.write_deck(str(tmp_path / "r.in"), 64, 64, decks.seeded_states(1, 2, False),
                                solver="tl_use_ppcg", preconditioner="jac_block", halo_depth=2)
    else:
        path = decks.write_deck(str(tmp_path / "r.in"), 64, 64, decks.seeded_states(1, 2, False),
                                extra=["tl_preconditioner_type=jac_block profiler_on"])
    assert oracle.parse_deck_values(path) is None
    with pytest.raises(hiplib.TealeafError):
        hiplib.parse_deck(path)
