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


BLOCKS_DECK = [
    "text before any block is not read: x_cells=5",
    "*tea",
    "state 1 density=100.0 energy=0.0001",
    "state 2 density=0.1 energy=25.0 geometry=rectangle xmin=1.0 xmax=3.5 ymin=2.0 ymax=4.25",
    "x_cells=40 y_cells=30 xmin=0.0 ymin=0.0 xmax=10.0 ymax=10.0",
    "*notes",
    "x_cells=99 end_step=98",
    "*endnotes",
    "initial_timestep=0.004 end_step=3",
    "tl_use_cg halo_depth=1",
    "*endtea",
    "*tea",
    "x_cells=77 end_step=76",
    "*endtea",
]


def test_blocks_and_first_endtea(hiplib, oracle, tmp_path):
    """parse.f90:186-232: only records read while exactly the *tea block is
    open count (a nested block's records do not), and the reading stops at
    the *endtea that closes it, so a later *tea block is never read."""
    path = tmp_path / "b.in"
    path.write_text("\n".join(BLOCKS_DECK) + "\n")
    got, want = both(hiplib, oracle, str(path))
    assert_same(got, want)
    assert got[0][0] == 40 and got[0][2] == 3


def test_unmatched_end_rejected(hiplib, oracle, tmp_path):
    """An *end whose name is not the innermost open block is the
    'Unmatched */*end pair' error of parse.f90:208-214 in both readers."""
    path = tmp_path / "u.in"
    path.write_text("\n".join(["*tea", "state 1 density=1.0 energy=1.0", "*endnotes", "*endtea"]) + "\n")
    assert oracle.parse_deck_values(str(path)) is None
    with pytest.raises(hiplib.TealeafError, match="Unmatched"):
        hiplib.parse_deck(str(path))


# --- the library's reader against the reference's own (read_input.f90) ------
# The reference's Fortran driver, linked against the library
# (oracle/fortran_driver.mk), reads the deck with parse.f90/read_input.f90 and
# echoes every keyword it meets to tea.out (read_input.f90:130-296) before its
# first device call; without a GPU that call stops the run. The echo's
# values, written with the reference's own edit descriptors, are compared
# with tl_parse_deck's settings written with the same descriptors
# (tl_format_fortran, itself checked against amdflang).

def _gpu_present():
    import tealeaf_opencl_amd as tl
    try:
        return tl.device_count() > 0
    except Exception:
        return False


def _echo(tea_out):
    """{key: text} of the keyword echo, and [state dicts] in order."""
    lines = tea_out.splitlines()
    i0 = next(i for i, l in enumerate(lines) if l.strip() == "Reading input file")
    keys, states = {}, []
    for l in lines[i0 + 1:]:
        if l.strip() in ("Setting up initial geometry", "Input read finished."):
            break
        t = l.strip()
        if t.startswith("Reading specification for state"):
            states.append({"index": int(t.split()[-1])})
        elif t == "Profiler on":
            keys["profiler_on"] = "1"
        elif t.startswith("Diffusion coefficient"):
            keys["coefficient"] = "2" if "reciprocal" in t else "1"
        elif t.startswith("state geometry"):
            states[-1]["geometry"] = {"rectangular": 1, "circular": 2, "point": 3}[t.split()[-1]]
        elif t.startswith("state "):
            states[-1][l[1:26].split()[-1]] = l[26:]
        elif t.startswith("tl_preconditioner_type"):
            keys["tl_preconditioner_type"] = l[28:].strip()
        elif len(l) == 38 and l[1:26].strip():
            keys[l[1:26].strip()] = l[26:]
    return keys, states


def _check_against_echo(hiplib, deck_path, keys, states):
    import tealeaf_opencl_amd as tl
    ints, reals, st = hiplib.parse_deck(deck_path)
    e124 = lambda v: tl.format_fortran(0, v, 12, 4)
    i12 = lambda v: "%12d" % v
    want = {"initial_timestep": e124(reals[5]), "end_time": e124(reals[4]), "end_step": i12(ints[2]),
            "xmin": e124(reals[0]), "ymin": e124(reals[1]), "xmax": e124(reals[2]), "ymax": e124(reals[3]),
            "x_cells": i12(ints[0]), "y_cells": i12(ints[1]), "visit_frequency": i12(ints[4]),
            "summary_frequency": i12(ints[3]), "tl_ch_cg_presteps": i12(ints[7]),
            "tl_ch_cg_epslim": e124(reals[7]), "halo_depth": i12(ints[15]), "test_problem": i12(ints[16]),
            "tl_ppcg_inner_steps": i12(ints[9]), "profiler_on": str(ints[14]), "coefficient": str(ints[6]),
            "tl_preconditioner_type": {1: "none", 2: "jac_diag", 3: "jac_block"}.get(ints[12], "?")}
    checked = 0
    for k, text in keys.items():
        assert k in want, (k, text)
        assert want[k] == text, (k, want[k], text)
        checked += 1
    # states: density, energy, radius and geometry as read; the extents of
    # states 2.. are echoed before the dx/100, dy/100 shrink of
    # read_input.f90:320-327, so they are compared as numbers to the echo's
    # four digits
    assert len(states) == ints[18]
    dx = (reals[2] - reals[0]) / float(np.float32(ints[0]))
    dy = (reals[3] - reals[1]) / float(np.float32(ints[1]))
    for s, row in zip(states, st):
        n = s["index"]
        assert e124(row[0]) == s["density"] and e124(row[1]) == s["energy"]
        if "geometry" in s:
            assert int(row[7]) == s["geometry"]
        if "radius" in s:
            assert e124(row[6]) == s["radius"]
        for name, col, grow in (("xmin", 2, -dx), ("xmax", 3, dx), ("ymin", 4, -dy), ("ymax", 5, dy)):
            if name in s:
                v = row[col] + (grow / 100.0 if n >= 2 else 0.0)
                echo = float(s[name])
                assert abs(v - echo) <= 0.5001e-4 * max(abs(echo), 1e-300) * 10, (n, name, v, echo)
        checked += 1
    return checked


ECHO_DECKS = dict(SYNTHETIC)
ECHO_DECKS.update({"kat%d" % k: (lambda p, k=k: os.path.join(GOLDEN, "kat%d.in" % k)) for k in range(1, 7)})


@pytest.mark.parametrize("name", sorted(ECHO_DECKS) + ["blocks"])
def test_reader_against_reference_echo(hiplib, tmp_path, name):
    import _fortran_driver as fd
    if not fd.available():
        pytest.skip("the reference's Fortran driver is not built (no reference sources or toolchain)")
    if _gpu_present():
        pytest.skip("a GPU is present: the driver would run the whole deck")
    if name == "blocks":
        path = str(tmp_path / "b.in")
        open(path, "w").write("\n".join(BLOCKS_DECK) + "\n")
    else:
        path = ECHO_DECKS[name](str(tmp_path / (name + ".in")))
    rc, out, tea_out = fd.run(path, str(tmp_path / "run"), timeout=60)
    assert tea_out is not None and "Reading input file" in tea_out, out
    keys, states = _echo(tea_out)
    assert _check_against_echo(hiplib, path, keys, states) >= 3
    if name == "blocks":
        assert keys["x_cells"].strip() == "40" and keys["end_step"].strip() == "3"
