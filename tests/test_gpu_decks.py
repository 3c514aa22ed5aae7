"""Whole-deck parity on the GPU: field summaries, iteration counts and the
residual trajectory of the HIP driver (drop-in and native modes) against the
CPU oracle, plus the reference's own known-answer tests
(field_summary.f90:86-106: temp within 0.001 % of the published value).

The bar for FP64 is 1e-11 relative on every field-summary quantity; the residual trajectories agree to 1e-6 relative while the residual
is far above round-off (both are the same algorithm with dot products summed
in a different order).
"""
import os

import numpy as np
import pytest

from tealeaf_opencl_amd import decks

pytestmark = pytest.mark.gpu

SUMMARY_TOL = 1e-11


def _compare_summaries(lg, lo, tol=SUMMARY_TOL):
    assert lo["summary"] and len(lg["summary"]) == len(lo["summary"])
    for sg, so in zip(lg["summary"], lo["summary"]):
        assert sg[0] == so[0]
        for a, b in zip(sg[1:], so[1:]):
            assert abs(a - b) <= tol * abs(b), (sg, so)


def _compare_solves(lg, lo, iter_slack=1):
    assert lo["solve"] and len(lg["solve"]) == len(lo["solve"])
    for a, b in zip(lg["solve"], lo["solve"]):
        assert a[0] == b[0]
        assert abs(a[1] - b[1]) <= iter_slack, (a, b)


def _run_both(hiplib, oracle, deck, tmp_path, mode):
    lg = hiplib.run_deck(deck, tmp_path / ("gpu_%d.log" % mode), mode=mode)
    lo = oracle.run_deck(deck, tmp_path / "cpu.log")
    return lg, lo


@pytest.mark.parametrize("mode", [0, 1])
def test_kat1_jacobi(hiplib, oracle, decks_dir, tmp_path, mode):
    deck = os.path.join(decks_dir, "kat1.in")
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, mode)
    assert lg["qa"][2] == "PASSED", lg["qa"]
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo, 0)


@pytest.mark.parametrize("mode", [0, 1])
def test_kat6_ppcg_halo4(hiplib, oracle, decks_dir, tmp_path, mode):
    deck = os.path.join(decks_dir, "kat6.in")
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, mode)
    assert lg["qa"][2] == "PASSED", lg["qa"]
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo)


@pytest.mark.parametrize("mode", [0, 1])
def test_config1_cg_diag_512(hiplib, oracle, tmp_path, mode):
    """512^2 CG + diagonal Jacobi, two seeded rectangles, 10 steps, eps 1e-15."""
    deck = decks.config1_deck(str(tmp_path / "c1.in"), seed=1, extra=["verbose_on"])
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, mode)
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo)
    # residual trajectory of the first solve while far above round-off
    rg = [r[2] for r in lg["resid"] if r[0] == 1]
    ro = [r[2] for r in lo["resid"] if r[0] == 1]
    assert len(rg) > 10 and len(rg) == len(ro)
    r0 = ro[0]
    for i, (a, b) in enumerate(zip(rg, ro)):
        if b < 1e-9 * r0:
            break
        assert abs(a - b) <= 1e-6 * b, (i, a, b)


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
@pytest.mark.parametrize("solver", ["tl_use_cg", "tl_use_chebyshev", "tl_use_ppcg"])
@pytest.mark.parametrize("mode", [0, 1])
def test_solvers_small(hiplib, oracle, tmp_path, solver, prec, mode):
    extra = []
    if solver == "tl_use_chebyshev":
        extra = ["tl_ch_cg_presteps=20", "tl_ch_cg_epslim=1.0e300"]
    if solver == "tl_use_ppcg":
        extra = ["tl_ppcg_inner_steps=10"]
    deck = decks.write_deck(str(tmp_path / "s.in"), 200, 148, decks.seeded_states(2, 2, True), solver=solver,
                            preconditioner=prec, end_step=3, eps=1e-12, max_iters=4000, extra=extra)
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, mode)
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo, 2)


@pytest.mark.parametrize("prec,hd", [("jac_diag", 1), ("jac_block", 1), ("jac_diag", 2)])
@pytest.mark.parametrize("mode", [0, 1])
def test_ppcg_fresh_z(hiplib, oracle, tmp_path, prec, hd, mode):
    """The fresh-z PPCG option (deck keyword tl_ppcg_fresh_z: r.z formed with
    z = M^-1 r of the final r of the inner steps, as the Fortran TeaLeaf's
    calc_zrnorm does, instead of the reference's stale z) against the
    oracle's, in the reference call sequence and in the fused driver (whose
    last smoothing step stores the fresh z)."""
    extra = ["tl_ppcg_inner_steps=10", "tl_ppcg_fresh_z"]
    deck = decks.write_deck(str(tmp_path / "s.in"), 200, 148, decks.seeded_states(2, 2, True), solver="tl_use_ppcg",
                            preconditioner=prec, end_step=3, eps=1e-12, max_iters=4000, extra=extra, halo_depth=hd)
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, mode)
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo, 2)
    # the option changes the solve: the stale-z run takes other iteration counts
    stale = oracle.run_deck(decks.write_deck(str(tmp_path / "t.in"), 200, 148, decks.seeded_states(2, 2, True),
                                             solver="tl_use_ppcg", preconditioner=prec, end_step=3, eps=1e-12,
                                             max_iters=4000, extra=extra[:1], halo_depth=hd), tmp_path / "t.log")
    assert [r[1] for r in stale["solve"]] != [r[1] for r in lo["solve"]]


@pytest.mark.parametrize("case", ["none-3", "jac_diag-4", "jac_block-10", "jac_block-7-reflective",
                                  "jac_diag-5-halo2", "jac_block-10-fresh", "jac_diag-7-fresh"])
def test_ppcg_two_steps_per_launch_bitwise(hiplib, tmp_path, case):
    """The native PPCG with two inner steps per launch (ring of the tile
    recomputed, r and sd flipped between buffers) against one step per launch:
    bitwise the same field summaries and solve lines, on a grid whose tiles
    are partial in x and y, with odd and even inner counts, walls with
    non-zero faces and a deeper halo."""
    parts = case.split("-")
    prec, inner = parts[0], int(parts[1])
    extra = ["tl_ppcg_inner_steps=%d" % inner]
    kw = {}
    if "reflective" in parts:
        extra.append("reflective_boundary")
        kw["coefficient"] = "tl_coefficient_inverse_density"
    if "halo2" in parts:
        kw["halo_depth"] = 2
    if "fresh" in parts:
        extra.append("tl_ppcg_fresh_z")
    deck = decks.write_deck(str(tmp_path / "p.in"), 200, 148, decks.seeded_states(2, 2, True), solver="tl_use_ppcg",
                            preconditioner=prec, end_step=2, eps=1e-12, max_iters=2000, extra=extra, **kw)
    try:
        hiplib.set_tuning(hiplib.TUNE_PPCG_STEPS_PER_LAUNCH, 1)
        one = hiplib.run_deck(deck, tmp_path / "one.log", mode=1)
    finally:
        hiplib.set_tuning(hiplib.TUNE_PPCG_STEPS_PER_LAUNCH, 2)
    two = hiplib.run_deck(deck, tmp_path / "two.log", mode=1)
    # (the exact-error column is NaN without tl_check_result)
    assert one["solve"] and np.array_equal(np.array(one["solve"], float), np.array(two["solve"], float),
                                           equal_nan=True), (one["solve"], two["solve"])
    assert one["summary"] == two["summary"]


@pytest.mark.parametrize("case", ["none-20-4000", "jac_diag-20-4000", "jac_block-20-4000", "jac_block-20-61",
                                  "jac_block-20-62", "jac_block-21-4000", "jac_diag-30-4000-reflective"])
def test_cheby_two_iterations_per_launch_bitwise(hiplib, tmp_path, case):
    """The native Chebyshev with two iterations per launch (the tile's ring
    recomputed, u and p flipped between buffers, the second tail run by the
    same finisher) against one iteration per launch: bitwise the same field
    summaries and solve lines, converged and stopped by odd and even
    iteration caps; presteps 21 makes the first fused iteration even, which
    keeps one iteration per launch."""
    parts = case.split("-")
    prec, presteps, max_iters = parts[0], int(parts[1]), int(parts[2])
    extra = ["tl_ch_cg_presteps=%d" % presteps, "tl_ch_cg_epslim=1.0e300"]
    kw = {}
    if "reflective" in parts:
        extra.append("reflective_boundary")
        kw["coefficient"] = "tl_coefficient_inverse_density"
    deck = decks.write_deck(str(tmp_path / "c.in"), 200, 148, decks.seeded_states(2, 2, True),
                            solver="tl_use_chebyshev", preconditioner=prec, end_step=2, eps=1e-12,
                            max_iters=max_iters, extra=extra, **kw)
    try:
        hiplib.set_tuning(hiplib.TUNE_CHEBY_STEPS_PER_LAUNCH, 1)
        one = hiplib.run_deck(deck, tmp_path / "one.log", mode=1)
    finally:
        hiplib.set_tuning(hiplib.TUNE_CHEBY_STEPS_PER_LAUNCH, 2)
    two = hiplib.run_deck(deck, tmp_path / "two.log", mode=1)
    assert one["solve"] and np.array_equal(np.array(one["solve"], float), np.array(two["solve"], float),
                                           equal_nan=True), (one["solve"], two["solve"])
    assert one["summary"] == two["summary"]
    if max_iters < 100:
        assert [x[1] for x in two["solve"]] == [max_iters] * 2


@pytest.mark.parametrize("prec", ["none", "jac_block"])
@pytest.mark.parametrize("max_iters", [61, 62])
def test_cheby_native_iteration_cap(hiplib, oracle, tmp_path, prec, max_iters):
    """Native Chebyshev stopped by max_iters after an odd and an even number of
    double-buffered iterations: u must come back from whichever buffer the
    last iteration wrote, and the fields must match the oracle's."""
    deck = decks.write_deck(str(tmp_path / "m.in"), 160, 120, decks.seeded_states(4, 2, True),
                            solver="tl_use_chebyshev", preconditioner=prec, end_step=2, eps=1e-15,
                            max_iters=max_iters, extra=["tl_ch_cg_presteps=20", "tl_ch_cg_epslim=1.0e300"])
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, 1)
    assert [s[1] for s in lg["solve"]] == [max_iters] * 2
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo, 0)


@pytest.mark.parametrize("mode", [0, 1])
def test_reflective_and_inverse_density(hiplib, oracle, tmp_path, mode):
    deck = decks.write_deck(str(tmp_path / "r.in"), 130, 90, decks.seeded_states(3, 2, False),
                            solver="tl_use_cg", preconditioner="jac_block", end_step=3, eps=1e-13,
                            coefficient="tl_coefficient_inverse_density", extra=["reflective_boundary"])
    lg, lo = _run_both(hiplib, oracle, deck, tmp_path, mode)
    _compare_summaries(lg, lo)
    _compare_solves(lg, lo, 2)


def test_dropin_and_native_agree(hiplib, tmp_path):
    deck = decks.config1_deck(str(tmp_path / "c.in"), seed=2, n=256, steps=4, preconditioner="jac_block")
    a = hiplib.run_deck(deck, tmp_path / "a.log", mode=0)
    b = hiplib.run_deck(deck, tmp_path / "b.log", mode=1)
    _compare_summaries(b, a, 1e-12)
    _compare_solves(b, a, 1)


@pytest.mark.parametrize("solver", ["tl_use_cg", "tl_use_chebyshev", "tl_use_ppcg"])
def test_full_size_native_vs_dropin(hiplib, tmp_path, solver):
    """4096^2 deck (two rectangles and a circle, block Jacobi, eps 1e-15), one
    timestep with each solver: the fused device-resident kernels against the
    reference-granularity kernels of the drop-in boundary, which share no
    iteration kernel with them. At this size eps 1e-15 is below what the
    Chebyshev residual (a true residual) can reach, so Chebyshev stops at
    tl_max_iters 5000 in both modes; PPCG (about 3270 outer iterations) ends
    at the rounding floor of its stale-z norm, where the two modes' different
    summation orders can move the exit by some iterations; CG converges and
    conserves the internal energy (sum of vol*u) to its residual floor."""
    deck = decks.config2_deck(str(tmp_path / "f.in"), solver=solver, steps=1)
    a = hiplib.run_deck(deck, tmp_path / "a.log", mode=0)
    b = hiplib.run_deck(deck, tmp_path / "b.log", mode=1)
    _compare_summaries(b, a, 1e-11)
    # a PPCG exit in the stagnation region may come some iterations apart
    _compare_solves(b, a, {"tl_use_cg": 1, "tl_use_chebyshev": 0, "tl_use_ppcg": 50}[solver])
    if solver == "tl_use_cg":
        s0, s1 = b["summary"][0], b["summary"][-1]
        # sum(vol*u) changes by -sum(vol*r) of the true residual, whose floor at
        # 16.8M cells is a few 1e-12 of the total
        assert abs(s1[4] - s0[4]) <= 1e-11 * abs(s0[4]), (s0, s1)
        assert b["solve"][0][1] < 5000


@pytest.mark.parametrize("name,slack", [("cg", 2), ("chebyshev", 0), ("ppcg", 50)])
def test_config2_full_size_vs_oracle(hiplib, tmp_path, name, slack):
    """The 4096^2 deck (two rectangles and a circle, block Jacobi, eps 1e-15)
    on the device-resident driver against the FP64 CPU oracle's field
    summaries after every timestep (fixtures written by
    tests/golden/make_full_size.py: CG and Chebyshev over 20 timesteps, PPCG
    over 6), to the 1e-11 bar, with the iteration counts.

    PPCG: the first timestep, which starts from the same state in both, ends
    within `slack` iterations of the oracle's (measured: both at 3266). Later
    timesteps start from states that differ in their last digits (the field
    summaries agree to ~3e-12), and each PPCG solve spends its last decade
    above eps creeping along the rounding floor of its stale-z norm (on the
    GPU 400-900 iterations from 10 eps to eps), so the exit iteration depends
    on that rounding: measured 4548/4573, 4369/4762, 4695/4688, 4556/4866 and
    4821/4040 (GPU/oracle). For those the test requires that both solves end
    on the exit test, not at tl_max_iters."""
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "config2_4096_%s.json" % name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated" % os.path.basename(path))
    with open(path) as f:
        fx = json.load(f)
    deck = decks.config2_deck(str(tmp_path / "c2.in"), solver=fx["solver"], n=fx["n"], steps=fx["steps"])
    lg = hiplib.run_deck(deck, tmp_path / "c2.log", mode=1)
    lo = {"summary": fx["summary"], "solve": fx["solve"]}
    _compare_summaries(lg, lo)
    if name != "ppcg":
        _compare_solves(lg, lo, slack)
        return
    print("PPCG outer iterations (GPU, oracle):", [(a[1], b[1]) for a, b in zip(lg["solve"], lo["solve"])])
    _compare_solves({"solve": lg["solve"][:1]}, {"solve": lo["solve"][:1]}, slack)
    assert len(lg["solve"]) == len(lo["solve"])
    max_iters = 5000
    for a, b in zip(lg["solve"], lo["solve"]):
        assert a[0] == b[0] and a[1] < max_iters and b[1] < max_iters, (a, b)


@pytest.mark.parametrize("kat", [2, 3, 4])
def test_reference_kat_large(hiplib, decks_dir, tmp_path, kat):
    """960^2 (87 steps), 3840^2 (10 steps) and 960^2 (1000 steps) CG
    known-answer tests, native mode. (Test 5, 3840^2 over 100 steps, passes
    too but takes about a minute.)"""
    lg = hiplib.run_deck(os.path.join(decks_dir, "kat%d.in" % kat), tmp_path / "k.log", mode=1)
    assert lg["qa"][2] == "PASSED", lg["qa"]


def test_tea_out_lines(hiplib, tmp_path):
    """tl_run_deck_out writes the reference's tea.out lines with its Fortran
    edit descriptors: the step line (timestep.f90:58), the solve summary
    (tea_solve.f90:356-371) and the field-summary block with the known-answer
    verdict (field_summary.f90:41-104), KAT test 1."""
    import re
    deck = os.path.join(os.path.dirname(__file__), "golden", "decks", "kat1.in")
    out = tmp_path / "tea.out"
    lg = hiplib.run_deck(deck, str(tmp_path / "g.log"), tea_out=str(out))
    lines = out.read_text().splitlines()
    e2617 = r"( {2}-0\.\d{17}E[+-]\d{2}| {3}0\.\d{17}E[+-]\d{2})"
    steps = [l for l in lines if l.startswith(" step:")]
    assert len(steps) == len(lg["summary"])
    for l, s in zip(steps, lg["summary"]):
        assert re.fullmatch(r" step:[ \d]{7}" + e2617 * 5, l), l
        assert int(l[6:13]) == s[0]
        vals = [float(l[13 + 26 * i:13 + 26 * (i + 1)]) for i in range(5)]
        assert all(abs(a - b) <= 1e-16 * abs(b) for a, b in zip(vals, s[1:])), (vals, s)
    heads = [l for l in lines if "Volume" in l]
    assert heads and all(l == " " * 13 + "%26s%26s%26s%26s%26s" % ("Volume", "Mass", "Density", "Energy", "U")
                         for l in heads)
    step_lines = [l for l in lines if l.startswith(" Step ")]
    assert len(step_lines) == len(lg["solve"])
    assert re.fullmatch(r" Step       1 time   0\.0000000 timestep   \d\.\d\dE[+-]\d\d", step_lines[0]), step_lines[0]
    cond = [l for l in lines if l.startswith("Conduction error ")]
    assert len(cond) == len(lg["solve"])
    assert all(re.fullmatch(r"Conduction error  ?-?0\.\d{7}E[+-]\d{2}", l) for l in cond), cond[:2]
    its = [int(l[16:]) for l in lines if l.startswith("Iteration count ")]
    assert its == [s[1] for s in lg["solve"]]
    qa = [l for l in lines if l.startswith("Test problem")]
    assert len(qa) == 1 and re.fullmatch(r"Test problem   1 is within {2,3}-?0\.\d{7}E[+-]\d{2}% of the expected solution",
                                         qa[0]), qa
    assert " This test is considered PASSED" in lines
