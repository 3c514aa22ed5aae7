"""Benchmark of the implicit diffusion solve on one MI355X.

Workload: the 4096x4096 deck (two seeded rectangles and a circle on [0,10]^2,
density 0.1 / 100), FP64 CG with strip-4 block-Jacobi preconditioning,
tl_eps 1e-15, dt 0.004. A step is one implicit timestep (one tea_leaf solve to
convergence) run by the device-resident driver; the timed region covers
exactly `--steps` timesteps after `--warmup` untimed ones, with all fields
resident in HBM. `value` is CG iterations per second over the timed region.

Printed as one JSON line; see the fields for the solve time per step, the
per-kernel roofline of the dominant kernel (HIP events on the library's own
stream) and a CPU baseline (the FP64 oracle on the host cores, a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import tempfile
import time

# the CPU baseline's OpenMP threads stay on neighbouring cores (read by the
# OpenMP runtime when the oracle library is loaded)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12
METRIC = "solve s to 1e-10 rel resid @4096² (FP64 vs mixed); HBM GB/s % peak; iters/s"

# algorithmic FP64 bytes per cell (each distinct array read and/or written
# once); the face coefficients Kx, Ky are recomputed from the conductivity w
ALG_BYTES = {
    # p_new = z + beta p_old; w = A p_new; p.w : R z, p_old, w  W p_new
    "k_fused_cg_w": 4 * 8,
    # r -= a A p; z = M_blk^-1 r; r.z : R p, r, w  W r, z on odd iterations;
    # even ones also add a_{n-1} p_{n-1} + a_n p_n into u : + R u, p_{n-1}  W u.
    # The timed launches alternate between the two: (40 + 64) / 2 per launch.
    "k_fused_cg_ur": (5 * 8 + 8 * 8) // 2,
    # r = u0 - A u; p = ca p + cb M_blk^-1 r; u' = u + p : R u, u0, w, p  W p, u'
    "k_fused_cheby": 6 * 8,
    # two such iterations per launch : R u, u0, p, Kx, Ky  W u, p
    "k_fused_cheby2": 7 * 8,
    # r -= A sd; u += sd; sd' = ca sd + cb M_blk^-1 r : R u, r, w, sd  W u, r, sd'
    "k_fused_ppcg_inner": 7 * 8,
    # two such steps per launch : R u, r, sd, Kx, Ky  W u, r, sd
    "k_fused_ppcg_inner2": 8 * 8,
    # PPCG outer update (u every step, sd = z / theta) : R u, p, r, w  W u, r, z, sd
    "k_fused_cg_ur_ppcg": 8 * 8,
}
# the reference's bytes per cell for the same work with block Jacobi
# (kernel_files/tea_leaf_{cg,cheby,ppcg}_cl.cl, each array once per kernel)
REF_BYTES = {"cg": 144, "cheby": 104, "ppcg_inner": 112, "ppcg_outer": 208}
# the refined solve's inner kernels (tl_refine.hip): same streams as the fused
# FP64 kernels, with r, z, p, d and the face field stored in FP64 or FP32
# refined CG kernels: faces from one per-cell field (FP64: the conductivity,
# FP32: 1/(2w))
ALG_BYTES_RF = {
    # k_rf_ur: R p, r, w  W r, z, plus every second iteration R d, p_prev  W d
    # (the lagged update of d); the timed launches alternate: (5 + 8) / 2 values
    "fp64": {"k_rf_w": 4 * 8, "k_rf_ur": (5 * 8 + 8 * 8) // 2},
    "mixed": {"k_rf_w": 4 * 4, "k_rf_ur": (5 * 4 + 8 * 4) // 2},
}
REF_BYTES_PER_ITER = 144  # reference CG + strip-4 block Jacobi (calc_w + calc_ur + calc_p)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--cpu-iters", type=int, default=1000,
                    help="CG iterations of the CPU baseline sample (about 12 s on 16 host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", type=int, default=1, help="1 native (default), 0 drop-in")
    ap.add_argument("--max-iters", type=int, default=5000, help="tl_max_iters of the deck (short profiles)")
    ap.add_argument("--no-e10", action="store_true", help="skip the solve-to-1e-10 leg")
    ap.add_argument("--no-solvers", action="store_true", help="skip the Chebyshev and PPCG leg")
    ap.add_argument("--no-dropin", action="store_true", help="skip the leg in the other driver mode")
    ap.add_argument("--only-e10", action="store_true",
                    help="run only the solve-to-1e-10 leg (for profiling it in a process of its own)")
    ap.add_argument("--large", action="store_true",
                    help="run the large combined workload (8192^2 PPCG FP64 vs mixed, adjoint) instead")
    ap.add_argument("--large-n", type=int, default=8192)
    ap.add_argument("--large-steps", type=int, default=20)
    ap.add_argument("--large-fd", type=int, default=1, help="finite-difference cells of the large adjoint")
    ap.add_argument("--large-skip", default="",
                    help="comma list of large legs to skip: fp64,fp64_refined,mixed,adjoint")
    ap.add_argument("--large-out", default="", help="directory for the large legs' decks, logs and tea.out")
    ap.add_argument("--large-tol", type=float, default=1e-12, help="solve tolerance of the large adjoint")
    return ap.parse_args()


def host_cpu():
    """nproc, the CPU model and the socket count of the host (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count(), "model": None, "sockets": None,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], stdout=subprocess.PIPE, text=True, timeout=10).stdout)
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(deck, iters, d):
    """The FP64 oracle on the host cores: a bounded sample of the bench
    workload (the first `iters` CG iterations of its first timestep), and the
    512x512 CG + diagonal-Jacobi deck of two seeded rectangles timed in full
    (10 timesteps) beside the same deck on the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import tealeaf_opencl_amd as tl
    from tealeaf_opencl_amd import decks
    pyoracle.load()
    secs, done = pyoracle.time_cg_iterations(deck, iters)
    c1 = decks.config1_deck(os.path.join(d, "config1.in"), seed=1)
    t0 = time.perf_counter()
    lo = pyoracle.run_deck(c1, os.path.join(d, "config1_cpu.log"))
    cpu_s = time.perf_counter() - t0
    g = tl.bench_deck(c1, mode=tl.MODE_NATIVE, warmup=0, steps=10)
    its = int(sum(x[1] for x in lo["solve"]))
    return {"value": done / secs if secs > 0 else None, "unit": "iters/s", "cores": pyoracle.num_threads(),
            "kind": "port",
            "sample": "%d FP64 CG + block-Jacobi iterations of the first timestep at %s, oracle/tl_oracle.c "
                      "(OpenMP), timed after setup" % (done, os.path.basename(deck)),
            "host": host_cpu(),
            "config1_512": {"workload": "512x512, 2 rectangles (seed 1), CG + diagonal Jacobi, eps 1e-15, "
                                        "10 timesteps, whole deck",
                            "cpu_s": cpu_s, "cpu_iterations": its, "cpu_iters_per_s": its / cpu_s,
                            "gpu_solve_s": g["seconds"], "gpu_iterations": g["iterations"],
                            "gpu_iters_per_s": g["iterations"] / g["seconds"]}}


def refine_leg(n, tol=1e-10):
    """Solve seconds to a true relative residual of `tol` on the 4096^2
    inverse-density layout (two seeded rectangles, contrast 1e3, CG + block
    Jacobi, dt 0.004): the refined solve with FP64 inner vectors against the
    one with FP32 inner vectors (both with FP64 residual replacement and the
    same stopping test). Each is solved once untimed (workspace allocation and
    graph capture), then again from the same initial state; the library's
    device-event time of that second solve is reported, and the field summary
    after it is compared between the two precisions."""
    import numpy as np
    import tealeaf_opencl_amd as tl
    from tealeaf_opencl_amd import decks

    out = {}
    summaries = {}
    for name, precision in (("fp64", tl.PRECISION_FP64), ("mixed", tl.PRECISION_MIXED)):
        c = tl.Chunk(n, n, 1, tl.PREC_JAC_BLOCK)
        try:
            dx = 10.0 / float(np.float32(n))
            c.initialise_chunk(0.0, 0.0, dx, dx)
            c.generate_chunk(*decks.state_arrays(decks.seeded_states(1, 2, False), n, n))
            c.set_field()
            c.common_init(tl.COEF_RECIP_CONDUCTIVITY, 0.004)
            c.solve_refined(precision, tol=tol)
            c.common_init(tl.COEF_RECIP_CONDUCTIVITY, 0.004)
            r = c.solve_refined(precision, tol=tol)
            c.finalise()
            summaries[name] = np.array(c.field_summary())
            out[name] = r
        finally:
            c.close()
    rel = np.abs(summaries["mixed"] - summaries["fp64"]) / np.abs(summaries["fp64"])
    cells = n * n
    kernels = {}
    for name in ("fp64", "mixed"):
        r = out[name]
        kb = ALG_BYTES_RF[name]
        kernels[name] = {
            "k_rf_w": {"avg_ms": r["k1_ms"], "alg_bytes_per_launch": kb["k_rf_w"] * cells,
                       "achieved_GBs": kb["k_rf_w"] * cells / (r["k1_ms"] * 1e-3) / 1e9 if r["k1_ms"] > 0 else None},
            "k_rf_ur": {"avg_ms": r["k2_ms"], "alg_bytes_per_launch": kb["k_rf_ur"] * cells,
                        "achieved_GBs": kb["k_rf_ur"] * cells / (r["k2_ms"] * 1e-3) / 1e9 if r["k2_ms"] > 0 else None},
            "samples": r["k_samples"],
            "s_per_iteration": r["seconds"] / max(r["iterations"], 1),
        }
    return {
        "workload": "%dx%d, 2 rectangles (seed 1), conductivity 1/density, CG + strip-4 block Jacobi, "
                    "true ||u0-Au||/||u0-Au_init|| <= %g" % (n, n, tol),
        "fp64_s": out["fp64"]["seconds"], "fp64_iterations": out["fp64"]["iterations"],
        "fp64_updates": out["fp64"]["updates"], "fp64_rel_residual": out["fp64"]["rel_residual"],
        "mixed_s": out["mixed"]["seconds"], "mixed_iterations": out["mixed"]["iterations"],
        "mixed_updates": out["mixed"]["updates"], "mixed_rel_residual": out["mixed"]["rel_residual"],
        "speedup": out["fp64"]["seconds"] / out["mixed"]["seconds"],
        "summary_rel_diff_mixed_vs_fp64": dict(zip(["volume", "mass", "internal_energy", "temperature"],
                                                  rel.tolist())),
        "kernels": kernels,
    }


def dropin_leg(deck, mode, steps=2):
    """The same deck in the other driver mode: with the default native mode,
    the drop-in mode is the reference's own call sequence (tea_solve.f90 with
    host alpha/beta and two blocking scalars per CG iteration through the
    *_kernel_ocl entry points), one untimed and `steps` timed timesteps."""
    import tealeaf_opencl_amd as tl
    other = tl.MODE_DROPIN if mode == tl.MODE_NATIVE else tl.MODE_NATIVE
    r = tl.bench_deck(deck, mode=other, warmup=1, steps=steps)
    return {"mode": "drop-in" if other == tl.MODE_DROPIN else "native", "iterations": r["iterations"],
            "seconds": r["seconds"], "iters_per_s": r["iterations"] / r["seconds"],
            "final_summary": r["final_summary"]}


def solvers_leg(n, d, steps=1, max_iters=5000):
    """The other two solvers of the 4096^2 deck: Chebyshev after 20 CG
    eigenvalue-estimate steps and PPCG with 10 inner steps, both with strip-4
    block Jacobi, tl_eps 1e-15, run by the device-resident driver (one untimed
    timestep, then `steps` timed ones). Bytes per iteration are algorithmic:
    the fused Chebyshev iterations, two per launch, stream u, u0, p, Kx, Ky
    in and u, p out (56 B/cell per two iterations); a PPCG outer step is the two fused CG kernels (88 B/cell)
    plus 10 inner steps, two per launch, each launch streaming u, r, sd, Kx,
    Ky in and u, r, sd out (64 B/cell). `vs_reference_roofline` is the
    measured time against the time the reference's kernels would need for
    the same iterations at 100 % of peak HBM bandwidth with their own bytes
    (Chebyshev 104 B/cell, PPCG outer 208 and inner 112, CG 144):
    dimensionless, above 1 when fewer bytes beat the reference's
    roofline."""
    import tealeaf_opencl_amd as tl
    from tealeaf_opencl_amd import decks

    out = {}
    for name, solver in (("chebyshev", "tl_use_chebyshev"), ("ppcg", "tl_use_ppcg")):
        deck = decks.config2_deck(os.path.join(d, name + ".in"), solver=solver, n=n, steps=1 + steps,
                                  summary_frequency=0, max_iters=max_iters)
        r = tl.bench_deck(deck, mode=tl.MODE_NATIVE, warmup=1, steps=steps)
        its, inner, secs = r["iterations"], r["inner_iterations"], r["seconds"]
        cg_b = ALG_BYTES["k_fused_cg_w"] + ALG_BYTES["k_fused_cg_ur"]
        if name == "chebyshev":
            # about 20 CG steps (two fused CG kernels) then Chebyshev iterations
            cg = 20 * steps
            alg = (cg * cg_b + (its - cg) / 2 * ALG_BYTES["k_fused_cheby2"]) * n * n
            ref = (cg * REF_BYTES["cg"] + (its - cg) * REF_BYTES["cheby"]) * n * n
        else:
            alg = (its * (ALG_BYTES["k_fused_cg_w"] + ALG_BYTES["k_fused_cg_ur_ppcg"]) +
                   inner // 2 * ALG_BYTES["k_fused_ppcg_inner2"]) * n * n
            ref = (its * REF_BYTES["ppcg_outer"] + inner * REF_BYTES["ppcg_inner"]) * n * n
        out[name] = {"s_per_step": secs / steps, "iterations_per_step": its / steps,
                     "inner_iterations_per_step": inner / steps, "iterations_per_s": its / secs,
                     "alg_GBs": alg / secs / 1e9, "alg_frac_of_peak": alg / secs / PEAK_HBM,
                     "vs_reference_roofline": ref / PEAK_HBM / secs, "final_summary": r["final_summary"]}
        print("solvers leg %s: %.3f s per step, %d iterations" % (name, secs / steps, its // steps),
              file=sys.stderr, flush=True)
    return out


def large_leg(args):
    """The large combined workload: n^2 cells (8192^2 by default, 67M cells,
    537 MB per FP64 field), eight seeded rectangles with densities
    10^U[-1,2], PPCG + block Jacobi with 10 inner steps, tl_eps 1e-10, for
    `steps` timesteps, with the reference FP64 PPCG (device-resident driver,
    the reference's exit test) and with the refined PPCG to a true relative
    residual of 1e-10 in FP64 and in FP32-inner precision; then the density adjoint over the same timesteps
    (conductivity = density, CG + block Jacobi solves to --large-tol, checkpoints
    every 5 steps) with central differences on a few sensitive cells
    (h = 1e-6 and 1e-3 relative) and along one random direction over every
    cell (1e-6). The deck legs are timed around the whole deck run
    (chunk set-up included) by the host clock."""
    import numpy as np
    import tealeaf_opencl_amd as tl
    from tealeaf_opencl_amd import adjoint, decks

    n, steps = args.large_n, args.large_steps
    skip = set(x for x in args.large_skip.split(",") if x)
    out = {"workload": "%dx%d, 8 seeded rectangles (density 10^U[-1,2], seed 1), PPCG + strip-4 block Jacobi, "
                       "10 inner steps, dt 0.004, %d timesteps" % (n, n, steps), "cells": n * n}
    legs = {}
    outdir = args.large_out or tempfile.mkdtemp()
    os.makedirs(outdir, exist_ok=True)
    precisions = {"fp64": "reference", "fp64_refined": "fp64", "mixed": "mixed"}
    for name in ("fp64", "fp64_refined", "mixed"):
        if name in skip:
            continue
        deck = decks.config5_deck(os.path.join(outdir, name + ".in"), precision=precisions[name],
                                  n=n, steps=steps, summary_frequency=steps)
        # the whole deck run (chunk set-up and start included), with its log
        # and tea.out written step by step
        t0 = time.perf_counter()
        lg = tl.run_deck(deck, os.path.join(outdir, name + ".log"), mode=tl.MODE_NATIVE,
                         tea_out=os.path.join(outdir, name + ".tea.out"))
        secs = time.perf_counter() - t0
        legs[name] = {"seconds": secs, "outer_iterations": int(sum(x[1] for x in lg["solve"])),
                      "inner_iterations": int(sum(x[4] for x in lg["solve"])),
                      "final_summary": lg["summary"][-1][1:]}
        print("large %s: %.3f s" % (name, secs), file=sys.stderr, flush=True)
    out.update(legs)
    names = ["volume", "mass", "density", "internal_energy", "temperature"]
    for ref in ("fp64", "fp64_refined"):
        if ref in legs and "mixed" in legs:
            a, b = np.array(legs[ref]["final_summary"]), np.array(legs["mixed"]["final_summary"])
            out["speedup_mixed_vs_" + ref] = legs[ref]["seconds"] / legs["mixed"]["seconds"]
            out["summary_rel_diff_mixed_vs_" + ref] = dict(zip(names, (np.abs(b - a) / np.abs(a)).tolist()))
    if "adjoint" not in skip:
        dt, coef, tol, every = 0.004, tl.COEF_CONDUCTIVITY, args.large_tol, 5
        c = tl.Chunk(n, n, 1, tl.PREC_JAC_BLOCK)
        t_start = time.perf_counter()

        def progress(phase, step, its):
            print("large adjoint leg: %s solve of step %d done (%d iterations, %.0f s)"
                  % (phase, step, its, time.perf_counter() - t_start), file=sys.stderr, flush=True)

        c.set_progress(progress)
        try:
            dx = 10.0 / float(np.float32(n))
            c.initialise_chunk(0.0, 0.0, dx, dx)
            c.generate_chunk(*decks.state_arrays(decks.config5_states(2), n, n))
            c.set_field()
            e0 = c.get_field("energy1")
            u_obs = adjoint.forward_u(c, coef, dt, steps, tol, progress=progress)
            c.generate_chunk(*decks.state_arrays(decks.config5_states(1), n, n))
            c.set_field()
            rho, e0 = c.get_field("density"), c.get_field("energy1")
            L, grad, info = c.adjoint_density(coef, dt, steps, u_obs, tol=tol, checkpoint_every=every)
            print("large adjoint: forward %.3f s, adjoint %.3f s" % (info["forward_s"], info["adjoint_s"]),
                  file=sys.stderr, flush=True)
            def loss_at(r):
                c.put_field("density", r)
                c.put_field("energy1", e0)
                return c.forward_loss(coef, dt, steps, u_obs, tol=tol)

            # per cell: central differences with h = 1e-6 rho_c and, to show
            # where their error comes from, with h = 1e-3 rho_c (truncation
            # error ~h^2 is still far below 1e-6 there, while the forward
            # solves' rounding noise in L is divided by a 1000x larger step)
            fd = []
            s_ = np.abs(rho * grad)
            ks, js = np.nonzero(s_ >= 1e-2 * s_.max())
            rng = np.random.default_rng(7)
            for i in rng.choice(len(ks), size=min(args.large_fd, len(ks)), replace=False):
                k, j = int(ks[i]), int(js[i])
                cell = {"cell": [k, j], "adjoint": float(grad[k, j])}
                for hr in (1e-6, 1e-3):
                    h = hr * rho[k, j]
                    r1, r2 = rho.copy(), rho.copy()
                    r1[k, j] += h
                    r2[k, j] -= h
                    fdv = (loss_at(r1) - loss_at(r2)) / (2 * h)
                    cell["rel_err_h%.0e" % hr] = abs(fdv - grad[k, j]) / abs(grad[k, j])
                print("large fd cell %d %d done" % (k, j), file=sys.stderr, flush=True)
                fd.append(cell)
            # along a random direction v = rho * N(0,1) over every cell, eps 1e-6
            v = np.zeros_like(rho)
            v[1:-1, 1:-1] = np.random.default_rng(3).standard_normal((n, n)) * rho[1:-1, 1:-1]
            dd = float(np.sum(grad * v))
            fdd = (loss_at(rho + 1e-6 * v) - loss_at(rho - 1e-6 * v)) / 2e-6
            directional = {"eps": 1e-6, "fd": fdd, "adjoint": dd, "rel_err": abs(fdd - dd) / abs(dd)}
            out["adjoint"] = {"loss": L, "forward_s": info["forward_s"], "adjoint_s": info["adjoint_s"],
                              "adjoint_over_forward": info["adjoint_s"] / info["forward_s"],
                              "checkpoint_every": every, "checkpoints": info["checkpoints"],
                              "device_bytes": info["chunk_bytes"] + info["work_bytes"],
                              "work_bytes": info["work_bytes"], "chunk_bytes": info["chunk_bytes"],
                              "iterations": {"forward": info["forward_iterations"],
                                             "recompute": info["recompute_iterations"],
                                             "adjoint": info["adjoint_iterations"]},
                              "stalled_solves": info["stalled_solves"], "finite_differences": fd,
                              "directional_derivative": directional}
        finally:
            c.close()
    return out


def load_traffic(n):
    path = os.path.join(ROOT, "profiles", "traffic_k_fused_cg_ur_%d.json" % n)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the library runs one chunk on one GPU; --gpus %d is not supported "
              "(sharding over several GPUs is not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import tealeaf_opencl_amd as tl
    from tealeaf_opencl_amd import decks

    tl.load_library()
    if args.large:
        print(json.dumps({"metric": METRIC, "large": large_leg(args)}))
        return
    if args.only_e10:
        print(json.dumps({"metric": METRIC, "solve_to_1e-10": refine_leg(args.n)}))
        return
    n = args.n
    # attainable streaming bandwidth on the same layout (copy, and the stream
    # mix of the dominant kernel without its stencil)
    with tl.Chunk(n, n, 1, tl.PREC_JAC_BLOCK) as sc:
        stream = {"copy_GBs": sc.stream_bandwidth(0, 20) / 1e9, "mix_4r3w_GBs": sc.stream_bandwidth(1, 20) / 1e9}
    with tempfile.TemporaryDirectory() as d:
        deck = decks.config2_deck(os.path.join(d, "bench.in"), n=n, steps=args.warmup + args.steps,
                                  summary_frequency=0, max_iters=args.max_iters)
        r = tl.bench_deck(deck, mode=args.mode, warmup=args.warmup, steps=args.steps)
        # solve seconds to 1e-10 (one timestep, true residual reported)
        r10 = None if args.no_e10 else refine_leg(n)
        solvers = None if args.no_solvers else solvers_leg(n, d, max_iters=args.max_iters)
        dropin = None if args.no_dropin else dropin_leg(deck, args.mode)
        cpu = None if args.no_cpu else cpu_baseline(deck, args.cpu_iters, d)

    cells = n * n
    its = r["iterations"]
    secs = r["seconds"]
    it_s = its / secs
    k1, k2 = r["k1_ms"], r["k2_ms"]
    roof = None
    if k2 and k2 > 0:
        alg = ALG_BYTES["k_fused_cg_ur"] * cells
        ach = alg / (k2 * 1e-3)
        roof = {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "traffic": load_traffic(n), "kernel": "k_fused_cg_ur<block>",
                "measured_stream": stream, "frac_of_stream_mix": ach / 1e9 / stream["mix_4r3w_GBs"],
                "alg_bytes_per_launch": alg, "avg_ms": k2, "samples": r["k_samples"],
                "k_fused_cg_w": {"avg_ms": k1, "alg_bytes_per_launch": ALG_BYTES["k_fused_cg_w"] * cells,
                                 "achieved_GBs": ALG_BYTES["k_fused_cg_w"] * cells / (k1 * 1e-3) / 1e9}}
    per_iter_bytes = (ALG_BYTES["k_fused_cg_w"] + ALG_BYTES["k_fused_cg_ur"]) * cells
    out = {
        "metric": METRIC,
        "value": it_s,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "tealeaf %dx%d CG + strip-4 block Jacobi FP64, eps 1e-15, dt 0.004, "
                               "2 rectangles + circle (seed 1)" % (n, n),
                   "cells": cells, "mode": "native" if args.mode == 1 else "drop-in"},
        "iterations": its,
        "solve_s_per_step": secs / args.steps,
        "hbm_GBs_per_iteration": per_iter_bytes * it_s / 1e9,
        "hbm_frac_per_iteration": per_iter_bytes * it_s / PEAK_HBM,
        "vs_reference_roofline": REF_BYTES_PER_ITER * cells * it_s / PEAK_HBM,
        "solve_to_1e-10": r10,
        "other_mode": dropin,
        "other_solvers": solvers,
        "roofline": roof,
        "cpu_baseline": cpu,
        "final_summary": r["final_summary"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
