#!/usr/bin/env python3
"""Headline benchmark: implicit filament steps per second on one MI355X.

Workload: a rigid sphere (radius 40, 6,000 surface blobs, held fixed) carrying
1,300 elastic filaments x 20 segments (26,000 segments, 32,000 particles in all)
whose lowest point is 50 radii above a no-slip wall.  One step = one BDF2 implicit
step solved by Broyden's method; every Broyden iteration evaluates the wall-RPY
velocities of all segments, with the blob forces from a GMRES solve to a relative
FP64 residual of 1e-10.

Two paths are timed back to back on the same initial state:
  * plain   -- FP64 pair sums, FP64 GMRES with the reference's 6 pi mu a diagonal
               preconditioner, previous blob forces as initial guess;
  * mixed   -- FP32 compensated pair sums inside FP64 iterative refinement, the
               near-field block preconditioner on the blob solve (and the learned
               correction / initial guess when --learned is given).
`value` is the steps/s of the mixed path; `plain_steps_per_s` and the ratio are
reported beside it.

roofline: the pair kernel, timed with HIP events on the library's stream over the
timed steps of the mixed path; algorithmic flops per launch from the canonical
per-pair counts (CFM_FLOPS_PAIR_* in include/cufcm_fil.h).
cpu_baseline: the FP64 OpenMP oracle apply timed on a bounded sample of the same
particles on this host, extrapolated to one implicit step with the pair work the
plain GPU step performed.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (published)
PMC_SUMMARY = os.path.join(ROOT, "profiles", "rpy_pair_fp32c_sphere32k_pmc.md")


def measured_traffic(kernel):
    """HBM bytes per launch of the named FP32C pair kernel from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE runs calibrated on a device copy); None when the
    summary holds no figure for it."""
    key = "traffic_bytes_per_launch_sym:" if kernel.startswith("rpy_sym") else "traffic_bytes_per_launch:"
    try:
        with open(PMC_SUMMARY) as fh:
            for line in fh:
                if line.startswith(key):
                    return float(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


def perturbation(nfil, seed):
    """Tip perturbation of the first step: 1% of the follower force in a random direction
    per filament (seeded)."""
    import numpy as np

    from cufcm_fil_amd.filaments import DEFAULTS, end_force
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nfil, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ef = end_force(DEFAULTS["dimensionless_force"], DEFAULTS["kb"], DEFAULTS["dl"], 20)
    return (0.01 * ef * d).reshape(-1)


def build_system(args, precision, solve_mode, guess, precond, overlap=False, pert_seed=7):
    split = getattr(args, "tile_split", -1)
    from cufcm_fil_amd import geometry as geo
    from cufcm_fil_amd.filaments import FilamentSystem
    xs, xb, centre, dirs, tans = geo.sphere_over_wall(return_tangents=True)
    nfil = xs.size // 3 // 20
    pert = perturbation(nfil, pert_seed)
    return FilamentSystem(xs, tans, x_blob=xb, precision=precision, solve_mode=solve_mode,
                          guess=guess, pert=pert, blob_rtol=1e-10, blob_restart=50,
                          blob_max_iter=5000, blob_precond=precond, tile_split=split,
                          overlap=overlap)


def run_path(args, label, precision, solve_mode, guess, precond, overlap=False, symmetry=True):
    S = build_system(args, precision, solve_mode, guess, precond, overlap)
    S.mob.set_pair_symmetry(symmetry and not args.no_pair_symmetry)
    try:
        for nt in range(args.warmup):
            st = S.step(nt)
            print(f"[bench] {label} warmup step {nt}: {st['seconds']:.3f}s "
                  f"broyden={st['broyden_iterations']} gmres={st['gmres_iterations']}",
                  file=sys.stderr, flush=True)
        S.mob.profile(True)
        stats = []
        t0 = time.perf_counter()
        for nt in range(args.warmup, args.warmup + args.steps):
            st = S.step(nt)
            stats.append(st)
            print(f"[bench] {label} step {nt}: {st['seconds']:.3f}s "
                  f"broyden={st['broyden_iterations']} gmres={st['gmres_iterations']} "
                  f"res={st['max_gmres_rel_residual']:.2e}", file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        prof = S.mob.profile_read()
        S.mob.profile(False)
        final_x = S.state()["x"]
    finally:
        S.close()
    el = t1 - t0
    return {
        "steps_per_s": args.steps / el,
        "ms_per_step": 1e3 * el / args.steps,
        "broyden_iterations": sum(s["broyden_iterations"] for s in stats),
        "gmres_iterations": sum(s["gmres_iterations"] for s in stats),
        "max_gmres_rel_residual": max(s["max_gmres_rel_residual"] for s in stats),
        "all_converged": all(s["converged"] for s in stats),
        "broyden_per_step": [s["broyden_iterations"] for s in stats],
        "profile": prof,
        "final_x": final_x,
    }


def pair_roofline(prof, steps):
    """Roofline line of the dominant pair kernel from the library's HIP-event profile of
    the timed steps (events on the stream each launch ran on): the symmetric
    segment-segment product when it runs (kernel + finalisation between the events;
    flops of both pair directions), else the one-directional tiled kernel."""
    if prof["kernel_ms_sym"] >= prof["kernel_ms_fp32"]:
        launches, ms_tot, fl_tot = prof["launches_sym"], prof["kernel_ms_sym"], prof["alg_flops_sym"]
        kname = "rpy_sym_kernel + rpy_sym_finalize_kernel"
    else:
        launches, ms_tot, fl_tot = prof["launches_fp32"], prof["kernel_ms_fp32"], prof["alg_flops_fp32"]
        kname = "rpy_apply_kernel<float,6,true>"
    kms = ms_tot / max(1, launches)
    flops_launch = fl_tot / max(1, launches)
    achieved = flops_launch / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(kname),
            "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) + " (FETCH_SIZE and WRITE_SIZE "
                              "counter runs of this workload, copy-calibrated; not counted in "
                              "this run)",
            "kernel": kname,
            "launches": launches, "avg_ms": round(kms, 4),
            "alg_flops_per_launch": flops_launch,
            "other_pair_kernel_ms_per_step": round(
                (prof["kernel_ms_fp32"] if launches == prof["launches_sym"] else prof["kernel_ms_sym"])
                / steps, 3)}


def run_batched(args):
    """Batched ensemble: 64 independent wall systems of 100-4,000 segments (filaments
    x 20 on lattices as the 1,280-segment case, random tilts), every system's
    resistance problem M x = b (N(0,1) right-hand sides) solved together to a relative
    FP64 residual of 1e-10, after the autotuner has picked the pair-kernel
    decomposition.  Reported beside it: the same 64 solves one system at a time."""
    import statistics

    import numpy as np

    from cufcm_fil_amd import RpyMobilitySolver
    from cufcm_fil_amd import geometry as geo
    from cufcm_fil_amd.batch import BatchedMobility
    ens = geo.ensemble(n_sys=64, lo=100, hi=4000, seed=3)
    sizes = [n for n, _ in ens]
    x = np.concatenate([xs for _, xs in ens])
    rhs = geo.loads(sum(sizes), 6, 1)
    opts = dict(rtol=1e-10, precond="self", mode="mixed")
    with BatchedMobility(sizes) as B:
        B.set_positions(x)
        B.apply(rhs, "fp32c")
        best, tried = B.tune("fp32c")
        B.solve(rhs, **opts)                      # graph capture + warm-up
        walls = []
        for rep in range(args.batch_reps):
            t0 = time.perf_counter()
            sol, stats, ms = B.solve(rhs, **opts)
            walls.append(time.perf_counter() - t0)
            print(f"[bench] batched solve {rep}: {walls[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    wall = statistics.median(walls)
    lat = np.array([st["latency_ms"] for st in stats])
    its = [st["iterations"] for st in stats]
    # one system at a time through single-system contexts, same solver settings
    seq = 0.0
    for (n, xs), b in zip(ens, np.split(rhs, np.cumsum(sizes)[:-1] * 6)):
        with RpyMobilitySolver(n, 0, precision="fp32c") as m:
            m.set_positions(xs)
            m.solve("seg", b, **opts)
            t0 = time.perf_counter()
            m.solve("seg", b, **opts)
            seq += time.perf_counter() - t0
    print(f"[bench] one-at-a-time: {seq * 1e3:.1f} ms", file=sys.stderr, flush=True)
    return {
        "workload": "64 independent wall systems, 100-4000 segments each (filaments x 20, "
                    "tilted lattices), resistance solve per system, GMRES(50) rtol 1e-10, "
                    "self-block preconditioner, FP32C pair sums in FP64 refinement",
        "systems": len(sizes), "particles": int(sum(sizes)),
        "pairs": int(sum(n * n for n in sizes)),
        "solves_per_s": round(len(sizes) / wall, 3),
        "ms_per_batch": round(1e3 * wall, 3),
        "p50_latency_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_ms": round(float(np.percentile(lat, 99)), 3),
        "iterations_min_median_max": [min(its), int(statistics.median(its)), max(its)],
        "max_rel_residual": max(st["rel_residual"] for st in stats),
        "all_converged": all(st["converged"] for st in stats),
        "tuning": best, "tuning_candidates": len(tried),
        "one_at_a_time_solves_per_s": round(len(sizes) / seq, 3),
    }


def run_learned(args):
    """Learned coarse space (the network's coarse head) on the 32,000-particle workload.
    The (right-hand side, solution) pairs of the blob solves are recorded along implicit-
    step trajectories of the solver without the network (near-field preconditioner with
    the linear coarse fields, FP32C pair sums); the network is trained on the GPU, kept
    only if two validation solves get cheaper, and judged on three held-out blob solves
    (right-hand sides induced by random segment loads) against plain GMRES (diagonal
    preconditioner), the near-field inverse alone and the near-field inverse with the
    linear coarse fields; every solve reaches 1e-10.  Steps/s of the implicit step with
    and without the network follow, from the same initial state."""
    import numpy as np
    runs, steps = args.learned_runs, args.learned_steps
    per_step = 16                     # blob solves per implicit step of this workload (upper bound)
    cap = runs * steps * per_step
    S = build_system(args, "fp32c", "fp64", "previous", "near")
    out = {"workload": "3 held-out blob solves of the 32,000-particle system (18,000 unknowns), "
                       "rtol 1e-10; implicit steps with the near-field blob preconditioner",
           "trajectories": runs, "steps_per_trajectory": steps}
    try:
        m = S.mob
        m.nearfield_coarse("linear")
        m.gnn_init(seed=0, capacity=cap)
        m.gnn_record(True)
        t0 = time.perf_counter()
        for r in range(runs):
            if r > 0:   # a new trajectory from the initial state with another perturbation
                S.restart(perturbation(S.nfil, 7 + r))
            for nt in range(steps):
                S.step(nt)
                if (nt + 1) % 50 == 0:
                    print(f"[bench] learned section: trajectory {r + 1}/{runs} step {nt + 1}, "
                          f"{m.gnn_samples()} pairs recorded", file=sys.stderr, flush=True)
        m.gnn_record(False)
        out["record_seconds"] = round(time.perf_counter() - t0, 2)
        samples = m.gnn_samples()
        ns = S.n_seg
        rng = np.random.default_rng(11)
        rhs = [-m.evaluate_blob_segment_mobility(rng.standard_normal(6 * ns)) for _ in range(5)]
        val, held = rhs[:2], rhs[2:]

        def solves(precond, coarse):
            m.nearfield_coarse(coarse)
            # untimed: builds the preconditioner's one-time data (near-field blocks, coarse space)
            m.solve("blob", val[0], rtol=1e-10, precond=precond, mode="fp64", max_iter=5000)
            its, res = [], []
            t0 = time.perf_counter()
            for b in held:
                _, st = m.solve("blob", b, rtol=1e-10, precond=precond, mode="fp64", max_iter=5000)
                its.append(st["iterations"])
                res.append(st["rel_residual"])
            return its, max(res), round(1e3 * (time.perf_counter() - t0) / len(held), 3)

        for name, pc, co in (("diag", "diag", "none"), ("near", "near", "none"),
                             ("near_linear", "near", "linear")):
            its, res, ms = solves(pc, co)
            out[name] = {"iterations": its, "ms_per_solve": ms, "max_rel_residual": res}
            print(f"[bench] learned section: {name} {its}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        gate = m.gnn_train_coarse_gated(val, epochs=args.learned_epochs, batch=32, lr=2e-3)
        out["train_seconds"] = round(time.perf_counter() - t0, 2)
        out["training_samples"] = samples
        out["loss_before"], out["loss_after"] = gate["loss_before"], gate["loss_after"]
        out["kept_after_validation"] = gate["kept"]
        its, res, ms = solves("near", "learned" if gate["kept"] else "linear")
        out["near_learned"] = {"iterations": its, "ms_per_solve": ms, "max_rel_residual": res}
        print(f"[bench] learned section: near+learned {its} (kept {gate['kept']})", file=sys.stderr,
              flush=True)
        theta = m.gnn_params()
    finally:
        S.close()
    # implicit steps with the near-field preconditioner, without and with the network
    for name, coarse in (("steps_near_linear", "linear"), ("steps_near_learned", "learned")):
        T = build_system(args, "fp32c", "fp64", "learned" if coarse == "learned" else "projected", "near")
        try:
            if coarse == "learned":
                T.mob.gnn_init(seed=0, capacity=1)
                T.mob.gnn_set_params(theta)
            T.mob.nearfield_coarse(coarse if gate["kept"] or coarse == "linear" else "linear")
            T.step(0)
            st, t0 = [], time.perf_counter()
            for nt in range(1, 4):
                st.append(T.step(nt))
            el = time.perf_counter() - t0
        finally:
            T.close()
        out[name] = {"steps_per_s": round(3 / el, 3), "broyden_iterations": sum(x["broyden_iterations"] for x in st),
                     "gmres_iterations": sum(x["gmres_iterations"] for x in st),
                     "max_gmres_rel_residual": max(x["max_gmres_rel_residual"] for x in st)}
        print(f"[bench] learned section: {name} {out[name]}", file=sys.stderr, flush=True)
    return out


def run_cilia(args):
    """Prescribed-shape cilia: a 16 x 16 lattice of Fulford-Blake cilia (20 segments,
    spacing 2.6a, lattice spacing 60a) above the wall, beating in phase at 2 pi per
    unit time with 300 steps per beat; one wall-RPY resistance solve per step to a
    relative FP64 residual of 1e-10 and the generalised forces of every cilium."""
    import numpy as np

    from cufcm_fil_amd.cilia import PrescribedCilia, wall_lattice
    base = wall_lattice(16, 16, spacing=60.0)
    n = base.size // 3
    steps = 20
    with PrescribedCilia(base, np.zeros(n), precision="fp32c", solve_mode="mixed",
                         tile_split=-1) as C:
        for nt in range(3):
            C.step(nt)
        its, res = 0, 0.0
        t0 = time.perf_counter()
        for nt in range(3, 3 + steps):
            _, _, st = C.step(nt)
            its += st["iterations"]
            res = max(res, st["rel_residual"])
        el = time.perf_counter() - t0
    print(f"[bench] cilia: {1e3 * el / steps:.2f} ms/step", file=sys.stderr, flush=True)
    return {
        "workload": "16 x 16 lattice of Fulford-Blake cilia (20 segments, seg_sep 2.6a, "
                    "spacing 60a) above the wall, in phase, 300 steps per beat; one "
                    "wall-RPY resistance solve per step (rtol 1e-10, FP32C pair sums in "
                    "FP64 refinement, self-block preconditioner)",
        "cilia": n, "segments": 20 * n,
        "steps_per_s": round(steps / el, 3), "ms_per_step": round(1e3 * el / steps, 3),
        "gmres_iterations_per_step": its / steps, "max_rel_residual": res,
    }


def run_free_body(args):
    """Free rigid body: the 6,000-blob sphere of the headline workload (radius 40a,
    centre 90a above the wall) as a free body.  Times make_body_reference_matrices
    (six unit-load solves) and the free-body resistance solve (blob forces and body
    velocities for a body load plus a background velocity at the blobs) with the
    reference's block preconditioner in FP64 and with the exact Schur complement
    around the resident inverse."""
    import numpy as np

    from cufcm_fil_amd import RpyMobilitySolver
    from cufcm_fil_amd import geometry as geo
    R, h, nb = 40.0, 90.0, 6000
    ref = geo.fibonacci_sphere(nb, R)
    q = np.array([0.9, 0.1, -0.3, 0.2])
    q /= np.linalg.norm(q)
    rng = np.random.default_rng(3)
    rhs = np.concatenate([0.01 * rng.standard_normal(3 * nb), rng.standard_normal(6)])
    out = {"workload": "sphere R=40a of 6000 blobs, centre 90a above the wall, free body; "
                       "make_body_reference_matrices (6 solves) and the free-body "
                       "resistance solve (3 x 6000 + 6 unknowns), rtol 1e-10"}
    with RpyMobilitySolver(0, nb, wall=True, precision="fp32c", tile_split=-1) as m:
        m.body_init(ref, 1)
        m.body_set_state([0.0, 0.0, h], q)
        t0 = time.perf_counter()
        _, N, st = m.make_body_reference_matrices(0, precond="dense", mode="fp64", rtol=1e-10)
        out["reference_matrices_s"] = round(time.perf_counter() - t0, 3)
        out["reference_matrices_iterations"] = st["iterations"]
        # lab-frame translational mobility relative to 1/(6 pi mu R) beside Faxen's
        # (parallel) and Brenner's (normal) wall series; N comes back in the body frame
        w, a, b, c = q
        Rq = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - w * c), 2 * (a * c + w * b)],
                       [2 * (a * b + w * c), 1 - 2 * (a * a + c * c), 2 * (b * c - w * a)],
                       [2 * (a * c - w * b), 2 * (b * c + w * a), 1 - 2 * (a * a + b * b)]])
        Ntt = Rq @ N[:3, :3] @ Rq.T
        s6 = 6.0 * math.pi * R
        sr = R / h
        out["mobility_parallel_x6piR"] = round(0.5 * (Ntt[0, 0] + Ntt[1, 1]) * s6, 5)
        out["mobility_normal_x6piR"] = round(Ntt[2, 2] * s6, 5)
        out["faxen_parallel"] = round(1 - 9 / 16 * sr + sr ** 3 / 8 - 45 / 256 * sr ** 4 - sr ** 5 / 16, 5)
        out["brenner_normal"] = round(1 - 9 / 8 * sr + sr ** 3 / 2 - 0.57 * sr ** 4 + 0.2 * sr ** 5, 5)
        for name, pc, mode, reps in [("diag_fp64", "diag", "fp64", 2), ("dense_fp64", "dense", "fp64", 5),
                                     ("dense_mixed", "dense", "mixed", 5)]:
            m.solve("body", rhs, rtol=1e-10, precond=pc, mode=mode, max_iter=5000)  # warm-up / build
            its, res = 0, 0.0
            t0 = time.perf_counter()
            for _ in range(reps):
                _, st = m.solve("body", rhs, rtol=1e-10, precond=pc, mode=mode, max_iter=5000)
                its += st["iterations"]
                res = max(res, st["rel_residual"])
            el = (time.perf_counter() - t0) / reps
            out[name] = {"ms_per_solve": round(1e3 * el, 3), "iterations": its / reps,
                         "rel_residual": res}
            print(f"[bench] free body {name}: {1e3 * el:.1f} ms/solve, {its / reps:.0f} iterations",
                  file=sys.stderr, flush=True)
    return out


def run_free_swimmer(args):
    """Free swimmer: the headline workload with the sphere free to move (its position
    and orientation among the Broyden unknowns, the blob solve carrying the body
    velocities).  Plain: FP64 pair sums and the reference's block preconditioner (6 pi
    mu a on the blobs, (K^T K)^-1 and the body mobility).  Fast: FP32C pair sums and the
    dense Schur preconditioner built once in the body frame and applied rotated."""
    import numpy as np

    from cufcm_fil_amd import geometry as geo
    out = {"workload": "sphere R=40a, 6000 blobs + 1300x20 elastic filaments, wall 50a "
                       "below, body free (force- and torque-free swimmer); BDF2 step by "
                       "Broyden, blob + body GMRES(50) rtol 1e-10"}
    xs, xb, centre, dirs, tans = geo.sphere_over_wall(return_tangents=True)
    ref = xb.reshape(-1, 3) - centre
    mob = None
    for label, prec, pc in [("fast", "fp32c", "dense"), ("plain", "fp64", "diag")]:
        S = build_system(args, prec, "fp64", "previous", pc)
        try:
            t0 = time.perf_counter()
            S.attach_bodies(ref.reshape(-1), centre, [1.0, 0.0, 0.0, 0.0], mobility=mob)
            if mob is None:
                mob = S.mob.body_reference_matrices()[1].reshape(-1)
                out["reference_matrices_s"] = round(time.perf_counter() - t0, 3)
            steps, stats = 2, []
            S.step(0)
            t0 = time.perf_counter()
            for nt in range(1, 1 + steps):
                stats.append(S.step(nt))
            el = time.perf_counter() - t0
            b = S.body_state()
        finally:
            S.close()
        out[label] = {"steps_per_s": round(steps / el, 4), "ms_per_step": round(1e3 * el / steps, 2),
                      "broyden_iterations": sum(s["broyden_iterations"] for s in stats),
                      "gmres_iterations": sum(s["gmres_iterations"] for s in stats),
                      "max_gmres_rel_residual": max(s["max_gmres_rel_residual"] for s in stats),
                      "all_converged": all(s["converged"] for s in stats),
                      "body_velocity": [round(v, 6) for v in b["vel"]]}
        print(f"[bench] free swimmer {label}: {1e3 * el / steps:.1f} ms/step", file=sys.stderr, flush=True)
    out["speedup"] = round(out["fast"]["steps_per_s"] / out["plain"]["steps_per_s"], 3)
    return out


def run_orbit_ensemble(args):
    """Outer-loop flow maps (the periodic-orbit / bisection / Floquet drivers' unit of
    work): members of a 4 x 4 lattice of dynamic-phase Fulford-Blake cilia driven by
    their recorded generalised forces, 30 steps each (a tenth of a beat).  Eight members
    run batched (one CiliaEnsemble: one product launch and one lockstep GMRES per step for
    all of them), as a concurrent ensemble (one context and HIP stream per member) and one
    member at a time (near-synchronous phases); 32 members -- the central-difference states of the
    monodromy matrix around the first of them (phases +- 1e-4) -- batched and
    concurrent."""
    import numpy as np

    from cufcm_fil_amd.cilia import wall_lattice
    from cufcm_fil_amd.orbits import CiliaFlow
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    qp = np.loadtxt(os.path.join(gold, "fb_phase_generalised_forces_nseg20_sep2.6.dat"))[1:]
    flow = CiliaFlow(wall_lattice(4, 4, spacing=30.0), qp, rtol=1e-10)
    n = flow.nfil
    rng = np.random.default_rng(3)
    # near-synchronous states (a common phase +- 0.3): with independent random phases,
    # neighbours 30a apart (cilia 49a long) can overlap, where the phase speeds spike and
    # the flow map is not continuous at the scale of the solve tolerance
    xs = [np.concatenate([[1.0], rng.uniform(0, 2 * math.pi) + rng.uniform(-0.3, 0.3, n)])
          for _ in range(8)]
    # the monodromy matrix's 2 n = 32 central-difference states around the first one
    ms = []
    for j in range(n):
        for sgn in (1.0, -1.0):
            x = xs[0].copy()
            x[1 + j] += sgn * 1e-4
            ms.append(x)
    ndts = 30
    flow.ensemble(xs[:2], 2, workers=2, batched=False)   # warm-up (library load, first contexts)
    flow.ensemble(xs[:2], 2)

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return r, time.perf_counter() - t0

    bat, t_bat = timed(lambda: flow.ensemble(xs, ndts))
    conc, t_conc = timed(lambda: flow.ensemble(xs, ndts, workers=8, batched=False))
    seq, t_seq = timed(lambda: [flow.run(x, ndts) for x in xs])
    bat32, t_bat32 = timed(lambda: flow.ensemble(ms, ndts))
    conc32, t_conc32 = timed(lambda: flow.ensemble(ms, ndts, workers=8, batched=False))
    diff_c = max(float(np.abs(a - b).max()) for a, b in zip(conc, seq))
    diff_b = max(float(np.abs(a - b).max()) for a, b in zip(bat, seq))
    diff_b32 = max(float(np.abs(a - b).max()) for a, b in zip(bat32, conc32))
    print(f"[bench] ensemble: 8 members batched {t_bat:.2f}s, concurrent {t_conc:.2f}s, "
          f"sequential {t_seq:.2f}s; 32 members batched {t_bat32:.2f}s, concurrent {t_conc32:.2f}s",
          file=sys.stderr, flush=True)
    return {
        "workload": "flow maps of a 4 x 4 wall lattice of dynamic-phase Fulford-Blake cilia "
                    "(20 segments, spacing 30a) driven by recorded generalised forces, 30 "
                    "steps each, one coupled FP64 GMRES solve per step (rtol 1e-10)",
        "steps_per_member": ndts,
        "members_8": {
            "batched_flow_maps_per_s": round(8 / t_bat, 3),
            "concurrent_flow_maps_per_s": round(8 / t_conc, 3),
            "sequential_flow_maps_per_s": round(8 / t_seq, 3),
            "batched_speedup_vs_sequential": round(t_seq / t_bat, 3),
            "concurrent_speedup_vs_sequential": round(t_seq / t_conc, 3),
            "max_abs_diff_concurrent_vs_sequential": diff_c,
            "max_abs_diff_batched_vs_sequential": diff_b,
        },
        "members_32": {
            "states": "monodromy matrix: phases of one state +- 1e-4",
            "batched_flow_maps_per_s": round(32 / t_bat32, 3),
            "concurrent_flow_maps_per_s": round(32 / t_conc32, 3),
            "batched_speedup_vs_concurrent": round(t_conc32 / t_bat32, 3),
            "max_abs_diff_batched_vs_concurrent": diff_b32,
        },
    }


def run_filament_ensemble(args):
    """Ensembles of elastic-filament runs (parameter sweeps launch one executable per run in
    the reference, pyfile/driver/func.py:544): 16 members, each the 64 x 20-segment wall
    lattice of tilted filaments (8 x 8, spacing 10a, tilt <= 0.3 rad, its own seed and tip
    perturbation), FP32C pair sums; one warm-up step, then 4 timed implicit steps of every
    member.  Batched: one context with cfm_fil_set_members (one product launch for all
    members, a Broyden iteration per member); one member at a time in its own context;
    concurrent: one context and stream per member on 8 host threads.  Every member's state
    is compared with its lone run.  Only the timed steps are timed (contexts and warm-up
    steps are made first; the concurrent members start together, eight at a time)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    from cufcm_fil_amd import geometry as geo
    from cufcm_fil_amd.filaments import DEFAULTS, FilamentSystem, end_force
    E, steps, nseg = 16, 4, 20
    f_end = end_force(DEFAULTS["dimensionless_force"], DEFAULTS["kb"], DEFAULTS["dl"], nseg)
    xs, ts, ps = [], [], []
    for e in range(E):
        x, t = geo.tilted_lattice(nx=8, ny=8, spacing=10.0, nseg=nseg, seed=200 + e,
                                  return_tangents=True, min_sep=2.5)
        rng = np.random.default_rng(300 + e)
        xs.append(x)
        ts.append(np.asarray(t).reshape(-1, 3))
        ps.append((0.02 + 0.01 * (e % 4)) * f_end * rng.standard_normal(3 * 64))

    def lone(e, barrier=None):
        """Member e alone: (start, end) of its timed steps, iterations, positions; with a
        barrier, the timed steps start once every thread has its context and warm-up step."""
        with FilamentSystem(xs[e], ts[e], pert=ps[e], precision="fp32c") as S:
            S.step(0)
            if barrier is not None:
                barrier.wait()
            t0 = time.perf_counter()
            its = [S.step(nt)["broyden_iterations"] for nt in range(1, steps + 1)]
            return t0, time.perf_counter(), its, S.state()["x"]

    lone(0)   # warm-up: library and first context
    t_seq, its_seq, x_seq = 0.0, [], []
    for e in range(E):
        a, b, its, x = lone(e)
        t_seq += b - a
        its_seq.append(its)
        x_seq.append(x)
    import threading
    conc = []
    for g in range(0, E, 8):   # eight members at a time, one context and stream each
        bar = threading.Barrier(8)
        with ThreadPoolExecutor(8) as pool:
            conc += list(pool.map(lambda e: lone(e, bar), range(g, g + 8)))
    t_conc_wall = sum(max(r[1] for r in conc[g:g + 8]) - min(r[0] for r in conc[g:g + 8])
                      for g in range(0, E, 8))
    with FilamentSystem(np.concatenate(xs), np.concatenate(ts), pert=np.concatenate(ps),
                        precision="fp32c") as S:
        S.set_members(E)
        S.step(0)
        t0 = time.perf_counter()
        its_bat = []
        for nt in range(1, steps + 1):
            S.step(nt)
            its_bat.append([m["broyden_iterations"] for m in S.member_stats()])
        t_bat = time.perf_counter() - t0
        x_bat = S.state()["x"].reshape(E, -1)
    same_its = all([its_bat[k][e] for k in range(steps)] == its_seq[e] for e in range(E))
    diff = max(float(np.abs(x_bat[e] - x_seq[e]).max()) for e in range(E))
    diff_c = max(float(np.abs(conc[e][3] - x_seq[e]).max()) for e in range(E))
    member_steps = E * steps
    print(f"[bench] filament ensemble: batched {t_bat:.3f}s, concurrent {t_conc_wall:.3f}s, "
          f"one at a time {t_seq:.3f}s", file=sys.stderr, flush=True)
    return {
        "workload": "16 members, each 64 x 20-segment tilted filaments on an 8 x 8 wall lattice "
                    "(spacing 10a), FP32C pair sums, 4 implicit steps per member after one "
                    "warm-up step",
        "members": E,
        "max_broyden_iterations_per_step": [max(its_bat[k]) for k in range(steps)],
        "batched_member_steps_per_s": round(member_steps / t_bat, 2),
        "concurrent_member_steps_per_s": round(member_steps / t_conc_wall, 2),
        "sequential_member_steps_per_s": round(member_steps / t_seq, 2),
        "batched_speedup_vs_sequential": round(t_seq / t_bat, 3),
        "batched_speedup_vs_concurrent": round(t_conc_wall / t_bat, 3),
        "same_broyden_iterations_as_lone_runs": bool(same_its),
        "max_abs_diff_batched_vs_lone": diff,
        "max_abs_diff_concurrent_vs_lone": diff_c,
    }


def run_ciliated_swimmer(args):
    """Prescribed cilia on a free body (the reference's PRESCRIBED_CILIA swimmers): a
    sphere of 6,000 blobs (R = 40a) carrying 128 Fulford-Blake cilia of 20 segments at
    prescribed phase speed 2 pi, unbounded fluid; one resistance solve for the forces and
    the body velocity per step (rtol 1e-10), explicit body update.  Plain: FP64 pair sums
    and the reference's block preconditioner (6 pi mu a with make_precon_mat's 6 x 6
    Schur complement).  Fast: mixed precision (FP32C inner products, FP64 residuals) with
    each cilium's own mobility block inverted and the resident inverse of the blob block
    (frozen in the body frame) inside the Schur complement."""
    import math
    import numpy as np

    from cufcm_fil_amd import geometry as geo
    from cufcm_fil_amd.cilia import CiliatedBodies, cilium_frames
    R, nc = 40.0, 128
    out = {"workload": "sphere R=40a of 6000 blobs carrying 128x20 Fulford-Blake cilia, "
                       "prescribed phase speed, body force- and torque-free, unbounded; "
                       "GMRES(50) rtol 1e-10 per step"}
    blob_ref = geo.fibonacci_sphere(6000, R)
    dirs = geo.fibonacci_sphere(nc, 1.0).reshape(-1, 3)
    phase = np.random.default_rng(0).uniform(0.0, 2.0 * math.pi, nc)
    vel = {}
    for label, pc, mode in [("plain", "diag", "fp64"), ("fast", "dense", "mixed")]:
        with CiliatedBodies(blob_ref, (R + 1.3) * dirs, cilium_frames(dirs), [0.0, 0.0, 0.0],
                            [1.0, 0.0, 0.0, 0.0], phase, np.full(nc, 2.0 * math.pi), nseg=20,
                            dl=2.6, dt=1.0 / 300, wall=False, precond=pc, solve_mode=mode,
                            rtol=1e-10, max_iter=5000) as G:
            t0 = time.perf_counter()
            G.step(0)
            first = time.perf_counter() - t0
            steps, its, conv = 10, [], True
            t0 = time.perf_counter()
            for nt in range(1, 1 + steps):
                _, _, st = G.step(nt)
                its.append(st["iterations"])
                conv = conv and bool(st["converged"])
            el = time.perf_counter() - t0
            vel[label] = G.bodies()["vel"]
        out[label] = {"steps_per_s": round(steps / el, 3), "ms_per_step": round(1e3 * el / steps, 2),
                      "gmres_iterations_per_step": round(sum(its) / steps, 1),
                      "first_step_s": round(first, 3), "all_converged": conv,
                      "body_velocity": [round(v, 6) for v in vel[label]]}
        print(f"[bench] ciliated swimmer {label}: {1e3 * el / steps:.1f} ms/step", file=sys.stderr,
              flush=True)
    out["speedup"] = round(out["fast"]["steps_per_s"] / out["plain"]["steps_per_s"], 3)
    out["body_velocity_rel_diff"] = float(np.abs(vel["fast"] - vel["plain"]).max() /
                                          np.abs(vel["plain"]).max())
    return out


def cpu_baseline(args, flops_per_step):
    """FP64 OpenMP oracle apply on a bounded sample of targets, extrapolated."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, timed here as the CPU baseline
    from cufcm_fil_amd import geometry as geo
    xs, xb, c, d = geo.sphere_over_wall()
    ns, nb = xs.size // 3, xb.size // 3
    f = np.random.default_rng(1).standard_normal(6 * ns)
    threads = oracle.max_threads()

    def timed(nt):
        xt = xs[:3 * nt]
        t0 = oracle.wtime()
        oracle.apply(xt, 6, xs, 6, f, id_tgt=np.arange(nt), id_src=np.arange(ns))
        return oracle.wtime() - t0

    probe = 64
    tp = timed(probe)
    nt = int(min(ns, max(probe, probe * args.cpu_seconds / max(tp, 1e-6))))
    el = timed(nt)
    flops = nt * ns * 321.0
    rate = flops / el
    return {
        "value": rate / flops_per_step if flops_per_step > 0 else None,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP64 OpenMP oracle 6<-6 wall-RPY apply, {nt} target segments x {ns} "
                   f"sources ({el:.1f} s, {rate / 1e9:.2f} GFLOP/s algorithmic), extrapolated "
                   f"to one implicit step of {flops_per_step:.3e} algorithmic pair flops "
                   f"(the pair work of the plain FP64 GPU step)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--learned", action="store_true", help="use the learned GMRES initial guess")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-learned", action="store_true")
    ap.add_argument("--skip-cilia", action="store_true")
    ap.add_argument("--skip-body", action="store_true")
    ap.add_argument("--skip-swimmer", action="store_true")
    ap.add_argument("--skip-ciliated", action="store_true")
    ap.add_argument("--skip-ensemble", action="store_true")
    ap.add_argument("--skip-filament-ensemble", action="store_true")
    ap.add_argument("--skip-overlap", action="store_true",
                    help="skip the comparison run with every product on one stream")
    ap.add_argument("--serial", action="store_true",
                    help="headline path with every product on one stream (no second stream "
                         "beside the blob solve)")
    ap.add_argument("--no-pair-symmetry", action="store_true",
                    help="one-directional pair kernel for the segment-segment product")
    ap.add_argument("--symmetry-ab", action="store_true",
                    help="also time the fast path with the one-directional pair kernel")
    ap.add_argument("--only", choices=["batched", "learned", "cilia", "body", "swimmer", "ciliated",
                                       "ensemble", "filaments"],
                    help="run one side section alone and print its JSON (for profiling)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--learned-runs", type=int, default=1,
                    help="recorded trajectories of the learned section (each from its own perturbation)")
    ap.add_argument("--learned-steps", type=int, default=30, help="implicit steps per recorded trajectory")
    ap.add_argument("--learned-epochs", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--precond", default="dense", choices=["near", "dense"],
                    help="blob-solve preconditioner of the fast path")
    ap.add_argument("--mode", default="fp64", choices=["mixed", "fp64"],
                    help="blob-solve arithmetic of the fast path")
    ap.add_argument("--tile-split", type=int, default=-1,
                    help="source split of the pair kernel (-1: autotuned, 0: heuristic)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES bytes (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a copy of known size)")
    ap.add_argument("--issue-probe", action="store_true",
                    help="only measure the FP32 VALU issue rate of the instruction forms the "
                         "pair kernel is built from (packed FP32 disabled), lane-instructions/s")
    args = ap.parse_args()
    if args.issue_probe:
        from cufcm_fil_amd import _lib as L
        forms = ["v_fmac_f32", "v_mul_f32_sgpr", "v_add_f32_sgpr", "v_mul_f32", "v_fma_f32"]
        rates = {}
        for k, name in enumerate(forms):
            r = ctypes.c_double(0.0)
            rc = L.probe_lib().cfm_probe_valu_rate(0, k, ctypes.byref(r))
            if rc != 0:
                raise RuntimeError(f"cfm_probe_valu_rate({name}) failed: {rc}")
            rates[name] = r.value
        print(json.dumps({"valu_issue_rate_lane_instr_per_s": rates}), flush=True)
        return
    if args.copy_probe > 0:
        import torch
        a = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library runs the implicit step on a "
              "single MI355X (multi-GPU sharding is a design note, not a run path).",
              file=sys.stderr)
        sys.exit(2)

    if args.only:
        fn = {"batched": run_batched, "learned": run_learned, "cilia": run_cilia, "body": run_free_body,
              "swimmer": run_free_swimmer, "ciliated": run_ciliated_swimmer,
              "ensemble": run_orbit_ensemble, "filaments": run_filament_ensemble}[args.only]
        print(json.dumps({args.only: fn(args)}), flush=True)
        return
    guess = "learned" if args.learned else "projected"
    # the fast path: the segment-segment product on a second stream beside the blob solve
    # (cfm_fil_params.overlap), unless --serial
    mixed = run_path(args, "mixed", "fp32c", args.mode, guess, args.precond, overlap=not args.serial)
    plain = None if args.skip_plain else run_path(args, "plain", "fp64", "fp64", "previous", "diag")
    # the same fast path with every product on one stream: the same results, and the
    # dominant kernel's rate with nothing running beside it
    ser = ser_run = None
    if not args.skip_overlap and not args.serial:
        ser_run = run_path(args, "serial", "fp32c", args.mode, guess, args.precond)
        o = ser_run
        ser = {"steps_per_s": round(o["steps_per_s"], 4), "ms_per_step": round(o["ms_per_step"], 3),
               "broyden_iterations": o["broyden_iterations"], "gmres_iterations": o["gmres_iterations"],
               "max_gmres_rel_residual": o["max_gmres_rel_residual"],
               "max_abs_state_diff_vs_overlapped": float(abs(o["final_x"] - mixed["final_x"]).max())}

    if plain is not None:
        # both paths start from the same state and take the same steps: the mixed path's
        # segment positions after the timed steps against the plain FP64 path's
        import numpy as np
        mixed["max_abs_state_diff_vs_plain"] = float(np.abs(mixed["final_x"] - plain["final_x"]).max())
        mixed["same_broyden_iterations_as_plain"] = mixed["broyden_per_step"] == plain["broyden_per_step"]
    # The dominant kernel's rate is taken where it runs alone (the serial pass: the same
    # steps, every product on one stream); in the overlapped headline pass it shares the
    # device with the blob solve, so its launches last longer while the step gets shorter
    # (that pass's mean is reported beside it).
    if ser_run is not None:
        roofline = pair_roofline(ser_run["profile"], args.steps)
        roofline["measured_in"] = "serial pass (same steps, every product on one stream)"
        ovl_roof = pair_roofline(mixed["profile"], args.steps)
        roofline["avg_ms_in_overlapped_headline_pass"] = ovl_roof["avg_ms"]
    else:
        roofline = pair_roofline(mixed["profile"], args.steps)
        roofline["measured_in"] = "headline pass" + ("" if args.serial else " (overlapped)")
    sym_ab = None
    if args.symmetry_ab:
        o = run_path(args, "one-directional", "fp32c", args.mode, guess, args.precond, symmetry=False)
        sym_ab = {"steps_per_s": round(o["steps_per_s"], 4), "ms_per_step": round(o["ms_per_step"], 3),
                  "broyden_iterations": o["broyden_iterations"],
                  "pair_kernel_ms_per_step": round(o["profile"]["kernel_ms_fp32"] / args.steps, 3),
                  "max_abs_state_diff_vs_symmetric": float(abs(o["final_x"] - mixed["final_x"]).max())}
    batched = None if args.skip_batch else run_batched(args)
    learned = None if args.skip_learned else run_learned(args)
    cilia = None if args.skip_cilia else run_cilia(args)
    body = None if args.skip_body else run_free_body(args)
    swimmer = None if args.skip_swimmer else run_free_swimmer(args)
    ciliated = None if args.skip_ciliated else run_ciliated_swimmer(args)
    ensemble = None if args.skip_ensemble else run_orbit_ensemble(args)
    fil_ensemble = None if args.skip_filament_ensemble else run_filament_ensemble(args)
    cpu = None
    if not args.skip_cpu:
        ref = plain if plain else mixed
        fps = ref["profile"]["alg_flops"] / args.steps
        cpu = cpu_baseline(args, fps)

    out = {
        "metric": "implicit steps/sec at 32k blobs, 1e-10 residual: mixed+learned vs plain FP64",
        "value": round(mixed["steps_per_s"], 5),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32c+fp64",
        "data": "synthetic",
        "config": {
            "workload": "sphere R=40a, 6000 fixed blobs + 1300x20 elastic filaments, wall 50a "
                        "below (32000 particles); BDF2 step by Broyden, blob GMRES(50) rtol 1e-10",
            "particles": 32000,
            "learned_guess": bool(args.learned),
            "parallelism": "single GPU",
        },
        "mixed": {k: v for k, v in mixed.items() if k not in ("profile", "final_x")},
        "plain_fp64": None if plain is None else {k: v for k, v in plain.items()
                                                  if k not in ("profile", "final_x")},
        "speedup_vs_plain_fp64": None if plain is None else round(mixed["steps_per_s"] / plain["steps_per_s"], 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "learned": learned,
        "prescribed_cilia": cilia,
        "free_body": body,
        "free_swimmer": swimmer,
        "ciliated_swimmer": ciliated,
        "orbit_ensemble": ensemble,
        "filament_ensemble": fil_ensemble,
        "overlap": not args.serial,
        "serial": ser,
        "one_directional_pair_kernel": sym_ab,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
