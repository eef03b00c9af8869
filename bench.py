"""Benchmark: particle-steps/s of the 3D GSPH Sedov step at 1M particles.

One step = TimeStep, predict, tree build, PreInteraction, FluidForce,
(GravityForce), correct -- the reference's Solver::integrate
(src/solver.cpp:966-978) -- on a device-resident state.  The timed region is
exactly K steps of sphgpu_integrate, which returns only when the stream has
drained; inputs are resident in HBM before it starts.

Prints one JSON line (the driver's contract), with these extra objects:
  roofline      the dominant kernel's achieved algorithmic rate (HIP events on
                the library's stream over the timed region) against the
                MI355X peak of the roof that bounds it;
  kernels       the same figures for every priced kernel, with the measured
                HBM traffic per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE,
                profiles/traffic.json) and the achieved HBM GB/s;
  accuracy      the energy-drift gate at the benchmark size: the oracle runs
                the same initialisation and warm-up + timed steps on the host
                (FP64, plain guess) and |drift_GPU - drift_oracle| is reported
                against the 1e-4 bar;
  evrard_1m     the self-gravitating workload (3D Evrard collapse, 1M
                particles, DISPH + tree gravity): step time, gravity's kernel
                time and interactions, energy drift;
  dropin        the reference's driver and Simulation with the GPU Module
                adapters in place of its modules (the module substitution),
                per-phase step time at the benchmark size;
  cpu_baseline  the reference's own CPU modules (oracle/_ref, compiled from
                its unmodified sources) timed on a bounded sample of the same
                workload on this host's cores; cpu_port is the oracle's time
                on the gate run, with its ratio to the reference.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "sph-code-simulation_amd"))

import numpy as np  # noqa: E402

import ics  # noqa: E402
import learned  # noqa: E402
import sphgpu  # noqa: E402

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# Algorithmic FLOP model (3D, +,-,*,compare = 1, FMA = 2, div and sqrt = 1),
# with the measured per-step counters: N_p accepted force pairs, k_R Riemann
# Newton iterations per pair, N_c PreInteraction candidates, k_h smoothing-
# length Newton iterations, N_m members (neighbour counts), N_r candidate
# tests of a per-particle tree walk (members of every leaf whose box the
# particle's own search sphere reaches):
#   FluidForce (GSPH 2nd order)  N_p (270 + 35 k_R)
#   PreInteraction (GSPH)        N_c (50 k_h + 60) + N_m 60 (MUSCL gradients)
#   search (PreInteraction)      N_r 8 (3 differences, 3 products, 2 adds,
#                                plus the compare counted in the 8)
FORCE_PAIR_FLOP = 270.0
RIEMANN_ITER_FLOP = 35.0
PRE_NEWTON_FLOP = 50.0
PRE_FINAL_FLOP = 60.0
PRE_GRAD_FLOP = 60.0
SEARCH_TEST_FLOP = 8.0
# the Sedov workloads' background pressure (the workload text says 1e-5)
P_BACKGROUND_NOTE = ("0.1: at the workload's 1e-5 the reference's own second-order GSPH drives "
                     "the cold particles beside the blast to negative energy and NaN in the first "
                     "steps (aborting in Periodic::calc_r_ij); it runs at 0.1 from 32^3 to 100^3")
ENERGY_GATE = 1e-4
PROFILES = REPO / "profiles"


def make_workload(name: str, form: str):
    guess = ics.GUESS_LEARNED if form == "learned" else ics.GUESS_PLAIN
    if name == "sedov_3d_gsph_1m":
        p, f = ics.sedov_3d(n1=100, riemann_guess=guess)
    elif name == "sedov_3d_gsph_262k":
        p, f = ics.sedov_3d(n1=64, riemann_guess=guess)
    elif name == "sedov_3d_gsph_32k":
        p, f = ics.sedov_3d(n1=32, riemann_guess=guess)
    elif name == "evrard_3d_disph_1m":
        # the reference's own Evrard lattice (src/sample/evrard.cpp) at
        # n_side 124: 998,592 particles; on a seeded radial-CDF draw the
        # reference itself goes NaN within the first steps from 100k
        # particles on (tests/test_oracle_reference.py)
        p, f = ics.evrard_lattice_3d(124)
    else:
        raise SystemExit(f"unknown workload {name}")
    p["pair_precision"] = ics.PAIR_FP64 if form == "fp64" else ics.PAIR_FP32
    return p, f


def train_learned(log_workload: str = "sedov_3d_gsph_262k", log_steps: int = 4) -> tuple[np.ndarray, dict]:
    """Learned p* guess: log the pair states the GSPH solver sees on the 64^3
    Sedov run (plain guess, FP32 pairs; every 6th of the ~60M pair solves of 4
    steps, ~10M rows) and train the network on the GPU with the library's
    bf16 MFMA forward/backward kernels (untimed preparation)."""
    p, f = make_workload(log_workload, "fp32")
    t0 = time.perf_counter()
    rows = learned.log_pairs(p, f, log_steps, capacity=10_000_000, stride=6)
    t1 = time.perf_counter()
    w, info = learned.train(p, rows, holdout=0.2)
    info.update(log_workload=log_workload, log_steps=log_steps, log_stride=6, log_s=round(t1 - t0, 2))
    return w, info


def run_form(name: str, form: str, steps: int, warmup: int, mlp_params=None, kernel_table: bool = True):
    p, f = make_workload(name, form)
    n = f.shape[1]
    ctx = sphgpu.Context(p, n)
    ctx.upload(f)
    if mlp_params is not None:
        ctx.mlp_set_params(mlp_params)
    ctx.initialize()
    e_init = ctx.energy_momentum()
    if warmup:
        ctx.integrate(warmup)
    e0 = ctx.energy_momentum()
    ctx.reset_stats()
    # the timed steps carry HIP events around the pair kernels only (each
    # event adds a gap to the stream)
    ctx.enable_timing(2)
    t0 = time.perf_counter()
    ctx.integrate(steps)
    t1 = time.perf_counter()
    times = ctx.kernel_times()
    ctx.enable_timing(0)
    stats = ctx.stats()
    e1 = ctx.energy_momentum()
    sim_t = ctx.time
    # members per particle (the neighbour counts of the last PreInteraction)
    members = float(ctx.download()[ics.F_NEIGHBOR].sum())
    # a further pass of the same length with events around every kernel
    # group: the per-kernel table of the whole step (not the headline)
    times_all, t2, t3 = times, 0.0, 0.0
    if kernel_table:
        ctx.enable_timing(1)
        t2 = time.perf_counter()
        ctx.integrate(steps)
        t3 = time.perf_counter()
        times_all = ctx.kernel_times()
        ctx.enable_timing(0)
    ctx.close()
    wall = t1 - t0
    return dict(n=n, wall=wall, value=n * steps / wall, times=times, times_all=times_all, stats=stats, e0=e0, e1=e1,
                time=sim_t, members=members, ms_all_events=(t3 - t2) / steps * 1e3,
                drift=float((e1[3] - e_init[3]) / abs(e_init[3])))


def evrard_entry(steps: int, warmup: int, kernel_table: bool = True) -> dict:
    """The self-gravitating workload: 3D Evrard collapse, 998,592 particles
    on the reference's lattice,
    DISPH with time-dependent viscosity, Wendland C4, tree gravity at
    theta = 0.5 (FP32 pairs); whole steps with gravity's own kernel time and
    the energy drift over the warm-up and timed steps."""
    res = run_form("evrard_3d_disph_1m", "fp32", steps, warmup, kernel_table=kernel_table)
    st, t = res["stats"], res["times"]
    return {"workload": "evrard_3d_disph_1m", "particles": res["n"], "steps": steps, "warmup": warmup,
            "value": round(res["value"], 1), "unit": "particle-steps/s",
            "ms_per_step": round(res["wall"] / steps * 1e3, 4), "sim_time": res["time"],
            "gravity_ms": round(t["gravity"][0], 4) if "gravity" in t else None,
            "gravity_interactions_per_particle": round(st["gravity_interactions"] / (res["n"] * steps), 1),
            "energy_drift": res["drift"],
            "kernel_ms": {k: round(v[0], 4) for k, v in res["times_all"].items()}}


def evrard_full_run(t_end: float = 0.8, wall_cap_s: float = 120.0) -> dict:
    """The Evrard workload as a whole run: 998,592 particles from t = 0 to
    t_end (through the collapse, FP32 pairs, FP64 integration), with the
    conservation of total energy and momentum over the run.  Steps go in
    batches of 20 and the run stops early (reported) past wall_cap_s."""
    p, f = make_workload("evrard_3d_disph_1m", "fp32")
    ctx = sphgpu.Context(p, f.shape[1])
    ctx.upload(f)
    ctx.initialize()
    e0 = ctx.energy_momentum()
    steps = 0
    t0 = time.perf_counter()
    while ctx.time < t_end and time.perf_counter() - t0 < wall_cap_s:
        ctx.integrate(20)
        steps += 20
    e1 = ctx.energy_momentum()
    wall = time.perf_counter() - t0
    t_sim = ctx.time
    ctx.close()
    return {"t_end": t_end, "t_reached": t_sim, "steps": steps, "wall_s": round(wall, 2),
            "ms_per_step": round(wall / max(steps, 1) * 1e3, 3),
            "energy_start": [float(v) for v in e0[:4]], "energy_end": [float(v) for v in e1[:4]],
            "energy_drift": float((e1[3] - e0[3]) / abs(e0[3])),
            "momentum_end": [float(v) for v in e1[4:7]],
            "note": "energies are kinetic, thermal, potential, total; at 14,328 particles the same run "
                    "tracks the oracle to t = 0.8 (tests/test_gpu_long_runs.py)"}


def dropin_entry(workload: str, steps: int = 2) -> dict:
    """The module substitution a reference user makes: the reference's own
    Simulation and driver sequence (Solver::integrate, src/solver.cpp:966-978,
    with its host predict/correct and host BHTree build) with the GPU Module
    adapters (sph-code-simulation_amd/dropin) in place of TimeStep,
    PreInteraction, FluidForce and GravityForce (oracle/_ref/libsphdropin_d3.so,
    the reference sources compiled with the adapters).  Each phase of a step
    is timed on the host, host-device transfers of the adapters included."""
    checkers, cores = _checkers()
    if not checkers.dropin_available(3):
        return None
    p, f = make_workload(workload, "fp32")
    d = checkers.DropIn(p, f)
    t0 = time.perf_counter()
    d.initialize()
    t_init = time.perf_counter() - t0
    phases = ("timestep", "predict", "make_tree", "pre_interaction", "fluid_force", "gravity", "correct")
    acc = {k: 0.0 for k in phases}
    for _ in range(steps):
        for k in phases:
            t = time.perf_counter()
            getattr(d, k)()
            acc[k] += time.perf_counter() - t
    d.close()
    step_ms = sum(acc.values()) / steps * 1e3
    gpu_ms = sum(acc[k] for k in ("timestep", "pre_interaction", "fluid_force", "gravity")) / steps * 1e3
    n = f.shape[1]
    return {"workload": workload, "particles": n, "steps": steps, "form": "fp32 pairs, plain guess",
            "initialize_s": round(t_init, 2), "ms_per_step": round(step_ms, 1),
            "value": round(n / (step_ms * 1e-3), 1), "unit": "particle-steps/s",
            "gpu_modules_ms_per_step": round(gpu_ms, 1),
            "phase_ms": {k: round(v / steps * 1e3, 1) for k, v in acc.items()},
            "host_cores": cores,
            "note": "host phases (predict, make_tree, correct) are the reference's own code on the host"}


def measured_traffic(kernel: str, workload: str, form: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json), scaled by the
    copy-kernel calibration recorded there; None when absent."""
    f = PROFILES / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        e = d["kernels"][kernel]
        if d.get("workload") != workload or d.get("form") != form:
            return None
        return float(e["traffic_bytes"])
    except (KeyError, ValueError, TypeError):
        return None


def kernel_models(res: dict, steps: int, workload: str, form: str) -> dict:
    """Per priced kernel: average launch time (HIP events), algorithmic work
    per launch from the FLOP model and the run's own counters, the achieved
    rate against the FP32 (or FP64) vector peak, and the measured HBM traffic
    per launch with the achieved HBM bandwidth."""
    st = res["stats"]
    times = res["times"]
    n = res["n"]
    pairs = st["force_pairs"] / steps
    k_r = st["riemann_iters"] / max(1, st["riemann_calls"])
    cand = st["pre_candidates"] / steps
    k_h = st["pre_newton_iters"] / max(1, n * steps)
    work = {
        "fluid_force": pairs * (FORCE_PAIR_FLOP + RIEMANN_ITER_FLOP * k_r),
        "pre_interaction": cand * (PRE_NEWTON_FLOP * k_h + PRE_FINAL_FLOP) + res["members"] * PRE_GRAD_FLOP,
    }
    if st.get("search_reach_tests"):
        searches = max(1, times.get("search_pre", (0.0, 0))[1])
        work["search_pre"] = st["search_reach_tests"] / searches * SEARCH_TEST_FLOP
    peak = PEAK_FP32_TFLOPS if form != "fp64" else PEAK_FP64_TFLOPS
    out = {}
    for k, flop in work.items():
        if k not in times:
            continue
        avg_ms = times[k][0]
        achieved = flop / (avg_ms * 1e-3) / 1e12
        e = {"avg_ms": round(avg_ms, 4), "launches_per_step": round(times[k][1] / steps, 3),
             "bound": "valu", "flop_per_launch": flop, "achieved_tflops": round(achieved, 3),
             "peak_tflops": peak, "frac": round(achieved / peak, 4)}
        tr = measured_traffic(k, workload, form)
        if tr is not None:
            e["traffic_bytes"] = tr
            e["hbm_gbps"] = round(tr / (avg_ms * 1e-3) / 1e9, 1)
            e["hbm_frac"] = round(tr / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
        out[k] = e
    return out


# compulsory HBM bytes per particle-step of the whole step (each particle's
# fields moved once per pass: time step, predict, tree keys, sort and
# permutation, nodes, PreInteraction, FluidForce, correct), the upper end of
# the estimate 0.94-1.4 kB
STEP_COMPULSORY_BYTES = 1400.0


def path_roofline(kernels: dict, res: dict, steps: int, form: str) -> dict:
    """Whole-step bound: the priced pair kernels at the vector peak plus the
    step's compulsory HBM bytes at the HBM peak, over the measured step."""
    peak = PEAK_FP32_TFLOPS if form != "fp64" else PEAK_FP64_TFLOPS
    t_valu = sum(e["flop_per_launch"] * e["launches_per_step"] for e in kernels.values()) / (peak * 1e12)
    t_hbm = STEP_COMPULSORY_BYTES * res["n"] / (PEAK_HBM_GBS * 1e9)
    t_meas = res["wall"] / steps
    return {"t_bound_ms": round((t_valu + t_hbm) * 1e3, 4), "t_measured_ms": round(t_meas * 1e3, 4),
            "frac": round((t_valu + t_hbm) / t_meas, 4),
            "note": ("T_bound = sum over the pair kernels of algorithmic FLOP / vector peak + %d B per "
                     "particle-step of compulsory HBM traffic / 8 TB/s" % int(STEP_COMPULSORY_BYTES))}


def roofline(kernels: dict) -> dict:
    """The dominant priced kernel (largest time per step)."""
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches_per_step"])
    e = kernels[dom]
    return {
        "kernel": dom, "bound": "valu", "achieved": e["achieved_tflops"], "peak": e["peak_tflops"],
        "unit": "TFLOP/s", "frac": e["frac"], "traffic": e.get("traffic_bytes"),
        "flop_per_launch": e["flop_per_launch"], "avg_ms": e["avg_ms"],
        "note": ("VALU-issue-bound pair/search loop, algorithmic FLOP from the measured counters, priced "
                 "against the MI355X vector peak (packed FP32 instructions are not used, so the attainable "
                 "FP32 rate is half of it)"),
    }


def _checkers():
    sys.path.insert(0, str(REPO / "tests"))
    import checkers  # noqa: E402
    import ctypes
    lib = ctypes.CDLL(str(REPO / "oracle" / "liboracle.so"))
    lib.orc_num_threads.restype = ctypes.c_int
    return checkers, lib.orc_num_threads()


def accuracy_gate(workload: str, steps: int, drift_gpu: float) -> dict:
    """The oracle (FP64, plain guess, reference arithmetic) runs the same
    initialisation and `steps` steps of the workload on the host; the GPU's
    energy drift over the same steps must be within ENERGY_GATE of it."""
    checkers, cores = _checkers()
    p, f = make_workload(workload, "fp64")
    o = checkers.Oracle(p, f)
    t0 = time.perf_counter()
    o.initialize()
    e0 = o.energy()
    t1 = time.perf_counter()
    o.integrate(steps)
    t2 = time.perf_counter()
    e1 = o.energy()
    o.close()
    drift_o = float((e1[3] - e0[3]) / abs(e0[3]))
    diff = abs(drift_gpu - drift_o)
    return {"steps": steps, "drift_gpu": drift_gpu, "drift_oracle": drift_o, "abs_diff": diff,
            "bar": ENERGY_GATE, "pass": bool(diff <= ENERGY_GATE),
            "oracle_s_init": round(t1 - t0, 2), "oracle_s_per_step": round((t2 - t1) / steps, 3),
            "oracle_cores": cores, "n": f.shape[1]}


def cpu_reference(workload: str, steps: int = 3) -> dict:
    """The reference's own modules and Simulation (oracle/_ref/libsphref_d3.so:
    src/bhtree.cpp, src/pre_interaction.cpp, src/gsph/*.cpp, src/timestep.cpp,
    ... compiled unmodified with g++ -O2 -fopenmp -ffp-contract=off, driven
    by the Solver::initialize / Solver::integrate sequence), OpenMP on all
    host cores: initialisation, then `steps` timed steps."""
    checkers, cores = _checkers()
    if not checkers.ref_available(3):
        return None
    p, f = make_workload(workload, "fp64")
    r = checkers.Reference(p, f)
    t0 = time.perf_counter()
    r.initialize()
    t1 = time.perf_counter()
    r.integrate(steps)
    t2 = time.perf_counter()
    r.close()
    n = f.shape[1]
    return {"value": round(n * steps / (t2 - t1), 1), "unit": "particle-steps/s", "cores": cores,
            "kind": "reference",
            "sample": (f"reference modules (oracle/_ref/libsphref_d3.so, g++ -O2 -fopenmp -ffp-contract=off), "
                       f"{workload} ({n} particles), FP64 plain guess: initialise {t1 - t0:.1f} s, "
                       f"then {steps} steps timed, {t2 - t1:.1f} s")}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="sedov_3d_gsph_1m")
    ap.add_argument("--form", default="learned", choices=["fp32", "fp64", "learned"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-evrard", action="store_true")
    ap.add_argument("--no-evrard-run", action="store_true", help="skip the Evrard run from t = 0 to 0.8")
    ap.add_argument("--no-kernel-table", action="store_true",
                    help="skip the second pass that times every kernel group (profiler runs: the last "
                         "launches are then the timed steps)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        raise SystemExit(2)

    mlp_params, train_info = None, None
    if args.form == "learned":
        mlp_params, train_info = train_learned()
    res = run_form(args.workload, args.form, args.steps, args.warmup, mlp_params,
                   kernel_table=not args.no_kernel_table)
    kernels = kernel_models(res, args.steps, args.workload, args.form)
    rl = roofline(kernels)
    out = {
        "metric": "particle-steps/s at 1M 3D GSPH: learned-FP32 vs plain, energy drift <=1e-4",
        "value": round(res["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.form != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": res["n"], "form": args.form,
                   "sph": "gsph-2nd-order", "riemann": "iterative", "kernel": "cubic_spline",
                   "neighbor_number": 50, "leaf": 32, "periodic": True,
                   "p_background": ics.SEDOV_P_BACKGROUND, "p_background_reason": P_BACKGROUND_NOTE},
        "roofline": rl,
        "kernels": kernels,
        "path_roofline": path_roofline(kernels, res, args.steps, args.form),
    }
    st = res["stats"]
    # tree searches per step: the pre-interaction's, plus the FluidForce one
    # when its pair lists could not be derived from the pre-interaction's
    searches = sum(res["times"].get(k, (0.0, 0))[1] for k in ("search_pre", "search_force")) / args.steps
    per_wave = 64 / (max(searches, 1.0) * res["n"] * args.steps)
    out["counters"] = {
        "riemann_iters_per_call": round(st["riemann_iters"] / max(1, st["riemann_calls"]), 4),
        "riemann_wave_max_per_slot": round(st["riemann_wave_max_iters"] * 64 / max(1, st["riemann_calls"]), 4),
        "riemann_fallbacks": st["riemann_fallbacks"], "riemann_learned": st["riemann_learned"],
        "force_pairs_per_particle": round(st["force_pairs"] / (res["n"] * args.steps), 3),
        "pre_candidates_per_particle": round(st["pre_candidates"] / (res["n"] * args.steps), 3),
        "h_newton_iters_per_particle": round(st["pre_newton_iters"] / (res["n"] * args.steps), 3),
        "tree_nodes": st["nodes"],
        "search_node_visits_per_wave": round(st["search_node_visits"] * per_wave, 1),
        "search_member_tests_per_wave": round(st["search_member_tests"] * per_wave, 1),
        "search_windows_per_wave": round(st["search_windows"] * per_wave, 1),
        "search_leaves_per_wave": round(st["search_leaves"] * per_wave, 1),
        "search_fp64_decisions_per_particle": round(st["search_exact"] / (res["n"] * args.steps), 4),
        "searches_per_step": round(searches, 3),
        "force_list_searches": st["flist_searches"],
        "h_beyond_search_radius": st["flist_h_beyond"],
    }
    if st.get("search_reach_tests"):
        out["counters"]["search_reach_tests_per_particle"] = round(
            st["search_reach_tests"] / (max(searches, 1.0) * res["n"] * args.steps), 2)
    if train_info is not None:
        out["training"] = train_info
    out["energy_drift"] = res["drift"]
    out["kernel_ms"] = {k: round(v[0], 4) for k, v in res["times_all"].items()}
    if not args.no_kernel_table:
        out["kernel_ms_note"] = ("every kernel group, from a further pass of the same number of steps with events "
                                 "around each group (that pass ran at %.4f ms per step); the timed steps carry "
                                 "events around the pair kernels only" % res["ms_all_events"])
    if not args.no_fp64 and args.form != "fp64":
        r64 = run_form(args.workload, "fp64", args.steps, args.warmup)
        out["plain_fp64"] = {"value": round(r64["value"], 1), "ms_per_step": round(r64["wall"] / args.steps * 1e3, 4),
                             "ratio": round(res["value"] / r64["value"], 3), "energy_drift": r64["drift"],
                             "drift_difference": abs(res["drift"] - r64["drift"]),
                             "riemann_iters_per_call": round(r64["stats"]["riemann_iters"]
                                                             / max(1, r64["stats"]["riemann_calls"]), 4),
                             "riemann_wave_max_per_slot": round(r64["stats"]["riemann_wave_max_iters"] * 64
                                                                / max(1, r64["stats"]["riemann_calls"]), 4),
                             "kernel_ms": {k: round(v[0], 4) for k, v in r64["times_all"].items()}}
    if args.form == "learned" and not args.no_fp64:
        # the same FP32 pair kernels without the network: what the learned
        # start changes on this workload
        r32 = run_form(args.workload, "fp32", args.steps, args.warmup)
        st32 = r32["stats"]
        out["plain_fp32"] = {
            "value": round(r32["value"], 1), "ms_per_step": round(r32["wall"] / args.steps * 1e3, 4),
            "learned_over_plain": round(res["value"] / r32["value"], 4),
            "riemann_iters_per_call": round(st32["riemann_iters"] / max(1, st32["riemann_calls"]), 4),
            "riemann_wave_max_per_slot": round(st32["riemann_wave_max_iters"] * 64 / max(1, st32["riemann_calls"]), 4),
            "fluid_force_ms": round(r32["times"]["fluid_force"][0], 4) if "fluid_force" in r32["times"] else None,
            "note": ("the network starts the pairs its gate marks hard (pressure ratio above 4 or a velocity jump "
                     "above the mean sound speed): on this workload a few per mille of the pairs, which do not "
                     "set the wavefront's largest Newton count; its gain shows on the hard pairs themselves "
                     "(training.held_out) and not in the step time")}
    if not args.no_fp64 and args.workload == "sedov_3d_gsph_1m":
        # the 64^3 Sedov (262,144 particles), plain guess: FP32 against FP64
        c3 = {}
        for form in ("fp32", "fp64"):
            r = run_form("sedov_3d_gsph_262k", form, args.steps, args.warmup, kernel_table=False)
            c3[form] = {"value": round(r["value"], 1), "ms_per_step": round(r["wall"] / args.steps * 1e3, 4),
                        "energy_drift": r["drift"]}
        c3["ratio"] = round(c3["fp32"]["value"] / c3["fp64"]["value"], 3)
        c3["particles"] = 262144
        out["sedov_262k"] = c3
    if not args.no_evrard and args.workload != "evrard_3d_disph_1m":
        out["evrard_1m"] = evrard_entry(args.steps, args.warmup, not args.no_kernel_table)
        if not args.no_evrard_run:
            out["evrard_1m"]["whole_run"] = evrard_full_run()
    # plain copy of known size: attainable HBM bandwidth, and the calibration
    # kernel (k_copy_probe) for the FETCH_SIZE / WRITE_SIZE counters
    pc, _ = ics.sedov_3d(n1=4)
    probe = sphgpu.Context(pc, 64)
    nbytes = 1 << 30
    ms = probe.copy_probe(nbytes)
    probe.close()
    out["hbm_copy_probe"] = {"bytes_each_way": nbytes, "ms": round(ms, 4),
                             "GBps_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}
    if not args.no_cpu:
        # the energy gate at the benchmark size (the GPU drift covers the
        # warm-up and timed steps from the initialised state), then the
        # reference's own CPU path on a bounded sample of the same workload
        gate = accuracy_gate(args.workload, args.warmup + args.steps, res["drift"])
        out["accuracy"] = gate
        ref = cpu_reference(args.workload)
        port_value = round(gate["n"] / gate["oracle_s_per_step"], 1)
        out["cpu_port"] = {"value": port_value, "unit": "particle-steps/s", "cores": gate["oracle_cores"],
                           "kind": "port",
                           "sample": f"oracle (oracle/sph_oracle.cpp, OpenMP), {args.workload}, the "
                                     f"{gate['steps']} steps of the accuracy gate"}
        di = dropin_entry(args.workload)
        if di is not None:
            out["dropin"] = di
        if ref is not None:
            out["cpu_baseline"] = ref
            out["cpu_port"]["ratio_to_reference"] = round(port_value / ref["value"], 3)
            out["gpu_over_reference"] = round(res["value"] / ref["value"], 1)
        else:
            out["cpu_baseline"] = dict(out["cpu_port"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
