"""Benchmark on one MI355X: two-site DMRG steps of the C60 Heisenberg model
(HeisenbergSU2 with the full coupling matrix of the truncated icosahedron,
M = 4096 SU(2) multiplets), mixed precision against plain FP64, at the same
energies.

Workload (default, --workload c60): the 90 nearest-neighbour bonds of C60
(J = 1) in Cuthill-McKee chain order, the SU(2)-reduced MPO at its minimal
bond dimension (2 + rank of the couplings across each bond, at most 12
multiplets), total spin 0.  The state: DMRG from a random start, one full
sweep at each M of --warm (32, 64, 128, 256 multiplets), then every bond
expanded towards M = 4096 with orthonormal completions of the old basis (the
state unchanged; see vmps_amd.dmrg.expand_chain), environments rebuilt on
the device.  Synthetic data: no checkpoint is read.

One step = one two-site DMRG step sweeping right through the middle of the
chain (DmrgSolver::iteration_two): contract_AA2, the lowest eigenpair of
H_eff to tol_eigval 1e-11 / residual 1e-7 (--tol-eigval, --tol-state),
the split with truncation to M,
and the environment update.  --warmup W steps run first (untimed), then
--steps K steps are timed.  The sweep follows the mixed steps; before each,
the same step runs in plain FP64 from the same state on a copy of the chain
(timed on its own and dropped), so both legs solve the same effective
Hamiltonians from the same start vectors:
  mixed: split-BF16x3 Lanczos (FP64 vectors, dots and reorthogonalisation)
         refined to the FP64 tolerances by FP64-residual corrections;
  fp64:  plain FP64 Lanczos.
`value` is the mixed leg's throughput: algorithmic HxV flops (F_alg per HxV
call, every call of the step) per second of step wall time, split, host work
and environment update included.  `dmrg` holds both legs: per-step energies,
the largest |E_mixed - E_fp64| over the timed steps, HxV calls, and the step
time of each phase.

`roofline`: the dominant kernel of the mixed HxV on the first timed pivot,
its share of the staged work over its average launch time (HIP events on the
plan's stream), against 2.5 PFLOP/s / 3 (dense BF16 MFMA, three products per
FP32-equivalent one).  `cpu_baseline`: the oracle (the reference's literal
per-term (L.V).R loop in C, OpenMP over output blocks, all host cores) on
that pivot, one HxV, plus a single-core sample and the host's dgemm rate.

Other workloads (--workload NAME of vmps_amd.synth.WORKLOADS / BATCHES) time
the HxV alone on a random-environment pivot (one step = one HxV).

Prints one JSON line.  --gpus N with N > 1 is not supported (one GPU).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}
PEAK_HBM_GBS = 8000.0


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(pds, vins, falgs, budget_s, threads):
    """Time the oracle (CPU FP64 restatement of the reference loop) on the
    same pivots: HxV calls over the pivots in turn until about budget_s
    seconds have passed.  Returns the rate over the work actually done and
    the oracle's outputs for the pivots it reached."""
    import oracle
    from vmps_amd import synth
    lib = oracle.load()
    P = ctypes.POINTER(ctypes.c_double)
    outs, flops_done, flops_ref_done, calls = {}, 0.0, 0.0, 0
    handles, frefs = {}, {}
    t_used = 0.0
    i = 0
    while True:
        n = i % len(pds)
        pd, vin, falg = pds[n], vins[n], falgs[n]
        if n not in handles:  # each pivot is built once (the build is not timed)
            h = synth.feed(lib, "oracle_", pd, oracle.check)
            oracle.check(lib.oracle_pivot2_precalc(h))
            fref = ctypes.c_double()
            oracle.check(lib.oracle_flops_ref(h, ctypes.byref(fref)))
            handles[n], frefs[n] = h, fref.value
        out = np.zeros_like(vin)
        t0 = time.perf_counter()
        oracle.check(lib.oracle_hxv_omp(handles[n], vin.ctypes.data_as(P), out.ctypes.data_as(P), threads))
        t_used += time.perf_counter() - t0
        outs.setdefault(n, out)
        flops_done += falg
        flops_ref_done += frefs[n]
        calls += 1
        i += 1
        if t_used >= budget_s and calls >= 2:
            break
    for h in handles.values():
        lib.oracle_pivot2_destroy(h)
    return {
        "value": flops_done / t_used / 1e9,
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d HxV calls of the oracle (literal per-term (L.V).R loop, OpenMP over output blocks, %s) "
                  "over %d of the workload's %d pivot(s); %.3f s of CPU time; literal loop rate %.2f GFLOP/s"
                  % (calls, lib.variant, min(calls, len(pds)), len(pds), t_used, flops_ref_done / t_used / 1e9),
        "ms_per_hxv": t_used / calls * 1e3,
    }, outs


def sweep_bench(args, model="HeisenbergU1", L=32, M=64, sweeps=4):
    """Full two-site DMRG sweeps of a chain from the same random MPS, FP64 and
    mixed (BF16x3 Lanczos refined in FP64): energies, the energy difference
    and the wall time per sweep."""
    from vmps_amd import dmrg
    params = {"t": 1.0, "U": 8.0} if "Hubbard" in model else {"J": 1.0}
    out = {"model": model, "L": L, "M": M, "sweeps": sweeps}
    for prec in ("fp64", "bf16x3"):
        be = dmrg.GpuBackend(precision=prec, tol_eigval=1e-10, tol_state=1e-6)
        r = dmrg.run_dmrg(model, L, M, sweeps, seed=args.seed, backend=be, params=params)
        out[prec] = {"energy": r["energy"], "sweep_seconds": r["sweep_seconds"], "hxv_calls": be.hxv_calls,
                     "mixed_solves": be.mixed_solves, "max_truncation": max(r["truncation"][-(L - 1):])}
    out["dE_mixed_vs_fp64"] = out["bf16x3"]["energy"] - out["fp64"]["energy"]
    out["sweep_speedup_mixed_vs_fp64"] = out["fp64"]["sweep_seconds"][-1] / out["bf16x3"]["sweep_seconds"][-1]
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_pivot(pd, falg, vin, threads, single_seconds=4.0):
    """The oracle on one pivot: one HxV with OpenMP over output blocks on
    `threads` cores (the rate in F_alg units, so it compares with the GPU
    figures), a bounded single-core sample (the literal loop over the first
    output entries until single_seconds have passed, its rate and the
    projected single-core time of a whole HxV) and the host's single-core
    dgemm rate on 512^3.  Returns (record, the oracle's HxV output)."""
    import oracle
    from vmps_amd import synth
    import copy
    lib = oracle.load()
    P = ctypes.POINTER(ctypes.c_double)
    pd = copy.copy(pd)  # the oracle takes host environments
    pd.L = pd.L.to_host() if hasattr(pd.L, "handle") else pd.L
    pd.R = pd.R.to_host() if hasattr(pd.R, "handle") else pd.R
    h = synth.feed(lib, "oracle_", pd, oracle.check)
    oracle.check(lib.oracle_pivot2_precalc(h))
    fref = ctypes.c_double()
    oracle.check(lib.oracle_flops_ref(h, ctypes.byref(fref)))
    out = np.zeros_like(vin)
    t0 = time.perf_counter()
    oracle.check(lib.oracle_hxv_omp(h, vin.ctypes.data_as(P), out.ctypes.data_as(P), threads))
    t_omp = time.perf_counter() - t0
    # single core: output entries in order until the budget is spent
    nl = lib.oracle_nlhs(h)
    scratch = np.zeros_like(vin)
    q, t_single, f_single = 0, 0.0, 0.0
    while q < nl and t_single < single_seconds:
        q1 = min(nl, q + max(1, nl // 64))
        fr = ctypes.c_double()
        oracle.check(lib.oracle_flops_ref_range(h, q, q1, ctypes.byref(fr)))
        t0 = time.perf_counter()
        oracle.check(lib.oracle_hxv_range(h, vin.ctypes.data_as(P), scratch.ctypes.data_as(P), q, q1))
        t_single += time.perf_counter() - t0
        f_single += fr.value
        q = q1
    lib.oracle_pivot2_destroy(h)
    single_rate = f_single / t_single if t_single > 0 else 0.0
    dgemm = lib.oracle_dgemm_gflops(512, 3)
    return {
        "value": falg / t_omp / 1e9,
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "one HxV of the oracle (literal per-term (L.V).R loop of DmrgPivotMatrix2.h:161-231 in C, "
                  "OpenMP over output blocks, %s) on the first timed C60 pivot: %.2f s; rate in F_alg units "
                  "(the literal loop does %.2fx more flops: %.1f GFLOP/s of its own work)"
                  % (lib.variant, t_omp, fref.value / falg, fref.value / t_omp / 1e9),
        "seconds_per_hxv": t_omp,
        "cpu_model": cpu_model(),
        "single_core": {"literal_loop_gflops": single_rate / 1e9,
                        "sample": "%d of %d output entries, %.2f s" % (q, nl, t_single),
                        "projected_seconds_per_hxv": fref.value / single_rate if single_rate else None},
        "dgemm_512_gflops_single_core": dgemm,
    }, out


def lanczos_iteration_ms(workload, seed, iters):
    """Wall time of one FP64 Lanczos iteration (HxV, re-orthogonalisation,
    normalisation; replayed from a HIP graph, the scalars read by the host
    every few iterations) averaged over `iters` iterations with the
    tolerances at 0, from the same start vector as a first untimed solve that
    captures the graph.  It runs on the Hermitian form of the workload's
    pivot (environments symmetrised through the MPO, as an effective
    Hamiltonian is): the partial re-orthogonalisation reads the loss of
    orthogonality off the Lanczos coefficients, which holds for a symmetric
    operator only."""
    from vmps_amd import synth
    from vmps_amd.pivot import PivotMatrix2, lanczos
    pd = synth.make_workload(workload, seed=seed, hermitian=True)
    H = PivotMatrix2(pd, precision="fp64")
    try:
        H.autotune(iters=10)
        v0 = np.random.default_rng(1).uniform(-1, 1, pd.vsize)
        lanczos(H, v0=v0, max_iter=8, tol_eigval=0.0, tol_state=0.0)
        t0 = time.perf_counter()
        _, r = lanczos(H, v0=v0, max_iter=iters, tol_eigval=0.0, tol_state=0.0)
        return (time.perf_counter() - t0) / r["iterations"] * 1e3
    finally:
        H.close()


def _c2_hxv_lines(args):
    """HubbardU1xU1 L=48 M=512 random-environment pivot: HxV time and rate in
    FP64, FP32 and split-BF16x3 (autotuned), a few seconds in all."""
    from vmps_amd import synth
    from vmps_amd.pivot import PivotMatrix2
    pd = synth.make_workload("hubbard_u1xu1_L48_M512", seed=args.seed)
    out = {"workload": "hubbard_u1xu1_L48_M512"}
    for prec in ("fp64", "fp32", "bf16x3"):
        H = PivotMatrix2(pd, precision=prec)
        st = H.stats()
        H.autotune(iters=10)
        ms = H.time_hxv(warmup=20, iters=400)
        out[prec] = {"ms_per_hxv": ms, "gflops": st["flops_alg"] / (ms * 1e-3) / 1e9}
        if prec == "fp64":
            out["lanczos_ms_per_iteration"] = lanczos_iteration_ms("hubbard_u1xu1_L48_M512", args.seed, 60)
            out["lanczos_iterations"] = 60
        H.close()
    return out


def precision_report_c3(args, M=2048, steps=3):
    """HubbardSU2xU1 two-leg ladder 2 x 32 (snake order, t = 1, U = 6, N = 56,
    S = 0) at M = 2048 multiplets: two-site DMRG steps (from a state warmed up
    at small M and expanded to M) in FP64, FP32 and split-BF16x3, each mixed
    solve refined in FP64.  The sweep follows the FP64 steps; at every pivot
    the FP32 and split-BF16x3 steps start from the same state (copies of the
    chain), so that their differences are the precision's alone.  Per
    precision and step: the energy and its difference from FP64, the residual
    ||H psi - E psi|| (FP64 operator), the mixed solve's own Rayleigh quotient
    before refinement and its error, the truncation weight and its difference
    from FP64."""
    from vmps_amd import dmrg, models
    L, Q, p0 = 64, (1, 56), 30
    sites = models.chain_mpo("HubbardSU2xU1Ladder", L, t=1.0, U=6.0)
    tol = dict(tol_eigval=1e-11, tol_state=1e-7, max_iter=300)
    prep = dmrg.GpuBackend(precision="fp64", **tol)
    t0 = time.perf_counter()
    warm = dmrg.warm_chain(sites, Q, [32, 64, 128, 256], p0, prep, np.random.default_rng(args.seed + 3))
    big = dmrg.expand_chain(warm, M, p0, prep, np.random.default_rng(args.seed + 4))
    out = {"model": "HubbardSU2xU1 ladder 2x32 snake, t=1, U=6, N=56, S=0", "M": M, "pivots": [p0, p0 + steps - 1],
           "prep_seconds": time.perf_counter() - t0}
    legs = {prec: [] for prec in ("fp64", "fp32", "bf16x3")}
    be = {prec: dmrg.GpuBackend(precision=prec, mixed_min_flops=0, **tol) for prec in legs}
    c = big.copy()
    for i in range(steps):
        for prec in ("fp32", "bf16x3", "fp64"):  # the FP64 step last: the sweep continues from it
            chain = c if prec == "fp64" else c.copy()
            ts = time.perf_counter()
            E, tw, info = dmrg.dmrg_step(chain, be[prec], p0 + i, "right")
            legs[prec].append({"energy": E, "truncation": tw, "residual": info["residual"],
                               "seconds": time.perf_counter() - ts,
                               "energy_unrefined": info["energy_mixed"], "residual_unrefined": info["residual_mixed"],
                               "hxv_calls": info["hxv_calls"], "hxv_fp64": info["hxv_fp64"]})
    for prec in ("fp32", "bf16x3"):
        for a, b in zip(legs[prec], legs["fp64"]):
            a["abs_dE_vs_fp64"] = abs(a["energy"] - b["energy"])
            a["abs_dE_unrefined_vs_fp64"] = abs(a["energy_unrefined"] - b["energy"])
            a["d_truncation_vs_fp64"] = a["truncation"] - b["truncation"]
    out["legs"] = legs
    return out


def c2_sweeps(args, sweeps=3):
    """HubbardU1xU1 L=48 (t=1, U=8, half filling, Sz=0), M=512: full two-site
    DMRG sweeps in FP64 from a random MPS, energy per half sweep and time per
    sweep."""
    from vmps_amd import dmrg
    be = dmrg.GpuBackend(precision="fp64", tol_eigval=1e-10, tol_state=1e-6)
    r = dmrg.run_dmrg("HubbardU1xU1", 48, 512, sweeps, seed=args.seed, backend=be, params={"t": 1.0, "U": 8.0})
    return {"model": "HubbardU1xU1 L=48 t=1 U=8 N=48 Sz=0", "M": 512, "energy": r["energy"],
            "energies_per_half_sweep": r["energies"], "sweep_seconds": r["sweep_seconds"],
            "max_truncation_last_sweep": max(r["truncation"][-47:]), "hxv_calls": be.hxv_calls}


def dmrg_c60_bench(args):
    import torch
    from vmps_amd import dmrg, lattice, models, synth
    from vmps_amd.pivot import HxV, PivotMatrix2, clear_tile_cache

    assert torch.cuda.is_available(), "no HIP device"

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    L = 60
    J = lattice.c60_couplings()
    sites = models.chain_mpo("HeisenbergSU2Full", L, J=J)
    p0 = args.pivot
    nsteps = args.warmup + args.steps
    if p0 + nsteps > L - 1:
        raise SystemExit("bench.py: --pivot + --warmup + --steps must stay within the chain (<= %d)" % (L - 1))
    tol = dict(tol_eigval=args.tol_eigval, tol_state=args.tol_state, max_iter=300)
    prep = dmrg.GpuBackend(precision="fp64", **tol)
    t0 = time.perf_counter()
    warm = dmrg.warm_chain(sites, (1,), [int(x) for x in args.warm.split(",") if x], p0, prep,
                           np.random.default_rng(args.seed), log=log)
    t_warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    big = dmrg.expand_chain(warm, args.M, p0, prep, np.random.default_rng(args.seed + 1))
    t_expand = time.perf_counter() - t0
    log("state ready: warm-up %.1f s (E = %.10f), expansion to M=%d %.1f s" % (t_warm, warm.energy, args.M, t_expand))

    if args.hxv_reps:
        # profiling target: only HxV of the first timed pivot, mixed and FP64
        c = big.copy()
        for i in range(args.warmup):
            dmrg.dmrg_step(c, prep, p0 + i, "right")
        pd = c.pivot(p0 + args.warmup)
        v = synth.random_vector(pd, args.seed + 7)
        rec = {"workload": "c60_heisenberg_su2_M%d" % args.M, "pivot": p0 + args.warmup, "vsize": int(pd.vsize)}
        precs = {"both": (args.mixed, "fp64"), "mixed": (args.mixed,), "fp64": ("fp64",)}[args.hxv_precision]
        for prec in precs:
            H = PivotMatrix2(pd, precision=prec)
            if args.tiles:
                # a pinned configuration keeps the tuning candidates out of a profile
                we, bm, b64, xu = (int(x) for x in args.tiles.split(","))
                H.set_tiles(wave_edge=we, big_min=bm, big_min64=b64, explicit_u=xu)
                rec["tiles_" + prec] = H.tiles()
            else:
                rec["tiles_" + prec] = H.autotune(iters=2)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.hxv_reps):
                HxV(H, v)
            torch.cuda.synchronize()
            rec[prec + "_ms_per_hxv_incl_copies"] = (time.perf_counter() - t0) / args.hxv_reps * 1e3
            # device time of the stages (HIP events on the plan's stream): A, B, C, whole HxV
            rec[prec + "_stage_ms"] = [round(x, 4) for x in H.time_stages(warmup=2, iters=args.hxv_reps)]
            rec["flops_alg"] = H.stats()["flops_alg"]
            H.close()
        return rec

    # Each leg is one trajectory of consecutive steps from the same expanded
    # state, timed as a block: one device synchronisation before the first
    # timed step and one after the last, as in a sweep, where a step's
    # environment update still runs on the device while the host builds the
    # next pivot's plan.  A step's seconds run from its start to the next
    # step's start.  The precision check is separate: from the mixed
    # trajectory's state before each step, one FP64 control step (untimed)
    # solves the same effective Hamiltonian from the same start vector, so
    # |E_mixed - E_fp64| measures the precision alone.  Comparing the two
    # trajectories would not: at M = 4096 the kept set at each split is
    # sensitive to perturbations far below the truncation weight, and two
    # FP64 runs that differ only in the Lanczos tolerance drift apart by some
    # 1e-8 over 20 steps.
    def step_record(l, E, tw, info, tm, ts):
        return {"pivot": l, "timed": False, "energy": E, "truncation": tw, "start": ts,
                "phases": {k: round(v, 4) for k, v in tm.items()},
                "hxv_calls": info["hxv_calls"], "hxv_fp64": info["hxv_fp64"],
                "iterations": info["iterations"], "iterations_refine": info["iterations_refine"],
                "refine_steps": info["refine_steps"], "refine_cg": info["refine_cg"],
                "residual": info["residual"], "flops_alg": info["flops_alg"]}

    def run_leg(leg, chain, snapshots=None):
        clear_tile_cache()  # each leg pays for its own tile tuning, as a sweep of its own would
        recs = []
        for i in range(nsteps):
            l = p0 + i
            if snapshots is not None:
                snapshots.append(chain.copy())
            if i == args.warmup:
                torch.cuda.synchronize()
            tm = {}
            ts = time.perf_counter()
            E, tw, info = dmrg.dmrg_step(chain, be[leg], l, "right", timings=tm)
            recs.append(step_record(l, E, tw, info, tm, ts))
            recs[-1]["timed"] = i >= args.warmup
            log("%s step %d (pivot %d%s): E = %.12f  HxV %d (FP64 %d)  residual %.1e"
                % (leg, i, l, "" if i >= args.warmup else ", warm-up", E, info["hxv_calls"], info["hxv_fp64"],
                   info["residual"]))
        torch.cuda.synchronize()
        t_end = time.perf_counter()
        for a, b in zip(recs, recs[1:] + [None]):
            a["seconds"] = (b["start"] if b else t_end) - a["start"]
            del a["start"]
        return recs

    be = {"mixed": dmrg.GpuBackend(precision=args.mixed, mixed_min_flops=0, **tol)}
    if not args.no_fp64_leg:
        be["fp64"] = dmrg.GpuBackend(precision="fp64", **tol)
    snapshots = []
    steps = {"mixed": run_leg("mixed", big.copy(), snapshots)}
    first_pd = snapshots[args.warmup].pivot(p0 + args.warmup)
    control = []
    if "fp64" in be:
        steps["fp64"] = run_leg("fp64", big.copy())
        for i in range(args.warmup, nsteps):
            ctrl = snapshots[i]
            E, tw, info = dmrg.dmrg_step(ctrl, be["fp64"], p0 + i, "right")
            control.append({"pivot": p0 + i, "energy": E, "truncation": tw, "hxv_calls": info["hxv_calls"],
                            "residual": info["residual"]})
            snapshots[i] = None
            del ctrl
        torch.cuda.synchronize()
    del snapshots
    legs = {}
    for leg in be:
        timed = [st for st in steps[leg] if st["timed"]]
        t_timed = sum(st["seconds"] for st in timed)
        work = sum(st["flops_alg"] * st["hxv_calls"] for st in timed)
        phases = {}
        for st in timed:
            for k, v in st["phases"].items():
                phases[k] = phases.get(k, 0.0) + v
        legs[leg] = {"precision": be[leg].precision, "seconds_timed": t_timed,
                     "ms_per_step": t_timed / args.steps * 1e3,
                     "gflops": work / t_timed / 1e9, "hxv_calls": sum(st["hxv_calls"] for st in timed),
                     "hxv_fp64": sum(st["hxv_fp64"] for st in timed),
                     "phase_seconds": {k: round(v, 3) for k, v in phases.items()}, "steps": steps[leg]}
    mx = legs["mixed"]
    dmrg_rec = {"trajectory": "each leg one trajectory of consecutive steps from the same state, timed as a block; "
                              "the precision check from FP64 control steps (untimed) at the mixed trajectory's "
                              "state before each timed step",
                "legs": legs}
    if "fp64" in legs:
        f64 = legs["fp64"]
        mixed_timed = [st for st in mx["steps"] if st["timed"]]
        dE = [abs(a["energy"] - b["energy"]) for a, b in zip(mixed_timed, control)]
        dmrg_rec.update({"max_abs_dE_mixed_vs_fp64": max(dE), "dE_per_step": dE, "fp64_control_steps": control,
                         "step_speedup_mixed_vs_fp64": f64["seconds_timed"] / mx["seconds_timed"],
                         "fp64_ms_per_step": f64["ms_per_step"], "fp64_gflops": f64["gflops"]})

    if args.legs_only:  # profiling target: the DMRG steps and nothing after them
        return {"workload": "c60_heisenberg_su2_M%d" % args.M, "dmrg": dmrg_rec}

    # dominant kernel of the mixed HxV on the first timed pivot (HIP events on the plan's stream)
    H = PivotMatrix2(first_pd, precision=args.mixed)
    st = H.stats()
    tiles = H.autotune(iters=3)
    ms = H.time_stages(warmup=1, iters=5)
    stage_flops = {"A": st["flops_A"], "C": st["flops_C"]}
    dom = "A" if ms[0] >= ms[2] else "C"
    t_dom = ms[0] if dom == "A" else ms[2]
    achieved = stage_flops[dom] / (t_dom * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.mixed]
    big_tiles = tiles["big_min"] if args.mixed == "bf16x3" else tiles["big_min64"]
    kname = {"A": "k_stage_a", "C": "k_stage_c"}[dom] + ("_bf16x3" if args.mixed == "bf16x3" else "") + \
        ("_big (128x128 LDS tiles) + small-tile kernel" if big_tiles else "")
    hxv_mixed_ms = ms[3]
    H.close()
    H64 = PivotMatrix2(first_pd, precision="fp64")
    H64.autotune(iters=3)
    hxv_fp64_ms = H64.time_hxv(warmup=1, iters=5)
    vin = synth.random_vector(first_pd, args.seed + 7)
    out_gpu = HxV(H64, vin)
    H64.close()
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f).get("c60_M%d/%s" % (args.M, args.mixed), {})
        traffic = tr.get("k_stage_a" if dom == "A" else "k_stage_c")
    except (OSError, ValueError):
        pass

    res = {
        "metric": "DMRG HxV GFLOP/s + sweep time, mixed vs plain FP64 GPU, at |dE|<=1e-9: two-site DMRG step "
                  "throughput of the mixed path (F_alg of every HxV call per second of step wall time)",
        "value": round(mx["gflops"], 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (FP32 accumulation; FP64 vectors, dots and refinement)",
                  "fp32": "f32 (FP64 vectors, dots and refinement)"}[args.mixed],
        "data": "synthetic (C60 couplings; state from a random start, warm-up sweeps at small M, bonds expanded)",
        "config": {
            "workload": "c60_heisenberg_su2_M%d" % args.M,
            "model": "HeisenbergSU2, full coupling matrix of the C60 truncated icosahedron (90 bonds, J=1), "
                     "Cuthill-McKee chain order, S_total = 0",
            "L": L, "M": args.M, "D_W_max": max(W.dw_out for W in sites),
            "warm_up_M": args.warm, "first_pivot": p0, "timed_pivots": [p0 + args.warmup, p0 + nsteps - 1],
            "lanczos": "tol_eigval %g, residual %g (both legs; the mixed leg refined to them in FP64), "
                       "full reorthogonalisation" % (args.tol_eigval, args.tol_state),
            "first_timed_pivot": {"vsize": int(first_pd.vsize), "flops_alg": st["flops_alg"],
                                  "flops_ref": st["flops_ref"], "max_block_dim": st["max_block_dim"],
                                  "hxv_ms": {args.mixed: hxv_mixed_ms, "fp64": hxv_fp64_ms}},
            "parallelism": "single GPU",
        },
        "dmrg": dmrg_rec,
        "prep_seconds": {"warm_up": t_warm, "expansion": t_expand},
        "tiles": {k: tiles[k] for k in ("wave_edge", "big_min", "big_min64", "explicit_u")},
        "roofline": {
            "bound": "mfma",
            "kernel": kname,
            "stage": dom,
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "stage_ms": {"A": ms[0], "B": ms[1], "C": ms[2], "hxv": ms[3]},
            "peak_note": "dense BF16 MFMA 2.5 PFLOP/s / 3 products" if args.mixed == "bf16x3" else "FP32 MFMA",
        },
    }
    if not args.no_cpu:
        cb, out_cpu = cpu_baseline_pivot(first_pd, st["flops_alg"], vin, host_cores())
        res["cpu_baseline"] = cb
        res["parity_rel_err_vs_oracle"] = float(np.linalg.norm(out_gpu - out_cpu) / np.linalg.norm(out_cpu))
    extras = [x for x in args.extras.split(",") if x]
    if "c2" in extras:
        res["hxv_c2"] = _c2_hxv_lines(args)
    if "c1" in extras:
        res["sweeps_c1"] = sweep_bench(args, sweeps=4)
    if "c3" in extras:
        res["precision_c3"] = precision_report_c3(args)
    if "c2sweep" in extras:
        res["sweeps_c2"] = c2_sweeps(args)
    return res


def hxv_bench(args):
    """One step = one HxV of the workload's random-environment pivot (or of
    every pivot of a batched workload, one launch per stage) on
    device-resident vectors; the metric is F_alg / t."""
    import torch
    from vmps_amd import _lib, synth
    from vmps_amd.pivot import PivotBatch, PivotMatrix2

    assert torch.cuda.is_available(), "no HIP device"
    batched = args.workload in synth.BATCHES
    if batched:
        pds = synth.make_batch(args.workload, seed=args.seed)
        H = PivotBatch(pds, precision=args.precision)
    else:
        pds = [synth.make_workload(args.workload, seed=args.seed)]
        H = PivotMatrix2(pds[0], precision=args.precision)
    st = H.stats()
    if args.tiles:
        we, bm, b64, xu = (int(x) for x in args.tiles.split(","))
        H.set_tiles(wave_edge=we, big_min=bm, big_min64=b64, explicit_u=xu)
        args.no_tune = True
    tiles = H.autotune(iters=10) if not args.no_tune else H.tiles()
    vins = [synth.random_vector(pd, args.seed + 1 + i) for i, pd in enumerate(pds)]
    vin_h = np.concatenate(vins)
    vin = torch.from_numpy(vin_h).to("cuda")
    vout = torch.empty_like(vin)
    lib = H._lib

    def timed(plan, out):
        """Seconds per HxV over args.steps back-to-back calls on resident vectors."""
        def step():
            _lib.check(lib.vmps_hxv_device(plan, vin.data_ptr(), out.data_ptr(), None))
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    sec = timed(H._plan, vout)
    gflops = st["flops_alg"] / sec / 1e9
    out_gpu = vout.cpu().numpy()
    lanczos_ms = None
    if args.lanczos_iters and not batched:
        lanczos_ms = lanczos_iteration_ms(args.workload, args.seed, args.lanczos_iters)

    # the same workload in the mixed modes (split-BF16 x3 and FP32 MFMA with
    # FP32 accumulation), each autotuned, against the FP64 result
    mixed = {}
    if args.precision == "fp64" and not args.no_mixed:
        for prec in ("fp32", "bf16x3"):
            Hm = PivotBatch(pds, precision=prec) if batched else PivotMatrix2(pds[0], precision=prec)
            tm = Hm.autotune(iters=10) if not args.no_tune else Hm.tiles()
            om = torch.empty_like(vin)
            sm = timed(Hm._plan, om)
            o = om.cpu().numpy()
            mixed[prec] = {"gflops": st["flops_alg"] / sm / 1e9, "ms_per_hxv": sm * 1e3,
                           "speedup_vs_fp64": sec / sm,
                           "rel_err_vs_fp64": float(np.linalg.norm(o - out_gpu) / np.linalg.norm(out_gpu)),
                           "tiles": {k: tm[k] for k in ("wave_edge", "big_min", "big_min64", "explicit_u")}}
            Hm.close()

    # per-kernel device times (HIP events on the plan's stream)
    stage_iters = int(max(3, min(200, 0.5 / max(sec, 1e-9))))  # about half a second per stage
    ms = H.time_stages(warmup=2, iters=stage_iters)
    stage_flops = {"A": st["flops_A"], "C": st["flops_C"]}
    dom = "A" if ms[0] >= ms[2] else "C"
    t_dom = ms[0] if dom == "A" else ms[2]
    achieved = stage_flops[dom] / (t_dom * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.precision]
    ktype = {"fp64": "double", "fp32": "float", "bf16x3": "bf16x3"}[args.precision]
    kname = ("k_stage_a<%s> (T = L.V)" % ktype) if dom == "A" else ("k_stage_c<%s> ((sum c T).R)" % ktype)
    # HBM bytes per launch of that kernel, from the committed copy-calibrated
    # FETCH_SIZE / WRITE_SIZE profile of this workload (null when not profiled)
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f).get("%s/%s" % (args.workload, args.precision), {})
        traffic = tr.get("k_stage_a" if dom == "A" else "k_stage_c")
    except (OSError, ValueError):
        pass

    desc = (("%d independent two-site pivots (HeisenbergU1 / HubbardU1xU1 chains at random bonds, "
             "M log-uniform 128..2048), one plan, one launch per stage" % len(pds)) if batched else
            "two-site, middle of chain, random N(0,1) environments and wavefunction")
    res = {
        "metric": "DMRG HxV GFLOP/s (F_alg / t), %s GPU" % ("plain FP64" if args.precision == "fp64" else
                                                           "mixed " + args.precision),
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "pivot": desc,
            "n_pivots": len(pds),
            "vsize": int(H.vsize),
            "max_block_dim": st["max_block_dim"],
            "n_terms": st["n_terms"], "n_T": st["n_T"], "n_U": st["n_U"],
            "flops_alg": st["flops_alg"], "flops_ref": st["flops_ref"],
            "parallelism": "single GPU",
        },
        "stage_ms": {"A_LV": ms[0], "C_combine_UR": ms[2], "hxv": ms[3]},
        "tiles": {k: tiles[k] for k in ("wave_edge", "big_min", "big_min64", "explicit_u")} | {
                  "autotuned": not args.no_tune},
        "mixed_hxv": mixed,
        "lanczos_ms_per_iteration": lanczos_ms,
        "roofline": {
            "bound": "mfma",
            "kernel": kname,
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_hxv": st["bytes_alg_fp64"],
        },
    }
    if not batched:
        pd = pds[0]
        res["config"].update({
            "model": "HubbardU1xU1" if "hubbard" in args.workload else "HeisenbergU1",
            "bond_sectors": [len(pd.left), len(pd.right)],
            "D_W": pd.W12.dw_in,
        })
    if not args.no_cpu:
        # per-pivot algorithmic flops for the CPU rate (each pivot's own plan statistics)
        if batched:
            falgs = []
            for pd in pds:
                h = synth.feed(lib, "vmps_", pd, _lib.check)
                _lib.check(lib.vmps_pivot2_precalc(h))
                s2 = _lib.PlanStats()
                _lib.check(lib.vmps_pivot2_stats(h, ctypes.byref(s2)))
                lib.vmps_pivot2_destroy(h)
                falgs.append(s2.flops_alg)
        else:
            falgs = [st["flops_alg"]]
        cb, outs = cpu_baseline(pds, vins, falgs, args.cpu_seconds, host_cores())
        res["cpu_baseline"] = cb
        off = np.cumsum([0] + [pd.vsize for pd in pds])
        errs = [float(np.linalg.norm(out_gpu[off[i]:off[i + 1]] - ref) / np.linalg.norm(ref))
                for i, ref in sorted(outs.items())]
        res["parity_rel_err_vs_oracle"] = max(errs)
        res["parity_pivots_checked"] = len(errs)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (C60: DMRG steps, default 20; "
                                                            "HxV workloads: HxV calls, default 2000)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (C60 default 5, HxV 20)")
    ap.add_argument("--workload", default="c60", help="c60 (DMRG steps) or an HxV workload of vmps_amd.synth")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"],
                    help="HxV workloads: precision of the main line")
    ap.add_argument("--M", type=int, default=4096, help="C60: bond dimension (SU(2) multiplets)")
    ap.add_argument("--warm", default="32,64,128,256", help="C60: bond dimensions of the warm-up sweeps")
    ap.add_argument("--pivot", type=int, default=14, help="C60: site where the expanded state's steps start")
    ap.add_argument("--mixed", default="bf16x3", choices=["bf16x3", "fp32"], help="C60: mixed precision")
    ap.add_argument("--no-fp64-leg", action="store_true", help="C60: skip the FP64 comparison leg")
    ap.add_argument("--tol-eigval", type=float, default=1e-11, help="C60: Lanczos eigenvalue tolerance (both legs)")
    ap.add_argument("--tol-state", type=float, default=1e-7,
                    help="C60: residual tolerance ||Hv - Ev|| (both legs; the mixed leg refined to it in FP64)")
    ap.add_argument("--hxv-reps", type=int, default=0,
                    help="C60 profiling target: prepare the state, then only this many HxV of the first timed pivot "
                         "in the mixed precision and in FP64 (no DMRG legs)")
    ap.add_argument("--legs-only", action="store_true",
                    help="C60 profiling target: only the DMRG steps (no roofline timing, CPU baseline or extras)")
    ap.add_argument("--hxv-precision", default="both", choices=["both", "mixed", "fp64"],
                    help="C60 profiling target: which operator(s) --hxv-reps runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile configuration")
    ap.add_argument("--tiles", default=None,
                    help="fixed tile configuration wave_edge,big_min,big_min64,explicit_u (instead of autotuning)")
    ap.add_argument("--no-mixed", action="store_true", help="HxV workloads: skip the FP32 / split-BF16 x3 lines")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--extras", default="c2,c1,c3,c2sweep",
                    help="C60: secondary lines in the default run: c2 (HubbardU1xU1 M=512 HxV), "
                         "c1 (HeisenbergU1 L=32 M=64 sweeps), c3 (precision report, SU(2)xU(1) ladder M=2048), "
                         "c2sweep (HubbardU1xU1 L=48 M=512 sweeps); '' for none")
    ap.add_argument("--sweeps", type=int, default=4, help="HxV workloads: DMRG sweeps of the C1 chain (0: none)")
    ap.add_argument("--lanczos-iters", type=int, default=0,
                    help="HxV workloads: also time this many Lanczos iterations on the pivot (0: none)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run 5 device copies of this many MiB (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n * 8, "copies": 5}))
        return

    if args.gpus != 1:
        print("bench.py: the HxV runs on one GPU; --gpus %d is not supported "
              "(a pivot would shard over its output blocks across GPUs; no multi-GPU path is implemented)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "c60":
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 5 if args.warmup is None else args.warmup
        res = dmrg_c60_bench(args)
    else:
        args.steps = 2000 if args.steps is None else args.steps
        args.warmup = 20 if args.warmup is None else args.warmup
        res = hxv_bench(args)
        if args.sweeps > 0:
            res["sweeps"] = sweep_bench(args, sweeps=args.sweeps)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
