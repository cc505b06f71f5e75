"""Benchmark: effective-Hamiltonian application (HxV) on one MI355X.

One step = one HxV of the workload's two-site pivot (or of every pivot of a
batched workload, in one launch per stage) on device-resident vectors.
Default workload: HubbardU1xU1 chain, L=48, half filling, U=8, t=1, M=512,
D_W=6, pivot in the middle of the chain with random N(0,1) environment and
wavefunction blocks (synthetic data).  The metric value is the algorithmic
rate F_alg / t in GFLOP/s, where F_alg counts the staged contraction
(distinct L.V products plus distinct U.R products).

Each plan is autotuned first (its tile configuration is timed over the
candidates and the fastest kept; `tiles` key).  The same workload is also
timed in the mixed modes, FP32 MFMA and split-BF16x3 with FP32 accumulation
(`mixed_hxv`: rate, speedup over FP64 and the relative difference of the
result from the FP64 one).

The default run also times one large-block pivot (`--large`, default the
long-range SU(2) Heisenberg pivot with D_W = 40 and M = 4096 multiplets, 4.5
TFLOP per HxV) in FP64, FP32 and split-BF16x3 (`large_pivot`), where the
contraction is MFMA- rather than latency-bound.

Beside the HxV line the default run also times full two-site DMRG sweeps of
the HeisenbergU1 L=32 M=64 chain, FP64 and mixed (split-BF16x3 Lanczos refined
in FP64), and reports the energy difference (`sweeps` key).

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


def large_pivot_bench(args, name, iters=3):
    """One large-block pivot (the MFMA-bound regime) in FP64, FP32 and
    split-BF16x3, each autotuned: rate, share of the dtype's MFMA peak, and the
    mixed results' relative difference from the FP64 one."""
    from vmps_amd import synth
    from vmps_amd.pivot import HxV, PivotMatrix2
    pd = synth.make_workload(name, seed=args.seed)
    v = synth.random_vector(pd, args.seed + 1)
    out = {"workload": name}
    ref = None
    for prec in ("fp64", "fp32", "bf16x3"):
        H = PivotMatrix2(pd, precision=prec)
        st = H.stats()
        tiles = H.autotune(iters=iters)
        ms = H.time_hxv(warmup=1, iters=iters)
        o = HxV(H, v)
        if ref is None:
            ref = o
            out.update({"flops_alg": st["flops_alg"], "max_block_dim": st["max_block_dim"], "vsize": int(H.vsize)})
        tf = st["flops_alg"] / (ms * 1e-3) / 1e12
        out[prec] = {"ms_per_hxv": ms, "tflops": tf, "frac_of_peak": tf / PEAK_TFLOPS[prec],
                     "rel_err_vs_fp64": float(np.linalg.norm(o - ref) / np.linalg.norm(ref)),
                     "tiles": {k: tiles[k] for k in ("wave_edge", "big_min", "big_min64", "explicit_u")}}
        H.close()
    out["speedup_bf16x3_vs_fp64"] = out["fp64"]["ms_per_hxv"] / out["bf16x3"]["ms_per_hxv"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hubbard_u1xu1_L48_M512")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile configuration")
    ap.add_argument("--tiles", default=None,
                    help="fixed tile configuration wave_edge,big_min,big_min64,explicit_u (instead of autotuning)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32 / split-BF16 x3 HxV lines")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--large", default="heisenberg_su2_lr40_L60_M4096",
                    help="MFMA-regime pivot timed in FP64 and the mixed modes beside the main line ('' = skip)")
    ap.add_argument("--sweeps", type=int, default=4, help="DMRG sweeps of the HeisenbergU1 L=32 M=64 chain (0: none)")
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
    if args.sweeps > 0:
        res["sweeps"] = sweep_bench(args, sweeps=args.sweeps)
    if args.large and args.precision == "fp64":
        res["large_pivot"] = large_pivot_bench(args, args.large)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
