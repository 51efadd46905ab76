"""Benchmark: converged nonlinear static solves per second on one MI355X.

One step = one complete Static::Solve (all load steps converged) of the
workload, with the model resident in HBM (the state is reset on the device
between solves).  Default workload: the riser catenary of 10,000 Beam_1
elements (20,001 nodes), self-weight plus seeded N(0, 500 N) lateral nodal
loads over 10 load steps, Newton tolerance 1e-10.

Prints one JSON line (see the repository's measurement contract): value =
solves/s, p50/p99 latency, the roofline of the dominant device phase measured
live with HIP events on the model's stream, and the host-CPU baseline (the
oracle's FP64 restatement with MKL PARDISO, timed on this host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from giraffe_amd import abi, flop_model as FM, model as M  # noqa: E402

METRIC = "converged solves/s + p50/p99 latency, mixed vs FP64 GPU, rel.res ≤1e-10"


def workload(name):
    if name == "catenary":
        m, p, info = M.catenary(n_elem=10000)
        desc = "riser catenary: 10,000 Beam_1 (20,001 nodes), self-weight + N(0,500 N) lateral nodal loads, 10 load steps"
    elif name == "rollup":
        m, p, info = M.cantilever_rollup()
        desc = "cantilever roll-up: 100 Beam_1, follower tip moment to 2*pi*EI/L, 20 load steps"
    else:
        raise SystemExit(f"unknown workload {name}")
    return m, p, desc


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(m, p, kind):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ["GFO_VERBOSE"] = "2"  # the oracle prints a line per Newton iteration and per load step
    import oracle
    solver = oracle.SOLVER_BAND
    label = "oracle FP64 restatement, band LU"
    if kind in ("auto", "pardiso") and oracle.pardiso_available():
        solver = oracle.SOLVER_PARDISO
        label = "oracle FP64 restatement, MKL PARDISO phase 13 per iteration (as sparsesystem)"
    om = oracle.OracleModel(m, solver)
    log(f"cpu baseline: {label}")
    t0 = time.perf_counter()
    rep, iters = om.static_solve(p)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"one full solve of the same workload ({dt:.1f} s, {rep.total_iterations} Newton iterations); {label}",
            "iterations": int(rep.total_iterations)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="catenary")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "pardiso", "band", "none"])
    ap.add_argument("--no-timing", action="store_true", help="do not record per-phase HIP events")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(requests shard across GPUs with no collectives; see DESIGN.md).", file=sys.stderr)
        return 2
    log("importing torch")
    import torch
    from giraffe_amd import runtime

    log("torch imported")
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 2
    m, p, desc = workload(args.workload)
    prec = abi.GF_PREC_FP64 if args.precision == "fp64" else abi.GF_PREC_MIXED
    dm = runtime.DeviceModel(m, precision=prec)
    info = dm.info()
    log(f"model resident: {info}")
    for w in range(args.warmup):
        dm.reset()
        ts = time.perf_counter()
        dm.static_solve(p)
        log(f"warmup solve {w} {1e3 * (time.perf_counter() - ts):.1f} ms")
    if not args.no_timing:
        dm.timing(True)
        dm.timing_read(reset=True)
    lat = []
    reps = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        dm.reset()
        rep, iters = dm.static_solve(p)
        lat.append(time.perf_counter() - ts)
        reps.append((rep.total_iterations, rep.total_solves, rep.max_rel_residual, rep.aborted, list(iters)))
        log(f"timed solve {len(lat) - 1}: {1e3 * lat[-1]:.1f} ms, {rep.total_iterations} iterations")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ph = dm.timing_read(reset=True) if not args.no_timing else {}
    lat_ms = np.array(lat) * 1e3
    n_free, b = info["n_free"], info["half_band"]
    its, solves = reps[-1][0], reps[-1][1]
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 tangent/factor + f64 residual/refinement",
        "data": "synthetic (seeded generator, giraffe_amd/model.py)",
        "config": {"workload": desc, "precision": args.precision, "free_dofs": n_free, "half_bandwidth": b,
                   "block": info["block"], "newton_iterations": its, "tangent_solves": solves,
                   "step_iterations": [int(v) for v in reps[-1][4]]},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "max_rel_residual": max(r[2] for r in reps),
        "aborted": any(r[3] for r in reps),
    }
    if ph:
        out["phases_ms_per_solve"] = {k: v[0] / args.steps for k, v in ph.items()}
        out["phase_launches_per_solve"] = {k: v[1] / args.steps for k, v in ph.items()}
        # dominant device phase and its roofline (algorithmic work / measured time)
        dom = max(ph, key=lambda k: ph[k][0])
        ms, launches = ph[dom]
        if dom == "assemble":
            per_launch_elems = m.n_beams / max(1, info["colors"])
            unit_bytes = FM.BEAM_BYTES_FP64 * per_launch_elems
            avg_s = ms / launches * 1e-3
            ach = unit_bytes / avg_s / 1e9
            out["roofline"] = {"kernel": "beam_assemble_kernel<0>", "bound": "hbm", "achieved": ach,
                               "peak": FM.HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / FM.HBM_PEAK_GBS,
                               "traffic": None, "avg_launch_us": avg_s * 1e6,
                               "algorithmic_bytes_per_launch": unit_bytes}
        else:
            work = FM.band_lu(n_free, b) if dom == "factor" else FM.band_trsv(n_free, b)
            per = ms / max(1, solves * args.steps) * 1e-3  # one factorisation / one solve per tangent solve
            ach = work["bytes"] / per / 1e9
            out["roofline"] = {"kernel": f"block-tridiagonal {dom} (all levels)", "bound": "hbm", "achieved": ach,
                               "peak": FM.HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / FM.HBM_PEAK_GBS,
                               "traffic": None, "avg_launch_us": ms / launches * 1e3,
                               "algorithmic_bytes_per_unit": work["bytes"], "unit_of_work": f"one {dom}"}
    if args.cpu_baseline != "none":
        out["cpu_baseline"] = cpu_baseline(m, p, args.cpu_baseline)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
