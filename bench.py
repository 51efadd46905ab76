"""Benchmark: converged nonlinear static solves per second on one MI355X.

One step = one complete Static::Solve (all load steps converged) of the
workload, with the model resident in HBM (the state is reset on the device
between solves).  Default workload: the clamped Shell_1 plate of 64 x 64 quads
(8,192 elements, 16,641 nodes, 84,867 free DOFs) under a follower pressure
ramped over 15 load steps; the riser catenary of 10,000 Beam_1 elements and
the serving streams are behind --workload.  The line also carries the other
precision's figures (`compare`: mixed-precision solves/s next to FP64).

Prints one JSON line (see the repository's measurement contract): value =
solves/s, p50/p99 latency, the roofline of the dominant kernel measured live
with HIP event pairs around each of its launches on the model's stream, and
the host-CPU baseline (the
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
    elif name == "plate":
        m, p, info = M.shell_plate(nq=64)
        desc = ("clamped Shell_1 plate: 64x64 quads = 8,192 triangles (16,641 nodes), t = 5 mm, follower pressure "
                "ramped over 15 load steps, seeded N(0, 1e-4 m) imperfections")
    elif name == "rollup":
        m, p, info = M.cantilever_rollup()
        desc = "cantilever roll-up: 100 Beam_1, follower tip moment to 2*pi*EI/L, 20 load steps"
    else:
        raise SystemExit(f"unknown workload {name}")
    return m, p, desc


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(m, p, kind, gpu_iterations, budget_s=30.0):
    """The oracle (FP64 restatement, MKL PARDISO per iteration as sparsesystem
    does) on this host's cores.  Bounded sample: whole load steps are run until
    about budget_s has passed; solves/s is scaled from the time per Newton
    iteration and the full solve's iteration count on the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ["GFO_VERBOSE"] = "2"  # the oracle prints a line per Newton iteration and per load step
    import oracle
    solver = oracle.SOLVER_BAND
    label = "oracle FP64 restatement, band LU"
    if kind in ("auto", "pardiso") and oracle.pardiso_available():
        solver = oracle.SOLVER_PARDISO
        label = "oracle FP64 restatement, MKL PARDISO phase 13 per iteration (as sparsesystem)"
    import dataclasses
    n_steps = int(round((p.end_time - p.start_time) / p.time_step))

    def run(k):
        """The first k load steps of the same Static loop (load tables are in absolute time)."""
        q = dataclasses.replace(p, end_time=p.start_time + k * p.time_step) if k < n_steps else p
        om = oracle.OracleModel(m, solver)
        t0 = time.perf_counter()
        rep, _ = om.static_solve(q)
        return time.perf_counter() - t0, int(rep.total_iterations), k >= n_steps

    log(f"cpu baseline: {label}")
    dt, iters, full = run(1)
    k = max(1, min(n_steps, int(budget_s / max(dt, 1e-3))))
    if k > 1:
        dt, iters, full = run(k)
    per_iter = dt / max(1, iters)
    value = 1.0 / (per_iter * gpu_iterations)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = "the full solve" if full else f"the first {k} of {n_steps} load steps"
    return {"value": value, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {iters} Newton iterations in {dt:.1f} s ({per_iter * 1e3:.0f} ms/iteration), "
                      f"scaled to the {gpu_iterations} iterations of one full solve; {label}",
            "iterations_timed": iters}


SOLVERS = {"auto": abi.GF_SOLVER_AUTO, "band": abi.GF_SOLVER_BAND, "fronts": abi.GF_SOLVER_FRONTS}


def newton_iterations(m, p, precision, count, solver="auto"):
    """`count` Newton iterations of the first load step through the per-phase
    entry points (assembly, tangent solve, update), one line each: a short run
    of every kernel of the path for counter passes under rocprofv3."""
    from giraffe_amd import runtime

    prec = abi.GF_PREC_FP64 if precision == "fp64" else abi.GF_PREC_MIXED
    dm = runtime.DeviceModel(m, precision=prec, solver=SOLVERS[solver])
    dm.set_window(p.start_time, p.end_time)
    dm.begin_step(p.start_time, p.time_step)
    out = []
    for k in range(count):
        ts = time.perf_counter()
        res = dm.assemble()
        fail = dm.solve()
        inc = dm.update()
        out.append({"residual": [float(v) for v in res[:2]], "increment": [float(v) for v in inc[:2]],
                    "info_fail": int(fail), "ms": 1e3 * (time.perf_counter() - ts)})
        log(f"[{precision}] Newton iteration {k}: {out[-1]}")
    dm.close()
    return out


def measure(m, p, precision, steps, warmup, probe_steps, solver="auto"):
    """Warmup + `steps` timed full solves of one resident model (the state is
    reset on the device before each solve), with no instrumentation in the timed
    region.  Then `probe_steps` further solves with per-phase HIP events and HIP
    event pairs around every launch of the roofline candidates: the
    attribution of device time (recording events between the launches of the
    timed region would cost about a quarter of its throughput)."""
    import torch
    from giraffe_amd import runtime

    prec = abi.GF_PREC_FP64 if precision == "fp64" else abi.GF_PREC_MIXED
    dm = runtime.DeviceModel(m, precision=prec, solver=SOLVERS[solver])
    info = dm.info()
    log(f"[{precision}] model resident: {info}")
    for w in range(warmup):
        dm.reset()
        ts = time.perf_counter()
        dm.static_solve(p)
        log(f"[{precision}] warmup solve {w} {1e3 * (time.perf_counter() - ts):.1f} ms")
    st0 = dm.stats()
    lat, reps = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        dm.reset()
        rep, iters = dm.static_solve(p)
        lat.append(time.perf_counter() - ts)
        reps.append((rep.total_iterations, rep.total_solves, rep.max_rel_residual, rep.aborted, list(iters)))
        log(f"[{precision}] timed solve {len(lat) - 1}: {1e3 * lat[-1]:.1f} ms, {rep.total_iterations} iterations")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st1 = dm.stats()
    ph, kp = {}, {}
    if probe_steps > 0:
        dm.timing(True)
        dm.timing_read(reset=True)
        dm.kernel_probe(PROBES_FRONTS if info["layout"] == 2 else PROBES_BTD if info["block"] else PROBES_BAND)
        dm.kernel_probe_read(reset=True)
        for k in range(probe_steps):
            dm.reset()
            ts = time.perf_counter()
            dm.static_solve(p)
            log(f"[{precision}] attribution solve {k}: {1e3 * (time.perf_counter() - ts):.1f} ms (with events)")
        torch.cuda.synchronize()
        ph = dm.timing_read(reset=True)
        kp = dm.kernel_probe_read(reset=True)
        dm.kernel_probe([])
        dm.timing(False)
    lat_ms = np.array(lat) * 1e3
    return {"info": info, "wall": wall, "lat_ms": lat_ms, "reps": reps, "phases": ph, "kernels": kp,
            "refine_sweeps": st1["refine_sweeps"] - st0["refine_sweeps"],
            "fallbacks": st1["fallbacks"] - st0["fallbacks"], "solves": st1["solves"] - st0["solves"],
            "fp64_direct": st1["fp64_direct"] - st0["fp64_direct"],
            "fp32_factorizations": st1["fp32_factorizations"] - st0["fp32_factorizations"]}


# Kernel classes bracketed by HIP events in the timed region (per layout).
PROBES_BTD = ("btd_level", "btd_tail", "beam_assemble", "shell_assemble")
PROBES_BAND = ("band_panel", "band_update", "band_sweep", "beam_assemble", "shell_assemble")
PROBES_FRONTS = ("front_step", "front_assemble", "front_solve", "beam_assemble", "shell_assemble")

KERNEL_NAMES = {"btd_level": "btd_level_kernel", "btd_tail": "btd_tail_kernel",
                "band_panel": "band_batch_inv_kernel", "band_update": "band_step_kernel",
                "band_sweep": "band_batch_apply_kernel + band_couple_kernel",
                "beam_assemble": "beam_assemble_kernel<0>", "shell_assemble": "shell_element_kernel",
                "front_step": "front_step_kernel", "front_assemble": "front_assemble_kernel",
                "front_solve": "front_fwd_own/fwd_upd/bwd_upd/bwd_own_kernel"}


def kernel_work(cls, info):
    """Algorithmic work of one unit of a kernel class: (bound, amount, unit, what)."""
    n, b, mblk = info["n_free"], info["half_band"], max(1, info["block"])
    s = 8 if info["precision"] == abi.GF_PREC_FP64 else 4
    lu = FM.band_lu(n, b, s)
    if cls in ("btd_level", "btd_tail"):
        # the band factorisation's bytes apportioned per block row of m rows
        return "hbm", lu["bytes"] / n * mblk, "B", "block row eliminated: 2 m (2b+1) s bytes of the band LU"
    if cls == "band_update":
        return "mfma", lu["flops"] / n, "flop", "band column eliminated: 2 b^2 flops of the band LU"
    if cls == "band_panel":
        return "mfma", 1.0, "flop", "flop of the batch inverses of the diagonal blocks (64^3 products)"
    if cls == "front_step":
        return "mfma", 1.0, "flop", ("algorithmic flop of the fronts' partial LU: 2 (m - c - 1)^2 per eliminated "
                                     "column c of a front of order m")
    if cls == "front_assemble":
        return "hbm", s, "B", "front entry written (the write alone: a lower bound of the kernel's bytes)"
    if cls == "front_solve":
        return "hbm", 1.0, "B", "factor byte read by the triangular solves"
    if cls == "band_sweep":
        return "hbm", FM.band_trsv(n, b, s)["bytes"] / n / 2, "B", "row solved: half of the triangular-solve bytes"
    if cls == "beam_assemble":
        return "hbm", FM.BEAM_BYTES_FP64, "B", "element (FP64 tangent)"
    return "hbm", FM.SHELL_BYTES_FP64, "B", "element (FP64 tangent)"


def kernel_roofline(cls, rec, info):
    ms, launches, units = rec
    bound, per_unit, u, what = kernel_work(cls, info)
    avg_s = ms / launches * 1e-3
    per_launch = per_unit * units / launches
    if bound == "hbm":
        ach, peak, unit = per_launch / avg_s / 1e9, FM.HBM_PEAK_GBS, "GB/s"
    else:
        ach = per_launch / avg_s / 1e12
        peak = FM.FP64_PEAK_TFLOPS if info["precision"] == abi.GF_PREC_FP64 else FM.FP32_PEAK_TFLOPS
        unit = "TFLOP/s"
    return {"kernel": KERNEL_NAMES[cls], "bound": bound, "achieved": ach, "peak": peak, "unit": unit,
            "frac": ach / peak, "traffic": None, "avg_launch_us": avg_s * 1e6, "launches": launches,
            "units_per_launch": units / launches, "algorithmic_per_unit": per_unit,
            "algorithmic_per_launch": per_launch, "algorithmic_unit": u, "unit_of_work": what,
            "kernel_ms_total": ms}


def traffic_from_profile(kernel, workload):
    """HBM bytes per launch of `kernel` recorded under profiles/ from rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (line `traffic_bytes_per_launch: <kernel> <bytes>`)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel:
                return float(parts[2])
    return None


def roofline(m, r, workload_name):
    """Roofline of the dominant kernel: the algorithmic work of one launch
    (per-unit figure x the units it processed) over its average duration, both
    from the HIP event pairs around every launch of the class in the timed region."""
    kp, info = r["kernels"], r["info"]
    if not kp:
        return None
    dom = max(kp, key=lambda c: kp[c][0])
    out = kernel_roofline(dom, kp[dom], info)
    out["traffic"] = traffic_from_profile(out["kernel"], workload_name)
    out["other_kernels"] = {KERNEL_NAMES[c]: {k: v for k, v in kernel_roofline(c, rec, info).items()
                                              if k in ("achieved", "unit", "frac", "avg_launch_us", "launches")}
                            for c, rec in kp.items() if c != dom}
    return out


def serve(args):
    """Catenary request stream: riser requests (giraffe_amd.model.riser_request,
    seed 1004) with Poisson arrivals at `load` x the closed-loop capacity, served
    by gf_server (continuous admission by slot and device memory).  Latency =
    submit to final commit."""
    import torch
    from giraffe_amd import runtime

    mixed = args.workload == "serve-mixed"
    rng = np.random.default_rng(1005 if mixed else 1004)
    t0 = time.perf_counter()
    if mixed:  # 70 % riser chains, 30 % clamped plates
        reqs = [M.riser_request(rng)[:2] if rng.uniform() < 0.7 else M.plate_request(rng)[:2]
                for _ in range(args.requests)]
    else:
        reqs = [M.riser_request(rng)[:2] for _ in range(args.requests)]
    for m, _ in reqs:
        m.freeze()
    n_plates = sum(1 for m, _ in reqs if m.n_shells)
    log(f"generated {len(reqs)} requests in {time.perf_counter() - t0:.1f} s "
        f"({sum(m.n_beams for m, _ in reqs)} Beam_1, {n_plates} plates with "
        f"{sum(m.n_shells for m, _ in reqs)} Shell_1 in total)")
    prec = abi.GF_PREC_FP64 if args.precision == "fp64" else abi.GF_PREC_MIXED

    def drain(srv, pending, lat, t_sub):
        while pending:
            srv.pump(1000)
            for rid in list(pending):
                try:
                    rep, it, nodes, times = srv.take(rid)
                except runtime.GiraffeError:
                    continue
                pending.discard(rid)
                lat[rid] = (rep, times[1])

    # warm-up and closed-loop capacity on the first requests
    nwarm = min(len(reqs), 32)
    srv = runtime.Server(precision=prec, max_inflight=args.max_inflight)
    ids = {srv.submit(m, p) for m, p in reqs[:nwarm]}
    drain(srv, ids, {}, None)
    ncap = min(len(reqs), 200)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    ids = {srv.submit(m, p) for m, p in reqs[:ncap]}
    drain(srv, ids, {}, None)
    cap = ncap / (time.perf_counter() - tc)
    log(f"closed-loop capacity {cap:.1f} solves/s over {ncap} requests")
    # open loop: Poisson arrivals at load x capacity
    rate = args.load * cap
    gaps = rng.exponential(1.0 / rate, len(reqs))
    arrivals = np.cumsum(gaps)
    lat = {}
    pending = set()
    index_of = {}
    t_start = time.perf_counter()
    for i, (m, p) in enumerate(reqs):
        while time.perf_counter() - t_start < arrivals[i]:
            srv.pump(200)
            for rid in list(pending):
                try:
                    rep, it, nodes, times = srv.take(rid)
                except runtime.GiraffeError:
                    continue
                pending.discard(rid)
                lat[rid] = (rep, times[1])
        rid = srv.submit(m, p)
        index_of[rid] = i
        pending.add(rid)
        if (i + 1) % 200 == 0:
            log(f"submitted {i + 1} requests, {len(lat)} finished, {srv.counts()}")
    drain(srv, pending, lat, None)
    makespan = time.perf_counter() - t_start
    srv.close()
    lat_ms = np.array([v[1] for v in lat.values()])
    aborted = sorted(index_of[rid] for rid, v in lat.items() if v[0].aborted)
    finished = [v[0] for v in lat.values() if not v[0].aborted]
    out = {
        "metric": METRIC,
        "value": len(lat) / makespan,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": nwarm,
        "ms_per_step": makespan / len(lat) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 tangent/residual, f32 factor + f64 refinement",
        "data": "synthetic (seeded request generators, giraffe_amd/model.py)",
        "config": {"workload": (f"mixed serving stream: {len(reqs)} requests, {len(reqs) - n_plates} Beam_1 risers "
                                f"(E log-U[50,5000]) and {n_plates} clamped Shell_1 plates ({{16,24,32,48,64}}^2 "
                                f"quads, t U[2,10] mm, deflection U[5,25] t), seed 1005" if mixed else
                                f"serving stream: {len(reqs)} Beam_1 riser requests (E log-U[50,5000]), seed 1004")
                               + f"; Poisson arrivals at {args.load:.2f} x closed-loop capacity, "
                                 f"{args.max_inflight} resident",
                   "precision": args.precision, "arrival_rate_per_s": rate, "closed_loop_capacity": cap,
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "aborted": len(aborted),
        "aborted_requests": aborted[:20],
        "max_rel_residual": max((r.max_rel_residual for r in finished), default=0.0),
    }
    if args.cpu_baseline != "none":
        out["cpu_baseline"] = serve_cpu_baseline(reqs, args.cpu_budget)
    print(json.dumps(out))
    return 0


def serve_cpu_baseline(reqs, budget_s):
    """The oracle solving the same stream's requests one after another (closed loop)
    until about budget_s has passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ["GFO_VERBOSE"] = "1"
    import oracle
    solver = oracle.SOLVER_PARDISO if oracle.pardiso_available() else oracle.SOLVER_BAND
    t0 = time.perf_counter()
    n = 0
    for m, p in reqs:
        oracle.OracleModel(m, solver).static_solve(p)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    label = "MKL PARDISO" if solver == oracle.SOLVER_PARDISO else "band LU"
    return {"value": n / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"the first {n} requests of the stream solved one after another in {dt:.1f} s "
                      f"(oracle FP64 restatement, {label})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plate", choices=["catenary", "plate", "rollup", "serve", "serve-mixed"])
    ap.add_argument("--requests", type=int, default=2000, help="serve: requests in the stream")
    ap.add_argument("--max-inflight", type=int, default=16, help="serve: resident requests")
    ap.add_argument("--load", type=float, default=0.8, help="serve: arrival rate / closed-loop capacity")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="serve: HIP hardware queues for the request streams (GPU_MAX_HW_QUEUES, at most 32; "
                         "0 keeps the environment's setting)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--solver", default="auto", choices=sorted(SOLVERS),
                    help="tangent solver for wide bands (auto: the library's choice)")
    ap.add_argument("--compare-steps", type=int, default=2,
                    help="solves of the other precision, reported beside the main figure (0: skip)")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "pardiso", "band", "none"])
    ap.add_argument("--cpu-budget", type=float, default=30.0, help="seconds of CPU baseline sample")
    ap.add_argument("--no-timing", action="store_true", help="skip the attribution solves (phases, roofline)")
    ap.add_argument("--probe-steps", type=int, default=2,
                    help="solves after the timed region with per-phase and per-launch HIP events")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this many MiB once on the device (a plain copy "
                         "kernel of known size, the reference point for scaling HBM counter readings)")
    ap.add_argument("--newton-iterations", type=int, default=0,
                    help="instead of timed solves, run this many Newton iterations of the first load step "
                         "(assembly, tangent solve, update) and print them: a short run for counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(requests shard across GPUs with no collectives; see DESIGN.md).", file=sys.stderr)
        return 2
    if args.workload.startswith("serve") and args.hw_queues > 0:
        # one stream per resident request; HIP spreads streams over this many
        # hardware queues (its default is 4), set before the runtime starts
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, args.hw_queues))
    log("importing torch")
    import torch

    log("torch imported")
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 2
    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_mib} MiB device to device")
        del src, dst
    if args.workload.startswith("serve"):
        return serve(args)
    m, p, desc = workload(args.workload)
    if args.newton_iterations > 0:
        its = newton_iterations(m, p, args.precision, args.newton_iterations, args.solver)
        print(json.dumps({"mode": "newton-iterations", "workload": desc, "precision": args.precision,
                          "iterations": its}))
        return 0
    r = measure(m, p, args.precision, args.steps, args.warmup, 0 if args.no_timing else args.probe_steps, args.solver)
    info, reps = r["info"], r["reps"]
    its, solves = reps[-1][0], reps[-1][1]
    out = {
        "metric": METRIC,
        "value": args.steps / r["wall"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 tangent/residual, f32 factor + f64 refinement",
        "data": "synthetic (seeded generator, giraffe_amd/model.py)",
        "config": {"workload": desc, "precision": args.precision, "free_dofs": info["n_free"],
                   "half_bandwidth": info["half_band"], "block": info["block"],
                   "solver": {0: "block-tridiagonal cyclic reduction", 1: "band LU",
                              2: "multifrontal LU over nested dissection"}[info["layout"]],
                   "newton_iterations": its,
                   "tangent_solves": solves, "step_iterations": [int(v) for v in reps[-1][4]]},
        "p50_ms": float(np.percentile(r["lat_ms"], 50)),
        "p99_ms": float(np.percentile(r["lat_ms"], 99)),
        "max_rel_residual": max(x[2] for x in reps),
        "aborted": any(x[3] for x in reps),
    }
    if args.precision == "mixed":
        out["refinement"] = {"sweeps_per_solve": r["refine_sweeps"] / max(1, r["solves"]),
                             "fp64_fallbacks": r["fallbacks"], "fp64_direct": r["fp64_direct"],
                             "fp32_factorizations": r["fp32_factorizations"], "tangent_solves": r["solves"]}
    if r["phases"]:
        ph = r["phases"]
        n_att = args.probe_steps
        out["phases_ms_per_solve"] = {k: v[0] / n_att for k, v in ph.items()}
        out["phase_launches_per_solve"] = {k: v[1] / n_att for k, v in ph.items()}
        out["roofline"] = roofline(m, r, args.workload)
        out["roofline"]["measured"] = (f"HIP event pairs around every launch of the candidate kernels on the model's "
                                       f"stream, over {n_att} solves run after the timed region (events between the "
                                       f"timed launches would cost about a quarter of the throughput)")
    if args.compare_steps > 0:
        other = "mixed" if args.precision == "fp64" else "fp64"
        c = measure(m, p, other, args.compare_steps, 1, 0, args.solver)
        out["compare"] = {"precision": other, "solves_per_s": args.compare_steps / c["wall"],
                          "p50_ms": float(np.percentile(c["lat_ms"], 50)),
                          "p99_ms": float(np.percentile(c["lat_ms"], 99)),
                          "newton_iterations": c["reps"][-1][0],
                          "max_rel_residual": max(x[2] for x in c["reps"]),
                          "refine_sweeps_per_solve": c["refine_sweeps"] / max(1, c["solves"]),
                          "fp64_fallbacks": c["fallbacks"], "fp64_direct": c["fp64_direct"],
                          "fp32_factorizations": c["fp32_factorizations"], "tangent_solves": c["solves"]}
        out["compare"]["same_iterations"] = c["reps"][-1][4] == reps[-1][4]
        if other == "mixed":
            nd, ns = c["fp64_direct"], max(1, c["solves"])
            out["compare"]["paths"] = (
                f"{ns - nd} of {ns} tangent solves FP32-factored and refined in FP64, {nd} factored in FP64 directly "
                f"(the mixed mode's back-off: after FP32 refinement fell back, or cost more than an FP64 "
                f"factorisation and solve)")
    if args.cpu_baseline != "none":
        out["cpu_baseline"] = cpu_baseline(m, p, args.cpu_baseline, its, args.cpu_budget)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
