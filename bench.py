"""Benchmark: ms per plate solve to relative residual 1e-8 on one MI355X.

Default run: the whole metric, "ms/solve to 1e-8 rel. residual at 61k & 3M
DOF; learned vs block-Jacobi PCG".
  value  142 x 142 Q4 Mindlin plate (61,347 DOF), clamped, E = 210 GPa,
         nu = 0.3, t = 0.1 m, 10 kPa + 50 kPa Gaussian patch (sigma 1 m,
         centre from SplitMix64 seed 1): FP64 block-Jacobi PCG, whole loop on
         the device; a step is one full solve of the resident load vector.
  c2     the same plate with the other preconditioners: Chebyshev-polynomial
         block Jacobi, multigrid PCG, the learned preconditioner (trained
         first, off the clock) and the same preconditioner with a zero network.
  c5     1024 x 1024 plate (3,151,875 DOF) with a log-normal thickness
         field: block-Jacobi PCG, multigrid PCG, learned and zero-network.
Prints one JSON line.  roofline: the persistent block-Jacobi kernel of the
value, with HBM traffic measured in the same run by two rocprofv3 --pmc child
runs (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a plain copy;
cpu_baseline: the float32 restatement of solver.ts in Node on the host.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-gauntlet_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per DOF per FP64 PCG iteration: read p, x, r and the
# 3x3 block inverse (3 words per DOF), write p, x, r = 9 words of 8 bytes.
BYTES_PER_DOF_ITER_FP64 = 72
# FP32 vectors (6 words of 4 bytes) with the FP64 block inverse (3 x 8 bytes).
BYTES_PER_DOF_ITER_MIXED = 48


METRIC = "ms/solve to 1e-8 rel. residual at 61k & 3M DOF; learned vs block-Jacobi PCG"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c2: 142 x 142 FP64 (default); c3: 700 x 700 mixed precision; "
                         "c4: 256 batched Q4/DKT plates; c5: 1024 x 1024 variable thickness, "
                         "learned preconditioner vs block Jacobi")
    ap.add_argument("--validate", action="store_true",
                    help="c4: compare every plate with the single-plate FP64 solve")
    ap.add_argument("--train-loads", type=int, default=4,
                    help="c5: load cases for the solver-generated training pairs (8 pairs each)")
    ap.add_argument("--train-epochs", type=int, default=1, help="c5: training epochs")
    ap.add_argument("--train-source", default="solver", choices=["solver", "learned"],
                    help="c5: training pairs from block-Jacobi PCG iterates or from flexible "
                         "PCG with the network's current preconditioner")
    ap.add_argument("--cpu-sample-iters", type=int, default=-1,
                    help="iterations of the CPU baseline sample (0 = skip; default 1200, "
                         "150 for c3)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (profiler calibration)")
    ap.add_argument("--probe", choices=["copy", "pcg", "mg142", "pcg1024", "mg1024", "unet1024",
                                        "dkt400"],
                    help="only run a probe for the rocprofv3 runs: a 256 MiB copy, one "
                         "142 x 142 block-Jacobi solve, five 142 x 142 multigrid-preconditioned "
                         "solves, 1,500 block-Jacobi iterations on the 1024 x 1024 plate, its "
                         "multigrid-preconditioned solve, 20 applications of the learned "
                         "preconditioner there (U-net forward + block-Jacobi smoothing), or "
                         "the 400 x 400 split-quad DKT plate on the coloured path (three solves "
                         "with the persistent kernel, one with the launch-per-operation loop)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs (traffic: null)")
    ap.add_argument("--no-c5", action="store_true", help="default run without the 1024 x 1024 legs")
    ap.add_argument("--c5-steps", type=int, default=2,
                    help="timed solves per leg of the 1024 x 1024 plate (block Jacobi 3.7 s each)")
    ap.add_argument("--c2-train-loads", type=int, default=32,
                    help="load cases for the 142 x 142 network's training pairs")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain device-to-device copy of known size (the HIP runtime's copy
    kernel), for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = mib * 1024 * 1024
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), nbytes) or hip.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, nbytes)
    for _ in range(5):
        if hip.hipMemcpy(b, a, nbytes, 3):  # device to device
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": 5}))


def probe_pcg():
    """One 142 x 142 block-Jacobi solve (one launch of the persistent kernel)."""
    from platepcg import Context, Plan, _lib
    from platepcg import problem as P
    w = P.workload("c2")
    ctx = Context(0)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    plan.set_rhs(plan.pressure_load(w["q0"], w["patches"]))
    plan.solve_resident(plan.options(tolerance=1e-8, max_iterations=400000,
                                     flags=_lib.FLAG_NO_TRUE_RESIDUAL))
    plan.close()
    ctx.close()


def probe_mg142():
    """Five multigrid-preconditioned solves of the 142 x 142 plate to 1e-8."""
    from platepcg import Context, Plan, _lib
    from platepcg import problem as P
    w = P.workload("c2")
    ctx = Context(0)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    plan.set_rhs(plan.pressure_load(w["q0"], w["patches"]))
    for _ in range(5):
        plan.solve_resident(plan.options(tolerance=1e-8, max_iterations=5000,
                                         precond=_lib.PRECOND_MULTIGRID))
    plan.close()
    ctx.close()


def probe_unet1024():
    """20 applications of the learned preconditioner (block-Jacobi smoothing
    around the U-net, random parameters) on the 1024 x 1024 plate."""
    import numpy as np

    from platepcg import Context, Net, Plan, lognormal_thickness
    from platepcg import problem as P
    ctx = Context(0)
    _, t_e = lognormal_thickness(ctx, 1024, 1024, 10.0, 10.0, 0.5, 5, 0.1, 0.5)
    w = P.workload("c5", element_t=t_e)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"], element_t=w["element_t"])
    net = Net(plan, seed=5)
    r = np.random.default_rng(1).standard_normal(plan.n_dof)
    for _ in range(20):
        net.apply(r)
    net.close()
    plan.close()
    ctx.close()


def probe_dkt400():
    """The 400 x 400 clamped plate split into 320,000 DKT triangles (160,801
    nodes, greedy colouring), coloured path: three block-Jacobi solves to 1e-8
    with the persistent kernel, one with the launch-per-operation loop."""
    from platepcg import Context, Plan, _lib
    from platepcg import problem as P
    mesh = P.split_quads(P.rect_mesh(400))
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    ctx = Context(0)
    plan = Plan(ctx, mesh, 210e9, 0.3, 0.1, mask, P.colour_greedy(mesh))
    plan.set_rhs(plan.pressure_load(1e4))
    for flags in (0, 0, 0, _lib.FLAG_COLOURED_HOST_LOOP):
        st = plan.solve_resident(plan.options(tolerance=1e-8, max_iterations=5000,
                                              path=_lib.PATH_COLOURED, flags=flags))
        print(json.dumps({"flags": flags, "iterations": st.iterations,
                          "device_ms": st.device_ms}), flush=True)
    plan.close()
    ctx.close()


def probe_1024(multigrid):
    """The 1024 x 1024 plate (thickness field made on the device): 1,500
    iterations of the strided block-Jacobi kernel in one launch, or the whole
    multigrid-preconditioned solve to 1e-8."""
    from platepcg import Context, Plan, _lib, lognormal_thickness
    from platepcg import problem as P
    ctx = Context(0)
    _, t_e = lognormal_thickness(ctx, 1024, 1024, 10.0, 10.0, 0.5, 5, 0.1, 0.5)
    w = P.workload("c5", element_t=t_e)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"], element_t=w["element_t"])
    plan.set_rhs(plan.pressure_load(w["q0"], w["patches"]))
    if multigrid:
        plan.solve_resident(plan.options(tolerance=1e-8, max_iterations=5000,
                                         precond=_lib.PRECOND_MULTIGRID))
    else:
        plan.solve_resident(plan.options(tolerance=1e-30, max_iterations=1500,
                                         flags=_lib.FLAG_NO_TRUE_RESIDUAL))
    plan.close()
    ctx.close()


def measure_traffic(kernel_prefix="k_pcg_cg", probe_name="pcg"):
    """HBM bytes per launch of a kernel in a probe run (default: the
    persistent block-Jacobi kernel on the 142 x 142 plate): rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), one child run per counter,
    each scaled by what the same counter reads for a 256 MiB device-to-device
    copy (FETCH_SIZE reads 2x the copied bytes, WRITE_SIZE 1x on MI355X).
    Runs before this process touches the GPU.  Returns (bytes, details) or
    (None, reason)."""
    import csv
    import shutil
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, {"error": "rocprofv3 not found"}
    out = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            for probe in ("copy", probe_name):
                tag = f"{ctr}_{probe}"
                cmd = [exe, "--pmc", ctr, "--output-format", "csv", "-d", os.path.join(d, tag),
                       "-o", "r", "--", sys.executable, os.path.abspath(__file__), "--probe", probe]
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
                if r.returncode != 0:
                    return None, {"error": f"rocprofv3 {tag} exit {r.returncode}: "
                                           f"{(r.stderr or '')[-300:]}"}
                rows = []
                for root, _, files in os.walk(os.path.join(d, tag)):
                    for f in files:
                        if f.endswith("counter_collection.csv"):
                            with open(os.path.join(root, f)) as fh:
                                rows += list(csv.DictReader(fh))
                per = {}
                for row in rows:
                    if row.get("Counter_Name") != ctr:
                        continue
                    key = (row["Dispatch_Id"], row["Kernel_Name"])
                    per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
                if probe == "copy":
                    vals = [v for (_, k), v in per.items() if "copy" in k.lower()]
                else:
                    vals = [v for (_, k), v in per.items() if kernel_prefix in k]
                if not vals:
                    return None, {"error": f"no {tag} dispatch in the counter output"}
                out[tag] = max(vals)
    copy_kib = 256 * 1024
    fs = out["FETCH_SIZE_copy"] / copy_kib
    ws = out["WRITE_SIZE_copy"] / copy_kib
    fetch = out["FETCH_SIZE_" + probe_name] / fs * 1024
    write = out["WRITE_SIZE_" + probe_name] / ws * 1024
    return int(fetch + write), {"fetch_bytes": int(fetch), "write_bytes": int(write),
                                "fetch_scale": fs, "write_scale": ws,
                                "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, one child run "
                                          "each, scaled by the same counters on a 256 MiB copy"}


def cpu_baseline(w, F, full_iters, sample_iters, basis="iterations it needs to reach 1e-8"):
    """Float32 restatement of solver.ts (oracle/solver_f32.js) in Node, one
    thread, on the first `sample_iters` iterations of the same solve,
    extrapolated to `full_iters` iterations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O

    n = w["mesh"]["gridNx"]
    _, st = O.run_js_solver(n, F, w["E"], w["nu"], w["t"],
                            [{"type": "fixed", "location": "all_edges"}], tol=1e-8,
                            max_iter=sample_iters)
    ms_per_iter = st["solveTimeMs"] / max(1, st["iterations"])
    whole = st["iterations"] >= full_iters or st.get("converged")
    return {
        "value": st["solveTimeMs"] if whole else ms_per_iter * full_iters, "unit": "ms/solve",
        "cores": 1, "kind": "port", "host_cpus": st.get("cpus"),
        "sample": (f"float32 solver.ts restatement in Node {st['node']} on {st['cpuModel']} "
                   f"({st.get('cpus')} host CPUs, one thread used): "
                   + (f"the whole solve, {st['iterations']} PCG iterations to 1e-8 "
                      f"({st['solveTimeMs']:.0f} ms)" if whole else
                      f"{st['iterations']} PCG iterations timed ({st['solveTimeMs']:.0f} ms), "
                      f"extrapolated x {full_iters} {basis}")),
        "ms_per_iteration": ms_per_iter, "iterations": st["iterations"],
    }


def js_ms_per_dof_iteration(n=64, iters=300):
    """Float32 solver.ts restatement in Node: ms per PCG iteration per DOF on
    an n x n clamped Q4 plate (one thread)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    import numpy as np

    from platepcg import problem as P
    mesh = P.rect_mesh(n)
    op = O.OracleProblem(mesh, 210e9, 0.3, 0.1, P.constrained_mask(
        mesh, [{"type": "fixed", "location": "all_edges"}]))
    F = op.pressure_load(1e4).astype(np.float32)
    _, st = O.run_js_solver(n, F, 210e9, 0.3, 0.1, [{"type": "fixed", "location": "all_edges"}],
                            tol=1e-30, max_iter=iters)
    return st["solveTimeMs"] / max(1, st["iterations"]) / (3 * (n + 1) ** 2), st


def bench_batch(a):
    """256 plates (Q4 and DKT, random sizes, materials, supports and loads):
    one autotuned persistent launch per step, every plate to 1e-8 (FP64 true
    residual, restarts included)."""
    import numpy as np

    from platepcg import Batch, Context, Plan
    from platepcg import problem as P
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    plates = P.batch_plates()
    ctx = Context(0)
    b = Batch(ctx, plates)
    # every plate's load integrated on the device straight into the batch
    t0 = time.perf_counter()
    b.pressure_loads([p["q0"] for p in plates], [p["patches"] for p in plates])
    load_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    sched, tuned, trials = b.autotune(tolerance=1e-8)
    tune_s = time.perf_counter() - t0
    for _ in range(a.warmup):
        b.solve(tolerance=1e-8)
    dev, t0 = [], time.perf_counter()
    for _ in range(a.steps):
        st = b.solve(tolerance=1e-8)
        dev.append(st["device_ms"])
    wall = (time.perf_counter() - t0) * 1e3 / a.steps
    its = np.array([b.result(i)[1]["iterations"] for i in range(len(plates))])
    dofs = np.array(b.n_dof)
    dof_its = float(np.sum(its * dofs))
    kern_ms = float(np.mean(dev))
    achieved = BYTES_PER_DOF_ITER_FP64 * dof_its / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": "ms per batch of 256 plates to 1e-8 rel. residual (Q4 + DKT, FP64 block-Jacobi PCG)",
        "value": wall, "unit": "ms/batch", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c4: 256 plates, n log-uniform in [16, 200], alternating Q4 / "
                   "split-quad DKT, t U[0.01, 0.3], nu U[0.25, 0.35], edges fixed/pinned, "
                   "10 kPa + 1-3 Gaussian patches (seed 4)",
                   "dof": int(dofs.sum()), "plates": len(plates),
                   "q4_plates": sum(p["kind"] == "q4" for p in plates),
                   "max_iterations": int(its.max()), "total_iterations": int(its.sum()),
                   "precond": "block-jacobi-3x3", "precision": "fp64"},
        "device_ms_mean": kern_ms, "worst_true_rel_residual": st["worst_true_rel_residual"],
        "load_ms": load_ms,
        "converged": st["converged"],
        "autotune": {"trials": trials, "seconds": tune_s,
                     "schedule": {k: sched[k] for k in ("workgroups_per_cu", "tile_shape",
                                                        "balance_iterations", "resident_plates")},
                     "workgroups": st["workgroups"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch_pcg (persistent, all plates in one launch)",
                     "algorithmic_bytes_per_launch": BYTES_PER_DOF_ITER_FP64 * dof_its},
    }
    if a.validate:
        worst = 0.0
        import oracle_api as O
        for i, p in enumerate(plates):
            col = P.colour_split_quads(p["mesh"]) if p["kind"] == "dkt" else None
            plan = Plan(ctx, p["mesh"], p["E"], p["nu"], p["t"], p["mask"], col)
            F = b.pressure_load(i, p["q0"], p["patches"])
            x1, _, _ = plan.solve(F, tolerance=1e-8, max_iterations=400000)
            worst = max(worst, O.max_nodal_w_error(b.result(i)[0], x1))
            plan.close()
        out["validation"] = {"reference": "single-plate FP64 solve of each plate",
                             "worst_max_nodal_w_error": worst}
    if a.cpu_sample_iters != 0:
        try:
            per, st_js = js_ms_per_dof_iteration()
            out["cpu_baseline"] = {
                "value": per * dof_its, "unit": "ms/batch", "cores": 1, "kind": "port",
                "sample": (f"float32 solver.ts restatement in Node {st_js['node']} on "
                           f"{st_js['cpuModel']}: 300 PCG iterations of a 64 x 64 Q4 plate timed, "
                           f"{per * 1e6:.3f} ns per DOF-iteration, extrapolated x the batch's "
                           f"{dof_its:.3g} DOF-iterations (extrapolation)")}
        except (OSError, subprocess.CalledProcessError) as e:
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))
    b.close()
    ctx.close()


def bench_learned(a):
    """--workload c5: the 1024 x 1024 legs alone, the learned solve as value."""
    res = c5_legs(a)
    ln = res["learned"]
    out = {
        "metric": METRIC, "value": ln["ms"], "unit": "ms/solve", "n_gpus": 1,
        "steps": 3, "warmup": 1, "ms_per_step": ln["ms"], "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64+bf16", "data": "synthetic",
        "config": {"workload": res.pop("workload"), "dof": res["dof"],
                   "precond": "learned U-net (16/32/64, BF16 MFMA) + Galerkin multigrid V-cycle "
                              "smoothing, flexible PCG (FCG(4))",
                   "precision": "fp64 outer, bf16 network"},
    }
    out.update(res)
    print(json.dumps(out))


def timed_leg(plan, o, warmup, steps):
    """warmup + steps solves of the resident load with options o; wall and
    device ms per solve, iterations, true residual of the last."""
    import numpy as np

    for _ in range(warmup):
        plan.solve_resident(o)
    t0, dev = time.perf_counter(), []
    for _ in range(steps):
        st = plan.solve_resident(o)
        dev.append(st.device_ms)
    return {"ms": (time.perf_counter() - t0) * 1e3 / steps, "device_ms": float(np.mean(dev)),
            "iterations": st.iterations, "true_rel_residual": st.true_rel_residual,
            "converged": bool(st.converged), "sweeps": st.sweeps}


def learned_legs(plan, F, train_loads, epochs, warmup, steps, x_ref, seed=5, source="solver"):
    """The learned preconditioner (3-level U-net with a multigrid V-cycle
    around it, flexible PCG) trained first on solver-generated pairs (8 per
    independent load case, not timed): iterates of block-Jacobi PCG
    (source "solver") or of flexible PCG with the network's current
    preconditioner (source "learned"); then timed, and the same
    preconditioner with every network parameter zero."""
    import math

    import numpy as np

    from platepcg import Net, _lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    net = Net(plan, seed=seed, multigrid=True)
    src = {"solver": _lib.TRAIN_SOLVER, "learned": _lib.TRAIN_LEARNED}[source]
    t0 = time.perf_counter()
    first = last = None
    # a few load cases per call, so progress stays visible; each call runs one
    # piece of a single cosine learning-rate decay 1e-3 -> 1e-5 over all calls
    chunk = 1 if source == "solver" else 4
    calls = (train_loads + chunk - 1) // chunk
    lr_at = lambda f: 1e-5 + 0.5 * (1e-3 - 1e-5) * (1.0 + math.cos(math.pi * f))
    for c in range(calls):
        k0, k1 = c * chunk, min(train_loads, (c + 1) * chunk)
        tr = net.train(pairs=k1 - k0, snapshots=8, epochs=epochs, seed0=1000 + k0,
                       lr=lr_at(c / calls), lr_final=lr_at((c + 1) / calls), max_grad_norm=1.0,
                       source=src)
        first = tr["loss_first"] if first is None else first
        last = tr["loss_last"]
        if k1 % 8 == 0 or k1 == train_loads:
            print(f"training {plan.n_dof} DOF: {k1}/{train_loads} load cases, loss "
                  f"{tr['loss_first']:.3f} -> {tr['loss_last']:.3f}, "
                  f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    train_s = time.perf_counter() - t0
    plan.set_rhs(F)  # training replaced the resident load
    net.attach()
    lo = plan.options(tolerance=1e-8, max_iterations=40000, precond=_lib.PRECOND_LEARNED)
    trained = net.params()
    out = {}
    for tag, params in (("learned", trained), ("zero_network", np.zeros_like(trained))):
        net.set_params(params)
        out[tag] = timed_leg(plan, lo, warmup, steps)
        out[tag]["max_nodal_w_error_vs_block_jacobi"] = O.max_nodal_w_error(plan.solution()[0],
                                                                            x_ref)
        print(f"{tag} {plan.n_dof} DOF: {out[tag]['iterations']} iterations, "
              f"{out[tag]['ms']:.1f} ms", file=sys.stderr, flush=True)
    net.detach()
    net.close()
    ok = out["learned"]["converged"] and out["zero_network"]["converged"]
    out["network_iteration_gain"] = (out["zero_network"]["iterations"] /
                                     max(1, out["learned"]["iterations"]) if ok else None)
    out["network_time_gain"] = (out["zero_network"]["ms"] / out["learned"]["ms"] if ok else None)
    out["training"] = {"source": ("block-Jacobi PCG" if source == "solver" else
                                  "flexible PCG (the network's current preconditioner)") +
                                 " iterates (x* - x_k) of independent seeded pressure loads",
                       "pairs": train_loads * 8, "epochs": epochs,
                       "steps": train_loads * 8 * epochs, "seconds": train_s,
                       "loss_first": first, "loss_last": last}
    return out


def c5_legs(a):
    """The 1024 x 1024 variable-thickness plate (3.15M DOF): block-Jacobi PCG,
    multigrid PCG, the learned preconditioner and its zero-network form, each
    to 1e-8 true relative residual."""
    import numpy as np

    from platepcg import Context, Plan, _lib, lognormal_thickness
    from platepcg import problem as P
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    ctx = Context(0)
    t0 = time.perf_counter()
    _, t_e = lognormal_thickness(ctx, 1024, 1024, 10.0, 10.0, 0.5, 5, 0.1, 0.5)
    field_ms = (time.perf_counter() - t0) * 1e3
    w = P.workload("c5", element_t=t_e)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"], element_t=w["element_t"])
    F = plan.pressure_load(w["q0"], w["patches"])
    plan.set_rhs(F)
    res = {"workload": "c5: 1024 x 1024 Q4 plate, t_e = 0.1 exp(0.5 g), g unit Gaussian field "
                       "with correlation length 0.5 m (seed 5), clamped, 10 kPa + 3 Gaussian "
                       "patches", "dof": plan.n_dof,
           "thickness_field": {"where": "device (pp_lognormal_thickness)", "ms": field_ms}}
    bj = timed_leg(plan, plan.options(tolerance=1e-8, max_iterations=1000000), 1, a.c5_steps)
    x_bj = plan.solution()[0]
    res["block_jacobi"] = bj
    print(f"block_jacobi 3.15M DOF: {bj['iterations']} iterations, {bj['ms']:.0f} ms",
          file=sys.stderr, flush=True)
    mg = timed_leg(plan, plan.options(tolerance=1e-8, max_iterations=40000,
                                      precond=_lib.PRECOND_MULTIGRID), 1, 5)
    mg["max_nodal_w_error_vs_block_jacobi"] = O.max_nodal_w_error(plan.solution()[0], x_bj)
    res["multigrid"] = mg
    print(f"multigrid 3.15M DOF: {mg['iterations']} iterations, {mg['ms']:.0f} ms",
          file=sys.stderr, flush=True)
    res.update(learned_legs(plan, F, a.train_loads, a.train_epochs, 1, 3, x_bj,
                            source=a.train_source))
    if res["learned"]["converged"]:
        res["speedup_learned_vs_block_jacobi"] = bj["ms"] / res["learned"]["ms"]
    res["speedup_multigrid_vs_block_jacobi"] = bj["ms"] / mg["ms"]
    n_dof = plan.n_dof
    it_bj = bj["iterations"]
    byt = BYTES_PER_DOF_ITER_FP64 * n_dof * it_bj
    ach = byt / (bj["device_ms"] * 1e-3) / 1e9
    res["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "k_pcg (block-Jacobi FP64 PCG, tiles strided over the CUs)",
                       "algorithmic_bytes_per_launch": byt}
    tr = getattr(a, "c5_traffic", None)
    if tr and tr[0]:
        # the probe launch runs 1,500 iterations of the same plate; its
        # counter bytes per iteration times this solve's iterations
        res["roofline"]["traffic"] = tr[0] / 1500 * it_bj
        res["roofline"]["traffic_measurement"] = dict(tr[1], probe="1,500 iterations in one "
                                                      "launch (bench.py --probe pcg1024)",
                                                      bytes_per_iteration=tr[0] / 1500)
    elif tr:
        res["roofline"]["traffic_measurement"] = tr[1]
    if a.cpu_sample_iters != 0:
        try:
            per, st_js = js_ms_per_dof_iteration()
            res["cpu_baseline"] = {
                "value": per * n_dof * it_bj, "unit": "ms/solve", "cores": 1, "kind": "port",
                "host_cpus": st_js.get("cpus"),
                "sample": (f"float32 solver.ts restatement in Node {st_js['node']} on "
                           f"{st_js['cpuModel']}: 300 PCG iterations of a 64 x 64 Q4 plate timed, "
                           f"{per * 1e6:.3f} ns per DOF-iteration, extrapolated x {n_dof} DOF x "
                           f"{it_bj} block-Jacobi iterations (the reference takes one thickness "
                           f"per plate and cannot run this plate itself; extrapolation)")}
        except (OSError, subprocess.CalledProcessError) as e:
            res["cpu_baseline"] = {"error": str(e)}
    plan.close()
    ctx.close()
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.probe == "copy":
        calibrate_copy(256)
        return
    if a.probe == "pcg":
        probe_pcg()
        return
    if a.probe in ("pcg1024", "mg1024"):
        probe_1024(a.probe == "mg1024")
        return
    if a.probe == "mg142":
        probe_mg142()
        return
    if a.probe == "unet1024":
        probe_unet1024()
        return
    if a.probe == "dkt400":
        probe_dkt400()
        return
    if a.workload == "c4":
        bench_batch(a)
        return
    if a.workload == "c5":
        bench_learned(a)
        return
    default_run = a.workload == "c2"
    traffic, traffic_info = None, None
    if default_run and not a.no_traffic:
        # before this process initialises the GPU: the counter runs are children
        try:
            traffic, traffic_info = measure_traffic()
        except (OSError, subprocess.SubprocessError, ValueError, KeyError) as e:
            traffic, traffic_info = None, {"error": str(e)}
        print(f"traffic: {traffic} ({traffic_info})", file=sys.stderr, flush=True)
        if not a.no_c5:
            try:
                a.c5_traffic = measure_traffic("k_pcg<", "pcg1024")
            except (OSError, subprocess.SubprocessError, ValueError, KeyError) as e:
                a.c5_traffic = (None, {"error": str(e)})
            print(f"c5 traffic: {a.c5_traffic}", file=sys.stderr, flush=True)
    import numpy as np

    from platepcg import Context, Plan, _lib
    from platepcg import problem as P

    w = P.workload(a.workload)
    mixed = a.workload == "c3"
    if a.cpu_sample_iters < 0 and not default_run:
        a.cpu_sample_iters = 150 if mixed else 1200
    ctx = Context(0)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"], w["patches"])
    plan.set_rhs(F)
    prec = _lib.PRECISION_MIXED if mixed else _lib.PRECISION_FP64
    opts = plan.options(tolerance=1e-8, max_iterations=400000, precision=prec,
                        flags=_lib.FLAG_NO_TRUE_RESIDUAL)
    for _ in range(a.warmup):
        plan.solve_resident(opts)
    dev_ms, iters = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st = plan.solve_resident(opts)
        dev_ms.append(st.device_ms)
        iters.append(st.iterations)
    wall = (time.perf_counter() - t0) * 1e3
    chk = plan.solve_resident(plan.options(tolerance=1e-8, max_iterations=400000, precision=prec))
    # PCIe-inclusive: host F in, solve, host x (FP32 and FP64) out
    t1 = time.perf_counter()
    n_e2e = 1 if mixed else 3
    for _ in range(n_e2e):
        plan.solve(F, tolerance=1e-8, max_iterations=400000, precision=prec,
                   flags=_lib.FLAG_NO_TRUE_RESIDUAL)
    e2e_ms = (time.perf_counter() - t1) * 1e3 / n_e2e
    # post-processing of the resident solution (nodal moments, float32 out)
    mom = []
    for _ in range(6):
        t2 = time.perf_counter()
        plan.moments()
        mom.append((time.perf_counter() - t2) * 1e3)
    moments_ms = float(np.median(mom[1:]))
    ms_per_step = wall / a.steps
    n_dof = plan.n_dof
    it = int(np.median(iters))
    kern_ms = float(np.mean(dev_ms))
    bpd = BYTES_PER_DOF_ITER_MIXED if mixed else BYTES_PER_DOF_ITER_FP64
    bytes_per_launch = bpd * n_dof * it
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    if mixed:
        metric = "ms/solve to 1e-8 rel. residual (1.47M DOF, FP32-inner / FP64-refined PCG)"
        kernel = "k_pcg32 (persistent FP32 PCG, MFMA element apply) + FP64 refinement"
    else:
        metric = METRIC if default_run else "ms/solve to 1e-8 rel. residual (FP64 block-Jacobi PCG)"
        kernel = "k_pcg_cg (persistent single-reduction FP64 PCG)"
    out = {
        "metric": metric,
        "value": ms_per_step, "unit": "ms/solve", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {w['mesh']['gridNx']}x{w['mesh']['gridNy']} Q4 "
                   f"Mindlin plate, t={w['t']} m, {n_dof} DOF, clamped, 10 kPa"
                   + (" + Gaussian patch" if len(w["patches"]) else ""),
                   "dof": n_dof, "iterations": it, "precond": "block-jacobi-3x3",
                   "precision": "mixed" if mixed else "fp64"},
        "device_ms_mean": kern_ms, "device_ms_min": float(np.min(dev_ms)),
        "pcie_inclusive_ms": e2e_ms,
        "moments_ms": moments_ms,
        "true_rel_residual": chk.true_rel_residual, "converged": bool(chk.converged),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    if traffic_info is not None:
        out["roofline"]["traffic_measurement"] = traffic_info
    if mixed:
        out["sweeps"] = chk.sweeps
        n = min(chk.sweeps, _lib.MAX_SWEEPS)
        out["sweep_iterations"] = [chk.sweep_iterations[i] for i in range(n)]
        out["sweep_true_rel_residual"] = [chk.sweep_residual[i] for i in range(n)]
    basis = "iterations it needs to reach 1e-8"
    full = it
    if mixed:
        st64 = plan.solve_resident(plan.options(tolerance=1e-8, max_iterations=400000))
        out["fp64_same_plate"] = {"device_ms": st64.device_ms, "iterations": st64.iterations,
                                  "true_rel_residual": st64.true_rel_residual}
        full = st64.iterations
        basis = "iterations of FP64 block-Jacobi PCG to 1e-8 (extrapolation)"
    elif a.workload == "c2":
        full = int(np.load(os.path.join(ROOT, "tests", "golden", "c2_142.npz"))["it_js"])
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    x_bj = plan.solution()[0]  # the block-Jacobi answer (the solves above)
    # the same solve in FP64 with block Jacobi inside a degree-8 Chebyshev
    # polynomial (PP_PRECOND_CHEBYSHEV; FP64 workloads) and with one
    # multigrid V-cycle per iteration (PP_PRECOND_MULTIGRID, hierarchy built
    # in the warm-up): reported alternatives, not the value above
    alts = (() if mixed else (("chebyshev8", _lib.PRECOND_CHEBYSHEV),)) + \
        (("multigrid", _lib.PRECOND_MULTIGRID),)
    legs = {"block_jacobi": {"ms": ms_per_step, "device_ms": kern_ms, "iterations": it,
                             "true_rel_residual": chk.true_rel_residual,
                             "converged": bool(chk.converged)}}
    for key, pc in alts:
        try:
            oc = plan.options(tolerance=1e-8, max_iterations=400000, precond=pc,
                              poly_degree=8 if key == "chebyshev8" else 0)
            legs[key] = timed_leg(plan, oc, max(1, a.warmup), a.steps)
            legs[key]["max_nodal_w_error_vs_block_jacobi"] = O.max_nodal_w_error(
                plan.solution()[0], x_bj)
            legs[key]["speedup_vs_block_jacobi"] = ms_per_step / legs[key]["ms"]
        except Exception as e:  # plates too large for the resident kernel / coarsest level
            legs[key] = {"error": str(e)}
    if default_run:
        legs.update(learned_legs(plan, F, a.c2_train_loads, 2, max(1, a.warmup), a.steps, x_bj))
        if legs["learned"]["converged"]:
            legs["speedup_learned_vs_block_jacobi"] = ms_per_step / legs["learned"]["ms"]
        out["c2"] = legs
    else:
        out.update({k: v for k, v in legs.items() if k != "block_jacobi"})
    if a.cpu_sample_iters != 0:
        if default_run and a.cpu_sample_iters < 0:
            a.cpu_sample_iters = full  # the whole solve, nothing extrapolated
        try:
            out["cpu_baseline"] = cpu_baseline(w, F, full, a.cpu_sample_iters, basis)
        except (OSError, subprocess.CalledProcessError) as e:
            out["cpu_baseline"] = {"error": str(e)}
    plan.close()
    ctx.close()
    if default_run and not a.no_c5:
        out["c5"] = c5_legs(a)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
