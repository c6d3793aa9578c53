"""Newton-GMRES throughput of the implicit-DG step on one MI355X.

Workload: 2D compressible Navier-Stokes, DG p = 3 (10 basis functions, 4 equations,
40 x 40 Jacobian blocks), Re = 5000, M = 0.5, Pr = 0.72 on a seeded perturbed
structured triangulation of [-5,5]^2 around a unit cylinder, 200,000 triangles
(8.0 M DOFs).  One step = one BDF1 pseudo-time step with one Newton iteration:
residual + exact element Jacobian assembly, block-ILU0 factorisation (RCM order),
left-preconditioned GMRES(30) to the reference's default tolerance 1e-2, update
with the positivity limiter.  Synthetic data: the freestream state of the model.

The same steps run twice from the same state: with FP64 preconditioner factors
and sweeps, and mixed (FP32 factors/sweeps under the FP64 residual, Jacobian and
GMRES).  `value` is the mixed-precision DOF-updates/s; the FP64 rate, their ratio
and both normalised residual histories are reported beside it.

`--workload euler3d_tet` measures the 3D configuration instead: compressible Euler on the
Kuhn-split unit cube (64^3 cubes, 1,572,864 tetrahedra, 5 % seeded vertex jitter), M = 0.8 with a
Gaussian density and pressure pulse (amplitudes 0.05, width 0.1), element order 2 (P1: the highest order
the reference's basis3Dtetra builds), BDF1 Newton steps with blockilu0 GMRES.

`--workload batched_ns` measures many independent small solves sharing the GPU: 64 seeded 2D
Navier-Stokes cylinder instances (5k-50k triangles each, element order 2-4, Re in [1e3, 1e4],
M in [0.2, 0.6]), each with its own context and stream, driven concurrently; `value` is the
aggregate DOF-updates/s, and the lift coefficient with its exact gradient is evaluated per
instance after the steps.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the FP32 ILU level kernel on the 400 x 250 cylinder (200,000 triangles):
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes, scaled by the copy kernel of known size (reads
# 2.00 x FETCH_SIZE, writes 1.00 x WRITE_SIZE); profiles/ilu_factor_rows_ns2d_p3_cylinder200k_pmc.md
FACTOR_TRAFFIC_400x250 = 2.00 * 6620.3 * 1024 + 4460.7 * 1024


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_case(n_theta, n_r):
    from flowpro_v2_amd import cases
    return cases.cylinder_mesh(n_theta, n_r)


def run_solver(case, props, bits, warmup, steps, state=None):
    from flowpro_v2_amd import solver
    t0 = time.perf_counter()
    dev = solver.DeviceCase(case, props, precond_bits=bits)
    dev.set_state(dev.initial_state() if state is None else state(dev))
    log(f"fp{bits}: context ready in {time.perf_counter() - t0:.1f} s, dofs {dev.dofs}, "
        f"ILU levels {dev.ilu_levels}, block nonzeros {dev.nnzb}")
    s = solver.GpuImplicitSolver(dev, props)
    s.solve(warmup)
    log(f"fp{bits}: warmup {warmup} steps done")
    t0 = time.perf_counter()
    recs = s.solve(steps)[warmup:]
    elapsed = time.perf_counter() - t0
    updates = sum(r.newton for r in recs)
    iters = [it for r in recs for it in r.gmres_iters]
    log(f"fp{bits}: {steps} steps in {elapsed:.3f} s, GMRES iterations {iters}, residuals "
        + " ".join(f"{r.residuum:.4e}" for r in recs))
    return dev, dict(elapsed=elapsed, updates=updates, iters=iters, residuals=[r.residuum for r in recs],
                     failures=s.linear_failures)


def autotune_summary(dev):
    """The factor kernel's launch shape picked by the library's autotuner, with every shape's time
    per factorisation on this case's matrix."""
    v, shapes, ms = dev.autotune_report()
    if v < 0:
        return None
    return {"kernel": "k_ilu_factor_rows", "chosen": {"task_waves": shapes[v][0], "inversion_waves": shapes[v][1]},
            "factor_ms": {f"{nw}x{ng}": round(t, 3) for (nw, ng), t in zip(shapes, ms)}}


def kernel_table(dev, reps=5):
    names = {0: "residual", 1: "jacobian", 2: "spmv", 3: "ilu_factor", 4: "ilu_apply", 5: "jtv",
             6: "ilu_factor_levels"}
    return {names[k]: dev.time_kernel(k, reps) for k in (0, 1, 2, 3, 4, 6)}


def roofline_jacobian(dev, ms):
    """Algorithmic bytes of one Jacobian-assembly launch: per element row 8 B^2 (1 + n_nbr)
    block bytes written (plus 4 B^2 (1 + n_nbr) for the FP32 copy the block-ILU0 factorisation
    starts from, which the kernel writes with FP32 factors) + residual-input bytes 8 B (2 + 1)
    + geometry and neighbour indices (112 + 12 B for a triangle, 208 + 16 B for a tetrahedron)."""
    B = dev.B
    blocks_bytes = (8.0 + (4.0 if dev.precond_bits == 32 else 0.0)) * B * B * dev.nnzb
    geometry = 112 + 12 if dev.dim == 2 else 208 + 16   # affine map, faces, neighbour indices
    elem_bytes = dev.n_elem * (8.0 * B * 3 + geometry)
    traffic = blocks_bytes + elem_bytes
    achieved = traffic / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, kernel="k_jacobian_team",
                algorithmic_bytes=traffic, avg_launch_ms=ms)


def factor_algorithmic_bytes(case, dev, TT=None):
    """Minimum HBM bytes of one block-ILU0 factorisation in DOF order (rows processed in
    increasing order, lower neighbours k of row i eliminated in increasing order): every block
    of row i read and written once, D_k^-1 read once per lower neighbour k, and U_kj read once
    per update A_ij -= L_ik U_kj (j = i and the common neighbours of i and k above k)."""
    from flowpro_v2_amd import native
    if TT is None:
        TT = native.neighbor_matrix(case.TP, case.bnd)
    order = dev.element_order()
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size, dtype=order.dtype)
    nbr = [set(int(t) for t in row if t >= 0) for row in TT]
    blocks = 0
    for e in range(case.n_elem):
        blocks += 2 * (1 + len(nbr[e]))
        for k in nbr[e]:
            if rank[k] < rank[e]:
                blocks += 1 + 1 + sum(1 for j in nbr[e] if j in nbr[k] and rank[j] > rank[k])
    s = 4 if dev.precond_bits == 32 else 8
    return float(blocks) * dev.B * dev.B * s


def roofline_factor(case, dev, ms_levels, traffic=None, TT=None):
    """The dominant kernel of the step: the level-scheduled block-ILU0 factor kernel
    (k_ilu_factor_rows).  `avg_launch_ms` = HIP-event time of all level launches / number of levels."""
    levels = dev.ilu_levels[0]
    total = factor_algorithmic_bytes(case, dev, TT)
    per_launch = total / levels
    avg = ms_levels / levels
    achieved = per_launch / (avg * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel="k_ilu_factor_rows",
                algorithmic_bytes=per_launch, avg_launch_ms=round(avg, 5), launches=levels)


def roofline_spmv(dev, ms):
    B = dev.B
    traffic = 8.0 * B * B * dev.nnzb + 8.0 * B * 2 * dev.n_elem + 4.0 * dev.nnzb
    achieved = traffic / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel=f"k_bsr_spmv_b<{B}>" if B in (20, 24, 30, 40, 64) else "k_bsr_spmv",
                algorithmic_bytes=traffic, avg_launch_ms=ms)


def cpu_baseline(threads, n_theta=192, n_r=64, nsteps=3):
    """The oracle (C++ restatement of the reference's Java path: FD Jacobian, CRS, sequential
    block ILU0, MGS GMRES, OpenMP element loops) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from flowpro_v2_amd import cases
    case = cases.cylinder_mesh(n_theta, n_r)
    props = cases.ns_props(order=4, CFL=10.0)
    o = O.Oracle(case, props, threads=threads)
    o.run(1)     # first step (setup of CRS / ILU structure happens here as in LinearSolver.factory)
    t0 = time.perf_counter()
    o.run(nsteps)
    el = time.perf_counter() - t0
    rate = o.dofs * nsteps / el
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=round(rate, 1), unit="DOF-updates/s", cores=threads, kind="port",
                host={"nproc": os.cpu_count(), "cpu": model},
                sample=f"{nsteps} Newton steps (FD Jacobian h=1e-8 as the reference, blockilu0 GMRES) of the same "
                       f"NS p=3 Re=5000 M=0.5 case on a {case.n_elem}-triangle cylinder mesh, {el:.1f} s")


def bench_batched(args):
    """Independent instances solved concurrently (flowpro_v2_amd.batch), FP64 then mixed."""
    from flowpro_v2_amd import batch
    t0 = time.perf_counter()
    inst = batch.seeded_ns_instances(n=args.instances, cfl=args.cfl, per_element_orders=args.per_element_orders,
                                     orders=(2, 3, 4, 5) if args.per_element_orders else (2, 3, 4))
    n_elem = sum(c.n_elem for c, _ in inst)
    log(f"{len(inst)} instances, {n_elem} triangles in total, meshes built in {time.perf_counter() - t0:.1f} s")
    out = {}
    for bits in ([64, 32] if args.only is None else ([64] if args.only == "fp64" else [32])):
        t0 = time.perf_counter()
        b = batch.BatchedSolver(inst, precond_bits=bits)
        dofs = b.dofs
        log(f"fp{bits}: {len(inst)} contexts ready in {time.perf_counter() - t0:.1f} s, {dofs} DOFs")
        b.solve(args.warmup)
        t0 = time.perf_counter()
        recs = b.solve(args.steps)
        elapsed = time.perf_counter() - t0
        updates = sum(it.dev.dofs * sum(r.newton for r in rs) for it, rs in zip(b.instances, recs))
        t1 = time.perf_counter()
        lifts = b.functionals(1)
        t_lift = time.perf_counter() - t1
        iters = [sum(it for r in rs for it in r.gmres_iters) for rs in recs]
        out[bits] = dict(rate=updates / elapsed, elapsed=elapsed, updates=updates, dofs=dofs, iters=iters,
                         final=[rs[-1].residuum for rs in recs], lift_ms=1e3 * t_lift,
                         lift=[v for v, _ in lifts])
        log(f"fp{bits}: {args.steps} steps of every instance in {elapsed:.3f} s, {updates / elapsed:.4g} DOF-updates/s; "
            f"lift and gradient of all instances in {1e3 * t_lift:.1f} ms")
        b.close()
    main_bits = 32 if 32 in out else 64
    r = out[main_bits]
    # the same steps with the instances one after another (one host thread): the gain of sharing the GPU
    b = batch.BatchedSolver(inst, precond_bits=main_bits, workers=1)
    b.solve(args.warmup)
    t0 = time.perf_counter()
    recs = b.solve(args.steps)
    serial = sum(it.dev.dofs * sum(x.newton for x in rs) for it, rs in zip(b.instances, recs)) / (time.perf_counter() - t0)
    b.close()
    log(f"fp{main_bits}, instances one after another: {serial:.4g} DOF-updates/s")
    result = {
        "metric": "Newton-GMRES DOF-updates/s, mixed vs FP64 GPU at equal residual; % HBM peak",
        "value": round(r["rate"], 1),
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/Jacobian/GMRES, f32 preconditioner" if main_bits == 32 else "f64",
        "data": "synthetic (seeded cylinder meshes, freestream states, seeded Re and M per instance)",
        "config": {"workload": f"{len(inst)} independent NS2D cylinder instances, 5k-50k triangles each, element order "
                               f"{'2-5 per element' if args.per_element_orders else '2-4 per instance'}, Re U[1e3,1e4], "
                               f"M U[0.2,0.6], BDF1 CFL {args.cfl} ramp, blockilu0 (RCM) GMRES(30) "
                               "tol 1e-2, one context and stream per instance",
                   "instances": len(inst), "elements": n_elem, "dofs": r["dofs"], "precond_bits": main_bits},
        "fp64_value": round(out[64]["rate"], 1) if 64 in out else None,
        "mixed_over_fp64": round(out[32]["rate"] / out[64]["rate"], 3) if (32 in out and 64 in out) else None,
        "serial_value": round(serial, 1),
        "final_residual_max_rel_diff": (max(abs(a / b - 1) for a, b in zip(out[32]["final"], out[64]["final"]))
                                        if (32 in out and 64 in out) else None),
        "lift_and_gradient_ms": round(r["lift_ms"], 2),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def bench_steady(args):
    """Time to a steady state: the 200k-triangle cylinder at Re = 40 (steady flow), NS p=3, M = 0.5,
    CFL varied by the residual (State/CFLSetup with CFL <= 0, maxCFL 1e3) until the relative residual
    (State.residuum) falls below --tol, from the freestream; FP64 preconditioner, then FP32.  Each
    context factors once before the timed solve so that the autotuner's trials are not timed."""
    from flowpro_v2_amd import cases, solver
    case = make_case(args.n_theta, args.n_r)
    props = cases.ns_props(order=4, reynolds=40.0, CFL=-1.0, maxCFL=1e3, residuum=args.tol)
    log(f"mesh {case.name}: {case.n_elem} triangles, Re 40, to residuum {args.tol:g}")
    out, states = {}, {}
    for bits in (64, 32):
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        w0 = dev.initial_state()
        dev.set_state(w0)
        dev.assemble(1e-3, -1.0, 1, True)
        dev.linear_solve()
        dev.set_state(w0)
        s = solver.GpuImplicitSolver(dev, props)
        t0 = time.perf_counter()
        while len(s.history) < args.max_steps and (not s.history or s.history[-1].residuum > args.tol):
            n0 = len(s.history)
            s.solve(min(10, args.max_steps - n0))   # GpuImplicitSolver.solve continues and returns the whole history
            if len(s.history) == n0:
                break
            log(f"fp{bits}: step {len(s.history)}, residuum {s.history[-1].residuum:.3e}, CFL {s.state.current_cfl:.3g}, "
                f"{time.perf_counter() - t0:.1f} s")
        elapsed = time.perf_counter() - t0
        recs = s.history
        iters = sum(it for r in recs for it in r.gmres_iters)
        out[bits] = dict(seconds=elapsed, steps=len(recs), residuum=recs[-1].residuum, gmres=iters,
                         updates=sum(r.newton for r in recs), dofs=dev.dofs)
        states[bits] = dev.get_state()
        log(f"fp{bits}: {len(recs)} steps, {iters} GMRES iterations, residuum {recs[-1].residuum:.3e} in {elapsed:.2f} s")
        dev.close()
    diff = float(np.abs(states[32] - states[64]).max() / np.abs(states[64]).max())
    reached = all(out[b]["residuum"] <= args.tol for b in out)
    result = {
        "metric": f"Newton-GMRES wall time to a {args.tol:g} relative steady residual, mixed vs FP64 preconditioner",
        "value": round(out[32]["seconds"], 3),
        "unit": "s",
        "n_gpus": 1,
        "steps": out[32]["steps"],
        "warmup": 0,
        "ms_per_step": round(1e3 * out[32]["seconds"] / max(out[32]["steps"], 1), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/Jacobian/GMRES, f32 preconditioner",
        "data": "synthetic (freestream start, seeded mesh)",
        "config": {"workload": f"NS2D DG p=3 Re=40 M=0.5 cylinder (steady), {case.n_elem} triangles, {out[32]['dofs']} DOFs, "
                               f"BDF1, CFL varied by the residual (maxCFL 1e3), blockilu0 (RCM) GMRES(30) tol 1e-2, "
                               f"until residuum <= {args.tol:g}",
                   "elements": case.n_elem, "dofs": out[32]["dofs"], "order": 4},
        "reached_tolerance": reached,
        "fp64_seconds": round(out[64]["seconds"], 3),
        "speedup_mixed_over_fp64": round(out[64]["seconds"] / out[32]["seconds"], 3),
        "steps_fp64": out[64]["steps"],
        "gmres_iterations": {"fp64": out[64]["gmres"], "fp32": out[32]["gmres"]},
        "final_residuum": {"fp64": out[64]["residuum"], "fp32": out[32]["residuum"]},
        "state_max_rel_diff_mixed_vs_fp64": diff,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-theta", type=int, default=400)
    ap.add_argument("--n-r", type=int, default=250)
    ap.add_argument("--cfl", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--workload", choices=["ns2d_cylinder", "euler3d_tet", "batched_ns", "mixed_ns", "ns2d_steady"],
                    default="ns2d_cylinder")
    ap.add_argument("--tol", type=float, default=1e-8, help="ns2d_steady: the relative steady residual to reach")
    ap.add_argument("--max-steps", type=int, default=400, help="ns2d_steady: step limit")
    ap.add_argument("--mx", type=int, default=520, help="channel cells along x of the mixed_ns workload")
    ap.add_argument("--my", type=int, default=256, help="channel cells along y of the mixed_ns workload")
    ap.add_argument("--instances", type=int, default=64, help="instances of the batched_ns workload")
    ap.add_argument("--per-element-orders", action="store_true",
                    help="batched_ns: every element of an instance draws its order from 2-5 (order.txt)")
    ap.add_argument("--cube", type=int, default=64, help="cubes per edge of the euler3d_tet workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU design is "
              "element partitioning with a halo layer, not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched_ns":
        bench_batched(args)
        return
    if args.workload == "ns2d_steady":
        bench_steady(args)
        return
    from flowpro_v2_amd import cases, native
    tet = args.workload == "euler3d_tet"
    mixed = args.workload == "mixed_ns"
    state = None
    TT = None
    if mixed:
        # squares and triangles (elementType.txt) with per-element orders 2-4 (order.txt)
        case = cases.mixed_channel(args.mx, args.my, jitter=0.2, seed=7)
        case.orders = cases.seeded_orders(case.n_elem, (2, 3, 4), 3)
        props = cases.ns_props(order=4, CFL=args.cfl)
        TT = native.neighbor_matrix_mixed(case.TP, case.types, case.bnd)
        log(f"mesh {case.name}: {case.n_elem} elements ({int((case.types == 4).sum())} squares), orders 2-4")
    elif tet:
        case = cases.kuhn_cube(args.cube, 0.05, 4)
        props = cases.euler3d_props(mach=0.8, order=2, CFL=args.cfl)
        TT = native.neighbor_matrix_tet(case.TP, case.bnd)

        def state(dev):
            return cases.density_pulse_state(case, dev.nb, dev.initial_state()[0][:: dev.nb], amp=0.05, width=0.1,
                                             pressure_amp=0.05)
        log(f"mesh {case.name}: {case.n_elem} tetrahedra")
    else:
        case = make_case(args.n_theta, args.n_r)
        props = cases.ns_props(order=4, CFL=args.cfl)
        log(f"mesh {case.name}: {case.n_elem} triangles")
    out = {}
    devs = {}
    for bits in ([64, 32] if args.only is None else ([64] if args.only == "fp64" else [32])):
        dev, res = run_solver(case, props, bits, args.warmup, args.steps, state)
        res["rate"] = dev.dofs * res["updates"] / res["elapsed"]
        out[bits] = res
        if bits == 32 or args.only is not None:
            kt = kernel_table(dev)
            log("kernel ms per launch: " + json.dumps({k: round(v, 4) for k, v in kt.items()}))
            devs["kt"] = kt
            devs["dev"] = dev
        else:
            dev.close()
    dev = devs["dev"]
    kt = devs["kt"]
    main_bits = 32 if 32 in out else 64
    r = out[main_bits]
    default_mesh = not tet and not mixed and (args.n_theta, args.n_r) == (400, 250) and main_bits == 32
    roof = roofline_factor(case, dev, kt["ilu_factor_levels"], FACTOR_TRAFFIC_400x250 if default_mesh else None, TT)
    others = {"k_jacobian_team": roofline_jacobian(dev, kt["jacobian"]), "k_bsr_spmv": roofline_spmv(dev, kt["spmv"])}
    if tet:
        others = {"k_jacobian_tet_team": dict(roofline_jacobian(dev, kt["jacobian"]), kernel="k_jacobian_tet_team"),
                  "k_bsr_spmv": roofline_spmv(dev, kt["spmv"])}
    if mixed:
        # blocks of the padded size B = 16 nEqs; the factor runs the per-level block kernels at B = 64
        roof["kernel"] = "k_ilu_factor_level"
        others = {"k_jacobian_mixed": dict(roofline_jacobian(dev, kt["jacobian"]), kernel="k_jacobian_mixed"),
                  "k_bsr_spmv": roofline_spmv(dev, kt["spmv"])}
    result = {
        "metric": "Newton-GMRES DOF-updates/s, mixed vs FP64 GPU at equal residual; % HBM peak",
        "value": round(r["rate"], 1),
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/Jacobian/GMRES, f32 preconditioner" if main_bits == 32 else "f64",
        "data": "synthetic (freestream state, seeded mesh)",
        "config": ({"workload": f"NS2D DG Re=5000 M=0.5 channel with a bump, {case.n_elem} elements mixing squares and "
                                f"triangles, per-element orders 2-4 (padded to 16 basis functions), {dev.dofs} DOFs, "
                                f"BDF1 CFL {args.cfl} ramp, blockilu0 (RCM) GMRES(30) tol 1e-2",
                    "elements": case.n_elem, "dofs": dev.dofs, "order": "2-4 per element", "precond_bits": main_bits}
                   if mixed else
                   {"workload": f"Euler3D DG P1 (order 2) Kuhn cube {args.cube}^3 x 6 = {case.n_elem} tetrahedra, "
                                f"{dev.dofs} DOFs, M=0.8 density and pressure pulse, BDF1 CFL {args.cfl} ramp, blockilu0 (RCM) "
                                "GMRES(30) tol 1e-2",
                    "elements": case.n_elem, "dofs": dev.dofs, "order": 2, "precond_bits": main_bits} if tet else
                   {"workload": f"NS2D DG p=3 Re=5000 M=0.5 cylinder, {case.n_elem} triangles, "
                                f"{dev.dofs} DOFs, BDF1 CFL {args.cfl} ramp, blockilu0 (RCM) GMRES(30) tol 1e-2",
                    "elements": case.n_elem, "dofs": dev.dofs, "order": 4, "precond_bits": main_bits}),
        "fp64_value": round(out[64]["rate"], 1) if 64 in out else None,
        "mixed_over_fp64": round(out[32]["rate"] / out[64]["rate"], 3) if (32 in out and 64 in out) else None,
        "residuals": {f"fp{b}": [float(f"{v:.6e}") for v in out[b]["residuals"]] for b in out},
        "gmres_iterations": {f"fp{b}": out[b]["iters"] for b in out},
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
        "roofline": roof,
        "other_kernels": others,
        "autotune": autotune_summary(dev),
    }
    dev.close()
    if tet:
        result["data"] = "synthetic (freestream with a Gaussian density and pressure pulse, seeded jittered mesh)"
        result["cpu_baseline"] = None   # the CPU restatement of the tetrahedra covers residual and Jacobian only
    elif mixed:
        result["data"] = "synthetic (freestream state, seeded mesh, element types and orders)"
        result["cpu_baseline"] = None
    elif not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        result["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
