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
# HBM bytes of one launch of the FP32 dataflow factor kernel (the whole factorisation) on the 400 x 250
# cylinder (200,000 triangles): rocprofv3 FETCH_SIZE and WRITE_SIZE passes, scaled by the copy kernel of
# known size (reads 2.00 x FETCH_SIZE, writes 1.00 x WRITE_SIZE): 4,259,063.5 KiB fetched, 3,135,064 KiB
# written per dispatch; profiles/ilu_factor_flow_ns2d_p3_cylinder200k_pmc.md
FACTOR_TRAFFIC_400x250 = 2.00 * 4259063.5 * 1024 + 3135064.0 * 1024


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
                     failures=s.linear_failures, last_dt=recs[-1].dt)


def autotune_summary(dev):
    """The factor kernel's launch shape picked by the library's autotuner, with every shape's time
    per factorisation on this case's matrix."""
    v, shapes, ms = dev.autotune_report()
    if v < 0:
        return None
    def name(nw, ng):   # ng < 0: the task waves split the block's columns in the inversion
        return f"{nw}x{ng}" if ng > 0 else f"{nw}x{-ng}split"
    kernel = "k_ilu_factor_flow" if factor_kernel_name(dev)[1] else "k_ilu_factor_rows"
    return {"kernel": kernel, "chosen": {"task_waves": shapes[v][0], "inversion_waves": abs(shapes[v][1]),
                                                      "split_inversion": shapes[v][1] < 0},
            "factor_ms": {name(nw, ng): round(t, 3) for (nw, ng), t in zip(shapes, ms)}}


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


def factor_kernel_name(dev):
    """The factor kernel the context runs, from the autotuner's report: the dataflow row kernel
    (one launch per factorisation) when the row kernel serves the block size, else the per-level
    block kernels."""
    v, shapes, _ = dev.autotune_report()
    if v < 0:
        return "k_ilu_factor_level_4w" if dev.B >= 24 else "k_ilu_factor_level_w", False
    nw, ng = shapes[v]
    flow = os.environ.get("FLOWPRO_ILU_FLOW", "1") != "0"
    mr = 5 if dev.B in (20, 64) else 4
    tf = "float" if dev.precond_bits == 32 else "double"
    return f"k_ilu_factor_{'flow' if flow else 'rows'}<{tf}, {dev.B}, {mr}, {nw}, {ng}>", flow


def roofline_factor(case, dev, ms_factor, traffic=None, TT=None):
    """The dominant kernel of the step: the block-ILU0 factorisation.  With the dataflow kernel one
    launch factors the whole matrix (`avg_launch_ms` = the HIP-event time of the launch); with the
    per-level kernels the launch is one level (time / number of levels).  Algorithmic bytes per
    row: 2 w B^2 (1 + n_nbr), every block of the row read and written once; the count that adds one
    read of D_k^-1 and of U_kj per use is reported beside it."""
    name, flow = factor_kernel_name(dev)
    launches = 1 if flow else dev.ilu_levels[0]
    w = 4 if dev.precond_bits == 32 else 8
    total = 2.0 * w * dev.B * dev.B * dev.nnzb
    wide = factor_algorithmic_bytes(case, dev, TT)
    avg = ms_factor / launches
    achieved = total / launches / (avg * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=name,
                algorithmic_bytes=total / launches, algorithmic_bytes_with_operand_reads=wide / launches,
                frac_with_operand_reads=round(wide / launches / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                avg_launch_ms=round(avg, 5), launches=launches)


def roofline_apply(dev, ms):
    """One block-ILU0 application (forward and backward dataflow sweeps): every factor block read
    once, w B^2 (1 + n_nbr) per row, plus the right-hand side, the two published vectors and the
    result (8 B + 2 x 8 B + 8 B per DOF in FP32; the packed values carry their tag)."""
    w = 4 if dev.precond_bits == 32 else 8
    traffic = w * dev.B * dev.B * dev.nnzb + 8.0 * dev.dofs * 4
    achieved = traffic / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel=f"k_ilu_sweep_flow<{'float' if w == 4 else 'double'}, {dev.B}, 4, 0|1>",
                algorithmic_bytes=traffic, avg_launch_ms=round(ms, 4), launches=2)


def roofline_spmv(dev, ms):
    B = dev.B
    traffic = 8.0 * B * B * dev.nnzb + 8.0 * B * 2 * dev.n_elem + 4.0 * dev.nnzb
    achieved = traffic / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel=f"k_bsr_spmv_b<{B}>" if B in (20, 24, 30, 40, 64) else "k_bsr_spmv",
                algorithmic_bytes=traffic, avg_launch_ms=ms)


def adjoint_measurement(dev, kt, last_dt):
    """The adjoint path at the workload's size, at the state the timed steps reached: the
    matrix-free J^T v launch, and drag adjoints lambda = (a1 M - dR/dW)^-T dI/dW by GMRES on J^T v
    preconditioned with the transposed block-ILU0 of a1 M - dR/dW (assembly and factorisation
    included in the time): a1 = 1/dt of the last step (the adjoint of that BDF1 step) and a1 = 0
    (the steady form of FlowPro/toolbox/optimisation/getFunctionalGradient.m:5-12, which the
    reference applies at a converged state; this state is a transient one)."""
    jtv_ms = dev.time_kernel(5, 5)
    res = {"jtv_ms": round(jtv_ms, 4), "functional": "DragCoefficient",
           "preconditioner": "transposed block-ILU0 of a1 M - dR/dW (U^T then L^T dataflow sweeps)"}
    for name, a1 in (("step_adjoint", 1.0 / last_dt), ("steady_adjoint", 0.0)):
        t0 = time.perf_counter()
        _, g = dev.functional(0)
        rc, lam, it, err = dev.adjoint_solve(g, a1=a1)
        el = time.perf_counter() - t0
        log(f"adjoint ({name}, a1 = {a1:.4g}): rc {rc}, {it} GMRES iterations, rel. residual {err:.2e}, {el:.3f} s")
        res[name] = {"a1": a1, "seconds": round(el, 3), "gmres_iterations": it, "relative_residual": err,
                     "converged": rc == 0}
    log(f"adjoint: J^T v {jtv_ms:.3f} ms")
    return res


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
    """Independent instances solved concurrently (flowpro_v2_amd.batch), FP64 then mixed.  --mixed-types:
    meshes of squares and triangles; --steady: every instance runs the SER CFL law until its residuum
    reaches --tol or --max-steps steps; then the lift, its exact gradient and its adjoint per instance
    (a1 = 0 after steady runs, else a1 = 1/dt of the instance's last step)."""
    from flowpro_v2_amd import batch
    t0 = time.perf_counter()
    steady = {"maxCFL": args.max_cfl, "residuum": args.tol} if args.steady else None
    inst = batch.seeded_ns_instances(n=args.instances, cfl=args.cfl, per_element_orders=args.per_element_orders,
                                     orders=(2, 3, 4, 5) if args.per_element_orders else (2, 3, 4),
                                     mixed_types=args.mixed_types, steady=steady)
    n_elem = sum(c.n_elem for c, _ in inst)
    n_sq = sum(int((c.types == 4).sum()) for c, _ in inst if getattr(c, "types", None) is not None)
    log(f"{len(inst)} instances, {n_elem} elements ({n_sq} squares) in total, meshes built in {time.perf_counter() - t0:.1f} s")
    out = {}
    steps = args.max_steps if args.steady else args.steps
    warm = 0 if args.steady else args.warmup   # a steady run is timed from the freestream
    for bits in ([64, 32] if args.only is None else ([64] if args.only == "fp64" else [32])):
        t0 = time.perf_counter()
        b = batch.BatchedSolver(inst, precond_bits=bits)
        dofs = b.dofs
        log(f"fp{bits}: {len(inst)} contexts ready in {time.perf_counter() - t0:.1f} s, {dofs} DOFs")
        b.solve(warm)
        t0 = time.perf_counter()
        recs = b.solve(steps)
        elapsed = time.perf_counter() - t0
        recs = [rs[warm:] for rs in recs]   # the solver returns the whole history, warm-up steps included
        updates = sum(it.dev.dofs * sum(r.newton for r in rs) for it, rs in zip(b.instances, recs))
        t1 = time.perf_counter()
        lifts = b.functionals(1)
        t_lift = time.perf_counter() - t1
        t1 = time.perf_counter()
        a1 = [0.0 if args.steady else 1.0 / rs[-1].dt for rs in recs]
        adj = b.adjoints(1, a1)
        t_adj = time.perf_counter() - t1
        iters = [sum(it for r in rs for it in r.gmres_iters) for rs in recs]
        out[bits] = dict(rate=updates / elapsed, elapsed=elapsed, updates=updates, dofs=dofs, iters=iters,
                         final=[rs[-1].residuum for rs in recs], lift_ms=1e3 * t_lift,
                         lift=[v for v, _ in lifts], steps=[len(rs) for rs in recs],
                         adjoint=dict(seconds=round(t_adj, 3), a1="0 (steady)" if args.steady else "1/dt of the last step",
                                      converged=sum(1 for rc, _, _, _ in adj if rc == 0),
                                      gmres_iterations=[it for _, _, it, _ in adj],
                                      max_relative_residual=max(err for _, _, _, err in adj)))
        log(f"fp{bits}: {elapsed:.3f} s for {sum(len(rs) for rs in recs)} instance steps, {updates / elapsed:.4g} DOF-updates/s; "
            f"lift and gradient of all instances in {1e3 * t_lift:.1f} ms; lift adjoints in {t_adj:.2f} s "
            f"({out[bits]['adjoint']['converged']} of {len(adj)} converged)")
        if args.steady:
            reached = sum(1 for rs in recs if rs[-1].residuum <= args.tol)
            log(f"fp{bits}: {reached} of {len(recs)} instances reached residuum {args.tol:g}; final residua "
                + " ".join(f"{rs[-1].residuum:.1e}" for rs in recs))
            out[bits]["reached"] = reached
        b.close()
    main_bits = 32 if 32 in out else 64
    r = out[main_bits]
    # the same steps with the instances one after another (one host thread): the gain of sharing the GPU
    b = batch.BatchedSolver(inst, precond_bits=main_bits, workers=1)
    b.solve(warm)
    t0 = time.perf_counter()
    recs = b.solve(steps)
    dt_serial = time.perf_counter() - t0
    # GpuImplicitSolver.solve returns the whole history: the warm-up steps are not counted
    serial = sum(it.dev.dofs * sum(x.newton for x in rs[warm:]) for it, rs in zip(b.instances, recs)) / dt_serial
    b.close()
    log(f"fp{main_bits}, instances one after another: {serial:.4g} DOF-updates/s")
    result = {
        "metric": "Newton-GMRES DOF-updates/s, mixed vs FP64 GPU at equal residual; % HBM peak",
        "value": round(r["rate"], 1),
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": round(1e3 * r["elapsed"] / max(1, max(r["steps"])), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/Jacobian/GMRES, f32 preconditioner" if main_bits == 32 else "f64",
        "data": "synthetic (seeded cylinder meshes, freestream states, seeded Re and M per instance)",
        "config": {"workload": f"{len(inst)} independent NS2D cylinder instances, 5k-50k "
                               f"{'squares and triangles' if args.mixed_types else 'triangles'} each, element order "
                               f"{'2-5 per element' if args.per_element_orders else '2-4 per instance'}, Re U[1e3,1e4], "
                               f"M U[0.2,0.6], BDF1 "
                               + (f"SER CFL (maxCFL {args.max_cfl:g}) to residuum {args.tol:g} or {steps} steps"
                                  if args.steady else f"CFL {args.cfl} ramp")
                               + ", blockilu0 (RCM) GMRES(30) tol 1e-2, one context and stream per instance",
                   "instances": len(inst), "elements": n_elem, "dofs": r["dofs"], "precond_bits": main_bits},
        "fp64_value": round(out[64]["rate"], 1) if 64 in out else None,
        "mixed_over_fp64": round(out[32]["rate"] / out[64]["rate"], 3) if (32 in out and 64 in out) else None,
        "serial_value": round(serial, 1),
        "final_residual_max_rel_diff": (max(abs(a / b - 1) for a, b in zip(out[32]["final"], out[64]["final"]))
                                        if (32 in out and 64 in out) else None),
        "lift_and_gradient_ms": round(r["lift_ms"], 2),
        "lift_adjoint": r["adjoint"],
        "steps_per_instance": r["steps"],
        "steady_reached": r.get("reached"),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def bench_steady(args):
    """Time to a steady state: the 200k-triangle cylinder at Re = 40 (steady flow), NS p=3, M = 0.5,
    CFL varied by the residual (State/CFLSetup with CFL <= 0, SER up to --max-cfl) until the relative
    residual (State.residuum) falls below --tol, from the freestream; FP64 preconditioner, then FP32.
    Each context factors once before the timed solve so that the autotuner's trials are not timed.
    The roofline entry is the block-ILU0 application (the two dataflow sweeps), which with the
    SpMV dominates the GMRES iterations of the high-CFL steps."""
    from flowpro_v2_amd import cases, solver
    case = make_case(args.n_theta, args.n_r)
    props = cases.ns_props(order=4, reynolds=40.0, CFL=-1.0, maxCFL=args.max_cfl, residuum=args.tol)
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
        if bits == 32:
            kt = kernel_table(dev)
            out[bits]["kernel_ms"] = {k: round(v, 4) for k, v in kt.items()}
            out[bits]["roofline"] = roofline_apply(dev, kt["ilu_apply"])
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
                               f"BDF1, CFL varied by the residual (SER, maxCFL {args.max_cfl:g}), blockilu0 (RCM) GMRES(30) tol 1e-2, "
                               f"until residuum <= {args.tol:g}",
                   "elements": case.n_elem, "dofs": out[32]["dofs"], "order": 4},
        "reached_tolerance": reached,
        "fp64_seconds": round(out[64]["seconds"], 3),
        "speedup_mixed_over_fp64": round(out[64]["seconds"] / out[32]["seconds"], 3),
        "steps_fp64": out[64]["steps"],
        "gmres_iterations": {"fp64": out[64]["gmres"], "fp32": out[32]["gmres"]},
        "final_residuum": {"fp64": out[64]["residuum"], "fp32": out[32]["residuum"]},
        "state_max_rel_diff_mixed_vs_fp64": diff,
        "kernel_ms": out[32]["kernel_ms"],
        "roofline": out[32]["roofline"],
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        cb = cpu_baseline(threads)
        cb["note"] = ("the CPU path's Newton-update rate on a bounded sample of the same NS p3 case; at that rate the "
                      f"{out[32]['updates']} updates of {out[32]['dofs']} DOFs the mixed run needed would take "
                      f"{out[32]['updates'] * out[32]['dofs'] / cb['value']:.0f} s")
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)


def bench_kernels(args):
    """Each kernel family of the default workload a few times on one FP32-preconditioned context (no Newton
    steps): the short program the counter passes of rocprofv3 run."""
    from flowpro_v2_amd import cases, solver
    case = make_case(args.n_theta, args.n_r)
    dev = solver.DeviceCase(case, cases.ns_props(order=4, CFL=args.cfl), precond_bits=32)
    w0 = dev.initial_state()
    dev.set_state(cases.perturbed_state(w0[0][:: dev.nb], case.n_elem, dev.nb, 0.01, 3))
    dev.assemble(1e-3, -1.0, 1, True)
    names = {0: "residual", 1: "jacobian", 2: "spmv", 6: "ilu_factor", 4: "ilu_apply", 5: "jtv"}
    res = {names[k]: round(dev.time_kernel(k, 3), 4) for k in (0, 1, 2, 6, 4, 5)}
    log("kernel ms per launch: " + json.dumps(res))
    dev.close()
    print(json.dumps({"workload": f"kernels of NS2D DG p=3 cylinder, {case.n_elem} triangles", "kernel_ms": res}), flush=True)


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
    ap.add_argument("--workload", choices=["ns2d_cylinder", "euler3d_tet", "batched_ns", "mixed_ns", "ns2d_steady", "sa_plate"],
                    default="ns2d_cylinder")
    ap.add_argument("--tol", type=float, default=1e-8, help="ns2d_steady: the relative steady residual to reach")
    ap.add_argument("--max-steps", type=int, default=1200, help="ns2d_steady: step limit")
    ap.add_argument("--max-cfl", type=float, default=300.0, help="ns2d_steady: maxCFL of the SER law")
    ap.add_argument("--mx", type=int, default=520, help="channel cells along x of the mixed_ns workload")
    ap.add_argument("--my", type=int, default=256, help="channel cells along y of the mixed_ns workload")
    ap.add_argument("--instances", type=int, default=64, help="instances of the batched_ns workload")
    ap.add_argument("--mixed-types", action="store_true",
                    help="batched_ns: meshes of squares and triangles (elementType.txt per instance)")
    ap.add_argument("--steady", action="store_true",
                    help="batched_ns: SER CFL law to --tol within --max-steps steps per instance")
    ap.add_argument("--per-element-orders", action="store_true",
                    help="batched_ns: every element of an instance draws its order from 2-5 (order.txt)")
    ap.add_argument("--cube", type=int, default=64, help="cubes per edge of the euler3d_tet workload")
    ap.add_argument("--plate-nx", type=int, default=3907, help="cells along the sa_plate channel (3907 x 64 x 2 = 500k)")
    ap.add_argument("--plate-ny", type=int, default=64, help="wall-normal cells of the sa_plate channel")
    ap.add_argument("--sa-cfl", type=float, default=2.0, help="CFL of the sa_plate workload (ramped over 20 steps)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time only the kernel families (residual, Jacobian, SpMV, factorisation, apply, J^T v) on the "
                         "default mesh at a seeded 1%% perturbation of the freestream, for profilers")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU design is "
              "element partitioning with a halo layer, not built)", file=sys.stderr)
        sys.exit(2)
    if args.kernels_only:
        bench_kernels(args)
        return
    if args.workload == "batched_ns":
        bench_batched(args)
        return
    if args.workload == "ns2d_steady":
        bench_steady(args)
        return
    from flowpro_v2_amd import cases, native
    tet = args.workload == "euler3d_tet"
    mixed = args.workload == "mixed_ns"
    sa = args.workload == "sa_plate"
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
    elif sa:
        # turbulent flat plate: stretched channel, first cell 1e-5, growth 1.15 (wall distance by
        # generateDistanceFunction), SpalartAllmaras p=2, Re 1e6, M 0.3, BDF2
        case = cases.plate_channel(args.plate_nx, args.plate_ny, first=1e-5, growth=1.15)
        props = cases.sa_props(mach=0.3, reynolds=1e6, order=3, CFL=args.sa_cfl, timeOrder=2)
        log(f"mesh {case.name}: {case.n_elem} triangles, wall spacing 1e-5, growth 1.15")
    else:
        case = make_case(args.n_theta, args.n_r)
        props = cases.ns_props(order=4, CFL=args.cfl)
        log(f"mesh {case.name}: {case.n_elem} triangles")
    out = {}
    devs = {}
    for bits in ([64, 32] if args.only is None else ([64] if args.only == "fp64" else [32])):
        dev, res = run_solver(case, props, bits, args.warmup, args.steps, state)
        res["rate"] = dev.dofs * res["updates"] / res["elapsed"]
        res["state"] = dev.get_state()
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
    default_mesh = not tet and not mixed and not sa and (args.n_theta, args.n_r) == (400, 250) and main_bits == 32
    roof = roofline_factor(case, dev, kt["ilu_factor_levels"], FACTOR_TRAFFIC_400x250 if default_mesh else None, TT)
    adjoint = None
    if not tet and not mixed and not sa:
        adjoint = adjoint_measurement(dev, kt, out[main_bits]["last_dt"])
    others = {"k_jacobian_team": roofline_jacobian(dev, kt["jacobian"]), "k_bsr_spmv": roofline_spmv(dev, kt["spmv"])}
    if sa:
        others = {"k_jacobian_team": dict(roofline_jacobian(dev, kt["jacobian"]), kernel="k_jacobian_team (SA)"),
                  "k_bsr_spmv": roofline_spmv(dev, kt["spmv"])}
    if tet:
        others = {"k_jacobian_tet_team": dict(roofline_jacobian(dev, kt["jacobian"]), kernel="k_jacobian_tet_team"),
                  "k_bsr_spmv": roofline_spmv(dev, kt["spmv"])}
    if mixed:
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
                   {"workload": f"RANS Spalart-Allmaras DG p=2 Re=1e6 M=0.3 flat plate, stretched channel {case.n_elem} "
                                f"triangles (first cell 1e-5, growth 1.15), {dev.dofs} DOFs, BDF2 CFL {args.sa_cfl} ramp, "
                                "blockilu0 (RCM) GMRES(30) tol 1e-2",
                    "elements": case.n_elem, "dofs": dev.dofs, "order": 3, "precond_bits": main_bits} if sa else
                   {"workload": f"NS2D DG p=3 Re=5000 M=0.5 cylinder, {case.n_elem} triangles, "
                                f"{dev.dofs} DOFs, BDF1 CFL {args.cfl} ramp, blockilu0 (RCM) GMRES(30) tol 1e-2",
                    "elements": case.n_elem, "dofs": dev.dofs, "order": 4, "precond_bits": main_bits}),
        "fp64_value": round(out[64]["rate"], 1) if 64 in out else None,
        "mixed_over_fp64": round(out[32]["rate"] / out[64]["rate"], 3) if (32 in out and 64 in out) else None,
        "state_max_rel_diff_mixed_vs_fp64": (float(np.abs(out[32]["state"] - out[64]["state"]).max()
                                                   / np.abs(out[64]["state"]).max()) if (32 in out and 64 in out) else None),
        "residuals": {f"fp{b}": [float(f"{v:.6e}") for v in out[b]["residuals"]] for b in out},
        "gmres_iterations": {f"fp{b}": out[b]["iters"] for b in out},
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
        "roofline": roof,
        "other_kernels": others,
        "autotune": autotune_summary(dev),
        "adjoint": adjoint,
    }
    dev.close()
    if tet:
        result["data"] = "synthetic (freestream with a Gaussian density and pressure pulse, seeded jittered mesh)"
        result["cpu_baseline"] = None   # the CPU restatement of the tetrahedra covers residual and Jacobian only
    elif mixed:
        result["data"] = "synthetic (freestream state, seeded mesh, element types and orders)"
        result["cpu_baseline"] = None
    elif sa:
        result["data"] = "synthetic (freestream state with inlet turbulence vtIn 3, seeded mesh)"
        result["cpu_baseline"] = None
    elif not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        result["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
