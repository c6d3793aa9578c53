"""Many independent cases solved concurrently on one GPU.

Each instance is its own `fp_ctx` (its own HIP stream, matrix, factors and Krylov space) and is
driven by its own host thread through `GpuImplicitSolver`, so the instances' kernels overlap on the
device while each instance runs exactly the arithmetic it runs alone (the ctypes calls release the
GIL; the library keeps no state outside its contexts, and graph capture is thread-local).  This is
the reference's way of running many cases, one LocalImplicitSolver per case
(FlowPro/src/main/java/flowpro/core/solver/LocalImplicitSolver.java:107-285), with the cases
sharing the GPU instead of queueing for it.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import cases, solver


@dataclass
class Instance:
    case: cases.Case
    props: dict
    dev: solver.DeviceCase | None = None
    run: solver.GpuImplicitSolver | None = None


class BatchedSolver:
    """Independent cases, one context and one host thread each."""

    def __init__(self, instances: list[tuple[cases.Case, dict]], precond_bits: int = 64, workers: int | None = None):
        self.instances = [Instance(c, p) for c, p in instances]
        self.workers = workers or len(self.instances)
        self.precond_bits = precond_bits
        with ThreadPoolExecutor(max_workers=self.workers) as pool:
            devs = list(pool.map(lambda it: solver.DeviceCase(it.case, it.props, precond_bits=precond_bits),
                                 self.instances))
        for it, dev in zip(self.instances, devs):
            it.dev = dev
            dev.set_state(dev.initial_state())
            it.run = solver.GpuImplicitSolver(dev, it.props)

    @property
    def dofs(self) -> int:
        return sum(it.dev.dofs for it in self.instances)

    def solve(self, steps: int) -> list[list[solver.StepRecord]]:
        """`steps` time steps of every instance, concurrently; the records of each."""
        with ThreadPoolExecutor(max_workers=self.workers) as pool:
            return list(pool.map(lambda it: it.run.solve(steps), self.instances))

    def functionals(self, which: int = 1) -> list[tuple[float, np.ndarray]]:
        """LiftCoefficient (1) or DragCoefficient (0) and its exact gradient dI/dW per instance."""
        with ThreadPoolExecutor(max_workers=self.workers) as pool:
            return list(pool.map(lambda it: it.dev.functional(which, gradient=True), self.instances))

    def adjoints(self, which: int = 1, a1: list[float] | None = None) -> list[tuple[int, np.ndarray, int, float]]:
        """Per instance the adjoint of LiftCoefficient (1) or DragCoefficient (0): lambda with
        (a1 M - dR/dW)^T lambda = dI/dW at the instance's current state (GMRES on J^T v; the transposed
        block-ILU0 where the dataflow sweeps serve the block size), a1 per instance (default 0, the steady
        adjoint).  Returns (return code, lambda, GMRES iterations, relative residual) per instance."""
        a1 = a1 or [0.0] * len(self.instances)

        def one(k):
            it = self.instances[k]
            _, g = it.dev.functional(which, gradient=True)
            return it.dev.adjoint_solve(g, a1[k])
        with ThreadPoolExecutor(max_workers=self.workers) as pool:
            return list(pool.map(one, range(len(self.instances))))

    def close(self):
        for it in self.instances:
            if it.dev is not None:
                it.dev.close()
                it.dev = None


def seeded_ns_instances(n: int = 64, seed: int = 11, min_elems: int = 5000, max_elems: int = 50000,
                        orders=(2, 3, 4), cfl: float = 10.0, per_element_orders: bool = False,
                        mixed_types: bool = False, steady: dict | None = None) -> list[tuple[cases.Case, dict]]:
    """n independent 2D NavierStokes cylinder cases: the number of elements drawn uniformly from
    [min_elems, max_elems], the element order from `orders`, Re from [1e3, 1e4] and M from
    [0.2, 0.6], all from numpy default_rng(seed).  Each mesh is the cylinder O-grid with
    n_theta ~ 1.6 n_r (n_theta a multiple of 8) and its own jitter seed.  per_element_orders:
    every element draws its order from `orders` (an order.txt per instance, seed 200 + i) and the
    instance's order is the largest.  mixed_types: every cell of the O-grid is a square element or two
    triangles at random (an elementType.txt per instance, about 1.5 elements per cell).  steady: the
    CFL is varied by the residual (CFL <= 0, the SER law up to steady["maxCFL"]) and an instance stops
    at State.residuum <= steady["residuum"]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = rng.uniform(min_elems, max_elems)
        order = int(rng.choice(orders))
        re = float(rng.uniform(1e3, 1e4))
        mach = float(rng.uniform(0.2, 0.6))
        per_cell = 1.5 if mixed_types else 2.0
        n_theta = max(16, 8 * int(round(math.sqrt(target / per_cell * 1.6) / 8)))
        n_r = max(4, int(round(target / (per_cell * n_theta))))
        if mixed_types:
            case = cases.mixed_cylinder(n_theta, n_r, seed=100 + i)
        else:
            case = cases.cylinder_mesh(n_theta, n_r, seed=100 + i)
        if per_element_orders:
            case.orders = cases.seeded_orders(case.n_elem, orders, 200 + i)
            order = int(case.orders.max())
        if steady:
            props = cases.ns_props(mach=mach, reynolds=re, order=order, CFL=-1.0, maxCFL=steady["maxCFL"],
                                   residuum=steady["residuum"])
        else:
            props = cases.ns_props(mach=mach, reynolds=re, order=order, CFL=cfl)
        out.append((case, props))
    return out
