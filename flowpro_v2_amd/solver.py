"""Host mirror of the reference's implicit solver around the device Newton step.

`DeviceCase` owns one `fp_ctx` (include/flowpro_hip.h).  `GpuImplicitSolver.solve`
restates the time-step loop of LocalImplicitSolver.solve
(FlowPro/src/main/java/flowpro/core/solver/LocalImplicitSolver.java:107-285) with
State (FlowPro/src/main/java/flowpro/core/State.java:145-240) and CFLSetup
(FlowPro/src/main/java/flowpro/core/solver/CFLSetup.java:34-63); every Newton step
(assembly, preconditioned GMRES, update) runs on the device.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import equation, native, quadrature
from .cases import Case

PRECONDITIONERS = {"none": 0, "blockjacobi": 1, "blockilu0": 2, "ilu0": 3}
MAX_VALUE = 1.7976931348623157e308


STOP_FILE_NAME = "STOP"   # MasterSolver.STOP_FILE_NAME


def first_time_save_index(t: float, time_save_rate: float) -> int:
    """MasterSolver.firstTimeSaveIndex (MasterSolver.java:76-81): index of the next time level
    at which results are saved."""
    if time_save_rate <= 0 or time_save_rate == MAX_VALUE:
        return 1
    return int(min(2 ** 31 - 1, math.floor(t / time_save_rate) + 1))


class UnsupportedCase(RuntimeError):
    """The device path has no kernel for this case (the reference's Java path applies)."""


class DeviceCase:
    """One case on the GPU: MasterSolver.factory + Mesh + LinearSolver.factory for the device."""

    def __init__(self, case: Case, props: dict, precond_bits: int = 64):
        L = native.lib()
        if props.get("spatialMethod", "DGpure") not in ("DGpure", "DG"):
            raise UnsupportedCase("spatial method")
        if props.get("movingMesh", False) or props.get("curvedBoundary", False):
            raise UnsupportedCase("moving or curved mesh")
        prec = props.get("preconditioner", "blockilu0")
        if prec not in PRECONDITIONERS:
            raise UnsupportedCase(f"preconditioner {prec}")
        self.mc = equation.init(props)
        order = int(props["order"])
        self.props = props
        self.case = case
        self.dim = int(self.mc.dim)
        types, orders = getattr(case, "types", None), getattr(case, "orders", None)
        # elementType.txt / order.txt: per-element types and orders (FlowProMain.java:520-560).  A uniform
        # mesh of order-5 triangles (15 functions) also runs on the mixed-mesh kernels, padded to 16.
        tri5 = self.dim == 2 and np.asarray(case.TP).shape[1] == 3 and order == 5
        self.mixed = self.dim == 2 and (types is not None or orders is not None or tri5)
        tp_in = np.asarray(case.TP)
        if self.mixed:
            if orders is not None:
                order = int(np.max(orders))   # Parameters.order: the largest element order (FlowProMain.java:545)
            ty = (np.full(case.n_elem, 4 if tp_in.shape[1] == 4 else 3, np.int32) if types is None
                  else np.where(np.asarray(types) == 4, 4, 3).astype(np.int32))
            od = (np.full(case.n_elem, order, np.int32) if orders is None
                  else np.ascontiguousarray(orders, np.int32))
            nbs = np.where(ty == 4, od * od, od * (od + 1) // 2)
            if tp_in.shape[1] == 3:
                tp_in = np.concatenate([tp_in, -np.ones((tp_in.shape[0], 1), tp_in.dtype)], 1)
        self.order = order
        if self.dim == 3:
            if self.mc.attack is not None:
                raise UnsupportedCase("attack angle on tetrahedra")
            vxi, vw = quadrature.tetra_rule(order)
            fxi, fw = quadrature.triangle_rule(order)
        elif case.TP.shape[1] == 4:   # squareElement
            vxi, vw = quadrature.square_rule(order)
            fxi, fw = quadrature.line_rule(order)
        else:
            vxi, vw = quadrature.triangle_rule(order)
            fxi, fw = quadrature.line_rule(order)
        self.quad = self.dim == 2 and case.TP.shape[1] == 4 and not self.mixed
        self._keep = [np.ascontiguousarray(case.PXY, np.float64), np.ascontiguousarray(tp_in, np.int32),
                      np.ascontiguousarray(case.bnd, np.int32), np.ascontiguousarray(vxi), np.ascontiguousarray(vw),
                      np.ascontiguousarray(fxi), np.ascontiguousarray(fw)]
        pxy, tp, bnd, vxi, vw, fxi, fw = self._keep
        wd = None
        if case.wall_distance is not None:
            wd = np.ascontiguousarray(case.wall_distance, np.float64)
            self._keep.append(wd)
        cs = native.FpCase()
        cs.n_points, cs.pxy = pxy.shape[0], native.dptr(pxy)
        cs.n_elems, cs.tp = tp.shape[0], native.iptr(tp)
        cs.n_bnd, cs.bnd = bnd.shape[0], native.iptr(bnd)
        cs.wall_distance = native.dptr(wd) if wd is not None else None
        cs.order = order
        cs.n_vol, cs.vol_xi, cs.vol_w = vw.shape[0], native.dptr(vxi), native.dptr(vw)
        cs.n_face, cs.face_xi, cs.face_w = fw.shape[0], native.dptr(fxi), native.dptr(fw)
        mc = self.mc
        cs.model, cs.viscous = mc.model_id, int(mc.diffusive)
        cs.kappa, cs.p_out, cs.reynolds, cs.prandtl = mc.kappa, mc.p_out, mc.reynolds, mc.prandtl
        cs.has_attack, cs.attack = int(mc.attack is not None), float(mc.attack or 0.0)
        cs.scalar_a0, cs.scalar_a1, cs.scalar_k = float(mc.a[0]), float(mc.a[1]), float(mc.k)
        cs.beta0 = float(props.get("beta0", 2.0))
        cs.penalty = float(props.get("penalty", 0.0))
        cs.ordering = 1 if str(props.get("elementOrdering", "rcm")).lower() == "rcm" else 0
        cs.preconditioner = PRECONDITIONERS[prec]
        cs.precond_bits = int(precond_bits)
        # Gmres(A, M, 30, 5, ...) in LinearSolver.java:53; the optional keys set the same arguments
        cs.gmres_restart = int(props.get("gmresRestart", 30))
        cs.gmres_cycles = int(props.get("gmresCycles", 5))
        cs.gmres_tol = float(props.get("iterativeSolverTol", 1e-2))
        cs.positivity_limiter = int(bool(props.get("positivityLimiter", True)))
        cs.dimension = self.dim
        cs.elem_type = 4 if self.quad else 3
        if self.mixed:
            self._types = ty
            self._keep += [ty, od, quadrature.RULES_DIR.encode()]
            cs.elem_types, cs.elem_orders = native.iptr(ty), native.iptr(od)
            cs.gauss_rules_dir = self._keep[-1]
        # DGpure.initMethod: dampTol and dampConst are required keys of the DG methods, the inner
        # damping keys optional
        cs.damp_tol = float(props.get("dampTol", 0.0))
        cs.damp_const = float(props.get("dampConst", 0.0))
        cs.damp_inner_tol = float(props.get("dampInnerTol", 0.0))
        flux = props.get("numericalFlux", "default")
        if flux not in ("default", "vanLeer"):
            raise UnsupportedCase(f"numerical flux {flux}")
        cs.numerical_flux = 1 if flux == "vanLeer" else 0
        coef = props.get("dampInnerCoef")
        if coef is not None:
            coef = [float(v) for v in (coef if hasattr(coef, "__len__") else [coef])]
            cs.n_damp_inner_coef = len(coef)
            for i, v in enumerate(coef):
                cs.damp_inner_coef[i] = v
        if mc.sa is not None:
            if wd is None:
                raise UnsupportedCase("SpalartAllmaras needs the vertex wall distance (Case.wall_distance)")
            cs.sa_vt_in = mc.vt_in
            cs.sa_sigma, cs.sa_cb1, cs.sa_cb2, cs.sa_ka = (mc.sa[k] for k in ("sigma", "cb1", "cb2", "ka"))
            cs.sa_cw2, cs.sa_cw3, cs.sa_cv1, cs.sa_prt = (mc.sa[k] for k in ("cw2", "cw3", "cv1", "Prt"))
        cs.cv, cs.p_ref, cs.rho_ref = mc.cv, mc.p_ref, mc.rho_ref
        if mc.kw is not None:
            if wd is None:
                raise UnsupportedCase("KOmegaSST needs the vertex wall distance (Case.wall_distance)")
            cs.kw_k_in, cs.kw_om_in, cs.kw_suth = mc.kw["kIn"], mc.kw["omIn"], mc.kw["SuthConst"]
            cs.kw_a1, cs.kw_sk1, cs.kw_som1, cs.kw_alpha1, cs.kw_beta1 = (mc.kw[k] for k in ("a1", "sk1", "som1", "alpha1", "beta1"))
            cs.kw_betast, cs.kw_sk2, cs.kw_som2, cs.kw_alpha2 = (mc.kw[k] for k in ("betast", "sk2", "som2", "alpha2"))
            cs.kw_beta2, cs.kw_prt = mc.kw["beta2"], mc.kw["Prt"]
        h = C.c_void_p()
        rc = L.fp_create(C.byref(cs), C.byref(h))
        if rc == native.FP_UNSUPPORTED:
            raise UnsupportedCase("no device kernel for this model / order")
        if rc != native.FP_OK:
            raise native.FatalError(f"fp_create failed with {rc}")
        self.h = h
        info = np.zeros(6, np.int64)
        L.fp_info(self.h, info.ctypes.data_as(C.POINTER(C.c_longlong)))
        self.nb, self.neq, self.dofs = int(info[0]), int(info[1]), int(info[2])
        self.ilu_levels = (int(info[3]), int(info[4]))
        self.nnzb = int(info[5])
        self.B = self.nb * self.neq
        self.n_elem = case.n_elem
        self.precond_bits = precond_bits
        if self.mixed:
            # the reference's layout: element after element, each [m][nb_e]
            self.sizes = nbs * self.neq
            self.off = np.concatenate([[0], np.cumsum(self.sizes)])
            self.elem_nb = nbs

    def _shape(self):
        return (self.dofs,) if self.mixed else (self.n_elem, self.B)

    def close(self):
        if getattr(self, "h", None):
            native.lib().fp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = native.lib().fp_last_error(self.h)
            raise native.FatalError(f"{what}: {rc} ({msg.decode() if msg else ''})")
        return rc

    # ---- state
    def set_state(self, W, Wo=None, Wo2=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64).reshape(self._shape())
                for a in (W, Wo, Wo2)]
        ptr = [native.dptr(a) if a is not None else None for a in arrs]
        self._check(native.lib().fp_set_state(self.h, *ptr), "fp_set_state")

    def get_state(self) -> np.ndarray:
        out = np.zeros(self._shape())
        self._check(native.lib().fp_get_state(self.h, native.dptr(out)), "fp_get_state")
        return out

    def next_level(self):
        self._check(native.lib().fp_next_level(self.h), "fp_next_level")

    def restore_level(self):
        self._check(native.lib().fp_restore_level(self.h), "fp_restore_level")

    # ---- Newton step
    def size_over_lambda(self) -> float:
        v = C.c_double()
        self._check(native.lib().fp_time_step(self.h, C.byref(v)), "fp_time_step")
        return v.value

    def assemble(self, dt: float, dto: float, time_order: int = 1, with_jacobian: bool = True):
        norms = np.zeros(self.neq)
        rhs = C.c_double()
        self._check(native.lib().fp_assemble(self.h, dt, dto, time_order, int(with_jacobian), native.dptr(norms),
                                             C.byref(rhs)), "fp_assemble")
        return norms, rhs.value

    def limiter(self, first_step: bool = False):
        """DGpure.limiter for all elements: the artificial viscosity of the current state."""
        self._check(native.lib().fp_limiter(self.h, int(first_step)), "fp_limiter")

    def artificial_viscosity(self):
        """(eps[n_elem], dampInner[n_elem, nEqs]) of the last limiter call, mesh order."""
        eps = np.zeros(self.n_elem)
        di = np.zeros((self.n_elem, self.neq))
        self._check(native.lib().fp_export_artificial_viscosity(self.h, native.dptr(eps), native.dptr(di)),
                    "fp_export_artificial_viscosity")
        return eps, di

    def averages(self) -> np.ndarray:
        """Element means of the state (Mesh.getAvgW), [n_elem, nEqs] in mesh order."""
        out = np.zeros((self.n_elem, self.neq))
        self._check(native.lib().fp_export_averages(self.h, native.dptr(out)), "fp_export_averages")
        return out

    def linear_solve(self):
        it, err = C.c_int(), C.c_double()
        rc = self._check(native.lib().fp_linear_solve(self.h, C.byref(it), C.byref(err)), "fp_linear_solve")
        return rc, it.value, err.value

    def linear_solve_jfnk(self):
        it, err = C.c_int(), C.c_double()
        rc = self._check(native.lib().fp_linear_solve_jfnk(self.h, C.byref(it), C.byref(err)), "fp_linear_solve_jfnk")
        return rc, it.value, err.value

    def functional(self, which: int = 0, gradient: bool = True):
        """DragCoefficient (0) / LiftCoefficient (1) value and exact dI/dW."""
        val = C.c_double()
        g = np.zeros(self._shape()) if gradient else None
        self._check(native.lib().fp_functional(self.h, which, C.byref(val), native.dptr(g) if g is not None else None),
                    "fp_functional")
        return val.value, g

    def adjoint_solve(self, rhs, a1: float = 0.0):
        """(a1 M - dR/dW)^T lambda = rhs with GMRES on the matrix-free J^T v."""
        rhs = np.ascontiguousarray(rhs, np.float64).reshape(self._shape())
        lam = np.zeros_like(rhs)
        it, err = C.c_int(), C.c_double()
        rc = self._check(native.lib().fp_adjoint_solve(self.h, a1, native.dptr(rhs), native.dptr(lam), C.byref(it),
                                                       C.byref(err)), "fp_adjoint_solve")
        return rc, lam, it.value, err.value

    def update(self) -> int:
        n = C.c_int()
        self._check(native.lib().fp_update(self.h, C.byref(n)), "fp_update")
        return n.value

    # ---- exports
    def residual(self):
        R = np.zeros(self._shape())
        RHS = np.zeros(self._shape())
        self._check(native.lib().fp_export_residual(self.h, native.dptr(R), native.dptr(RHS)), "fp_export_residual")
        return R, RHS

    def jacobian(self):
        if self.mixed:
            return self._jacobian_flat()
        d = np.zeros((self.n_elem, self.B, self.B))
        n = np.zeros((self.n_elem, self.case.TP.shape[1] if self.dim == 2 else 4, self.B, self.B))
        self._check(native.lib().fp_export_jacobian(self.h, native.dptr(d), native.dptr(n)), "fp_export_jacobian")
        return d, n

    def _jacobian_flat(self):
        """Mixed meshes: diag[e] (B_e x B_e) and neigh[(e, k)] (B_t x B_e, t the face-k neighbour),
        the layout of the reference's ADiag / ANeighs[k].A."""
        TT = native.neighbor_matrix_mixed(self._keep[1], self._types, self._keep[2])
        sz = self.sizes
        nd = int((sz ** 2).sum())
        nn = sum(int(sz[t] * sz[e]) for e in range(self.n_elem) for t in TT[e] if t > -1)
        d, n = np.zeros(nd), np.zeros(max(nn, 1))
        self._check(native.lib().fp_export_jacobian(self.h, native.dptr(d), native.dptr(n)), "fp_export_jacobian")
        diag, neigh, od, on = [], {}, 0, 0
        for e in range(self.n_elem):
            b = int(sz[e])
            diag.append(d[od:od + b * b].reshape(b, b))
            od += b * b
            for k, t in enumerate(TT[e]):
                if t > -1:
                    neigh[(e, k)] = n[on:on + int(sz[t]) * b].reshape(int(sz[t]), b)
                    on += int(sz[t]) * b
        return diag, neigh

    def solution(self) -> np.ndarray:
        x = np.zeros(self.dofs)
        self._check(native.lib().fp_export_solution(self.h, native.dptr(x)), "fp_export_solution")
        return x

    def element_order(self) -> np.ndarray:
        o = np.zeros(self.n_elem, np.int32)
        self._check(native.lib().fp_element_order(self.h, native.iptr(o)), "fp_element_order")
        return o

    def jv(self, a1: float, v):
        v = np.ascontiguousarray(v, np.float64).reshape(self._shape())
        out = np.zeros_like(v)
        self._check(native.lib().fp_jv(self.h, a1, native.dptr(v), native.dptr(out)), "fp_jv")
        return out

    def jtv(self, a1: float, v):
        v = np.ascontiguousarray(v, np.float64).reshape(self._shape())
        out = np.zeros_like(v)
        self._check(native.lib().fp_jtv(self.h, a1, native.dptr(v), native.dptr(out)), "fp_jtv")
        return out

    def autotune_report(self):
        """(chosen shape index or -1, [(task waves, inversion waves)], [ms per factorisation]) of the
        block-ILU0 factor kernel's launch-shape autotuner."""
        v = C.c_int(-1)
        cap = 16
        shapes = np.zeros((cap, 2), np.int32)
        ms = np.zeros(cap)
        n = self._check(native.lib().fp_autotune_report(self.h, C.byref(v), native.iptr(shapes), native.dptr(ms), cap),
                        "fp_autotune_report")
        n = min(n, cap)
        return v.value, [tuple(int(x) for x in r) for r in shapes[:n]], ms[:n].tolist()

    def time_kernel(self, which: int, reps: int = 10) -> float:
        ms = C.c_double()
        self._check(native.lib().fp_time_kernel(self.h, which, reps, C.byref(ms)), "fp_time_kernel")
        return ms.value

    def initial_state(self) -> np.ndarray:
        w0 = equation.const_init_condition(self.mc)
        if self.mixed:
            return np.concatenate([np.repeat(np.asarray(w0), nb) for nb in self.elem_nb])
        return np.tile(np.repeat(np.asarray(w0), self.nb), (self.n_elem, 1))


# ---------------------------------------------------------------- State / CFLSetup
@dataclass
class State:
    """State.java: CFL control, time step and the normalised residual history."""
    cfl: float
    vary_cfl: bool = False
    max_vary_cfl: float = 1e4
    fixed_dt: float | None = None
    steps: int = 0
    t: float = 0.0
    current_cfl: float = 0.0
    residuum: float = MAX_VALUE
    dt: float = 0.0
    normalisation: list = field(default_factory=list)
    n_norm_steps: int = 0
    res_old: float = -1.0
    equation_residuum: list = field(default_factory=list)

    def __post_init__(self):
        if self.fixed_dt is not None:
            self.dt = self.fixed_dt

    def update_cfl(self):                       # State.updateCFL + CFLSetup.getCFL
        if self.fixed_dt is not None:
            self.dt = min(2 * self.dt, self.fixed_dt)
            return
        if not self.vary_cfl:
            self.current_cfl = min(self.current_cfl + self.cfl / 20, self.cfl)
            return
        known = 0 < self.residuum < MAX_VALUE
        if self.current_cfl <= 0:
            self.res_old = self.residuum if known else -1
            self.current_cfl = self.cfl
            return
        factor = 1.0
        if known and self.res_old > 0:
            factor = max(1 / 2.0, min(self.res_old / self.residuum, 2.0))
        self.res_old = self.residuum if known else -1
        self.current_cfl = max(min(self.current_cfl, self.cfl), min(self.current_cfl * factor, self.max_vary_cfl))

    def reduce_cfl(self):                       # State.reduceCFL + CFLSetup.reduceCFL
        if self.fixed_dt is not None:
            self.dt /= 2
        else:
            self.res_old = -1
            self.current_cfl /= 1.5

    def convergence_hopeless(self) -> bool:     # State.isConvergenceHopeless (maxCFL = par.cfl)
        if self.fixed_dt is not None:
            return self.dt < self.fixed_dt / 1024
        return 0 < self.current_cfl < self.cfl / 20

    def time_step(self, size_over_lambda: float) -> float:   # State.getTimeStep
        if self.fixed_dt is not None:
            self.current_cfl = self.dt / size_over_lambda
        else:
            self.dt = self.current_cfl * size_over_lambda
        return self.dt

    def set_residuum(self, norms):              # State.setResiduum, N_NORMALISATION_STEPS = 5
        n = len(norms)
        if len(self.normalisation) != n:
            self.normalisation = [0.0] * n
            self.n_norm_steps = 0
        if self.n_norm_steps < 5:
            self.normalisation = [max(a, b) for a, b in zip(self.normalisation, norms)]
            self.n_norm_steps += 1
        mx = max(self.normalisation)
        self.equation_residuum = []
        res = 0.0
        for m in range(n):
            nrm = self.normalisation[m] if self.normalisation[m] > 0 else mx
            r = norms[m] / nrm if nrm > 0 else norms[m]
            self.equation_residuum.append(r)
            if math.isnan(r):
                res = math.nan
            elif not math.isnan(res):
                res = max(res, r)
        self.residuum = res


@dataclass
class StepRecord:
    step: int
    residuum: float
    dt: float
    cfl: float
    newton: int
    gmres_iters: list
    limited: int


class GpuImplicitSolver:
    """LocalImplicitSolver.solve with the Newton step on the device."""

    def __init__(self, dev: DeviceCase, props: dict, simulation_path: str | None = None):
        """simulation_path: a FlowPro simulation directory; residuum.txt gets a line per step
        (appendResiduum) and state.txt, W.txt, We.txt, ... are written every saveRate steps
        (saveData), in the Java tool chain's formats."""
        self.dev = dev
        self.props = props
        self.simulation_path = simulation_path
        self.save_rate = int(props.get("saveRate", 2 ** 31 - 1))   # Parameters.saveRate
        self.time_save_rate = float(props.get("timeSaveRate", MAX_VALUE))   # Parameters.timeSaveRate
        self.t_end = float(props.get("tEnd", MAX_VALUE))   # Parameters.tEnd (Parameters.java:143-147)
        self.execution_ms = 0.0
        # Parameters.java:207-221: CFL wins over dt, one of the two is required
        if "CFL" in props:
            cfl = float(props["CFL"])
            self.state = State(cfl=1.0 if cfl <= 0 else cfl, vary_cfl=cfl <= 0,
                               max_vary_cfl=float(props.get("maxCFL", 1e4)))
        elif "dt" in props:
            self.state = State(cfl=0.0, fixed_dt=float(props["dt"]))
        else:
            raise ValueError("either parameter CFL or dt must be defined")
        self.time_order = int(props.get("timeOrder", 1))
        self.newton_iters = int(props.get("newtonIters", 1))
        self.newton_tol = float(props.get("newtonIterTol", 1e-4))
        self.stop_residuum = float(props.get("residuum", -1.0))
        self.history: list[StepRecord] = []
        self.newton_history: list[float] = []
        self.linear_failures = 0

    def save(self) -> None:
        """LocalImplicitSolver.saveData into the simulation directory."""
        from . import results_io
        results_io.save_data(self.simulation_path, self.dev, self.state, int(self.execution_ms))

    def continue_from(self) -> None:
        """continueComputation: State.load and the initial W from We.txt (W.txt when the order
        changed) of the simulation directory."""
        from . import results_io
        results_io.load_initial_state(self.simulation_path, self.dev, self.state)

    def solve(self, steps: int) -> list[StepRecord]:
        import time
        from . import results_io
        t_start = time.perf_counter() - self.execution_ms / 1000.0
        st, dev = self.state, self.dev
        dto = -1.0
        converges = True
        total = st.steps + steps
        tsr = first_time_save_index(st.t, self.time_save_rate)
        st.steps += 1
        while st.steps <= total and st.residuum > self.stop_residuum and st.t < self.t_end:
            if converges:
                st.update_cfl()
            if st.convergence_hopeless():
                break
            dt = st.time_step(dev.size_over_lambda())
            # LocalImplicitSolver.java:151-159: tEnd reached up to the round-off of the summed time
            # (a remaining step of ~1e-12 dt would spoil the BDF solution); the last step is clipped
            if self.t_end - st.t <= 1e-6 * dt:
                break
            if st.t + dt > self.t_end:
                dt = self.t_end - st.t
            dev.limiter(first_step=not self.history)   # LocalImplicitSolver.java:170-172
            rhs0 = 0.0
            iters = []
            limited = 0
            newton_done = 0
            converges = True
            for s in range(self.newton_iters):
                if s > 0:
                    _, rhs = dev.assemble(dt, dto, self.time_order, with_jacobian=False)
                    rel = rhs / rhs0 if rhs0 > 0 else 0.0
                    self.newton_history.append(rel)
                    if rel < self.newton_tol:
                        break
                norms, rhs = dev.assemble(dt, dto, self.time_order, with_jacobian=True)
                if s == 0:
                    st.set_residuum(list(norms))
                    rhs0 = rhs
                rc, it, _ = dev.linear_solve()
                iters.append(it)
                if rc != native.FP_OK:
                    dev.restore_level()
                    st.reduce_cfl()
                    st.steps -= 1
                    self.linear_failures += 1
                    converges = False
                    break
                limited += dev.update()
                newton_done += 1
            if converges:
                if not math.isfinite(st.residuum):
                    break
                self.execution_ms = (time.perf_counter() - t_start) * 1000.0
                st.t += dt
                dto = dt
                if self.simulation_path:
                    results_io.append_residuum(os.path.join(self.simulation_path, "residuum.txt"), st.residuum, st.t,
                                               float(int(self.execution_ms)), st.equation_residuum)
                dev.next_level()
                self.history.append(StepRecord(st.steps, st.residuum, dt, st.current_cfl, newton_done, iters, limited))
                if st.steps % self.save_rate == 0 or st.t > tsr * self.time_save_rate:
                    if self.simulation_path:
                        self.save()
                    tsr = first_time_save_index(st.t, self.time_save_rate)
                # MasterSolver.isStopRequested (MasterSolver.java:61-68): a file STOP in the
                # simulation directory ends the computation after this step and is deleted
                if self.simulation_path and os.path.exists(os.path.join(self.simulation_path, STOP_FILE_NAME)):
                    os.remove(os.path.join(self.simulation_path, STOP_FILE_NAME))
                    st.steps += 1
                    break
            st.steps += 1
        st.steps -= 1
        return self.history
