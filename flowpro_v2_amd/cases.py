"""Synthetic cases in FlowPro's mesh format and the parameters of each workload.

A case is what ``FlowProMain.solverFactory`` reads from a simulation directory
(FlowPro/src/main/java/flowpro/core/FlowProMain.java:486-730): vertices
``PXY[nPts][2]``, triangles ``TP[nElem][3]`` (counter-clockwise), boundary faces
``boundaryType.txt`` rows ``(type, v0, v1)`` and the parameters.txt keys of the
model and solver.  The generators are seeded and deterministic.

Boundary codes follow Aerodynamics.BoundaryType
(ModuleProjects/Equations/Aerodynamics/src/main/java/flowpro/user/equation/Aerodynamics.java:11-19):
-1 WALL, -2 INLET, -3 OUTLET, -4 INVISCID_WALL.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

WALL, INLET, OUTLET, INVISCID_WALL = -1, -2, -3, -4


@dataclass
class Case:
    name: str
    PXY: np.ndarray            # (nPts, 2) float64
    TP: np.ndarray             # (nElem, 3) int32 triangles or (nElem, 4) quadrilaterals, counter-clockwise
    bnd: np.ndarray            # (nBnd, 3) int32: type, v0, v1
    wall_distance: np.ndarray | None = None
    props: dict = field(default_factory=dict)
    types: np.ndarray | None = None    # per-element type (elementType.txt: 3 triangle, 4 square); TP then has 4 columns
    orders: np.ndarray | None = None   # per-element order (order.txt)

    @property
    def n_elem(self) -> int:
        return int(self.TP.shape[0])


class JavaRandom:
    """java.util.Random restated from its published LCG specification.

    seed: (s ^ 0x5DEECE66D) & (2**48 - 1); next(b): s = s*0x5DEECE66D + 0xB mod 2**48,
    returns s >>> (48 - b); nextDouble = ((next(26) << 27) + next(27)) * 2**-53.
    The reference's tests draw their states with it (FlowPro/src/test/java/flowpro/core/
    JacobianTest.java:36-41 with seed 1, LinearSolverTest.java:32-38 with seed 2).
    """

    MULT = 0x5DEECE66D
    MASK = (1 << 48) - 1

    def __init__(self, seed: int):
        self.s = (seed ^ self.MULT) & self.MASK

    def next(self, bits: int) -> int:
        self.s = (self.s * self.MULT + 0xB) & self.MASK
        return self.s >> (48 - bits)

    def next_double(self) -> float:
        return ((self.next(26) << 27) + self.next(27)) * (1.0 / (1 << 53))


def _orient_ccw(PXY: np.ndarray, TP: np.ndarray) -> np.ndarray:
    a, b, c = PXY[TP[:, 0]], PXY[TP[:, 1]], PXY[TP[:, 2]]
    area2 = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    TP = TP.copy()
    neg = area2 < 0
    TP[neg, 1], TP[neg, 2] = TP[neg, 2].copy(), TP[neg, 1].copy()
    return TP


def _split_quads(vid: np.ndarray) -> np.ndarray:
    """Structured (ni+1) x (nj+1) vertex grid -> 2*ni*nj triangles with the
    alternating diagonals of the reference's test meshes
    (FlowPro/src/test/java/flowpro/core/TestSimulation.java:41-53)."""
    ni, nj = vid.shape[0] - 1, vid.shape[1] - 1
    I, J = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    I, J = I.T.ravel(), J.T.ravel()          # j outer, i inner as in the reference loop
    a, b = vid[I, J], vid[I + 1, J]
    c, d = vid[I + 1, J + 1], vid[I, J + 1]
    even = (I + J) % 2 == 0
    t1 = np.where(even[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t2 = np.where(even[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
    return np.stack([t1, t2], 1).reshape(-1, 3).astype(np.int32)


def _quads(vid: np.ndarray) -> np.ndarray:
    """Structured (ni+1) x (nj+1) vertex grid -> ni*nj quadrilaterals (a, b, c, d) =
    (i, j), (i+1, j), (i+1, j+1), (i, j+1), j outer and i inner."""
    ni, nj = vid.shape[0] - 1, vid.shape[1] - 1
    I, J = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    I, J = I.T.ravel(), J.T.ravel()
    return np.stack([vid[I, J], vid[I + 1, J], vid[I + 1, J + 1], vid[I, J + 1]], 1).astype(np.int32)


def _orient_ccw_quads(PXY: np.ndarray, TP: np.ndarray) -> np.ndarray:
    """Reverse (a, b, c, d) -> (a, d, c, b) where the shoelace area is negative."""
    X, Y = PXY[TP, 0], PXY[TP, 1]
    area2 = (X * np.roll(Y, -1, 1) - np.roll(X, -1, 1) * Y).sum(1)
    TP = TP.copy()
    neg = area2 < 0
    TP[neg, 1], TP[neg, 3] = TP[neg, 3].copy(), TP[neg, 1].copy()
    return TP


def square_mesh(n: int, size: float = 1.0, boundary_type: int = INLET, quads: bool = False) -> Case:
    """TestSimulation.writeSquareMesh (TestSimulation.java:28-80): [0,size]^2, n x n
    cells split with alternating diagonals, all boundary faces of one type (quads: the cells
    themselves, squareElement)."""
    h = size / n
    j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    PXY = np.stack([i.ravel() * h, j.ravel() * h], 1).astype(np.float64)

    def v(i, j):
        return j * (n + 1) + i

    vid = np.array([[v(ii, jj) for jj in range(n + 1)] for ii in range(n + 1)])
    TP = _quads(vid) if quads else _split_quads(vid)
    rows = []
    for k in range(n):
        rows += [(boundary_type, v(k, 0), v(k + 1, 0)), (boundary_type, v(n, k), v(n, k + 1)),
                 (boundary_type, v(k + 1, n), v(k, n)), (boundary_type, v(0, k + 1), v(0, k))]
    wd = np.array([min(min(ii, jj), min(n - ii, n - jj)) * h for jj in range(n + 1) for ii in range(n + 1)])
    return Case(f"square{n}" + ("q" if quads else ""), PXY, TP, np.array(rows, dtype=np.int32), wd)


def channel_bump(nx: int = 64, ny: int = 16, quads: bool = False, jitter: float = 0.0, seed: int = 5) -> Case:
    """Channel [0,3]x[0,1] with a 10%-height circular-arc bump on x in [1,2]
    (y_b = sqrt(1.69 - (x-1.5)^2) - 1.2, the GAMM geometry), (nx+1)x(ny+1) vertices,
    y = y_b(x) + eta (1 - y_b(x)); walls -1 bottom/top, inlet -2 left, outlet -3 right.
    quads: the cells as quadrilaterals (general bilinear maps over the bump); jitter moves the
    interior vertices by a seeded fraction of the spacing."""
    x = np.linspace(0.0, 3.0, nx + 1)
    yb = np.where((x >= 1.0) & (x <= 2.0), np.sqrt(np.maximum(1.69 - (x - 1.5) ** 2, 0.0)) - 1.2, 0.0)
    yb = np.maximum(yb, 0.0)
    eta = np.linspace(0.0, 1.0, ny + 1)
    X = np.repeat(x[:, None], ny + 1, 1)
    Y = yb[:, None] + eta[None, :] * (1.0 - yb[:, None])
    if jitter > 0:
        rng = np.random.default_rng(seed)
        hx, hy = 3.0 / nx, (1.0 - yb.max()) / ny
        X[1:-1, 1:-1] += jitter * hx * rng.uniform(-1, 1, size=X[1:-1, 1:-1].shape)
        Y[1:-1, 1:-1] += jitter * hy * rng.uniform(-1, 1, size=Y[1:-1, 1:-1].shape)
    vid = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    PXY = np.stack([X.ravel(), Y.ravel()], 1)
    TP = _orient_ccw_quads(PXY, _quads(vid)) if quads else _orient_ccw(PXY, _split_quads(vid))
    rows = []
    for i in range(nx):
        rows.append((WALL, vid[i, 0], vid[i + 1, 0]))
        rows.append((WALL, vid[i + 1, ny], vid[i, ny]))
    for j in range(ny):
        rows.append((INLET, vid[0, j + 1], vid[0, j]))
        rows.append((OUTLET, vid[nx, j], vid[nx, j + 1]))
    return Case(f"channel_bump_{nx}x{ny}" + ("q" if quads else ""), PXY, TP, np.array(rows, dtype=np.int32))


def mixed_channel(nx: int = 12, ny: int = 4, quad_fraction: float = 0.5, jitter: float = 0.2, seed: int = 7,
                  flat: bool = False) -> Case:
    """channel_bump's cells, each kept as a quadrilateral (squareElement, type 4) with probability
    quad_fraction or split into two triangles (type 3) with the alternating diagonals; TP has four
    columns (-1 in the fourth for triangles) and `types` the element types.  flat: no bump."""
    base = channel_bump(nx, ny, quads=True, jitter=jitter, seed=seed)
    PXY = base.PXY.copy()
    if flat:
        PXY[:, 1] = np.repeat(np.linspace(0.0, 1.0, ny + 1)[None, :], nx + 1, 0).ravel()
        if jitter > 0:
            rng = np.random.default_rng(seed)
            inner = (PXY[:, 0] > 1e-9) & (PXY[:, 0] < 3 - 1e-9) & (PXY[:, 1] > 1e-9) & (PXY[:, 1] < 1 - 1e-9)
            PXY[inner] += jitter * min(3.0 / nx, 1.0 / ny) * rng.uniform(-0.5, 0.5, size=(inner.sum(), 2))
    TP, types = _keep_some_quads(base.TP, quad_fraction, seed + 1)
    return Case(f"mixed_{nx}x{ny}", PXY, TP, base.bnd.copy(), types=types)


def _keep_some_quads(TPq: np.ndarray, quad_fraction: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Each counter-clockwise quadrilateral kept (type 4) with probability quad_fraction (numpy
    default_rng(seed), one draw per cell in order) or split into two counter-clockwise triangles (type 3)
    with alternating diagonals; rows of four vertices, -1 in the fourth for a triangle."""
    rng = np.random.default_rng(seed)
    rows, types = [], []
    for c, q in enumerate(TPq):
        if rng.uniform() < quad_fraction:
            rows.append(list(q))
            types.append(4)
        else:
            a, b, cc, d = q
            tris = [(a, b, cc), (a, cc, d)] if c % 2 == 0 else [(a, b, d), (b, cc, d)]
            for t in tris:
                rows.append(list(t) + [-1])
                types.append(3)
    return np.array(rows, dtype=np.int32), np.array(types, dtype=np.int32)


def mixed_cylinder(n_theta: int, n_r: int, quad_fraction: float = 0.5, seed: int = 2) -> Case:
    """cylinder_mesh's quadrilateral O-grid (jitter seed `seed`) with each cell kept as a square element
    or split into two triangles at random (seed + 1): an elementType.txt mesh of both types."""
    base = cylinder_mesh(n_theta, n_r, seed=seed, quads=True)
    TP, types = _keep_some_quads(base.TP, quad_fraction, seed + 1)
    return Case(f"mixed_cylinder_{n_theta}x{n_r}", base.PXY, TP, base.bnd.copy(), types=types)


def seeded_orders(n_elem: int, choices=(1, 2, 3, 4), seed: int = 3) -> np.ndarray:
    """Per-element orders (an order.txt) drawn from `choices` with numpy default_rng(seed)."""
    return np.random.default_rng(seed).choice(np.asarray(choices), size=n_elem).astype(np.int32)


def cylinder_mesh(n_theta: int = 400, n_r: int = 250, growth: float = 1.01, jitter: float = 0.15,
                  seed: int = 2, quads: bool = False) -> Case:
    """[-5,5]^2 around a unit-diameter cylinder: an O-grid from the circle r = 0.5 to
    the square boundary, geometric radial spacing, each cell split into two triangles
    (alternating diagonals), interior vertices jittered by a seeded fraction of the
    local spacing.  Cylinder -1 (no-slip), left -2, right -3, top/bottom -4."""
    assert n_theta % 8 == 0, "the square's corners must be vertices"
    th = 2 * np.pi * np.arange(n_theta) / n_theta
    circ = 0.5 * np.stack([np.cos(th), np.sin(th)], 1)
    sq = 5.0 * np.stack([np.cos(th), np.sin(th)], 1) / np.maximum(np.abs(np.cos(th)), np.abs(np.sin(th)))[:, None]
    s = (growth ** np.arange(n_r + 1) - 1.0) / (growth ** n_r - 1.0)
    P = circ[:, None, :] + s[None, :, None] * (sq - circ)[:, None, :]
    rng = np.random.default_rng(seed)
    if jitter > 0:
        ds = np.diff(s)
        h_r = np.minimum(ds[:-1], ds[1:])                     # interior radial spacing
        r_mid = np.linalg.norm(P[:, 1:-1, :], axis=2)
        h_t = 2 * np.pi * r_mid / n_theta
        h = np.minimum(h_r[None, :] * np.linalg.norm(sq - circ, axis=1)[:, None], h_t)
        P[:, 1:-1, :] += jitter * h[:, :, None] * rng.uniform(-1, 1, size=P[:, 1:-1, :].shape)
    PXY = P.reshape(-1, 2)
    vid = np.arange(n_theta * (n_r + 1)).reshape(n_theta, n_r + 1)
    vid = np.concatenate([vid, vid[:1]], 0)                # periodic in theta
    TP = _orient_ccw_quads(PXY, _quads(vid)) if quads else _orient_ccw(PXY, _split_quads(vid))
    rows = []
    for i in range(n_theta):
        a, b = vid[i, 0], vid[i + 1, 0]
        rows.append((WALL, a, b))
        a, b = vid[i, n_r], vid[i + 1, n_r]
        mx, my = 0.5 * (PXY[a] + PXY[b])
        if abs(mx + 5.0) < 1e-9:
            t = INLET
        elif abs(mx - 5.0) < 1e-9:
            t = OUTLET
        else:
            t = INVISCID_WALL
        rows.append((t, a, b))
    return Case(f"cylinder_{n_theta}x{n_r}" + ("q" if quads else ""), PXY, TP, np.array(rows, dtype=np.int32))


def plate_channel(nx: int = 64, ny: int = 24, x0: float = -0.5, x1: float = 2.0, first: float = 1e-5,
                  growth: float = 1.15, nx_lead: int | None = None, with_wall_distance: bool = True) -> Case:
    """Turbulent flat-plate channel of the reference's SA example (.claude/skills/flowpro-turbulence/
    examples/plate_sa.txt: -1 plate on x in [0, x1], -2 inlet, -3 outlet, -4 symmetry for x < 0
    and on top): [x0, x1] x [0, H], geometric wall-normal spacing with first cell `first` and
    ratio `growth` (H = first (growth^ny - 1) / (growth - 1)), nx_lead of the nx columns ahead of
    the plate, alternating diagonals.  The vertex wall distance is generateDistanceFunction's
    (fp_wall_distance) when `with_wall_distance`."""
    if nx_lead is None:
        nx_lead = max(1, nx // 5)
    x = np.concatenate([np.linspace(x0, 0.0, nx_lead + 1)[:-1], np.linspace(0.0, x1, nx - nx_lead + 1)])
    y = first * (growth ** np.arange(ny + 1) - 1.0) / (growth - 1.0)
    X = np.repeat(x[:, None], ny + 1, 1)
    Y = np.repeat(y[None, :], nx + 1, 0)
    vid = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    PXY = np.stack([X.ravel(), Y.ravel()], 1)
    TP = _orient_ccw(PXY, _split_quads(vid))
    rows = []
    for i in range(nx):
        rows.append((WALL if i >= nx_lead else INVISCID_WALL, vid[i, 0], vid[i + 1, 0]))
        rows.append((INVISCID_WALL, vid[i + 1, ny], vid[i, ny]))
    for j in range(ny):
        rows.append((INLET, vid[0, j + 1], vid[0, j]))
        rows.append((OUTLET, vid[nx, j], vid[nx, j + 1]))
    bnd = np.array(rows, dtype=np.int32)
    wd = None
    if with_wall_distance:
        from . import native
        TT = native.neighbor_matrix(TP, bnd)
        wd = native.wall_distance(PXY, TP, TT)
    return Case(f"plate_{nx}x{ny}", PXY, TP, bnd, wd)


# Kuhn split of a cube into 6 tetrahedra along the diagonal v000 -> v111: one per axis order
_KUHN_PATHS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def kuhn_cube(n: int = 4, jitter: float = 0.05, seed: int = 4, length: float = 1.0) -> Case:
    """Unit cube [0, length]^3, n^3 cubes each Kuhn-split into 6 tetrahedra, interior vertices
    jittered by a seeded fraction of the spacing.  Every tetrahedron is stored with a negative
    determinant of (P1-P0, P2-P0, P3-P0): faceTransformation3Dtriangle's normals
    normalize(Js2 x Js1) of the faces {0,2,1}, {0,1,3}, {1,2,3}, {2,0,3} then point outwards.
    Boundary rows (type, v0, v1, v2): x = 0 inlet -2, x = length outlet -3, the other faces
    inviscid walls -4."""
    h = length / n
    g = np.arange(n + 1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([I.ravel(), J.ravel(), K.ravel()], 1).astype(np.float64) * h
    vid = np.arange((n + 1) ** 3).reshape(n + 1, n + 1, n + 1)
    rng = np.random.default_rng(seed)
    inner = np.all((P > 1e-12) & (P < length - 1e-12), axis=1)
    P[inner] += jitter * h * rng.uniform(-1, 1, size=(int(inner.sum()), 3))
    # cube (i, j, k) in lexicographic order, then the six axis orders: the path from v000 adds
    # one axis at a time, so the vertex ids are cumulative sums of the axis strides
    stride = np.array([(n + 1) ** 2, n + 1, 1])
    base = vid[:n, :n, :n].reshape(-1)
    steps = stride[np.array(_KUHN_PATHS)]                                   # (6, 3)
    offs = np.concatenate([np.zeros((6, 1), np.int64), np.cumsum(steps, 1)], 1)   # (6, 4)
    TP = (base[:, None, None] + offs[None]).reshape(-1, 4).astype(np.int32)
    a, b, c, d = P[TP[:, 0]], P[TP[:, 1]], P[TP[:, 2]], P[TP[:, 3]]
    det = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    pos = det > 0
    TP[pos, 1], TP[pos, 2] = TP[pos, 2].copy(), TP[pos, 1].copy()
    # faces seen once, in the order of their first appearance (element, then local face)
    faces = np.array(((0, 2, 1), (0, 1, 3), (1, 2, 3), (2, 0, 3)))
    keys = np.sort(TP[:, faces].reshape(-1, 3), axis=1)
    uniq, first, cnt = np.unique(keys, axis=0, return_index=True, return_counts=True)
    once = np.argsort(first[cnt == 1], kind="stable")
    keys = uniq[cnt == 1][once]
    x = P[keys]                                                             # (nb, 3, 3)
    t = np.full(len(keys), INVISCID_WALL, np.int32)
    t[np.all(np.abs(x[:, :, 0]) < 1e-12, axis=1)] = INLET
    t[np.all(np.abs(x[:, :, 0] - length) < 1e-12, axis=1)] = OUTLET
    rows = np.concatenate([t[:, None], keys.astype(np.int32)], 1)
    return Case(f"kuhn_{n}", P, TP, np.ascontiguousarray(rows, dtype=np.int32))


def euler3d_props(mach: float = 0.8, order: int = 2, **kw) -> dict:
    p = euler_props(mach=mach, order=order, dimension=3)
    p.update(kw)
    return p


def density_pulse_state(case: Case, nb: int, w0, amp: float = 0.05, width: float = 0.1,
                        center=(0.5, 0.5, 0.5), pressure_amp: float = 0.0) -> np.ndarray:
    """Freestream w0 with a Gaussian density pulse (velocity kept; pressure kept, or raised by the
    same Gaussian times pressure_amp): nodal values of the P1 Lagrange basis at the vertices
    (nb = 4) or the element mean (nb = 1).  Without a pressure pulse the cross-stream momentum
    residuals of a flow along x stay at round-off; a pressure pulse sends an acoustic wave into
    every direction, so each equation's residual is of the order of the amplitudes."""
    X = case.PXY[case.TP]                                     # (nElem, 4, 3)
    if nb == 1:
        X = X.mean(1, keepdims=True)
    r2 = ((X - np.asarray(center)) ** 2).sum(-1)
    rho0 = w0[0]
    g = np.exp(-r2 / (width * width))
    rho = rho0 * (1 + amp * g)
    u = np.asarray(w0[1:4]) / rho0
    p_e = (w0[4] - 0.5 * rho0 * (u ** 2).sum()) * (1 + pressure_amp * g)   # pressure / (kappa - 1)
    W = np.zeros((case.n_elem, 5, nb))
    W[:, 0] = rho
    for d in range(3):
        W[:, 1 + d] = rho * u[d]
    W[:, 4] = p_e + 0.5 * rho * (u ** 2).sum()
    return W.reshape(case.n_elem, 5 * nb)


def min_signed_area(case: Case) -> float:
    P, T = case.PXY, case.TP
    a, b, c = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]
    return float((((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])) / 2).min())


def pout_for_mach(mach: float, kappa: float = 1.4) -> float:
    """Static/stagnation pressure ratio of the isentropic outlet for a Mach number."""
    return (1 + 0.5 * (kappa - 1) * mach * mach) ** (-kappa / (kappa - 1))


def euler_props(mach: float = 0.67, order: int = 2, **kw) -> dict:
    p = dict(model="NavierStokes", isFlowViscous=False, kappa=1.4, cv=1.0, pIn0=1.0, rhoIn0=1.0,
             pOut=pout_for_mach(mach), order=order, timeOrder=1, preconditioner="blockilu0",
             elementOrdering="rcm", newtonIters=1, CFL=1.0, penalty=0.0, beta0=2.0,
             iterativeSolverTol=1e-2, newtonIterTol=1e-4)
    p.update(kw)
    return p


def gamm_props(order: int = 2, **kw) -> dict:
    """Parameters of the reference's GAMM channel example (FlowPro/simulations/examples/GAMM/
    default/parameters.txt): AdiabaticEulers, pOut = 0.737, CFL 5, dampTol 3, ilu0."""
    p = dict(model="AdiabaticEulers", kappa=1.4, cv=1.0, pIn0=1.0, rhoIn0=1.0, pOut=0.737, attackAngle=0.0,
             order=order, timeOrder=1, CFL=5.0, newtonIters=1, penalty=1.0, dampTol=3.0, dampConst=0.0,
             preconditioner="ilu0", elementOrdering="rcm", iterativeSolverTol=1e-2, newtonIterTol=1e-4)
    p.update(kw)
    return p


def ns_props(mach: float = 0.5, reynolds: float = 5000.0, order: int = 4, **kw) -> dict:
    p = euler_props(mach=mach, order=order, isFlowViscous=True, reynolds=reynolds, prandtl=0.72, penalty=1.0)
    p.update(kw)
    return p


def sa_props(mach: float = 0.3, reynolds: float = 1e6, order: int = 3, vt_in: float = 3.0, **kw) -> dict:
    """SpalartAllmaras parameters of the flat-plate example with the template's model constants."""
    p = euler_props(mach=mach, order=order, model="SpalartAllmaras", isFlowViscous=True, reynolds=reynolds,
                    prandtl=0.72, penalty=1.0, vtIn=vt_in, sigma=0.6667, cb1=0.1355, cb2=0.622, ka=0.41,
                    cw2=0.3, cw3=2.0, cv1=7.1, Prt=0.9)
    p.update(kw)
    return p


def kw_props(mach: float = 0.3, reynolds: float = 1e5, order: int = 3, k_in: float = 1e-3, om_in: float = -4.0,
             **kw) -> dict:
    """KOmegaSST parameters with the template's model constants (Aerodynamics/src/main/resources/templates/
    KOmegaSST.txt); the inlet k and log(omega) give a turbulent viscosity of rho k exp(-omega) = 0.055 times
    the laminar one at the inlet."""
    p = euler_props(mach=mach, order=order, model="KOmegaSST", isFlowViscous=True, reynolds=reynolds,
                    prandtl=0.72, penalty=1.0, kIn=k_in, omIn=om_in, SuthConst=110.0, a1=0.31, sk1=0.85, som1=0.5,
                    alpha1=0.5532, beta1=0.075, betast=0.09, sk2=1.0, som2=0.856, alpha2=0.4404, beta2=0.0828,
                    Prt=0.9)
    p.update(kw)
    return p


def perturbed_state(w0: np.ndarray, n_elem: int, n_basis: int, rel: float = 0.01, seed: int = 1) -> np.ndarray:
    """Element-major Java layout W[e][m*Nb + j]: constant state w0 times (1 + rel*N(0,1))
    per coefficient (numpy default_rng(seed))."""
    n_eqs = len(w0)
    rng = np.random.default_rng(seed)
    base = np.repeat(np.asarray(w0, dtype=np.float64), n_basis)
    W = np.tile(base, (n_elem, 1))
    return W * (1.0 + rel * rng.standard_normal(W.shape))


def java_random_state(n_elem: int, b: int, seed: int) -> np.ndarray:
    """States as the reference's tests draw them: elements in mesh order, each W[i] in order."""
    r = JavaRandom(seed)
    return np.array([[r.next_double() for _ in range(b)] for _ in range(n_elem)], dtype=np.float64)
