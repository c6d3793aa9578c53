"""Files of a FlowPro simulation directory, written and read the way the Java tool chain does,
so that a device run can be continued (`continueComputation`) or post-processed (`ResultsPlot`)
by the reference and the other way round.

  W.txt, We.txt, artificialViscosity.txt   Mat.save(double[][]) / Mat.loadDoubleMatrix
                                           (FlowProAPI/src/main/java/flowpro/api/Mat.java:908-1020)
  state.txt, referenceValues.txt           java.util.Properties.store / load through
                                           FlowProProperties (State.save / load, State.java:80-137;
                                           Aerodynamics.saveReferenceValues, Aerodynamics.java:373-391)
  residuum.txt                             MasterSolver.appendResiduum (MasterSolver.java:94-107)
  LocalImplicitSolver.saveData             (LocalImplicitSolver.java:326-335)

Every double is written as java.lang.Double.toString writes it (JDK 19 and later, the JDKs the
reference is built with): the shortest decimal that rounds back to the double (when that has one
digit, the closest of the one- and two-digit ones), in plain notation for 1e-3 <= |x| < 1e7 and
computerised scientific notation ("1.0E-5") otherwise, always with a digit after the point.
"""
from __future__ import annotations

import datetime
import math
import os
from decimal import Decimal

import numpy as np


def _shortest_digits(x: float) -> tuple[str, int]:
    """Digits d1 d2 ... dn and exponent e with |x| = 0.d1...dn * 10^e, selected as Double.toString
    does: the shortest decimals that round to x (Python's repr gives the shortest, closest one);
    when that is one digit, the two-digit decimals join the candidates and the one closest to x
    wins (ties: even last digit)."""
    d = Decimal(repr(abs(x))).normalize()
    t = d.as_tuple()
    digits = "".join(map(str, t.digits))
    e = t.exponent + len(digits)
    if len(digits) == 1:
        exact = Decimal(abs(x))
        best, dist = digits, abs(d - exact)
        for cand in range(int(digits) * 10 - 9, int(digits) * 10 + 10):
            if not 10 <= cand <= 99 or cand % 10 == 0:
                continue
            v = Decimal(cand).scaleb(e - 2)
            if float(v) != abs(x):
                continue
            dv = abs(v - exact)
            if dv < dist or (dv == dist and len(best) == 2 and cand % 2 == 0):
                best, dist = str(cand), dv
        digits = best
    return digits, e


def java_double(x: float) -> str:
    """java.lang.Double.toString(x)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0.0"
    digits, e = _shortest_digits(x)
    ax = abs(x)
    if 1e-3 <= ax < 1e7:
        if e <= 0:
            return sign + "0." + "0" * (-e) + digits
        if e >= len(digits):
            return sign + digits + "0" * (e - len(digits)) + ".0"
        return sign + digits[:e] + "." + digits[e:]
    frac = digits[1:] or "0"
    return sign + digits[0] + "." + frac + "E" + str(e - 1)


# ------------------------------------------------------------------ Mat.save / loadDoubleMatrix
def save_matrix(path: str, A) -> None:
    """Mat.save(double[][]): each entry followed by a space, one row per line."""
    with open(path, "w") as f:
        for row in A:
            f.write("".join(java_double(v) + " " for v in row))
            f.write("\n")


def load_matrix(path: str) -> list[np.ndarray]:
    """Mat.loadDoubleMatrix: non-empty lines split at whitespace; rows may differ in length."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line:
                rows.append(np.array([_java_parse(t) for t in line.split()], dtype=np.float64))
    return rows


def _java_parse(tok: str) -> float:
    """Double.parseDouble for the tokens Double.toString writes (also "Infinity", "NaN")."""
    return float(tok)


# ------------------------------------------------------------------ java.util.Properties
_ESC = {"\\": "\\\\", "\t": "\\t", "\n": "\\n", "\r": "\\r", "\f": "\\f", "=": "\\=", ":": "\\:", "#": "\\#",
        "!": "\\!"}


def _escape(s: str, key: bool) -> str:
    out = []
    for i, ch in enumerate(s):
        if ch == " " and (key or i == 0):
            out.append("\\ ")
        else:
            out.append(_ESC.get(ch, ch))
    return "".join(out)


def save_properties(path: str, items) -> None:
    """Properties.store(out, null): a date comment line, then key=value lines (escaped as
    Properties.saveConvert does).  The reader does not depend on the order of the lines."""
    stamp = datetime.datetime.now(datetime.timezone.utc).strftime("%a %b %d %H:%M:%S UTC %Y")
    with open(path, "w", encoding="latin-1") as f:
        f.write("#" + stamp + "\n")
        for k, v in items:
            f.write(_escape(k, True) + "=" + _escape(v, False) + "\n")


def load_properties(path: str) -> dict:
    """Properties.load for single-line entries: '#'/'!' comments, the first unescaped '=' or ':'
    (or whitespace) separates key and value, backslash escapes."""
    out = {}
    with open(path, encoding="latin-1") as f:
        for raw in f:
            line = raw.lstrip(" \t\f").rstrip("\r\n")
            if not line or line[0] in "#!":
                continue
            key, i = [], 0
            while i < len(line):
                ch = line[i]
                if ch == "\\" and i + 1 < len(line):
                    key.append(_unescape_char(line[i + 1]))
                    i += 2
                    continue
                if ch in "=: \t\f":
                    break
                key.append(ch)
                i += 1
            while i < len(line) and line[i] in " \t\f":
                i += 1
            if i < len(line) and line[i] in "=:":
                i += 1
            while i < len(line) and line[i] in " \t\f":
                i += 1
            val, rest = [], line[i:]
            j = 0
            while j < len(rest):
                if rest[j] == "\\" and j + 1 < len(rest):
                    val.append(_unescape_char(rest[j + 1]))
                    j += 2
                else:
                    val.append(rest[j])
                    j += 1
            out["".join(key)] = "".join(val)
    return out


def _unescape_char(ch: str) -> str:
    return {"t": "\t", "n": "\n", "r": "\r", "f": "\f"}.get(ch, ch)


# ------------------------------------------------------------------ state.txt, residuum.txt
def save_state(path: str, state, order: int, cpu_ms: int = 0, transfer_ms: int = 0) -> None:
    """State.save: t, steps, CPU, transfer, residuum, CFL, dt, the residuum normalisation once
    it is complete (5 steps), order."""
    items = [("t", java_double(state.t)), ("steps", str(int(state.steps))), ("CPU", str(int(cpu_ms))),
             ("transfer", str(int(transfer_ms))), ("residuum", java_double(state.residuum)),
             ("CFL", java_double(state.current_cfl)), ("dt", java_double(state.dt))]
    if state.normalisation and state.n_norm_steps >= 5:
        items.append(("residuumNormalisation", ",".join(java_double(r) for r in state.normalisation)))
    items.append(("order", str(int(order))))
    save_properties(path, items)


def load_state(path: str, state) -> int:
    """State.load into a solver.State: t, steps, CFL and the residuum normalisation; returns the
    order the file was written with (State.hasOrderChanged compares it)."""
    p = load_properties(path)
    state.t = float(p["t"])
    state.steps = int(p["steps"])
    state.current_cfl = float(p["CFL"])
    if "residuumNormalisation" in p:
        vals = [v for v in p["residuumNormalisation"].replace(",", " ").split()]
        state.normalisation = [float(v) for v in vals]
        state.n_norm_steps = 5
    return int(p["order"])


def append_residuum(path: str, residuum: float, t: float, cpu: float, equation_residuum=None) -> None:
    """MasterSolver.appendResiduum: 'residuum t CPU r_0 ... r_(nEqs-1)'."""
    parts = [java_double(residuum), java_double(t), java_double(cpu)]
    if equation_residuum is not None:
        parts += [java_double(r) for r in equation_residuum]
    with open(path, "a") as f:
        f.write(" ".join(parts) + "\n")


def save_reference_values(path: str, mc) -> None:
    """Aerodynamics.saveReferenceValues: l, p, rho, v, t and the far field machInf, rhoInf, uInf."""
    kappa, pout = mc.kappa, mc.p_out
    mach_inf = math.sqrt(2 / (kappa - 1) * (math.pow(1 / pout, (kappa - 1) / kappa) - 1))
    rho_inf = math.pow(1 + ((kappa - 1) / 2) * mach_inf * mach_inf, 1 / (1 - kappa))
    u_inf = mach_inf * math.sqrt((kappa * pout) / rho_inf)
    l_ref = 1.0
    items = [("l", l_ref), ("p", mc.p_ref), ("rho", mc.rho_ref), ("v", mc.velocity_ref),
             ("t", l_ref / mc.velocity_ref), ("machInf", mach_inf), ("rhoInf", rho_inf), ("uInf", u_inf)]
    save_properties(path, [(k, java_double(v)) for k, v in items])


# ------------------------------------------------------------------ saveData / continueComputation
def save_data(simulation_path: str, dev, state, cpu_ms: int = 0) -> None:
    """LocalImplicitSolver.saveData: state.txt, referenceValues.txt (Aerodynamics models), W.txt
    (element means), We.txt (element coefficients), artificialViscosity.txt (eps + dampInner)."""
    save_state(os.path.join(simulation_path, "state.txt"), state, dev.order, cpu_ms)
    if dev.mc.model_id != 2:
        save_reference_values(os.path.join(simulation_path, "referenceValues.txt"), dev.mc)
    save_matrix(os.path.join(simulation_path, "W.txt"), dev.averages())
    W = dev.get_state()
    if getattr(dev, "mixed", False):   # one row per element, each of its own length nEqs nBasis_e
        W = [W[dev.off[e]:dev.off[e + 1]] for e in range(dev.n_elem)]
    save_matrix(os.path.join(simulation_path, "We.txt"), W)
    eps, di = dev.artificial_viscosity()
    save_matrix(os.path.join(simulation_path, "artificialViscosity.txt"), eps[:, None] + di)


def load_initial_state(simulation_path: str, dev, state) -> None:
    """FlowProMain's continueComputation branch (FlowProMain.java:651-663): State.load, then the
    element coefficients from We.txt (W.txt when the order has changed or We.txt is missing;
    W.txt holds element means, spread over the coefficients of the nodal basis)."""
    old_order = load_state(os.path.join(simulation_path, "state.txt"), state)
    we = os.path.join(simulation_path, "We.txt")
    mixed = getattr(dev, "mixed", False)
    if old_order == dev.order and os.path.exists(we):
        rows = load_matrix(we)
        W = np.concatenate(rows) if mixed else np.array(rows)
    else:
        means = np.array(load_matrix(os.path.join(simulation_path, "W.txt")))
        if mixed:
            W = np.concatenate([np.repeat(means[e], dev.elem_nb[e]) for e in range(dev.n_elem)])
        else:
            W = np.repeat(means, dev.nb, axis=1)
    dev.set_state(W)
