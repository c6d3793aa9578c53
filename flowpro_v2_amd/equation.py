"""Model plugins behind the reference's `Equation` contract (FlowProAPI/src/main/java/
flowpro/api/Equation.java:5-128), restricted to what the device path needs: the
derived constants of `init` and the constant initial condition.

NavierStokes      ModuleProjects/Equations/Aerodynamics/src/main/java/flowpro/user/equation/
                  Aerodynamics.java:119-312 (init), NavierStokes.java:18-21, 101-138
SpalartAllmaras   same directory, SpalartAllmaras.java:36-61 (init), 96-137 (constInitCondition)
KOmegaSST         same directory, KOmegaSST.java:45-71 (init), 109-151 (constInitCondition)
TestScalarEquation FlowPro/src/test/java/flowpro/user/equation/TestScalarEquation.java:28-37
Any other model name is unsupported on the device and runs on the reference's Java path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

MODEL_IDS = {"NavierStokes": 0, "SpalartAllmaras": 1, "TestScalarEquation": 2, "AdiabaticEulers": 3, "KOmegaSST": 4}
# model constants of the reference's SpalartAllmaras template
# (ModuleProjects/Equations/Aerodynamics/src/main/resources/templates/SpalartAllmaras.txt)
SA_DEFAULTS = {"sigma": 0.6667, "cb1": 0.1355, "cb2": 0.622, "ka": 0.41, "cw2": 0.3, "cw3": 2.0, "cv1": 7.1,
               "Prt": 0.9}
# the parameters.txt keys KOmegaSST.init reads (KOmegaSST.java:51-70)
KW_KEYS = ("kIn", "omIn", "SuthConst", "a1", "sk1", "som1", "alpha1", "beta1", "betast", "sk2", "som2", "alpha2",
           "beta2", "Prt")


@dataclass
class ModelConstants:
    model: str
    model_id: int
    n_eqs: int
    dim: int
    diffusive: bool
    kappa: float = 1.4
    cv: float = 1.0
    p_ref: float = 1.0
    rho_ref: float = 1.0
    velocity_ref: float = 1.0
    p_out: float = 1.0        # normalised by p_ref
    reynolds: float = -1.0    # after the rescaling of Aerodynamics.init
    prandtl: float = -1.0
    attack: float | None = None
    a: tuple = (1.0, 0.5)
    k: float = 0.0
    vt_in: float = 0.0        # SpalartAllmaras inlet turbulence
    sa: dict | None = None    # SpalartAllmaras constants (sigma, cb1, cb2, ka, cw2, cw3, cv1, Prt)
    kw: dict | None = None    # KOmegaSST inlet values and constants (kIn, omIn, SuthConst, a1, sk1, ..., Prt)


class UnsupportedModel(ValueError):
    pass


def init(props: dict) -> ModelConstants:
    """Equation.init for the supported models (parameters.txt keys)."""
    model = props.get("model", "NavierStokes")
    if model == "TestScalarEquation":
        a = tuple(float(v) for v in props.get("a", (1.0, 0.5)))
        k = float(props.get("k", 0.0))
        return ModelConstants(model, 2, 1, 2, k > 0, a=a, k=k)
    if model not in ("NavierStokes", "SpalartAllmaras", "AdiabaticEulers", "KOmegaSST"):
        raise UnsupportedModel(model)
    dim = int(props.get("dimension", 2))
    sa = model == "SpalartAllmaras"
    adiabatic = model == "AdiabaticEulers"
    kw = model == "KOmegaSST"
    if dim not in (2, 3) or (dim == 3 and (sa or adiabatic or kw or bool(props.get("isFlowViscous", False)))):
        raise UnsupportedModel(f"{model} in {dim}D")
    if (adiabatic or kw) and bool(props.get("isInletSupersonic", False)):
        raise UnsupportedModel(f"{model} with a supersonic inlet")
    kappa, cv = float(props["kappa"]), float(props["cv"])
    # subsonic stagnation inlet: pIn0, rhoIn0, pOut (Aerodynamics.java:169-182)
    p_ref, rho_ref = float(props["pIn0"]), float(props["rhoIn0"])
    p_out = float(props["pOut"]) / p_ref
    velocity_ref = math.sqrt(p_ref / rho_ref)
    # SpalartAllmaras.init and KOmegaSST.init pass isDiffusive = true (SpalartAllmaras.java:38,
    # KOmegaSST.java:47); AdiabaticEulers.init: Aerodynamics.init(props, dim, dim + 1, false)
    # (AdiabaticEulers.java:13-16)
    diffusive = True if (sa or kw) else (False if adiabatic else bool(props.get("isFlowViscous", False)))
    re = pr = -1.0
    if diffusive:
        # far-field Reynolds number rescaled to the reference values (Aerodynamics.java:262-281)
        pin0 = 1.0
        mach_inf = math.sqrt(2 / (kappa - 1) * (math.pow(pin0 / p_out, (kappa - 1) / kappa) - 1))
        base = 1 + (kappa - 1) / 2 * mach_inf * mach_inf
        rho_inf = rho_ref / math.pow(base, 1 / (kappa - 1))
        p_inf = p_ref / math.pow(base, kappa / (kappa - 1))
        u_inf = mach_inf * math.sqrt(kappa * p_inf / rho_inf)
        re = float(props["reynolds"]) * (rho_ref * velocity_ref) / (rho_inf * u_inf)
        pr = float(props["prandtl"])
    attack = float(props["attackAngle"]) if "attackAngle" in props else None
    extra = {}
    if sa:
        extra = dict(vt_in=float(props["vtIn"]), sa={k: float(props.get(k, v)) for k, v in SA_DEFAULTS.items()})
    if kw:   # KOmegaSST.init reads every key (KOmegaSST.java:51-70)
        extra = dict(kw={k: float(props[k]) for k in KW_KEYS})
    n_eqs = dim + (3 if sa else (1 if adiabatic else (4 if kw else 2)))
    return ModelConstants(model, MODEL_IDS[model], n_eqs, dim, diffusive, kappa=kappa, cv=cv, p_ref=p_ref,
                          rho_ref=rho_ref, velocity_ref=velocity_ref, p_out=p_out, reynolds=re, prandtl=pr,
                          attack=attack, **extra)


def const_init_condition(mc: ModelConstants) -> list[float]:
    """NavierStokes / SpalartAllmaras / KOmegaSST constInitCondition, subsonic inlet branch
    (NavierStokes.java:105-137, SpalartAllmaras.java:100-137: rho nu~ = rho vtIn appended;
    KOmegaSST.java:109-151: rho kIn, rho omIn appended)."""
    if mc.model_id == 2:
        return [0.0]
    kappa = mc.kappa
    mach_in = math.sqrt(2 / (kappa - 1) * (math.pow(1 / mc.p_out, (kappa - 1) / kappa) - 1))
    rho_in = math.pow(1 + ((kappa - 1) / 2) * mach_in * mach_in, 1 / (1 - kappa))
    v_in = mach_in * math.sqrt((kappa * mc.p_out) / rho_in)
    alpha = mc.attack or 0.0
    w = [rho_in, rho_in * v_in * math.cos(alpha), rho_in * v_in * math.sin(alpha)]
    if mc.dim == 3:
        w.append(0.0)
    if mc.model_id == 3:   # AdiabaticEulers.constInitCondition: no energy (AdiabaticEulers.java:97-126)
        return w
    w.append(mc.p_out / (kappa - 1) + 0.5 * rho_in * v_in * v_in)
    if mc.model_id == 1:
        w.append(rho_in * mc.vt_in)
    if mc.model_id == 4:   # KOmegaSST.java:146-147
        w += [rho_in * mc.kw["kIn"], rho_in * mc.kw["omIn"]]
    return w
