// Oracle (test infrastructure only): the model functions of the reference's equation
// classes (NavierStokes / Euler in 2D and 3D, SpalartAllmaras, TestScalarEquation), shared by
// the triangle and the tetrahedron restatements.  Loops run over M.dim and M.nEqs.
#pragma once
#include <cmath>

#include "oracle_core.hpp"

namespace oracle {

static const double RHO_TOL = 1e-2;   // AERO/Aerodynamics.java:22
static const double P_TOL = 1e-2;     // AERO/NavierStokes.java:15
static const double P_TOL_SA = 1e-1;  // AERO/SpalartAllmaras.java:15, AERO/KOmegaSST.java:15
enum { WALL = -1, INLET = -2, OUTLET = -3, INVISCID_WALL = -4 };

struct QPData {            // ElementData fields used on the path (API/ElementData.java)
    double currentX[3];
    double meshVelocity[3];
    double wallDistance;
};

// ------------------------------------------------------------- NavierStokes
// AERO/Aerodynamics.java:359-365
template <class T> static T limiteRho(T rho) {
    if (val(rho) < RHO_TOL) return RHO_TOL * dexp((rho - RHO_TOL) / RHO_TOL);
    return rho;
}
// AERO/NavierStokes.java:23-37, AERO/SpalartAllmaras.java:63-76 (its own P_TOL)
template <class T> static T pressure(const ModelConst& M, T* W) {
    if (M.kind == 3) {   // AERO/AdiabaticEulers.java:25-33: isentropic p = rho^kappa, P_TOL = 0.1
        W[0] = limiteRho(W[0]);
        T p = dpow(W[0], M.kapa);
        if (val(p) < 1e-1) p = 1e-1 * dexp((p - 1e-1) / 1e-1);
        return p;
    }
    const double ptol = (M.kind == 1 || M.kind == 4) ? P_TOL_SA : P_TOL;   // AERO/KOmegaSST.java:75-88
    W[0] = limiteRho(W[0]);
    T momentum2 = 0.0;
    for (int d = 0; d < M.dim; ++d) momentum2 += W[d + 1] * W[d + 1];
    T p = (M.kapa - 1) * (W[M.dim + 1] - momentum2 / (2 * W[0]));
    if (val(p) < ptol) p = ptol * dexp((p - ptol) / ptol);
    return p;
}
// AERO/NavierStokes.java:194-214
template <class T> static void convectiveFlux(const ModelConst& M, T* W, const double* n, T* f) {
    if (M.kind == 2) {   // T/user/equation/TestScalarEquation.java:74-81
        double an = 0;
        for (int d = 0; d < M.dim; d++) an += M.a[d] * n[d];
        f[0] = an * W[0];
        return;
    }
    W[0] = limiteRho(W[0]);
    T V = 0.0;
    for (int d = 0; d < M.dim; ++d) V += W[d + 1] * n[d];
    V /= W[0];
    T p = pressure(M, W);
    f[0] = W[0] * V;
    for (int d = 0; d < M.dim; ++d) f[d + 1] = W[d + 1] * V + p * n[d];
    if (M.kind == 3) return;   // AERO/AdiabaticEulers.java:177-193: no energy equation
    f[M.dim + 1] = (W[M.dim + 1] + p) * V;
    if (M.kind == 1) f[M.dim + 2] = W[M.dim + 2] * V;   // AERO/SpalartAllmaras.java:196
    if (M.kind == 4) {   // AERO/KOmegaSST.java:210-211
        f[M.dim + 2] = W[M.dim + 2] * V;
        f[M.dim + 3] = W[M.dim + 3] * V;
    }
}
// AERO/Aerodynamics.java:323-337
template <class T> static T maxEigenvalue(const ModelConst& M, T* W, const QPData& q) {
    if (M.kind == 2) return T(M.lambda);
    W[0] = limiteRho(W[0]);
    T p = pressure(M, W);
    T a = dsqrt(M.kapa * p / W[0]);
    T v = 0.0;
    for (int d = 0; d < M.dim; ++d) {
        T vRel = W[d + 1] / W[0] - q.meshVelocity[d];
        v += vRel * vRel;
    }
    v = dsqrt(v);
    return v + a;
}
// AERO/NavierStokes.java:407-465 fplus (sign +1) / fminus (sign -1): van Leer's split fluxes;
// inside |M| < 1 the mass flux is +-rho a (M +- 1)^2 / 4, beyond the sonic point one side
// carries the whole flux and the other none
template <class T> static void vanLeerSplit(const ModelConst& M, const T* W, T p, const double* n, int sign, T* f) {
    const int dim = M.dim, nEqs = M.nEqs;
    const double kapa = M.kapa;
    for (int j = 0; j < nEqs; j++) f[j] = 0.0;
    T vel[3];
    T Vn = 0.0, q2 = 0.0;
    for (int d = 0; d < dim; ++d) {
        vel[d] = W[d + 1] / W[0];
        Vn += vel[d] * n[d];
        q2 += vel[d] * vel[d];
    }
    T a = dsqrt(kapa * p / W[0]);
    T Mn = Vn / a;
    if (std::fabs(val(Mn)) < 1) {
        T fm = sign > 0 ? W[0] * a * ((Mn + 1) * (Mn + 1)) / 4 : -W[0] * a * ((Mn - 1) * (Mn - 1)) / 4;
        f[0] = fm;
        for (int d = 0; d < dim; ++d) f[d + 1] = fm * (vel[d] + (-Vn + sign * 2 * a) / kapa * n[d]);
        T c = (kapa - 1) * Vn + sign * 2 * a;
        f[dim + 1] = fm * ((q2 - Vn * Vn) / 2 + c * c / (2 * (kapa * kapa - 1)));
    } else if (sign > 0 ? val(Mn) >= 1 : val(Mn) <= -1) {
        f[0] = W[0] * Vn;
        for (int d = 0; d < dim; ++d) f[d + 1] = W[d + 1] * Vn + p * n[d];
        f[dim + 1] = (W[dim + 1] + p) * Vn;
    }
}
// AERO/NavierStokes.java:141-192 (default = Lax-Friedrichs, numericalFlux = vanLeer)
template <class T>
static void numericalConvectiveFlux(const ModelConst& M, T* WL, T* WR, const double* n, int TT, const QPData& q, T* f) {
    const int nEqs = M.nEqs, dim = M.dim;
    if (M.kind == 2) {   // T/user/equation/TestScalarEquation.java:84-88
        T fL[1], fR[1];
        convectiveFlux(M, WL, n, fL);
        convectiveFlux(M, WR, n, fR);
        f[0] = (fL[0] + fR[0]) / 2 - M.lambda * (WR[0] - WL[0]) / 2;
        return;
    }
    WL[0] = limiteRho(WL[0]);
    WR[0] = limiteRho(WR[0]);
    for (int j = 0; j < nEqs; j++) f[j] = 0.0;
    switch (TT) {
        case WALL:
        case INVISCID_WALL: {
            // NavierStokes: p(WL); SpalartAllmaras: p(WR) (AERO/SpalartAllmaras.java:146)
            // AdiabaticEulers: p(WR) and no energy component (AERO/AdiabaticEulers.java:136-149)
            // KOmegaSST: p(WR) (AERO/KOmegaSST.java:161-175)
            T p = pressure(M, (M.kind == 1 || M.kind == 3 || M.kind == 4) ? WR : WL);
            f[0] = 0.0;
            double V = .0;
            for (int d = 0; d < dim; ++d) {
                f[d + 1] = p * n[d];
                V += q.meshVelocity[d] * n[d];
            }
            if (M.kind != 3) f[dim + 1] = p * V;
            for (int j = 0; j < nEqs; j++) f[j] += V * WR[j];
            break;
        }
        case INLET:
        case OUTLET:
            convectiveFlux(M, WR, n, f);
            break;
        default: {
            if (M.vanLeer) {   // AERO/NavierStokes.java:174-179
                T pL = pressure(M, WL);
                T fp[MAXEQ], fm[MAXEQ];
                vanLeerSplit(M, WL, pL, n, +1, fp);
                T pR = pressure(M, WR);
                vanLeerSplit(M, WR, pR, n, -1, fm);
                for (int j = 0; j < nEqs; j++) f[j] = fp[j] + fm[j];
                break;
            }
            T fL[MAXEQ], fR[MAXEQ];
            convectiveFlux(M, WL, n, fL);
            convectiveFlux(M, WR, n, fR);
            T lamL = maxEigenvalue(M, WL, q);
            T lamR = maxEigenvalue(M, WR, q);
            T maxEigenValue = dmax(lamL, lamR);
            for (int j = 0; j < nEqs; j++) f[j] = (fL[j] + fR[j] - maxEigenValue * (WR[j] - WL[j])) / 2;
        }
    }
}
// AERO/AdiabaticEulers.java:36-95 (subsonic inlet): the wall state adds n (m . n), as
// SpalartAllmaras does; the outlet density is pOut^(1/kappa)
template <class T> static void boundaryValueAdiabatic(const ModelConst& M, T* WL, const double* n, int TT, T* WR) {
    const int nEqs = M.nEqs, dim = M.dim;
    const double kapa = M.kapa;
    WL[0] = limiteRho(WL[0]);
    for (int j = 0; j < nEqs; j++) WR[j] = 0.0;
    switch (TT) {
        case WALL:
        case INVISCID_WALL: {
            for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
            T nu = 0.0;
            for (int d = 0; d < dim; ++d) nu += WL[d + 1] * n[d];
            for (int d = 0; d < dim; ++d) WR[d + 1] = WL[d + 1] + n[d] * nu;
            break;
        }
        case INLET: {
            const double pIn0 = 1, rhoIn0 = 1;
            T p = pressure(M, WL);
            T mach;
            if (val(p) > pIn0) mach = 0.0;
            else mach = dsqrt((2 / (kapa - 1)) * (-1 + dpow(pIn0 / p, (kapa - 1) / kapa)));
            T rhoIn = rhoIn0 * dpow((1 + ((kapa - 1) / 2) * mach * mach), 1 / (1 - kapa));
            T normalVelocity = mach * dsqrt((kapa * p) / rhoIn);
            WR[0] = rhoIn;
            for (int d = 0; d < dim; ++d) WR[d + 1] = -rhoIn * normalVelocity * n[d];
            break;
        }
        case OUTLET: {
            T absVelocity = .0;
            for (int d = 0; d < dim; ++d) absVelocity += WL[d + 1] * WL[d + 1];
            absVelocity = dsqrt(absVelocity) / WL[0];
            T a = dsqrt(kapa * pressure(M, WL) / WL[0]);
            T mach = absVelocity / a;
            if (val(mach) < 1) {
                const double rhoOut = std::pow(M.pOut, 1. / kapa);
                WR[0] = rhoOut;
                for (int d = 0; d < dim; ++d) WR[d + 1] = rhoOut * WL[d + 1] / WL[0];
            } else {
                for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
            }
            break;
        }
        default:
            break;
    }
}
// AERO/KOmegaSST.java:218-306 (subsonic inlet): the wall keeps log(omega) and zeroes k, the inviscid
// wall copies the inner state (no reflection), inlet k and omega from kIn and omIn
template <class T> static void boundaryValueKW(const ModelConst& M, T* WL, const double* n, int TT, const QPData& q, T* WR) {
    const int nEqs = M.nEqs, dim = M.dim;
    const double kapa = M.kapa;
    WL[0] = limiteRho(WL[0]);
    for (int j = 0; j < nEqs; j++) WR[j] = 0.0;
    T p = pressure(M, WL);
    switch (TT) {
        case WALL: {   // isDiffusive is always true for this model
            const double* u = q.meshVelocity;
            double absVelocity2 = .0;
            for (int d = 0; d < dim; ++d) absVelocity2 += u[d] * u[d];
            WR[0] = WL[0];
            for (int d = 0; d < dim; ++d) WR[d + 1] = WR[0] * u[d];
            WR[dim + 1] = p / (kapa - 1) + WR[0] * absVelocity2 / 2;
            WR[dim + 2] = 0.0;
            WR[dim + 3] = WL[dim + 3];
            break;
        }
        case INVISCID_WALL:
            for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
            break;
        case INLET: {
            const double pIn0 = 1, rhoIn0 = 1;
            T mach;
            if (val(p) > pIn0) mach = 0.0;
            else mach = dsqrt((2 / (kapa - 1)) * (-1 + dpow(pIn0 / p, (kapa - 1) / kapa)));
            T rhoIn = rhoIn0 * dpow((1 + ((kapa - 1) / 2) * mach * mach), 1 / (1 - kapa));
            T inletVelocity = mach * dsqrt((kapa * p) / rhoIn);
            T E = p / (kapa - 1) + 0.5 * rhoIn * inletVelocity * inletVelocity;
            WR[0] = rhoIn;
            if (!M.hasAttack) {
                for (int d = 0; d < dim; ++d) WR[d + 1] = -rhoIn * inletVelocity * n[d];
            } else {
                double dir[2] = {std::cos(M.attack), std::sin(M.attack)};
                for (int d = 0; d < dim; ++d) WR[d + 1] = rhoIn * inletVelocity * dir[d];
            }
            WR[dim + 1] = E;
            WR[dim + 2] = rhoIn * M.kIn;
            WR[dim + 3] = rhoIn * M.omIn;
            break;
        }
        case OUTLET: {
            T rhoOut = WL[0];
            T absVelocity2 = .0;
            for (int d = 0; d < dim; ++d) absVelocity2 += WL[d + 1] * WL[d + 1];
            absVelocity2 /= rhoOut * rhoOut;
            T a = dsqrt(kapa * pressure(M, WL) / WL[0]);
            T mach = dsqrt(absVelocity2) / a;
            if (val(mach) < 1) {
                WR[0] = rhoOut;
                for (int d = 0; d < dim; ++d) WR[d + 1] = rhoOut * WL[d + 1] / WL[0];
                WR[dim + 1] = M.pOut / (kapa - 1) + rhoOut * absVelocity2 / 2;
                WR[dim + 2] = WL[dim + 2];
                WR[dim + 3] = WL[dim + 3];
            } else {
                for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
            }
            break;
        }
        default:
            break;
    }
}
// AERO/NavierStokes.java:216-315 (subsonic inlet, no attack angle unless given)
template <class T> static void boundaryValue(const ModelConst& M, T* WL, const double* n, int TT, const QPData& q, T* WR) {
    const int nEqs = M.nEqs, dim = M.dim;
    if (M.kind == 2) {   // exact(currentX)
        WR[0] = std::sin(2 * q.currentX[0] + q.currentX[1]);
        return;
    }
    if (M.kind == 3) {
        boundaryValueAdiabatic(M, WL, n, TT, WR);
        return;
    }
    if (M.kind == 4) {
        boundaryValueKW(M, WL, n, TT, q, WR);
        return;
    }
    WL[0] = limiteRho(WL[0]);
    for (int j = 0; j < nEqs; j++) WR[j] = 0.0;
    T p = pressure(M, WL);
    const double kapa = M.kapa;
    switch (TT) {
        case WALL:
            if (M.diffusive) {
                const double* u = q.meshVelocity;
                double absVelocity2 = .0;
                for (int d = 0; d < dim; ++d) absVelocity2 += u[d] * u[d];
                WR[0] = WL[0];
                for (int d = 0; d < dim; ++d) WR[d + 1] = WR[0] * u[d];
                WR[dim + 1] = p / (kapa - 1) + WR[0] * absVelocity2 / 2;
                // SpalartAllmaras: WR[dim + 2] = 0 (AERO/SpalartAllmaras.java:216)
            } else {
                for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
                T nu = 0.0;
                for (int d = 0; d < dim; ++d) nu += WL[d + 1] * n[d];
                for (int d = 0; d < dim; ++d) WR[d + 1] = WL[d + 1] - n[d] * nu;
            }
            break;
        case INVISCID_WALL: {
            for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
            T nu = 0.0;
            for (int d = 0; d < dim; ++d) nu += WL[d + 1] * n[d];
            // SpalartAllmaras adds n nu (AERO/SpalartAllmaras.java:229-231), NavierStokes subtracts
            if (M.kind == 1)
                for (int d = 0; d < dim; ++d) WR[d + 1] = WL[d + 1] + n[d] * nu;
            else
                for (int d = 0; d < dim; ++d) WR[d + 1] = WL[d + 1] - n[d] * nu;
            break;
        }
        case INLET: {
            const double pIn0 = 1, rhoIn0 = 1;
            T mach;
            if (val(p) > pIn0) mach = 0.0;
            else mach = dsqrt((2 / (kapa - 1)) * (-1 + dpow(pIn0 / p, (kapa - 1) / kapa)));
            T rhoIn = rhoIn0 * dpow((1 + ((kapa - 1) / 2) * mach * mach), 1 / (1 - kapa));
            T inletVelocity = mach * dsqrt((kapa * p) / rhoIn);
            T E = p / (kapa - 1) + 0.5 * rhoIn * inletVelocity * inletVelocity;
            WR[0] = rhoIn;
            if (!M.hasAttack) {
                for (int d = 0; d < dim; ++d) WR[d + 1] = -rhoIn * inletVelocity * n[d];
            } else {
                double dir[3] = {std::cos(M.attack), std::sin(M.attack), 0.0};   // 2D: one angle
                for (int d = 0; d < dim; ++d) WR[d + 1] = rhoIn * inletVelocity * dir[d];
            }
            WR[dim + 1] = E;
            if (M.kind == 1) WR[dim + 2] = rhoIn * M.vtIn;   // AERO/SpalartAllmaras.java:262
            break;
        }
        case OUTLET: {
            T rhoOut = WL[0];
            T absVelocity2 = .0;
            for (int d = 0; d < dim; ++d) absVelocity2 += WL[d + 1] * WL[d + 1];
            absVelocity2 /= rhoOut * rhoOut;
            T a = dsqrt(kapa * pressure(M, WL) / WL[0]);
            T mach = dsqrt(absVelocity2) / a;
            if (val(mach) < 1) {
                WR[0] = rhoOut;
                for (int d = 0; d < dim; ++d) WR[d + 1] = rhoOut * WL[d + 1] / WL[0];
                WR[dim + 1] = M.pOut / (kapa - 1) + rhoOut * absVelocity2 / 2;
                if (M.kind == 1) WR[dim + 2] = WL[dim + 2];   // AERO/SpalartAllmaras.java:283
            } else {
                for (int j = 0; j < nEqs; j++) WR[j] = WL[j];
            }
            break;
        }
        default:
            break;
    }
}
// SpalartAllmaras.max (AERO/SpalartAllmaras.java:453-459): a > b ? a : b
template <class T> static T saMax(T a, T b) { return val(a) > val(b) ? a : b; }
// java.lang.Math.min for ordinary values: a <= b ? a : b
template <class T> static T jmin(T a, T b) { return val(a) <= val(b) ? a : b; }

// AERO/SpalartAllmaras.java:292-351 diffusiveFlux
template <class T> static void diffusiveFluxSA(const ModelConst& M, T* W, const T* dW, const double* n, T* flux) {
    const int nEqs = M.nEqs, dim = M.dim;
    W[0] = limiteRho(W[0]);
    T rho = W[0];
    const double lam = -2. / 3;
    T velocity[3];
    T velocity2 = .0;
    for (int d = 0; d < dim; ++d) {
        velocity[d] = W[d + 1] / rho;
        velocity2 += velocity[d] * velocity[d];
    }
    T velocityJac[9];
    for (int d = 0; d < dim; ++d)
        for (int f = 0; f < dim; ++f) velocityJac[dim * d + f] = (dW[f * nEqs + d + 1] - dW[f * nEqs] * velocity[d]) / rho;
    T p = pressure(M, W);
    T pOverRhoDer[3];
    for (int d = 0; d < dim; ++d) {
        T temp = .0;
        for (int f = 0; f < dim; ++f) temp += velocity[f] * velocityJac[dim * f + d];
        T pDer = (M.kapa - 1) * (dW[d * nEqs + dim + 1] - dW[d * nEqs] * velocity2 / 2 - rho * temp);
        pOverRhoDer[d] = (rho * pDer - p * dW[d * nEqs]) / (rho * rho);
    }
    T stress[9];
    T trace = .0;
    for (int d = 0; d < dim; ++d) {
        trace += velocityJac[dim * d + d];
        for (int f = 0; f < dim; ++f) stress[dim * d + f] = velocityJac[dim * d + f] + velocityJac[dim * f + d];
    }
    for (int d = 0; d < dim; ++d) stress[dim * d + d] += lam * trace;
    T vt = saMax(T(0.0), W[dim + 2] / rho);
    T vtDer[3];
    for (int d = 0; d < dim; ++d) vtDer[d] = 1 / rho * (dW[d * nEqs + dim + 2] - dW[d * nEqs] * vt);
    T xi = rho * vt;
    T fv1 = xi * xi * xi / (xi * xi * xi + M.cv1 * M.cv1 * M.cv1);
    T mut = rho * vt * fv1;
    T constant = M.kapa / (M.kapa - 1) * (1 / M.Pr + mut / M.Prt);
    for (int j = 0; j < nEqs; j++) flux[j] = 0.0;
    for (int d = 0; d < dim; ++d) {
        T tmp = .0;
        for (int f = 0; f < dim; ++f) {
            flux[f + 1] += (1 + mut) * stress[dim * d + f] * n[d] / M.Re;
            tmp += velocity[f] * (1 + mut) * stress[dim * d + f];
        }
        flux[dim + 1] += (tmp + constant * pOverRhoDer[d]) * n[d] / M.Re;
        flux[dim + 2] += 1 / (M.Re * M.sigma) * (1 + rho * vt) * vtDer[d] * n[d];
    }
}

// KOmegaSST.max / min (AERO/KOmegaSST.java:572-585): a > b ? a : b, a < b ? a : b
template <class T> static T kwMax(T a, T b) { return val(a) > val(b) ? a : b; }
template <class T> static T kwMin(T a, T b) { return val(a) < val(b) ? a : b; }
// AERO/KOmegaSST.java:566-570 sutherland: the temperature p / rho, TRef = pRef / (rhoRef cv (kappa - 1))
template <class T> static T sutherlandKW(const ModelConst& M, T rho, T p) {
    T Tm = p / rho;
    const double TRef = 1 / (M.cv * (M.kapa - 1)) * M.pRef / M.rhoRef;
    return dpow(Tm, 1.5) * (1 + M.suth / TRef) / (Tm + M.suth / TRef);
}
// the wall distance of the point as KOmegaSST takes it: |d|, at least 1e-5 (AERO/KOmegaSST.java:366-369)
static double kwWallDistance(const QPData& q) {
    double walldist = std::fabs(q.wallDistance);
    if (walldist < 1e-5) walldist = 1e-5;
    return walldist;
}
// AERO/KOmegaSST.java:309-419 diffusiveFlux (2D: the vorticity |dv_1/dx_2 - dv_2/dx_1|)
template <class T> static void diffusiveFluxKW(const ModelConst& M, T* W, const T* dW, const double* n, const QPData& q, T* flux) {
    const int nEqs = M.nEqs, dim = M.dim;
    W[0] = limiteRho(W[0]);
    T rho = W[0];
    const double lam = -2. / 3;
    T velocity[3];
    T velocity2 = .0;
    for (int d = 0; d < dim; ++d) {
        velocity[d] = W[d + 1] / rho;
        velocity2 += velocity[d] * velocity[d];
    }
    T velocityJac[9];
    for (int d = 0; d < dim; ++d)
        for (int f = 0; f < dim; ++f) velocityJac[dim * d + f] = (dW[f * nEqs + d + 1] - dW[f * nEqs] * velocity[d]) / rho;
    T p = pressure(M, W);
    T pOverRhoDer[3];
    for (int d = 0; d < dim; ++d) {
        T temp = .0;
        for (int f = 0; f < dim; ++f) temp += velocity[f] * velocityJac[dim * f + d];
        T pDer = (M.kapa - 1) * (dW[d * nEqs + dim + 1] - dW[d * nEqs] * velocity2 / 2 - rho * temp);
        pOverRhoDer[d] = (rho * pDer - p * dW[d * nEqs]) / (rho * rho);
    }
    T stress[9];
    T trace = .0;
    for (int d = 0; d < dim; ++d) {
        trace += velocityJac[dim * d + d];
        for (int f = 0; f < dim; ++f) stress[dim * d + f] = velocityJac[dim * d + f] + velocityJac[dim * f + d];
    }
    for (int d = 0; d < dim; ++d) stress[dim * d + d] += lam * trace;
    T k = kwMax(W[dim + 2] / rho, T(1e-10));
    T om = W[dim + 3] / rho;
    T etaTemp = sutherlandKW(M, rho, p);
    T vortic = dabs(velocityJac[1] - velocityJac[2]);
    const double walldist = kwWallDistance(q);
    T arg2 = kwMax(2 * dsqrt(k) / (0.09 * dexp(om) * walldist) / std::sqrt(M.Re),
                   500 * etaTemp / rho / (walldist * walldist * dexp(om)) / M.Re);
    T F2 = dtanh(arg2 * arg2);
    T mut = kwMax(T(0.0), rho * M.a1kw * k / kwMax(M.a1kw * dexp(om), vortic * F2));
    T turbulentStress[9];
    for (int i = 0; i < dim * dim; i++) turbulentStress[i] = stress[i] * mut;
    T kMax = 2.0 / 3 * kwMax(T(0.0), rho * k);
    for (int d = 0; d < dim; ++d) turbulentStress[dim * d + d] -= kMax;
    T kDer[3], omDer[3];
    for (int d = 0; d < dim; ++d) {
        kDer[d] = (dW[d * nEqs + dim + 2] - dW[d * nEqs] * k) / rho;
        omDer[d] = (dW[d * nEqs + dim + 3] - dW[d * nEqs] * om) / rho;
    }
    T constant = M.kapa / (M.kapa - 1) * (etaTemp / M.Pr + mut / M.Prt);
    T omkDer = 0.0;
    for (int d = 0; d < dim; ++d) omkDer += omDer[d] * kDer[d];
    T CD = kwMax(2 * M.som2 * rho * omkDer, T(1e-10));
    T arg11 = kwMax(dsqrt(k) / (0.09 * dexp(om) * walldist) / std::sqrt(M.Re),
                    500 * etaTemp / rho / (dexp(om) * walldist * walldist) / M.Re);
    T arg12 = 4 * rho * M.som2 * k / (CD * walldist * walldist);
    T arg1 = kwMin(arg11, arg12);
    T F1 = dtanh(dpow(arg1, 4.0));
    T sk0 = F1 * M.sk1 + (1 - F1) * M.sk2;
    T som0 = F1 * M.som1 + (1 - F1) * M.som2;
    for (int j = 0; j < nEqs; j++) flux[j] = 0.0;
    for (int d = 0; d < dim; ++d) {
        T tmp = .0;
        for (int f = 0; f < dim; ++f) {
            flux[f + 1] += (etaTemp * stress[dim * d + f] + turbulentStress[dim * d + f]) * n[d] / M.Re;
            tmp += velocity[f] * (etaTemp * stress[dim * d + f] + turbulentStress[dim * d + f]);
        }
        flux[dim + 1] += (tmp + constant * pOverRhoDer[d]) * n[d] / M.Re;
        flux[dim + 2] += (etaTemp + sk0 * mut) / M.Re * kDer[d] * n[d];
        flux[dim + 3] += (etaTemp + som0 * mut) / M.Re * omDer[d] * n[d];
    }
}

// AERO/NavierStokes.java:317-374
template <class T> static void diffusiveFlux(const ModelConst& M, T* W, const T* dW, const double* n, const QPData& q, T* flux) {
    const int nEqs = M.nEqs, dim = M.dim;
    if (M.kind == 1) {
        diffusiveFluxSA(M, W, dW, n, flux);
        return;
    }
    if (M.kind == 4) {
        diffusiveFluxKW(M, W, dW, n, q, flux);
        return;
    }
    if (M.kind == 2) {   // T/user/equation/TestScalarEquation.java:108-114
        T fl = 0.0;
        for (int d = 0; d < dim; d++) fl += M.k * dW[d] * n[d];
        flux[0] = fl;
        return;
    }
    W[0] = limiteRho(W[0]);
    T rho = W[0];
    const double lam = -2. / 3;
    T velocity[3];
    T velocity2 = .0;
    for (int d = 0; d < dim; ++d) {
        velocity[d] = W[d + 1] / rho;
        velocity2 += velocity[d] * velocity[d];
    }
    T velocityJac[9];
    for (int d = 0; d < dim; ++d)
        for (int f = 0; f < dim; ++f) velocityJac[dim * d + f] = (dW[f * nEqs + d + 1] - dW[f * nEqs] * velocity[d]) / rho;
    T p = pressure(M, W);
    T pOverRhoDer[3];
    for (int d = 0; d < dim; ++d) {
        T temp = .0;
        for (int f = 0; f < dim; ++f) temp += velocity[f] * velocityJac[dim * f + d];
        T pDer = (M.kapa - 1) * (dW[d * nEqs + dim + 1] - dW[d * nEqs] * velocity2 / 2 - rho * temp);
        pOverRhoDer[d] = (rho * pDer - p * dW[d * nEqs]) / (rho * rho);
    }
    T stress[9];
    T trace = .0;
    for (int d = 0; d < dim; ++d) {
        trace += velocityJac[dim * d + d];
        for (int f = 0; f < dim; ++f) stress[dim * d + f] = velocityJac[dim * d + f] + velocityJac[dim * f + d];
    }
    for (int d = 0; d < dim; ++d) stress[dim * d + d] += lam * trace;
    const double constant = M.kapa / (M.kapa - 1) / M.Pr;
    for (int j = 0; j < nEqs; j++) flux[j] = 0.0;
    for (int d = 0; d < dim; ++d) {
        T tmp = .0;
        for (int f = 0; f < dim; ++f) {
            flux[f + 1] += stress[dim * d + f] * n[d] / M.Re;
            tmp += velocity[f] * stress[dim * d + f];
        }
        flux[dim + 1] += (tmp + constant * pOverRhoDer[d]) * n[d] / M.Re;
    }
}
// AERO/NavierStokes.java:376-395
template <class T>
static void numericalDiffusiveFlux(const ModelConst& M, T* Wc, const T* dWc, const double* n, int TT, const QPData& q, T* flux) {
    if (M.kind == 2) {
        diffusiveFlux(M, Wc, dWc, n, q, flux);
        return;
    }
    Wc[0] = limiteRho(Wc[0]);
    diffusiveFlux(M, Wc, dWc, n, q, flux);
    if (TT < 0) {
        if (TT == WALL) {
            T work = .0;
            for (int d = 0; d < M.dim; ++d) work += q.meshVelocity[d] * flux[d + 1];
            flux[M.dim + 1] = work;
        } else {
            for (int j = 0; j < M.nEqs; j++) flux[j] = 0.0;
        }
    }
}
static bool isIPFace(const ModelConst& M, int TT) { return M.kind == 2 ? true : TT == WALL; }
static bool isSourcePresent(const ModelConst& M) { return M.kind == 2 || M.kind == 1 || M.kind == 4; }
// AERO/SpalartAllmaras.java:367-451 sourceTerm (2D: rotationMagnitude = |U[1] - U[2]|)
template <class T> static void sourceTermSA(const ModelConst& M, T* W, const T* dW, const QPData& q, T* source) {
    const int nEqs = M.nEqs, dim = M.dim;
    W[0] = limiteRho(W[0]);
    T rho = W[0];
    T velocity[3];
    for (int d = 0; d < dim; ++d) velocity[d] = W[d + 1] / rho;
    T velocityJac[9];
    for (int d = 0; d < dim; ++d)
        for (int f = 0; f < dim; ++f) velocityJac[dim * d + f] = (dW[f * nEqs + d + 1] - dW[f * nEqs] * velocity[d]) / rho;
    // the stress magnitude Smag only enters the commented production limiter: not evaluated
    T vt = saMax(T(0.0), W[dim + 2] / rho);
    T vtDerMag = 0.0;
    for (int d = 0; d < dim; ++d) {
        T vtDer = 1 / rho * (dW[d * nEqs + dim + 2] - dW[d * nEqs] * vt);
        vtDerMag += vtDer * vtDer;
    }
    if (val(vt) < 0) vt = 0.0;
    const double D = q.wallDistance;
    const double ka = M.ka;
    T xi = rho * vt;
    const double ft2 = 0;
    T fv1 = xi * xi * xi / (xi * xi * xi + M.cv1 * M.cv1 * M.cv1);
    T fv2 = 1 - xi / (1 + xi * fv1);
    T Om = dabs(velocityJac[1] - velocityJac[2]);
    T S = Om + 1 / M.Re * vt / (ka * ka * D * D) * fv2;
    T rt = jmin(vt / (M.Re * S * ka * ka * D * D), T(10.0));
    T g = rt + M.cw2 * (dpow(rt, 6.0) - rt);
    T fw = g * dpow((1 + std::pow(M.cw3, 6.0)) / (dpow(g, 6.0) + std::pow(M.cw3, 6.0)), 1.0 / 6);
    for (int j = 0; j < nEqs; j++) source[j] = 0.0;
    T d = 1 / M.Re * rho * M.cb2 * vtDerMag + M.cb1 * (1 - ft2) * rho * S * vt -
          1 / M.Re * (M.cw1 * fw - M.cb1 / (ka * ka) * ft2) / rho * (rho * vt / D) * (rho * vt / D);
    if (val(d) < -100) d = -100.0;   // limitDestruction
    source[dim + 2] = d;
}
// AERO/KOmegaSST.java:447-537 sourceTerm: k and log(omega) production (limited) and destruction,
// the cross-diffusion and the omega-gradient terms, each clipped at zero
template <class T> static void sourceTermKW(const ModelConst& M, T* W, const T* dW, const QPData& q, T* source) {
    const int nEqs = M.nEqs, dim = M.dim;
    W[0] = limiteRho(W[0]);
    T rho = W[0];
    const double lam = -2. / 3;
    T velocity[3];
    for (int d = 0; d < dim; ++d) velocity[d] = W[d + 1] / rho;
    T velocityJac[9];
    for (int d = 0; d < dim; ++d)
        for (int f = 0; f < dim; ++f) velocityJac[dim * d + f] = (dW[f * nEqs + d + 1] - dW[f * nEqs] * velocity[d]) / rho;
    T p = pressure(M, W);
    T k = kwMax(W[dim + 2] / rho, T(1e-10));
    T om = W[dim + 3] / rho;
    T expOm = dexp(om);
    T etaTemp = sutherlandKW(M, rho, p);
    T vortic = dabs(velocityJac[1] - velocityJac[2]);
    const double walldist = kwWallDistance(q);
    T arg2 = kwMax(2 * dsqrt(k) / (0.09 * dexp(om) * walldist) / std::sqrt(M.Re),
                   500 * etaTemp / rho / (walldist * walldist * dexp(om)) / M.Re);
    T F2 = dtanh(arg2 * arg2);
    T mut = kwMax(T(0.0), rho * M.a1kw * k / kwMax(M.a1kw * dexp(om), vortic * F2));
    T turbulentStress[9];
    T trace = .0;
    for (int d = 0; d < dim; ++d) {
        trace += velocityJac[dim * d + d];
        for (int f = 0; f < dim; ++f) turbulentStress[dim * d + f] = mut * (velocityJac[dim * d + f] + velocityJac[dim * f + d]);
    }
    T kMax = 2.0 / 3 * kwMax(T(0.0), rho * k);
    for (int d = 0; d < dim; ++d) turbulentStress[dim * d + d] += mut * lam * trace - kMax;
    T omkDer = 0.0, omDerSqr = 0.0;
    for (int d = 0; d < dim; ++d) {
        T kDer = (dW[d * nEqs + dim + 2] - dW[d * nEqs] * k) / rho;
        T omDer = (dW[d * nEqs + dim + 3] - dW[d * nEqs] * om) / rho;
        omkDer += omDer * kDer;
        omDerSqr += omDer * omDer;
    }
    T Tv = 0.0;
    for (int d = 0; d < dim; d++)
        for (int f = 0; f < dim; f++) Tv += turbulentStress[dim * d + f] * velocityJac[dim * d + f];
    T CD = kwMax(2 * M.som2 * rho * omkDer, T(1e-10));
    T arg11 = kwMax(dsqrt(k) / (0.09 * dexp(om) * walldist) / std::sqrt(M.Re),
                    500 * etaTemp / rho / (dexp(om) * walldist * walldist) / M.Re);
    T arg12 = 4 * rho * M.som2 * k / (CD * walldist * walldist);
    T arg1 = kwMin(arg11, arg12);
    T F1 = dtanh(dpow(arg1, 4.0));
    T som0 = F1 * M.som1 + (1 - F1) * M.som2;
    T alpha0 = F1 * M.alpha1 + (1 - F1) * M.alpha2;
    T beta0 = F1 * M.beta1 + (1 - F1) * M.beta2;
    if (val(Tv) > val(10 * M.betast * rho * k * expOm)) Tv = 10 * M.betast * rho * k * expOm;
    T Tvom = alpha0 * rho / mut / expOm * Tv;
    if (val(Tvom) > val(10 * beta0 * rho * expOm)) Tvom = 10 * beta0 * rho * expOm;
    for (int j = 0; j < nEqs; j++) source[j] = 0.0;
    source[dim + 2] = kwMax(Tv - M.betast * rho * k * expOm, T(0.0));
    source[dim + 3] = kwMax(Tvom - beta0 * rho * expOm + 1 / M.Re * M.som2 * rho / expOm * omkDer * 2 * (1 - F1) +
                                1 / M.Re * (etaTemp + som0 * mut) * omDerSqr,
                            T(0.0));
}
// T/user/equation/TestScalarEquation.java:131-136
template <class T> static void sourceTerm(const ModelConst& M, T* W, const T* dW, const QPData& q, T* s) {
    if (M.kind == 1) {
        sourceTermSA(M, W, dW, q, s);
        return;
    }
    if (M.kind == 4) {
        sourceTermKW(M, W, dW, q, s);
        return;
    }
    double phase = 2 * q.currentX[0] + q.currentX[1];
    s[0] = (2 * M.a[0] + M.a[1]) * std::cos(phase) + 5 * M.k * std::sin(phase);
}


}  // namespace oracle
