// Oracle (test infrastructure only): models, DGpure residual, BDF assembly
// (FD as in the reference, or exact with dual numbers), J*v and J^T*v.
#include <array>
#include <cstring>

#include "oracle_core.hpp"

namespace oracle {

static const double RHO_TOL = 1e-2;   // AERO/Aerodynamics.java:22
static const double P_TOL = 1e-2;     // AERO/NavierStokes.java:15
static const double P_TOL_SA = 1e-1;  // AERO/SpalartAllmaras.java:15
enum { WALL = -1, INLET = -2, OUTLET = -3, INVISCID_WALL = -4 };

struct QPData {            // ElementData fields used on the path (API/ElementData.java)
    double currentX[2];
    double meshVelocity[2];
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
    const double ptol = M.kind == 1 ? P_TOL_SA : P_TOL;
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
    f[M.dim + 1] = (W[M.dim + 1] + p) * V;
    if (M.kind == 1) f[M.dim + 2] = W[M.dim + 2] * V;   // AERO/SpalartAllmaras.java:196
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
// AERO/NavierStokes.java:141-192 (default = Lax-Friedrichs)
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
            T p = pressure(M, M.kind == 1 ? WR : WL);
            f[0] = 0.0;
            double V = .0;
            for (int d = 0; d < dim; ++d) {
                f[d + 1] = p * n[d];
                V += q.meshVelocity[d] * n[d];
            }
            f[dim + 1] = p * V;
            for (int j = 0; j < nEqs; j++) f[j] += V * WR[j];
            break;
        }
        case INLET:
        case OUTLET:
            convectiveFlux(M, WR, n, f);
            break;
        default: {
            T fL[5], fR[5];
            convectiveFlux(M, WL, n, fL);
            convectiveFlux(M, WR, n, fR);
            T lamL = maxEigenvalue(M, WL, q);
            T lamR = maxEigenvalue(M, WR, q);
            T maxEigenValue = dmax(lamL, lamR);
            for (int j = 0; j < nEqs; j++) f[j] = (fL[j] + fR[j] - maxEigenValue * (WR[j] - WL[j])) / 2;
        }
    }
}
// AERO/NavierStokes.java:216-315 (subsonic inlet, no attack angle unless given)
template <class T> static void boundaryValue(const ModelConst& M, T* WL, const double* n, int TT, const QPData& q, T* WR) {
    const int nEqs = M.nEqs, dim = M.dim;
    if (M.kind == 2) {   // exact(currentX)
        WR[0] = std::sin(2 * q.currentX[0] + q.currentX[1]);
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
                double dir[2] = {std::cos(M.attack), std::sin(M.attack)};
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

// AERO/NavierStokes.java:317-374
template <class T> static void diffusiveFlux(const ModelConst& M, T* W, const T* dW, const double* n, T* flux) {
    const int nEqs = M.nEqs, dim = M.dim;
    if (M.kind == 1) {
        diffusiveFluxSA(M, W, dW, n, flux);
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
        diffusiveFlux(M, Wc, dWc, n, flux);
        return;
    }
    Wc[0] = limiteRho(Wc[0]);
    diffusiveFlux(M, Wc, dWc, n, flux);
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
static bool isSourcePresent(const ModelConst& M) { return M.kind == 2 || M.kind == 1; }
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
// T/user/equation/TestScalarEquation.java:131-136
template <class T> static void sourceTerm(const ModelConst& M, T* W, const T* dW, const QPData& q, T* s) {
    if (M.kind == 1) {
        sourceTermSA(M, W, dW, q, s);
        return;
    }
    double phase = 2 * q.currentX[0] + q.currentX[1];
    s[0] = (2 * M.a[0] + M.a[1]) * std::cos(phase) + 5 * M.k * std::sin(phase);
}

// ------------------------------------------------------------- DGpure residual
// C/element/DGpure.java:195-355 residuumWall.  Wown = W + V of this element,
// Wnb = neighbour state (nullptr on boundary faces).  KR may be nullptr.
template <class T>
static void residuumWall(const Case& c, const Elem& e, int k, const T* Wown, const T* Wnb, T* K, T* KR) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs, dim = 2, nBasis = e.nBasis;
    const int TT = e.TT[k];
    const int nR = nBasis;
    if (KR) for (int j = 0; j < nR * nEqs; j++) KR[j] = 0.0;
    double beta0 = c.P.beta0;
    if (c.P.order == 1) beta0 = 0;   // DGpure.java:287-289
    const bool convective = true;
    const bool diffusive = M.diffusive;
    for (int p = 0; p < e.nqf[k]; p++) {
        const double* baseLeft = &e.baseL[k][p * nBasis];
        const double* dBaseLeft = &e.dbaseL[k][p * nBasis * 2];
        const double Jac = e.JacF[k][p];
        const double weight = e.wF[k][p];
        const double* baseRight = TT > -1 ? &e.baseR[k][p * nR] : nullptr;
        const double* dBaseRight = TT > -1 ? &e.dbaseR[k][p * nR * 2] : nullptr;
        const double* nn = e.n[k][p];
        double dL = 0;
        for (int d = 0; d < dim; d++) {
            if (TT > -1) dL += (e.Xs[d] - c.E[TT].Xs[d]) * nn[d];
            else dL += (e.Xs[d] - e.Xes[k][d]) * nn[d];
        }
        dL = std::fabs(dL);
        // interpolateVelocityAndFillElementDataObjectOnFace (Element.java:300-332)
        QPData q;
        {
            int ei[2] = {k, (k + 1) % 3};
            static const double CX[3][2] = {{0, 0}, {1, 0}, {0, 1}};
            double cXi[2] = {0, 0};
            q.meshVelocity[0] = q.meshVelocity[1] = 0;
            q.wallDistance = 0;
            for (int j = 0; j < 2; j++) {
                double ip = e.interpF[k][2 * p + j];
                for (int d = 0; d < dim; d++) cXi[d] += ip * CX[ei[j]][d];
                q.wallDistance += ip * e.wallDist[ei[j]];
            }
            q.currentX[0] = e.A[0] * cXi[0] + e.A[1] * cXi[1] + e.A[2];
            q.currentX[1] = e.B[0] * cXi[0] + e.B[1] * cXi[1] + e.B[2];
        }
        T WL[5], WR[5], dWL[10], dWR[10];
        for (int m = 0; m < nEqs; m++) { WL[m] = 0.0; WR[m] = 0.0; }
        for (int m = 0; m < dim * nEqs; m++) { dWL[m] = 0.0; dWR[m] = 0.0; }
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nBasis; j++) {
                WL[m] += Wown[m * nBasis + j] * baseLeft[j];
                for (int d = 0; d < dim; d++) dWL[nEqs * d + m] += Wown[m * nBasis + j] * dBaseLeft[j * 2 + d];
            }
        if (TT > -1) {
            for (int m = 0; m < nEqs; m++)
                for (int j = 0; j < nR; j++) {
                    WR[m] += Wnb[m * nR + j] * baseRight[j];
                    for (int d = 0; d < dim; d++) dWR[nEqs * d + m] += Wnb[m * nR + j] * dBaseRight[j * 2 + d];
                }
        } else {
            boundaryValue(M, WL, nn, TT, q, WR);
            for (int m = 0; m < dim * nEqs; m++) dWR[m] = dWL[m];
        }
        double vn = 0;
        T Wale[5];
        T fn[5];
        if (convective) {
            for (int d = 0; d < dim; d++) vn = vn + q.meshVelocity[d] * nn[d];
            if (TT > -1)
                for (int j = 0; j < nEqs; j++) Wale[j] = (WL[j] + WR[j]) / 2;
            else
                for (int j = 0; j < nEqs; j++) Wale[j] = WR[j];
            numericalConvectiveFlux(M, WL, WR, nn, TT, q, fn);
        }
        T fvn[5];
        if (diffusive) {
            T Wc[5], dWc[10];
            for (int m = 0; m < nEqs; m++) {
                if (TT > -1) Wc[m] = (WL[m] + WR[m]) / 2;
                else Wc[m] = WR[m];
                for (int d = 0; d < dim; d++)
                    dWc[nEqs * d + m] = (dWL[nEqs * d + m] + dWR[nEqs * d + m]) / 2 + beta0 * (WR[m] - WL[m]) / dL * nn[d];
            }
            numericalDiffusiveFlux(M, Wc, dWc, nn, TT, q, fvn);
        }
        for (int m = 0; m < nEqs; m++) {
            T dWsum = 0.0;
            if (TT > -1)
                for (int d = 0; d < dim; d++) dWsum += (dWL[nEqs * d + m] + dWR[nEqs * d + m]) / 2 * nn[d];
            T avFlux = 0.0;
            if (TT > -1) {
                const Elem& er = c.E[TT];
                double epsFace = 0.5 * (e.eps + er.eps) + 0.0 /*dampConst*/ + 0.5 * (e.dampInner[m] + er.dampInner[m]);
                avFlux = epsFace * (dWsum + beta0 * (WR[m] - WL[m]) / dL);
            }
            for (int j = 0; j < nBasis; j++) {
                double jwb = Jac * weight * baseLeft[j];
                if (convective) {
                    K[nBasis * m + j] -= jwb * (fn[m] - vn * Wale[m]);
                    if (TT > -1) K[nBasis * m + j] += avFlux * jwb;
                }
                if (diffusive) {
                    K[nBasis * m + j] += jwb * fvn[m];
                    if (TT > -1) K[nBasis * m + j] -= e.c_IP * jwb * (WL[m] - WR[m]);
                    else if (isIPFace(M, TT)) K[nBasis * m + j] -= e.c_IP * jwb * (WL[m] - WR[m]);
                }
            }
            if (KR) {
                for (int j = 0; j < nR; j++) {
                    double jwb = Jac * weight * baseRight[j];
                    if (convective) {
                        KR[nR * m + j] -= jwb * (fn[m] - vn * Wale[m]);
                        KR[nR * m + j] += avFlux * jwb;
                    }
                    if (diffusive) {
                        KR[nR * m + j] += jwb * fvn[m];
                        KR[nR * m + j] -= e.c_IP * jwb * (WL[m] - WR[m]);
                    }
                }
            }
        }
    }
}

// C/element/DGpure.java:108-192 residuum (faces first, then the volume)
template <class T>
static void residuum(const Case& c, const Elem& e, const T* Wown, const T* const* Wnb, T* K, T* const* KR) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs, dim = 2, nBasis = e.nBasis;
    for (int k = 0; k < 3; k++) residuumWall(c, e, k, Wown, Wnb[k], K, KR[k]);
    const bool volumeFluxes = e.order > 1;
    if (!(volumeFluxes || isSourcePresent(M))) return;
    static const double CX[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    for (int p = 0; p < e.nq; p++) {
        const double* base = &e.baseV[p * nBasis];
        const double* dBase = &e.dbaseV[p * nBasis * 2];
        const double Jac = e.JacV[p];
        const double weight = e.wV[p];
        QPData q;   // Element.java:263-298
        {
            double cXi[2] = {0, 0};
            q.meshVelocity[0] = q.meshVelocity[1] = 0;
            q.wallDistance = 0;
            for (int j = 0; j < 3; j++) {
                double ip = e.interpV[3 * p + j];
                for (int d = 0; d < dim; d++) cXi[d] += ip * CX[j][d];
                q.wallDistance += ip * e.wallDist[j];
            }
            q.currentX[0] = e.A[0] * cXi[0] + e.A[1] * cXi[1] + e.A[2];
            q.currentX[1] = e.B[0] * cXi[0] + e.B[1] * cXi[1] + e.B[2];
        }
        const double u[2] = {0, 0};
        T WInt[5], dWInt[10];
        for (int m = 0; m < nEqs; m++) WInt[m] = 0.0;
        for (int m = 0; m < dim * nEqs; m++) dWInt[m] = 0.0;
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nBasis; j++) {
                WInt[m] += Wown[m * nBasis + j] * base[j];
                for (int d = 0; d < dim; d++) dWInt[nEqs * d + m] += Wown[m * nBasis + j] * dBase[j * 2 + d];
            }
        T f[2][5], fv[2][5], product[5];
        double nor[2] = {0, 0};
        if (volumeFluxes)
            for (int d = 0; d < dim; d++) {
                nor[d] = 1;
                convectiveFlux(M, WInt, nor, f[d]);
                nor[d] = 0;
            }
        if (volumeFluxes && M.diffusive)
            for (int d = 0; d < dim; d++) {
                nor[d] = 1;
                diffusiveFlux(M, WInt, dWInt, nor, fv[d]);
                nor[d] = 0;
            }
        if (isSourcePresent(M)) sourceTerm(M, WInt, dWInt, q, product);
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nBasis; j++) {
                if (volumeFluxes) {
                    T fsum = 0.0, dWsum = 0.0;
                    for (int d = 0; d < dim; d++) {
                        fsum += (f[d][m] - u[d] * WInt[m]) * dBase[j * 2 + d];
                        dWsum += dWInt[nEqs * d + m] * dBase[j * 2 + d];
                    }
                    K[nBasis * m + j] += Jac * weight * fsum - (e.eps + 0.0 + e.dampInner[m]) * Jac * weight * dWsum;
                }
                if (volumeFluxes && M.diffusive) {
                    T fvsum = 0.0;
                    for (int d = 0; d < dim; d++) fvsum += fv[d][m] * dBase[j * 2 + d];
                    K[nBasis * m + j] -= Jac * weight * fvsum;
                }
                if (isSourcePresent(M)) K[nBasis * m + j] += Jac * weight * product[m] * base[j];
            }
    }
}

// ------------------------------------------------------------- limiter
// C/element/DGpure.java:358-400 for dampTol = dampInnerTol = 0: eps = 0, c_IP = penalty
void limiter_all(Case& c) {
    for (auto& e : c.E) {
        e.eps = 0;
        e.c_IP = 0;
        if (e.order > 1 && c.Mc.diffusive) e.c_IP = c.P.penalty;
    }
}

// ------------------------------------------------------------- residual
void residual_all(Case& c, double* R) {
    const int nEqs = c.Mc.nEqs;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        const int B = e.nBasis * nEqs;
        const double* Wnb[3];
        double* KR[3] = {nullptr, nullptr, nullptr};
        for (int k = 0; k < 3; k++) Wnb[k] = e.TT[k] > -1 ? c.E[e.TT[k]].W.data() : nullptr;
        double* K = R + (size_t)i * B;
        for (int j = 0; j < B; j++) K[j] = 0;
        residuum<double>(c, e, e.W.data(), Wnb, K, KR);
    }
}

// BDF2.coefficients (C/element/BDF2.java:53-61); BDF1 uses {1/dt, -1/dt, 0}
static void bdf_coeffs(const Case& c, double dt, double dto, double* a) {
    if (c.P.time_order == 2 && dto > 0) {
        a[0] = (2 * dt + dto) / (dt * (dt + dto));
        a[1] = -(dt + dto) / (dt * dto);
        a[2] = dt / (dto * (dt + dto));
    } else {
        a[0] = 1 / dt;
        a[1] = -1 / dt;
        a[2] = 0;
    }
}

// BDF1.assembleRHS / BDF2.assembleRHS (C/element/BDF1.java:34-49, BDF2.java:32-51)
static void assemble_rhs(const Case& c, Elem& e, const double* Rw, const double* a) {
    const int nEqs = c.Mc.nEqs, Nb = e.nBasis, B = Nb * nEqs;
    for (int j = 0; j < B; j++) { e.RHS[j] = Rw[j]; e.spatialR[j] = Rw[j]; }
    for (int m = 0; m < nEqs; m++)
        for (int i = 0; i < Nb; i++)
            for (int j = 0; j < Nb; j++) {
                if (c.P.time_order == 2)
                    e.RHS[Nb * m + i] -= e.M[i * Nb + j] * a[0] * e.W[m * Nb + j] + e.M[i * Nb + j] * a[1] * e.Wo[m * Nb + j] +
                                         e.M[i * Nb + j] * a[2] * e.Wo2[m * Nb + j];
                else
                    e.RHS[Nb * m + i] -= e.M[i * Nb + j] * a[0] * e.W[m * Nb + j] + e.M[i * Nb + j] * a[1] * e.Wo[m * Nb + j];
            }
}

void assemble_residuum(Case& c, double dt, double dto) {
    double a[3];
    bdf_coeffs(c, dt, dto, a);
    const int nEqs = c.Mc.nEqs;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        Elem& e = c.E[i];
        const int B = e.nBasis * nEqs;
        std::vector<double> Rw(B, 0.0);
        const double* Wnb[3];
        double* KR[3] = {nullptr, nullptr, nullptr};
        for (int k = 0; k < 3; k++) Wnb[k] = e.TT[k] > -1 ? c.E[e.TT[k]].W.data() : nullptr;
        residuum<double>(c, e, e.W.data(), Wnb, Rw.data(), KR);
        assemble_rhs(c, e, Rw.data(), a);
    }
}

// BDF1/BDF2.assembleJacobiMatrix (C/element/BDF1.java:60-119, BDF2.java:72-132)
void assemble(Case& c, double dt, double dto) {
    double a[3];
    bdf_coeffs(c, dt, dto, a);
    const int nEqs = c.Mc.nEqs;
    const double h = c.P.h;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        Elem& e = c.E[i];
        const int Nb = e.nBasis, B = Nb * nEqs;
        std::vector<double> Rw(B, 0.0);
        std::vector<double> RwNeigh[3], RwNeighH[3];
        double* KR[3] = {nullptr, nullptr, nullptr};
        double* KRh[3] = {nullptr, nullptr, nullptr};
        const double* Wnb[3];
        for (int k = 0; k < 3; k++) {
            Wnb[k] = e.TT[k] > -1 ? c.E[e.TT[k]].W.data() : nullptr;
            if (e.TT[k] > -1) {
                RwNeigh[k].assign(B, 0.0);
                RwNeighH[k].assign(B, 0.0);
                KR[k] = RwNeigh[k].data();
                KRh[k] = RwNeighH[k].data();
            }
        }
        residuum<double>(c, e, e.W.data(), Wnb, Rw.data(), KR);
        assemble_rhs(c, e, Rw.data(), a);
        if (c.P.jacobian_mode == 0) {
            // one-sided FD, the perturbed residual accumulates onto -Rw (BDF1.java:89-107)
            std::vector<double> Wh(B);
            for (int ii = 0; ii < B; ii++) {
                double* Ai = &e.ADiag[(size_t)ii * B];
                for (int j = 0; j < B; j++) Ai[j] = -Rw[j];
                for (int j = 0; j < B; j++) Wh[j] = e.W[j] + (j == ii ? h : 0.0);
                residuum<double>(c, e, Wh.data(), Wnb, Ai, KRh);
                for (int k = 0; k < 3; k++) {
                    if (e.TT[k] > -1) {
                        Elem& nb = c.E[e.TT[k]];
                        double* Aaux = &nb.ANeigh[e.faceIndexReverse[k]][(size_t)ii * B];
                        for (int j = 0; j < B; j++) Aaux[j] = (RwNeighH[k][j] - RwNeigh[k][j]) / h;
                    }
                }
            }
            for (int j = 0; j < B * B; j++) e.ADiag[j] /= -h;
        } else {
            // exact derivative: forward mode, one direction per perturbed DOF
            std::vector<Dual> Wd(B), K(B), KRd[3];
            std::vector<Dual> Wnbd[3];
            const Dual* Wnbp[3];
            Dual* KRp[3];
            for (int k = 0; k < 3; k++) {
                Wnbp[k] = nullptr;
                KRp[k] = nullptr;
                if (e.TT[k] > -1) {
                    const auto& Wn = c.E[e.TT[k]].W;
                    Wnbd[k].resize(B);
                    for (int j = 0; j < B; j++) Wnbd[k][j] = Dual(Wn[j], 0.0);
                    KRd[k].resize(B);
                    Wnbp[k] = Wnbd[k].data();
                    KRp[k] = KRd[k].data();
                }
            }
            for (int ii = 0; ii < B; ii++) {
                for (int j = 0; j < B; j++) { Wd[j] = Dual(e.W[j], j == ii ? 1.0 : 0.0); K[j] = Dual(0.0, 0.0); }
                residuum<Dual>(c, e, Wd.data(), Wnbp, K.data(), KRp);
                double* Ai = &e.ADiag[(size_t)ii * B];
                for (int j = 0; j < B; j++) Ai[j] = -K[j].d;
                for (int k = 0; k < 3; k++) {
                    if (e.TT[k] > -1) {
                        Elem& nb = c.E[e.TT[k]];
                        double* Aaux = &nb.ANeigh[e.faceIndexReverse[k]][(size_t)ii * B];
                        for (int j = 0; j < B; j++) Aaux[j] = KRd[k][j].d;
                    }
                }
            }
        }
        for (int m = 0; m < nEqs; m++)
            for (int r = 0; r < Nb; r++)
                for (int s = 0; s < Nb; s++) e.ADiag[(size_t)(Nb * m + r) * B + Nb * m + s] += a[0] * e.M[r * Nb + s];
    }
}

// (a1 M - dR/dW) v with the full tangent (own and neighbour W carry v): J*v.
void jvp(Case& c, double a1, const double* v, double* out) {
    const int nEqs = c.Mc.nEqs;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        const int Nb = e.nBasis, B = Nb * nEqs;
        std::vector<Dual> Wd(B), K(B);
        std::vector<Dual> Wnbd[3];
        const Dual* Wnbp[3];
        Dual* KRp[3] = {nullptr, nullptr, nullptr};
        for (int j = 0; j < B; j++) { Wd[j] = Dual(e.W[j], v[(size_t)i * B + j]); K[j] = Dual(0.0, 0.0); }
        for (int k = 0; k < 3; k++) {
            Wnbp[k] = nullptr;
            int t = e.TT[k];
            if (t > -1) {
                Wnbd[k].resize(B);
                for (int j = 0; j < B; j++) Wnbd[k][j] = Dual(c.E[t].W[j], v[(size_t)t * B + j]);
                Wnbp[k] = Wnbd[k].data();
            }
        }
        residuum<Dual>(c, e, Wd.data(), Wnbp, K.data(), KRp);
        for (int m = 0; m < nEqs; m++)
            for (int r = 0; r < Nb; r++) {
                double mv = 0;
                for (int s = 0; s < Nb; s++) mv += e.M[r * Nb + s] * v[(size_t)i * B + m * Nb + s];
                out[(size_t)i * B + m * Nb + r] = a1 * mv - K[m * Nb + r].d;
            }
    }
}

// (a1 M - dR/dW)^T v, matrix free: per element, per own DOF, one tangent pass;
// the neighbour rows are reached through KR (the neighbour's residual as seen
// from this element, the same convention as the reference's neighbour blocks).
void vjp(Case& c, double a1, const double* v, double* out) {
    const int nEqs = c.Mc.nEqs;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        const int Nb = e.nBasis, B = Nb * nEqs;
        std::vector<Dual> Wd(B), K(B), KRd[3];
        std::vector<Dual> Wnbd[3];
        const Dual* Wnbp[3];
        Dual* KRp[3];
        for (int k = 0; k < 3; k++) {
            Wnbp[k] = nullptr;
            KRp[k] = nullptr;
            int t = e.TT[k];
            if (t > -1) {
                Wnbd[k].resize(B);
                for (int j = 0; j < B; j++) Wnbd[k][j] = Dual(c.E[t].W[j], 0.0);
                KRd[k].resize(B);
                Wnbp[k] = Wnbd[k].data();
                KRp[k] = KRd[k].data();
            }
        }
        for (int ii = 0; ii < B; ii++) {
            for (int j = 0; j < B; j++) { Wd[j] = Dual(e.W[j], j == ii ? 1.0 : 0.0); K[j] = Dual(0.0, 0.0); }
            residuum<Dual>(c, e, Wd.data(), Wnbp, K.data(), KRp);
            double s = 0;
            for (int j = 0; j < B; j++) s += -K[j].d * v[(size_t)i * B + j];
            for (int k = 0; k < 3; k++) {
                int t = e.TT[k];
                if (t > -1)
                    for (int j = 0; j < B; j++) s += KRd[k][j].d * v[(size_t)t * B + j];
            }
            int m = ii / Nb, r = ii % Nb;
            double mv = 0;
            for (int q = 0; q < Nb; q++) mv += e.M[q * Nb + r] * v[(size_t)i * B + m * Nb + q];
            out[(size_t)i * B + ii] = s + a1 * mv;
        }
    }
}

// ------------------------------------------------------------- functionals
// Element.computeFunctional / computeFunctionalsWall (C/element/Element.java:789-862) with
// FOI/DragCoefficient.java:31-48 (which = 0, p n_x) and FOI/LiftCoefficient.java:32-34 (which = 1, p n_y).
static double functional_elem(const Case& c, const Elem& e, int which, const double* W) {
    const int nEqs = c.Mc.nEqs, Nb = e.nBasis;
    double f = 0;
    for (int k = 0; k < 3; k++) {
        double fk = 0;
        for (int p = 0; p < e.nqf[k]; p++) {
            const double* baseLeft = &e.baseL[k][p * Nb];
            double Jac = e.JacF[k][p], weight = e.wF[k][p];
            double WL[5] = {0, 0, 0, 0, 0};
            for (int j = 0; j < nEqs; j++)
                for (int m = 0; m < Nb; m++) WL[j] += W[j * Nb + m] * baseLeft[m];
            double aux = 0;
            if (e.TT[k] == -1) {
                double momentum2 = .0;
                for (int d = 0; d < 2; ++d) momentum2 += WL[d + 1] * WL[d + 1];
                double pr = (c.P.kappa - 1) * (WL[3] - momentum2 / (2 * WL[0]));
                aux = pr * e.n[k][p][which];
            }
            fk += Jac * weight * aux;
        }
        f += fk;
    }
    return f;   // insideValue is 0
}

double functional(const Case& c, int which, const double* W) {
    double I = 0;
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        const int B = e.nBasis * c.Mc.nEqs;
        I += functional_elem(c, e, which, W ? W + (size_t)i * B : e.W.data());
    }
    return I;
}

// OptimisationToolExport.exportFunctionalDerivative via Element.exportLocalFunctionalDerivative
// (C/element/Element.java:720-731): forward FD with h per DOF on the element sum.
void functional_grad_fd(Case& c, int which, double* dIdW) {
    const double h = c.P.h;
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        const int B = e.nBasis * c.Mc.nEqs;
        double Iw = functional_elem(c, e, which, e.W.data());
        std::vector<double> Wh(e.W);
        for (int j = 0; j < B; j++) {
            Wh[j] = e.W[j] + h;
            double Iwh = functional_elem(c, e, which, Wh.data());
            dIdW[(size_t)i * B + j] = (Iwh - Iw) / h;
            Wh[j] = e.W[j];
        }
    }
}


// Pointwise model functions for the tests: source term and the diffusive flux in x and y of
// the case's model at one state (W, dW[d*nEqs + m]) and wall distance D.
void model_pointwise(const Case& c, const double* W, const double* dW, double D, double* source, double* fvx, double* fvy) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs;
    double w[5], dw[10];
    QPData q;
    q.currentX[0] = q.currentX[1] = 0;
    q.meshVelocity[0] = q.meshVelocity[1] = 0;
    q.wallDistance = D;
    for (int m = 0; m < nEqs; m++) w[m] = W[m];
    for (int m = 0; m < 2 * nEqs; m++) dw[m] = dW[m];
    for (int m = 0; m < nEqs; m++) source[m] = 0;
    if (isSourcePresent(M)) sourceTerm(M, w, dw, q, source);
    const double ex[2] = {1, 0}, ey[2] = {0, 1};
    for (int m = 0; m < nEqs; m++) w[m] = W[m];
    diffusiveFlux(M, w, dw, ex, fvx);
    for (int m = 0; m < nEqs; m++) w[m] = W[m];
    diffusiveFlux(M, w, dw, ey, fvy);
}
}  // namespace oracle
