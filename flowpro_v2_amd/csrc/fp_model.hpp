// Device models of the implicit-DG path: compressible Euler / Navier-Stokes, the
// Spalart-Allmaras and k-omega SST RANS models, the isentropic Euler equations and the
// linear scalar test equation, written once over a
// scalar type T that is
// either double (residual) or a forward-mode dual number (exact Jacobian
// columns, J*v, J^T*v).  The call sequence of each function follows the
// reference model classes so that the density clipping `limiteRho`, which
// overwrites its argument, acts on the same values:
//   ModuleProjects/Equations/Aerodynamics/src/main/java/flowpro/user/equation/
//     Aerodynamics.java:323-370, NavierStokes.java:23-395, SpalartAllmaras.java:36-459,
//     AdiabaticEulers.java:9-211 (isentropic Euler: rho, rho u, rho v with p = rho^kappa),
//     KOmegaSST.java:45-585 (rho, rho u, rho v, rho E, rho k, rho log(omega))
//   FlowPro/src/test/java/flowpro/user/equation/TestScalarEquation.java:28-136
#pragma once
#include <hip/hip_runtime.h>

namespace fp {

struct dual {
    double v, d;
    __host__ __device__ dual() : v(0), d(0) {}
    __host__ __device__ dual(double a) : v(a), d(0) {}
    __host__ __device__ dual(double a, double b) : v(a), d(b) {}
};
__host__ __device__ inline dual operator+(dual a, dual b) { return dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline dual operator-(dual a, dual b) { return dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline dual operator-(dual a) { return dual(-a.v, -a.d); }
__host__ __device__ inline dual operator*(dual a, dual b) { return dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline dual operator/(dual a, dual b) {
    double q = a.v / b.v;
    return dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline dual& operator+=(dual& a, dual b) { a = a + b; return a; }
__host__ __device__ inline dual& operator-=(dual& a, dual b) { a = a - b; return a; }
__host__ __device__ inline dual& operator/=(dual& a, dual b) { a = a / b; return a; }
__host__ __device__ inline double value(double a) { return a; }
__host__ __device__ inline double value(dual a) { return a.v; }
__host__ __device__ inline double tangent(double) { return 0.0; }
__host__ __device__ inline double tangent(dual a) { return a.d; }
__device__ inline double t_sqrt(double a) { return sqrt(a); }
__device__ inline dual t_sqrt(dual a) {
    double s = sqrt(a.v);
    return dual(s, a.d == 0.0 ? 0.0 : a.d / (2.0 * s));
}
__device__ inline double t_exp(double a) { return exp(a); }
__device__ inline dual t_exp(dual a) { double e = exp(a.v); return dual(e, e * a.d); }
__device__ inline double t_pow(double a, double b) { return pow(a, b); }
__device__ inline dual t_pow(dual a, double b) {
    double p = pow(a.v, b);
    return dual(p, a.d == 0.0 ? 0.0 : b * pow(a.v, b - 1.0) * a.d);
}
// java.lang.Math.max for ordinary values: (a >= b) ? a : b
__device__ inline double t_max(double a, double b) { return (a >= b) ? a : b; }
__device__ inline dual t_max(dual a, dual b) { return (a.v >= b.v) ? a : b; }
// SpalartAllmaras.max (SpalartAllmaras.java:453-459): (a > b) ? a : b
__device__ inline double t_sa_max(double a, double b) { return (a > b) ? a : b; }
__device__ inline dual t_sa_max(dual a, dual b) { return (a.v > b.v) ? a : b; }
// KOmegaSST.min (KOmegaSST.java:579-585): (a < b) ? a : b
__device__ inline double t_kw_min(double a, double b) { return (a < b) ? a : b; }
__device__ inline dual t_kw_min(dual a, dual b) { return (a.v < b.v) ? a : b; }
__device__ inline double t_tanh(double a) { return tanh(a); }
__device__ inline dual t_tanh(dual a) { double t = tanh(a.v); return dual(t, (1.0 - t * t) * a.d); }
// java.lang.Math.min for ordinary values: (a <= b) ? a : b
__device__ inline double t_min(double a, double b) { return (a <= b) ? a : b; }
__device__ inline dual t_min(dual a, dual b) { return (a.v <= b.v) ? a : b; }
__device__ inline double t_abs(double a) { return fabs(a); }
__device__ inline dual t_abs(dual a) { return a.v < 0.0 ? -a : a; }

enum BoundaryType { BT_WALL = -1, BT_INLET = -2, BT_OUTLET = -3, BT_INVISCID_WALL = -4 };
enum ModelKind { MODEL_NAVIER_STOKES = 0, MODEL_SA = 1, MODEL_SCALAR = 2, MODEL_ADIABATIC = 3, MODEL_KOMEGA = 4 };
// equations per model: rho, rho u, rho v, rho E (+ rho nu~ for Spalart-Allmaras, + rho k, rho log(omega)
// for k-omega SST); 1 for the scalar; rho, rho u, rho v for AdiabaticEulers
template <int MK> constexpr int neq_of() {
    return MK == MODEL_SCALAR ? 1 : (MK == MODEL_SA ? 5 : (MK == MODEL_ADIABATIC ? 3 : (MK == MODEL_KOMEGA ? 6 : 4)));
}
// models with a state-dependent source term contracted with the test function itself
template <int MK> constexpr bool has_rans_source() { return MK == MODEL_SA || MK == MODEL_KOMEGA; }

// Derived model constants (the host computes them as Aerodynamics.init does).
struct ModelParams {
    int kind;          // ModelKind
    int nEqs;
    int diffusive;
    int has_attack;
    double kapa;
    double pOut;       // normalised outlet pressure
    double Re, Pr;
    double attack;
    double a0, a1, k, lambda;   // scalar model
    // Spalart-Allmaras: inlet turbulence and model constants, cw1 = cb1/ka^2 + (1 + cb2)/sigma
    double vtIn, sigma, cb1, cb2, ka, cw1, cw2, cw3, cv1, Prt;
    int flux;          // NavierStokes interior flux: 0 Lax-Friedrichs, 1 van Leer splitting
    // k-omega SST (KOmegaSST.java:45-71): inlet k and log(omega), Suth / TRef of the Sutherland law
    // (TRef = pRef / (rhoRef cv (kappa - 1))) and the two blended constant sets; Prt above is shared
    double kIn, omIn, suthRatio, kwA1, sk1, som1, alpha1, beta1, betast, sk2, som2, alpha2, beta2;
};

constexpr double RHO_TOL = 1e-2;   // Aerodynamics.java:22
constexpr double P_TOL = 1e-2;     // NavierStokes.java:15
constexpr double P_TOL_SA = 1e-1;  // SpalartAllmaras.java:15, KOmegaSST.java:15

template <class T> __device__ inline T limite_rho(T rho) {
    if (value(rho) < RHO_TOL) return RHO_TOL * t_exp((rho - RHO_TOL) / RHO_TOL);
    return rho;
}

// NavierStokes.pressure / SpalartAllmaras.pressure (same formula, own P_TOL)
template <int MK, class T> __device__ inline T pressure_m(const ModelParams& P, T* W) {
    if constexpr (MK == MODEL_ADIABATIC) {   // AdiabaticEulers.pressure: rho^kappa, P_TOL = 0.1
        W[0] = limite_rho(W[0]);
        T p = t_pow(W[0], P.kapa);
        if (value(p) < 1e-1) p = 1e-1 * t_exp((p - 1e-1) / 1e-1);
        return p;
    }
    constexpr double ptol = (MK == MODEL_SA || MK == MODEL_KOMEGA) ? P_TOL_SA : P_TOL;
    W[0] = limite_rho(W[0]);
    T mom2 = W[1] * W[1];
    mom2 += W[2] * W[2];
    T p = (P.kapa - 1.0) * (W[3] - mom2 / (2.0 * W[0]));
    if (value(p) < ptol) p = ptol * t_exp((p - ptol) / ptol);
    return p;
}
template <class T> __device__ inline T pressure(const ModelParams& P, T* W) { return pressure_m<MODEL_NAVIER_STOKES>(P, W); }

// physical convective flux in direction n
template <int MK, class T> __device__ inline void convective_flux(const ModelParams& P, T* W, const double* n, T* f) {
    if constexpr (MK == MODEL_SCALAR) {
        double an = P.a0 * n[0] + P.a1 * n[1];
        f[0] = an * W[0];
        return;
    }
    W[0] = limite_rho(W[0]);
    T V = W[1] * n[0];
    V += W[2] * n[1];
    V /= W[0];
    T p = pressure_m<MK>(P, W);
    f[0] = W[0] * V;
    f[1] = W[1] * V + p * n[0];
    f[2] = W[2] * V + p * n[1];
    if constexpr (MK != MODEL_ADIABATIC) f[3] = (W[3] + p) * V;
    if constexpr (MK == MODEL_SA) f[4] = W[4] * V;
    if constexpr (MK == MODEL_KOMEGA) {   // KOmegaSST.java:210-211
        f[4] = W[4] * V;
        f[5] = W[5] * V;
    }
}

// NavierStokes.fplus (sign +1) / fminus (sign -1) (ModuleProjects/Equations/Aerodynamics/src/
// main/java/flowpro/user/equation/NavierStokes.java:407-465) in DIM dimensions: van Leer's split
// fluxes, mass flux +-rho a (M +- 1)^2 / 4 for |M| < 1, the whole flux on one side beyond
template <int DIM, class T>
__device__ inline void van_leer_split(const ModelParams& P, const T* W, T p, const double* n, int sign, T* f) {
    const double kapa = P.kapa;
#pragma unroll
    for (int j = 0; j < DIM + 2; j++) f[j] = 0.0;
    T vel[DIM];
    T Vn = 0.0, q2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        vel[d] = W[d + 1] / W[0];
        Vn += vel[d] * n[d];
        q2 += vel[d] * vel[d];
    }
    T a = t_sqrt(kapa * p / W[0]);
    T M = Vn / a;
    if (fabs(value(M)) < 1) {
        T fm = sign > 0 ? W[0] * a * ((M + 1.0) * (M + 1.0)) / 4.0 : -W[0] * a * ((M - 1.0) * (M - 1.0)) / 4.0;
        f[0] = fm;
#pragma unroll
        for (int d = 0; d < DIM; ++d) f[d + 1] = fm * (vel[d] + (-Vn + (sign * 2.0) * a) / kapa * n[d]);
        T c = (kapa - 1.0) * Vn + (sign * 2.0) * a;
        f[DIM + 1] = fm * ((q2 - Vn * Vn) / 2.0 + c * c / (2.0 * (kapa * kapa - 1.0)));
    } else if (sign > 0 ? value(M) >= 1 : value(M) <= -1) {
        f[0] = W[0] * Vn;
#pragma unroll
        for (int d = 0; d < DIM; ++d) f[d + 1] = W[d + 1] * Vn + p * n[d];
        f[DIM + 1] = (W[DIM + 1] + p) * Vn;
    }
}

template <int MK, class T> __device__ inline T max_eigenvalue(const ModelParams& P, T* W) {
    if constexpr (MK == MODEL_SCALAR) return T(P.lambda);
    W[0] = limite_rho(W[0]);
    T p = pressure_m<MK>(P, W);
    T a = t_sqrt(P.kapa * p / W[0]);
    T v1 = W[1] / W[0], v2 = W[2] / W[0];
    T v = v1 * v1;
    v += v2 * v2;
    return t_sqrt(v) + a;
}

// numerical convective flux (Lax-Friedrichs inside, boundary fluxes per code); static mesh
template <int MK, class T>
__device__ inline void numerical_convective_flux(const ModelParams& P, T* WL, T* WR, const double* n, int TT, T* f) {
    if constexpr (MK == MODEL_SCALAR) {
        T fL[1], fR[1];
        convective_flux<MK>(P, WL, n, fL);
        convective_flux<MK>(P, WR, n, fR);
        f[0] = (fL[0] + fR[0]) / 2.0 - P.lambda * (WR[0] - WL[0]) / 2.0;
        return;
    }
    constexpr int NE = neq_of<MK>();
    WL[0] = limite_rho(WL[0]);
    WR[0] = limite_rho(WR[0]);
    if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        // NavierStokes takes p(WL), SpalartAllmaras, AdiabaticEulers and KOmegaSST p(WR)
        // (SpalartAllmaras.java:146, AdiabaticEulers.java:136-149, KOmegaSST.java:161-175)
        T p = pressure_m<MK>(P, (MK == MODEL_SA || MK == MODEL_ADIABATIC || MK == MODEL_KOMEGA) ? WR : WL);
        f[0] = 0.0;
        f[1] = p * n[0];
        f[2] = p * n[1];
        if constexpr (MK != MODEL_ADIABATIC) f[3] = 0.0;   // p * (u_mesh . n) with u_mesh = 0
        if constexpr (MK == MODEL_SA) f[4] = 0.0;
        if constexpr (MK == MODEL_KOMEGA) { f[4] = 0.0; f[5] = 0.0; }
        return;
    }
    if (TT == BT_INLET || TT == BT_OUTLET) {
        convective_flux<MK>(P, WR, n, f);
        return;
    }
    if constexpr (MK == MODEL_NAVIER_STOKES) {
        if (P.flux == 1) {   // NavierStokes.java:174-179: fplus(WL, pressure(WL)) + fminus(WR, pressure(WR))
            T pL = pressure_m<MK>(P, WL);
            T fp[NE], fm[NE];
            van_leer_split<2>(P, WL, pL, n, +1, fp);
            T pR = pressure_m<MK>(P, WR);
            van_leer_split<2>(P, WR, pR, n, -1, fm);
#pragma unroll
            for (int j = 0; j < NE; j++) f[j] = fp[j] + fm[j];
            return;
        }
    }
    T fL[NE], fR[NE];
    convective_flux<MK>(P, WL, n, fL);
    convective_flux<MK>(P, WR, n, fR);
    T lL = max_eigenvalue<MK>(P, WL);
    T lR = max_eigenvalue<MK>(P, WR);
    T lam = t_max(lL, lR);
#pragma unroll
    for (int j = 0; j < NE; j++) f[j] = (fL[j] + fR[j] - lam * (WR[j] - WL[j])) / 2.0;
}

// AdiabaticEulers.boundaryValue (AdiabaticEulers.java:36-95), subsonic inlet: walls add n (m . n)
// (as SpalartAllmaras does), the subsonic outlet takes the density pOut^(1/kappa)
template <class T> __device__ inline void boundary_value_adiabatic(const ModelParams& P, T* WL, const double* n, int TT, T* WR) {
    const double kapa = P.kapa;
    WL[0] = limite_rho(WL[0]);
    if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        T nu = WL[1] * n[0];
        nu += WL[2] * n[1];
        WR[0] = WL[0];
        WR[1] = WL[1] + n[0] * nu;
        WR[2] = WL[2] + n[1] * nu;
    } else if (TT == BT_INLET) {
        const double pIn0 = 1.0, rhoIn0 = 1.0;
        T p = pressure_m<MODEL_ADIABATIC>(P, WL);
        T mach;
        if (value(p) > pIn0) mach = 0.0;
        else mach = t_sqrt((2.0 / (kapa - 1.0)) * (-1.0 + t_pow(pIn0 / p, (kapa - 1.0) / kapa)));
        T rhoIn = rhoIn0 * t_pow((1.0 + ((kapa - 1.0) / 2.0) * mach * mach), 1.0 / (1.0 - kapa));
        T vel = mach * t_sqrt((kapa * p) / rhoIn);
        WR[0] = rhoIn;
        WR[1] = -rhoIn * vel * n[0];
        WR[2] = -rhoIn * vel * n[1];
    } else if (TT == BT_OUTLET) {
        T av = WL[1] * WL[1];
        av += WL[2] * WL[2];
        av = t_sqrt(av) / WL[0];
        T a = t_sqrt(kapa * pressure_m<MODEL_ADIABATIC>(P, WL) / WL[0]);
        T mach = av / a;
        if (value(mach) < 1.0) {
            const double rhoOut = pow(P.pOut, 1.0 / kapa);
            WR[0] = rhoOut;
            WR[1] = rhoOut * WL[1] / WL[0];
            WR[2] = rhoOut * WL[2] / WL[0];
        } else {
            WR[0] = WL[0];
            WR[1] = WL[1];
            WR[2] = WL[2];
        }
    } else {
        WR[0] = 0.0;
        WR[1] = 0.0;
        WR[2] = 0.0;
    }
}

// KOmegaSST.boundaryValue (KOmegaSST.java:218-306), subsonic inlet: the wall zeroes k and keeps
// log(omega), the inviscid wall copies the inner state, the inlet takes kIn and omIn
template <class T> __device__ inline void boundary_value_kw(const ModelParams& P, T* WL, const double* n, int TT, T* WR) {
    const double kapa = P.kapa;
    WL[0] = limite_rho(WL[0]);
    T p = pressure_m<MODEL_KOMEGA>(P, WL);
    if (TT == BT_WALL) {   // the model is always diffusive; static mesh: zero wall velocity
        WR[0] = WL[0];
        WR[1] = 0.0;
        WR[2] = 0.0;
        WR[3] = p / (kapa - 1.0);
        WR[4] = 0.0;
        WR[5] = WL[5];
    } else if (TT == BT_INVISCID_WALL) {
#pragma unroll
        for (int m = 0; m < 6; m++) WR[m] = WL[m];
    } else if (TT == BT_INLET) {
        const double pIn0 = 1.0, rhoIn0 = 1.0;
        T mach;
        if (value(p) > pIn0) mach = 0.0;
        else mach = t_sqrt((2.0 / (kapa - 1.0)) * (-1.0 + t_pow(pIn0 / p, (kapa - 1.0) / kapa)));
        T rhoIn = rhoIn0 * t_pow((1.0 + ((kapa - 1.0) / 2.0) * mach * mach), 1.0 / (1.0 - kapa));
        T vel = mach * t_sqrt((kapa * p) / rhoIn);
        T E = p / (kapa - 1.0) + 0.5 * rhoIn * vel * vel;
        WR[0] = rhoIn;
        if (!P.has_attack) {
            WR[1] = -rhoIn * vel * n[0];
            WR[2] = -rhoIn * vel * n[1];
        } else {
            WR[1] = rhoIn * vel * cos(P.attack);
            WR[2] = rhoIn * vel * sin(P.attack);
        }
        WR[3] = E;
        WR[4] = rhoIn * P.kIn;
        WR[5] = rhoIn * P.omIn;
    } else if (TT == BT_OUTLET) {
        T rhoOut = WL[0];
        T av2 = WL[1] * WL[1];
        av2 += WL[2] * WL[2];
        av2 /= rhoOut * rhoOut;
        T a = t_sqrt(kapa * pressure_m<MODEL_KOMEGA>(P, WL) / WL[0]);
        T mach = t_sqrt(av2) / a;
        if (value(mach) < 1.0) {
            WR[0] = rhoOut;
            WR[1] = rhoOut * WL[1] / WL[0];
            WR[2] = rhoOut * WL[2] / WL[0];
            WR[3] = P.pOut / (kapa - 1.0) + rhoOut * av2 / 2.0;
            WR[4] = WL[4];
            WR[5] = WL[5];
        } else {
#pragma unroll
            for (int m = 0; m < 6; m++) WR[m] = WL[m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < 6; m++) WR[m] = 0.0;
    }
}

// boundary state from the inner state (subsonic stagnation inlet, pressure outlet, walls)
template <int MK, class T>
__device__ inline void boundary_value(const ModelParams& P, T* WL, const double* n, int TT, const double* X, T* WR) {
    if constexpr (MK == MODEL_SCALAR) {
        WR[0] = sin(2.0 * X[0] + X[1]);
        return;
    }
    if constexpr (MK == MODEL_ADIABATIC) {
        boundary_value_adiabatic(P, WL, n, TT, WR);
        return;
    }
    if constexpr (MK == MODEL_KOMEGA) {
        boundary_value_kw(P, WL, n, TT, WR);
        return;
    }
    constexpr int NE = neq_of<MK>();
    WL[0] = limite_rho(WL[0]);
    T p = pressure_m<MK>(P, WL);
    const double kapa = P.kapa;
    if (TT == BT_WALL && P.diffusive) {
        WR[0] = WL[0];
        WR[1] = 0.0;
        WR[2] = 0.0;
        WR[3] = p / (kapa - 1.0);   // + rho |u_mesh|^2 / 2 with u_mesh = 0
        if constexpr (MK == MODEL_SA) WR[4] = 0.0;
    } else if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        T nu = WL[1] * n[0];
        nu += WL[2] * n[1];
        WR[0] = WL[0];
        if constexpr (MK == MODEL_SA) {   // SpalartAllmaras adds n nu (SpalartAllmaras.java:229-231)
            WR[1] = WL[1] + n[0] * nu;
            WR[2] = WL[2] + n[1] * nu;
            WR[4] = WL[4];
        } else {
            WR[1] = WL[1] - n[0] * nu;
            WR[2] = WL[2] - n[1] * nu;
        }
        WR[3] = WL[3];
    } else if (TT == BT_INLET) {
        const double pIn0 = 1.0, rhoIn0 = 1.0;
        T mach;
        if (value(p) > pIn0) mach = 0.0;
        else mach = t_sqrt((2.0 / (kapa - 1.0)) * (-1.0 + t_pow(pIn0 / p, (kapa - 1.0) / kapa)));
        T rhoIn = rhoIn0 * t_pow((1.0 + ((kapa - 1.0) / 2.0) * mach * mach), 1.0 / (1.0 - kapa));
        T vel = mach * t_sqrt((kapa * p) / rhoIn);
        T E = p / (kapa - 1.0) + 0.5 * rhoIn * vel * vel;
        WR[0] = rhoIn;
        if (!P.has_attack) {
            WR[1] = -rhoIn * vel * n[0];
            WR[2] = -rhoIn * vel * n[1];
        } else {
            WR[1] = rhoIn * vel * cos(P.attack);
            WR[2] = rhoIn * vel * sin(P.attack);
        }
        WR[3] = E;
        if constexpr (MK == MODEL_SA) WR[4] = rhoIn * P.vtIn;
    } else if (TT == BT_OUTLET) {
        T rhoOut = WL[0];
        T av2 = WL[1] * WL[1];
        av2 += WL[2] * WL[2];
        av2 /= rhoOut * rhoOut;
        T a = t_sqrt(kapa * pressure_m<MK>(P, WL) / WL[0]);
        T mach = t_sqrt(av2) / a;
        if (value(mach) < 1.0) {
            WR[0] = rhoOut;
            WR[1] = rhoOut * WL[1] / WL[0];
            WR[2] = rhoOut * WL[2] / WL[0];
            WR[3] = P.pOut / (kapa - 1.0) + rhoOut * av2 / 2.0;
            if constexpr (MK == MODEL_SA) WR[4] = WL[4];
        } else {
#pragma unroll
            for (int m = 0; m < NE; m++) WR[m] = WL[m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < NE; m++) WR[m] = 0.0;
    }
}

// KOmegaSST.sutherland (KOmegaSST.java:566-570): T = p / rho, suthRatio = Suth / TRef
template <class T> __device__ inline T kw_sutherland(const ModelParams& P, T rho, T p) {
    T Tm = p / rho;
    return t_pow(Tm, 1.5) * (1.0 + P.suthRatio) / (Tm + P.suthRatio);
}
// the wall distance as KOmegaSST takes it: |d|, at least 1e-5 (KOmegaSST.java:366-369)
__device__ inline double kw_wall_distance(double D) {
    double w = fabs(D);
    return w < 1e-5 ? 1e-5 : w;
}
// the blending quantities shared by the k-omega SST flux and source (KOmegaSST.java:364-375, 398-406):
// mut = max(0, rho a1 k / max(a1 e^om, |vorticity| F2)) and F1 = tanh(arg1^4)
template <class T>
__device__ inline void kw_blend(const ModelParams& P, T rho, T k, T om, T eta, T vortic, double wd, T omkDer, T& mut, T& F1) {
    const double sRe = sqrt(P.Re);
    T arg2 = t_sa_max(2.0 * t_sqrt(k) / (0.09 * t_exp(om) * wd) / sRe, 500.0 * eta / rho / (wd * wd * t_exp(om)) / P.Re);
    T F2 = t_tanh(arg2 * arg2);
    mut = t_sa_max(T(0.0), rho * P.kwA1 * k / t_sa_max(P.kwA1 * t_exp(om), vortic * F2));
    T CD = t_sa_max(2.0 * P.som2 * rho * omkDer, T(1e-10));
    T arg11 = t_sa_max(t_sqrt(k) / (0.09 * t_exp(om) * wd) / sRe, 500.0 * eta / rho / (t_exp(om) * wd * wd) / P.Re);
    T arg12 = 4.0 * rho * P.som2 * k / (CD * wd * wd);
    T arg1 = t_kw_min(arg11, arg12);
    F1 = t_tanh(t_pow(arg1, 4.0));
}

// viscous flux in direction n; dW[d*NE + m] = d W_m / d x_d; D: the wall distance at the point
// (k-omega SST only)
template <int MK, class T>
__device__ inline void diffusive_flux(const ModelParams& P, T* W, const T* dW, const double* n, T* flux, double D = 0.0) {
    constexpr int NE = neq_of<MK>();
    if constexpr (MK == MODEL_SCALAR) {
        T fl = P.k * dW[0] * n[0];
        fl += P.k * dW[NE] * n[1];
        flux[0] = fl;
        return;
    }
    W[0] = limite_rho(W[0]);
    T rho = W[0];
    const double lam = -2.0 / 3.0;
    T vel[2] = {W[1] / rho, W[2] / rho};
    T vel2 = vel[0] * vel[0];
    vel2 += vel[1] * vel[1];
    T vj[4];   // vj[2d + f] = d v_d / d x_f
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int f = 0; f < 2; ++f) vj[2 * d + f] = (dW[f * NE + d + 1] - dW[f * NE] * vel[d]) / rho;
    T p = pressure_m<MK>(P, W);
    T prd[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        T temp = vel[0] * vj[d];
        temp += vel[1] * vj[2 + d];
        T pDer = (P.kapa - 1.0) * (dW[d * NE + 3] - dW[d * NE] * vel2 / 2.0 - rho * temp);
        prd[d] = (rho * pDer - p * dW[d * NE]) / (rho * rho);
    }
    T st[4];
    T trace = vj[0];
    trace += vj[3];
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int f = 0; f < 2; ++f) st[2 * d + f] = vj[2 * d + f] + vj[2 * f + d];
    st[0] += lam * trace;
    st[3] += lam * trace;
#pragma unroll
    for (int m = 0; m < NE; m++) flux[m] = 0.0;
    if constexpr (MK == MODEL_SA) {   // SpalartAllmaras.java:325-349
        T vt = t_sa_max(T(0.0), W[4] / rho);
        T vtDer[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) vtDer[d] = 1.0 / rho * (dW[d * NE + 4] - dW[d * NE] * vt);
        T xi = rho * vt;
        T fv1 = xi * xi * xi / (xi * xi * xi + P.cv1 * P.cv1 * P.cv1);
        T mut = rho * vt * fv1;
        T cst = P.kapa / (P.kapa - 1.0) * (1.0 / P.Pr + mut / P.Prt);
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            T tmp = 0.0;
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                flux[f + 1] += (1.0 + mut) * st[2 * d + f] * n[d] / P.Re;
                tmp += vel[f] * (1.0 + mut) * st[2 * d + f];
            }
            flux[3] += (tmp + cst * prd[d]) * n[d] / P.Re;
            flux[4] += 1.0 / (P.Re * P.sigma) * (1.0 + rho * vt) * vtDer[d] * n[d];
        }
    } else if constexpr (MK == MODEL_KOMEGA) {   // KOmegaSST.java:309-419
        T k = t_sa_max(W[4] / rho, T(1e-10));
        T om = W[5] / rho;
        T eta = kw_sutherland(P, rho, p);
        T vortic = t_abs(vj[1] - vj[2]);
        const double wd = kw_wall_distance(D);
        T kDer[2], omDer[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            kDer[d] = (dW[d * NE + 4] - dW[d * NE] * k) / rho;
            omDer[d] = (dW[d * NE + 5] - dW[d * NE] * om) / rho;
        }
        T omkDer = omDer[0] * kDer[0];
        omkDer += omDer[1] * kDer[1];
        T mut, F1;
        kw_blend(P, rho, k, om, eta, vortic, wd, omkDer, mut, F1);
        T ts[4];
#pragma unroll
        for (int i = 0; i < 4; i++) ts[i] = st[i] * mut;
        T kMax = 2.0 / 3.0 * t_sa_max(T(0.0), rho * k);
        ts[0] -= kMax;
        ts[3] -= kMax;
        T cst = P.kapa / (P.kapa - 1.0) * (eta / P.Pr + mut / P.Prt);
        T sk0 = F1 * P.sk1 + (1.0 - F1) * P.sk2;
        T som0 = F1 * P.som1 + (1.0 - F1) * P.som2;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            T tmp = 0.0;
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                flux[f + 1] += (eta * st[2 * d + f] + ts[2 * d + f]) * n[d] / P.Re;
                tmp += vel[f] * (eta * st[2 * d + f] + ts[2 * d + f]);
            }
            flux[3] += (tmp + cst * prd[d]) * n[d] / P.Re;
            flux[4] += (eta + sk0 * mut) / P.Re * kDer[d] * n[d];
            flux[5] += (eta + som0 * mut) / P.Re * omDer[d] * n[d];
        }
    } else {
        const double cst = P.kapa / (P.kapa - 1.0) / P.Pr;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            T tmp = 0.0;
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                flux[f + 1] += st[2 * d + f] * n[d] / P.Re;
                tmp += vel[f] * st[2 * d + f];
            }
            flux[3] += (tmp + cst * prd[d]) * n[d] / P.Re;
        }
    }
}

template <int MK, class T>
__device__ inline void numerical_diffusive_flux(const ModelParams& P, T* Wc, const T* dWc, const double* n, int TT, T* flux,
                                                double D = 0.0) {
    if constexpr (MK == MODEL_SCALAR) {
        diffusive_flux<MK>(P, Wc, dWc, n, flux);
        return;
    }
    constexpr int NE = neq_of<MK>();
    Wc[0] = limite_rho(Wc[0]);
    diffusive_flux<MK>(P, Wc, dWc, n, flux, D);
    if (TT < 0) {
        if (TT == BT_WALL) {
            flux[3] = 0.0;   // u_mesh . (tau n) with u_mesh = 0
        } else {
#pragma unroll
            for (int m = 0; m < NE; m++) flux[m] = 0.0;
        }
    }
}

template <int MK> __device__ inline bool is_ip_face(int TT) { return MK == MODEL_SCALAR ? true : TT == BT_WALL; }

template <class T> __device__ inline void source_term(const ModelParams& P, const double* X, T* s) {
    double phase = 2.0 * X[0] + X[1];
    s[0] = (2.0 * P.a0 + P.a1) * cos(phase) + 5.0 * P.k * sin(phase);
}

// SpalartAllmaras.sourceTerm (SpalartAllmaras.java:367-451) in 2D with ft2 = 0; D is the wall
// distance interpolated at the point.  The stress magnitude only enters a commented-out
// production limiter of the reference and is not evaluated.
template <class T> __device__ inline void sa_source(const ModelParams& P, T* W, const T* dW, double D, T* s) {
    constexpr int NE = 5;
    W[0] = limite_rho(W[0]);
    T rho = W[0];
    T vel[2] = {W[1] / rho, W[2] / rho};
    T vj[4];   // vj[2d + f] = d v_d / d x_f
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int f = 0; f < 2; ++f) vj[2 * d + f] = (dW[f * NE + d + 1] - dW[f * NE] * vel[d]) / rho;
    T vt = t_sa_max(T(0.0), W[4] / rho);
    T vtDerMag = 0.0;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        T vtDer = 1.0 / rho * (dW[d * NE + 4] - dW[d * NE] * vt);
        vtDerMag += vtDer * vtDer;
    }
    if (value(vt) < 0.0) vt = 0.0;
    const double ka = P.ka;
    T xi = rho * vt;
    T fv1 = xi * xi * xi / (xi * xi * xi + P.cv1 * P.cv1 * P.cv1);
    T fv2 = 1.0 - xi / (1.0 + xi * fv1);
    T Om = t_abs(vj[1] - vj[2]);
    T S = Om + 1.0 / P.Re * vt / (ka * ka * D * D) * fv2;
    T rt = t_min(vt / (P.Re * S * ka * ka * D * D), T(10.0));
    T g = rt + P.cw2 * (t_pow(rt, 6.0) - rt);
    const double cw36 = pow(P.cw3, 6.0);
    T fw = g * t_pow((1.0 + cw36) / (t_pow(g, 6.0) + cw36), 1.0 / 6);
    T d = 1.0 / P.Re * rho * P.cb2 * vtDerMag + P.cb1 * rho * S * vt -
          1.0 / P.Re * (P.cw1 * fw) / rho * (rho * vt / D) * (rho * vt / D);
    if (value(d) < -100.0) d = -100.0;   // limitDestruction
    s[0] = 0.0;
    s[1] = 0.0;
    s[2] = 0.0;
    s[3] = 0.0;
    s[4] = d;
}

// KOmegaSST.sourceTerm (KOmegaSST.java:447-537) in 2D: the limited production Tv of k, its omega
// counterpart, the destruction terms, the cross-diffusion and the |grad omega|^2 term, each
// equation's source clipped at zero; D is the wall distance interpolated at the point.
template <class T> __device__ inline void kw_source(const ModelParams& P, T* W, const T* dW, double D, T* s) {
    constexpr int NE = 6;
    const double lam = -2.0 / 3.0;
    W[0] = limite_rho(W[0]);
    T rho = W[0];
    T vel[2] = {W[1] / rho, W[2] / rho};
    T vj[4];   // vj[2d + f] = d v_d / d x_f
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int f = 0; f < 2; ++f) vj[2 * d + f] = (dW[f * NE + d + 1] - dW[f * NE] * vel[d]) / rho;
    T p = pressure_m<MODEL_KOMEGA>(P, W);
    T k = t_sa_max(W[4] / rho, T(1e-10));
    T om = W[5] / rho;
    T expOm = t_exp(om);
    T eta = kw_sutherland(P, rho, p);
    T vortic = t_abs(vj[1] - vj[2]);
    const double wd = kw_wall_distance(D);
    T omkDer = 0.0, omDerSqr = 0.0;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        T kDer = (dW[d * NE + 4] - dW[d * NE] * k) / rho;
        T omDer = (dW[d * NE + 5] - dW[d * NE] * om) / rho;
        omkDer += omDer * kDer;
        omDerSqr += omDer * omDer;
    }
    T mut, F1;
    kw_blend(P, rho, k, om, eta, vortic, wd, omkDer, mut, F1);
    T trace = vj[0];
    trace += vj[3];
    T ts[4];
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int f = 0; f < 2; ++f) ts[2 * d + f] = mut * (vj[2 * d + f] + vj[2 * f + d]);
    T kMax = 2.0 / 3.0 * t_sa_max(T(0.0), rho * k);
    ts[0] += mut * lam * trace - kMax;
    ts[3] += mut * lam * trace - kMax;
    T Tv = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) Tv += ts[i] * vj[i];
    T som0 = F1 * P.som1 + (1.0 - F1) * P.som2;
    T alpha0 = F1 * P.alpha1 + (1.0 - F1) * P.alpha2;
    T beta0 = F1 * P.beta1 + (1.0 - F1) * P.beta2;
    T plim = 10.0 * P.betast * rho * k * expOm;
    if (value(Tv) > value(plim)) Tv = plim;
    T Tvom = alpha0 * rho / mut / expOm * Tv;
    T olim = 10.0 * beta0 * rho * expOm;
    if (value(Tvom) > value(olim)) Tvom = olim;
    s[0] = 0.0;
    s[1] = 0.0;
    s[2] = 0.0;
    s[3] = 0.0;
    s[4] = t_sa_max(Tv - P.betast * rho * k * expOm, T(0.0));
    s[5] = t_sa_max(Tvom - beta0 * rho * expOm + 1.0 / P.Re * P.som2 * rho / expOm * omkDer * 2.0 * (1.0 - F1) +
                        1.0 / P.Re * (eta + som0 * mut) * omDerSqr,
                    T(0.0));
}

}  // namespace fp
