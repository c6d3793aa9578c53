// Compressible Euler in 3D (NavierStokes with dimension = 3 and isFlowViscous = false) over
// a scalar type T (double or dual), following ModuleProjects/Equations/Aerodynamics/src/main/
// java/flowpro/user/equation/NavierStokes.java:23-315 and Aerodynamics.java:323-370 with the
// reference's call order, so that the density clipping (which overwrites its argument) acts
// on the same values.  State W = (rho, rho u, rho v, rho w, rho E).
#pragma once
#include "fp_model.hpp"

namespace fp {

template <class T> __device__ inline T pressure3(const ModelParams& P, T* W) {
    W[0] = limite_rho(W[0]);
    T mom2 = W[1] * W[1];
    mom2 += W[2] * W[2];
    mom2 += W[3] * W[3];
    T p = (P.kapa - 1.0) * (W[4] - mom2 / (2.0 * W[0]));
    if (value(p) < P_TOL) p = P_TOL * t_exp((p - P_TOL) / P_TOL);
    return p;
}

template <class T> __device__ inline void convective_flux3(const ModelParams& P, T* W, const double* n, T* f) {
    W[0] = limite_rho(W[0]);
    T V = W[1] * n[0];
    V += W[2] * n[1];
    V += W[3] * n[2];
    V /= W[0];
    T p = pressure3(P, W);
    f[0] = W[0] * V;
    f[1] = W[1] * V + p * n[0];
    f[2] = W[2] * V + p * n[1];
    f[3] = W[3] * V + p * n[2];
    f[4] = (W[4] + p) * V;
}

template <class T> __device__ inline T max_eigenvalue3(const ModelParams& P, T* W) {
    W[0] = limite_rho(W[0]);
    T p = pressure3(P, W);
    T a = t_sqrt(P.kapa * p / W[0]);
    T v1 = W[1] / W[0], v2 = W[2] / W[0], v3 = W[3] / W[0];
    T v = v1 * v1;
    v += v2 * v2;
    v += v3 * v3;
    return t_sqrt(v) + a;
}

template <class T>
__device__ inline void numerical_convective_flux3(const ModelParams& P, T* WL, T* WR, const double* n, int TT, T* f) {
    WL[0] = limite_rho(WL[0]);
    WR[0] = limite_rho(WR[0]);
    if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        T p = pressure3(P, WL);
        f[0] = 0.0;
        f[1] = p * n[0];
        f[2] = p * n[1];
        f[3] = p * n[2];
        f[4] = 0.0;   // p (u_mesh . n) with u_mesh = 0
        return;
    }
    if (TT == BT_INLET || TT == BT_OUTLET) {
        convective_flux3(P, WR, n, f);
        return;
    }
    if (P.flux == 1) {   // van Leer splitting, NavierStokes.java:174-179
        T pL = pressure3(P, WL);
        T fp[5], fm[5];
        van_leer_split<3>(P, WL, pL, n, +1, fp);
        T pR = pressure3(P, WR);
        van_leer_split<3>(P, WR, pR, n, -1, fm);
#pragma unroll
        for (int j = 0; j < 5; j++) f[j] = fp[j] + fm[j];
        return;
    }
    T fL[5], fR[5];
    convective_flux3(P, WL, n, fL);
    convective_flux3(P, WR, n, fR);
    T lL = max_eigenvalue3(P, WL);
    T lR = max_eigenvalue3(P, WR);
    T lam = t_max(lL, lR);
#pragma unroll
    for (int j = 0; j < 5; j++) f[j] = (fL[j] + fR[j] - lam * (WR[j] - WL[j])) / 2.0;
}

// boundary state of the inviscid model: slip reflection on walls, subsonic stagnation inlet
// (flow along -n; the 3D attack angles are not supported on the device), pressure outlet
template <class T> __device__ inline void boundary_value3(const ModelParams& P, T* WL, const double* n, int TT, T* WR) {
    WL[0] = limite_rho(WL[0]);
    T p = pressure3(P, WL);
    const double kapa = P.kapa;
    if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        T nu = WL[1] * n[0];
        nu += WL[2] * n[1];
        nu += WL[3] * n[2];
        WR[0] = WL[0];
        WR[1] = WL[1] - n[0] * nu;
        WR[2] = WL[2] - n[1] * nu;
        WR[3] = WL[3] - n[2] * nu;
        WR[4] = WL[4];
    } else if (TT == BT_INLET) {
        const double pIn0 = 1.0, rhoIn0 = 1.0;
        T mach;
        if (value(p) > pIn0) mach = 0.0;
        else mach = t_sqrt((2.0 / (kapa - 1.0)) * (-1.0 + t_pow(pIn0 / p, (kapa - 1.0) / kapa)));
        T rhoIn = rhoIn0 * t_pow((1.0 + ((kapa - 1.0) / 2.0) * mach * mach), 1.0 / (1.0 - kapa));
        T vel = mach * t_sqrt((kapa * p) / rhoIn);
        T E = p / (kapa - 1.0) + 0.5 * rhoIn * vel * vel;
        WR[0] = rhoIn;
        WR[1] = -rhoIn * vel * n[0];
        WR[2] = -rhoIn * vel * n[1];
        WR[3] = -rhoIn * vel * n[2];
        WR[4] = E;
    } else if (TT == BT_OUTLET) {
        T rhoOut = WL[0];
        T av2 = WL[1] * WL[1];
        av2 += WL[2] * WL[2];
        av2 += WL[3] * WL[3];
        av2 /= rhoOut * rhoOut;
        T a = t_sqrt(kapa * pressure3(P, WL) / WL[0]);
        T mach = t_sqrt(av2) / a;
        if (value(mach) < 1.0) {
            WR[0] = rhoOut;
            WR[1] = rhoOut * WL[1] / WL[0];
            WR[2] = rhoOut * WL[2] / WL[0];
            WR[3] = rhoOut * WL[3] / WL[0];
            WR[4] = P.pOut / (kapa - 1.0) + rhoOut * av2 / 2.0;
        } else {
#pragma unroll
            for (int m = 0; m < 5; m++) WR[m] = WL[m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) WR[m] = 0.0;
    }
}

}  // namespace fp
