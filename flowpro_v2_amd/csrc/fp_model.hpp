// Device models of the implicit-DG path: compressible Euler / Navier-Stokes and
// the linear scalar test equation, written once over a scalar type T that is
// either double (residual) or a forward-mode dual number (exact Jacobian
// columns, J*v, J^T*v).  The call sequence of each function follows the
// reference model classes so that the density clipping `limiteRho`, which
// overwrites its argument, acts on the same values:
//   ModuleProjects/Equations/Aerodynamics/src/main/java/flowpro/user/equation/
//     Aerodynamics.java:323-370, NavierStokes.java:23-395
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

enum BoundaryType { BT_WALL = -1, BT_INLET = -2, BT_OUTLET = -3, BT_INVISCID_WALL = -4 };
enum ModelKind { MODEL_NAVIER_STOKES = 0, MODEL_SCALAR = 2 };

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
};

constexpr double RHO_TOL = 1e-2;   // Aerodynamics.java:22
constexpr double P_TOL = 1e-2;     // NavierStokes.java:15

template <class T> __device__ inline T limite_rho(T rho) {
    if (value(rho) < RHO_TOL) return RHO_TOL * t_exp((rho - RHO_TOL) / RHO_TOL);
    return rho;
}

template <class T> __device__ inline T pressure(const ModelParams& P, T* W) {
    W[0] = limite_rho(W[0]);
    T mom2 = W[1] * W[1];
    mom2 += W[2] * W[2];
    T p = (P.kapa - 1.0) * (W[3] - mom2 / (2.0 * W[0]));
    if (value(p) < P_TOL) p = P_TOL * t_exp((p - P_TOL) / P_TOL);
    return p;
}

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
    T p = pressure(P, W);
    f[0] = W[0] * V;
    f[1] = W[1] * V + p * n[0];
    f[2] = W[2] * V + p * n[1];
    f[3] = (W[3] + p) * V;
}

template <int MK, class T> __device__ inline T max_eigenvalue(const ModelParams& P, T* W) {
    if constexpr (MK == MODEL_SCALAR) return T(P.lambda);
    W[0] = limite_rho(W[0]);
    T p = pressure(P, W);
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
    WL[0] = limite_rho(WL[0]);
    WR[0] = limite_rho(WR[0]);
    if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        T p = pressure(P, WL);
        f[0] = 0.0;
        f[1] = p * n[0];
        f[2] = p * n[1];
        f[3] = 0.0;   // p * (u_mesh . n) with u_mesh = 0
        return;
    }
    if (TT == BT_INLET || TT == BT_OUTLET) {
        convective_flux<MK>(P, WR, n, f);
        return;
    }
    T fL[4], fR[4];
    convective_flux<MK>(P, WL, n, fL);
    convective_flux<MK>(P, WR, n, fR);
    T lL = max_eigenvalue<MK>(P, WL);
    T lR = max_eigenvalue<MK>(P, WR);
    T lam = t_max(lL, lR);
#pragma unroll
    for (int j = 0; j < 4; j++) f[j] = (fL[j] + fR[j] - lam * (WR[j] - WL[j])) / 2.0;
}

// boundary state from the inner state (subsonic stagnation inlet, pressure outlet, walls)
template <int MK, class T>
__device__ inline void boundary_value(const ModelParams& P, T* WL, const double* n, int TT, const double* X, T* WR) {
    if constexpr (MK == MODEL_SCALAR) {
        WR[0] = sin(2.0 * X[0] + X[1]);
        return;
    }
    WL[0] = limite_rho(WL[0]);
    T p = pressure(P, WL);
    const double kapa = P.kapa;
    if (TT == BT_WALL && P.diffusive) {
        WR[0] = WL[0];
        WR[1] = 0.0;
        WR[2] = 0.0;
        WR[3] = p / (kapa - 1.0);   // + rho |u_mesh|^2 / 2 with u_mesh = 0
    } else if (TT == BT_WALL || TT == BT_INVISCID_WALL) {
        T nu = WL[1] * n[0];
        nu += WL[2] * n[1];
        WR[0] = WL[0];
        WR[1] = WL[1] - n[0] * nu;
        WR[2] = WL[2] - n[1] * nu;
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
    } else if (TT == BT_OUTLET) {
        T rhoOut = WL[0];
        T av2 = WL[1] * WL[1];
        av2 += WL[2] * WL[2];
        av2 /= rhoOut * rhoOut;
        T a = t_sqrt(kapa * pressure(P, WL) / WL[0]);
        T mach = t_sqrt(av2) / a;
        if (value(mach) < 1.0) {
            WR[0] = rhoOut;
            WR[1] = rhoOut * WL[1] / WL[0];
            WR[2] = rhoOut * WL[2] / WL[0];
            WR[3] = P.pOut / (kapa - 1.0) + rhoOut * av2 / 2.0;
        } else {
            WR[0] = WL[0]; WR[1] = WL[1]; WR[2] = WL[2]; WR[3] = WL[3];
        }
    } else {
        WR[0] = 0.0; WR[1] = 0.0; WR[2] = 0.0; WR[3] = 0.0;
    }
}

// viscous flux in direction n; dW[d*4 + m] = d W_m / d x_d
template <int MK, class T> __device__ inline void diffusive_flux(const ModelParams& P, T* W, const T* dW, const double* n, T* flux) {
    if constexpr (MK == MODEL_SCALAR) {
        T fl = P.k * dW[0] * n[0];   // dW[d*4 + m]: x derivative at 0, y derivative at 4
        fl += P.k * dW[4] * n[1];
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
        for (int f = 0; f < 2; ++f) vj[2 * d + f] = (dW[f * 4 + d + 1] - dW[f * 4] * vel[d]) / rho;
    T p = pressure(P, W);
    T prd[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        T temp = vel[0] * vj[d];
        temp += vel[1] * vj[2 + d];
        T pDer = (P.kapa - 1.0) * (dW[d * 4 + 3] - dW[d * 4] * vel2 / 2.0 - rho * temp);
        prd[d] = (rho * pDer - p * dW[d * 4]) / (rho * rho);
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
    const double cst = P.kapa / (P.kapa - 1.0) / P.Pr;
    flux[0] = 0.0;
    flux[1] = 0.0;
    flux[2] = 0.0;
    flux[3] = 0.0;
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

template <int MK, class T>
__device__ inline void numerical_diffusive_flux(const ModelParams& P, T* Wc, const T* dWc, const double* n, int TT, T* flux) {
    if constexpr (MK == MODEL_SCALAR) {
        diffusive_flux<MK>(P, Wc, dWc, n, flux);
        return;
    }
    Wc[0] = limite_rho(Wc[0]);
    diffusive_flux<MK>(P, Wc, dWc, n, flux);
    if (TT < 0) {
        if (TT == BT_WALL) {
            flux[3] = 0.0;   // u_mesh . (tau n) with u_mesh = 0
        } else {
            flux[0] = 0.0; flux[1] = 0.0; flux[2] = 0.0; flux[3] = 0.0;
        }
    }
}

template <int MK> __device__ inline bool is_ip_face(int TT) { return MK == MODEL_SCALAR ? true : TT == BT_WALL; }

template <class T> __device__ inline void source_term(const ModelParams& P, const double* X, T* s) {
    double phase = 2.0 * X[0] + X[1];
    s[0] = (2.0 * P.a0 + P.a1) * cos(phase) + 5.0 * P.k * sin(phase);
}

}  // namespace fp
