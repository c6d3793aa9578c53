// DG kernels on quadrilaterals (squareElement: basis2Dsquare, QuadratureSquare with the tensor
// Gauss rule of the element order, transformation2Dsquare) for the NavierStokes (Euler when
// inviscid) and AdiabaticEulers models.
//
// The bilinear map's Jacobian varies over the element, so the physical basis gradients and the
// volume weights |det J| w are read per point from tables built at setup, and the mass matrix is
// the element's own.  One thread per element (residual, J v) or per (element, own DOF) (exact
// Jacobian columns and J^T v).  The arithmetic is the one of DGpure.residuum / residuumWall
// (FlowPro/src/main/java/flowpro/core/element/DGpure.java:108-355): faces first, then the volume;
// the neighbour-side values at a face point come from the neighbour's basis at the same physical
// point, found by the reference's Newton inverse of the neighbour's map (Integration.Face,
// FlowPro/src/main/java/flowpro/core/Integration.java:140-151, with transformation2Dsquare.getXi)
// and tabulated at setup.
#include <hip/hip_runtime.h>

#include "fp_common.hpp"

namespace fp {

// coefficient accessors: (element, local index) -> value of type T
struct QPlain {
    const double* W;
    int B;
    __device__ double operator()(int el, int i) const { return W[(size_t)el * B + i]; }
};
struct QTangent {   // W + eps V
    const double* W;
    const double* V;
    int B;
    __device__ dual operator()(int el, int i) const {
        const size_t o = (size_t)el * B + i;
        return dual(W[o], V[o]);
    }
};
struct QConst {     // W with a zero tangent
    const double* W;
    int B;
    __device__ dual operator()(int el, int i) const { return dual(W[(size_t)el * B + i], 0.0); }
};
struct QColumn {    // W + eps e_(e, dir)
    const double* W;
    int B, dir;
    __device__ dual operator()(int el, int i) const { return dual(W[(size_t)el * B + i], i == dir ? 1.0 : 0.0); }
};

// face integrand g = -fn + fvn - cIP [W] (+ the damping flux on interior faces: artificial
// viscosity of element e and its neighbour) (residuumWall's K update per point divided by the test
// function), in the reference's call order: boundary value, convective flux, diffusive flux, damping
template <int MK, int NE, bool VISC, bool AV, class T>
__device__ inline void quad_face_flux(const ModelParams& P, const DGParams& G, const double* av, int e, int TT, const double* n,
                                      double dL, const double* X, T* WL, T* dWL, T* WR, T* dWR, T* g) {
    if (TT < 0) {
        boundary_value<MK>(P, WL, n, TT, X, WR);
#pragma unroll
        for (int m = 0; m < 2 * NE; m++) dWR[m] = dWL[m];
    }
    T fn[NE];
    numerical_convective_flux<MK>(P, WL, WR, n, TT, fn);
#pragma unroll
    for (int m = 0; m < NE; m++) g[m] = -fn[m];
    if constexpr (VISC) {
        T Wc[NE], dWc[2 * NE], fvn[NE];
#pragma unroll
        for (int m = 0; m < NE; m++) {
            Wc[m] = TT > -1 ? (WL[m] + WR[m]) / 2.0 : WR[m];
            T jump = G.beta0 * (WR[m] - WL[m]) / dL;
            dWc[m] = (dWL[m] + dWR[m]) / 2.0 + jump * n[0];
            dWc[NE + m] = (dWL[NE + m] + dWR[NE + m]) / 2.0 + jump * n[1];
        }
        numerical_diffusive_flux<MK>(P, Wc, dWc, n, TT, fvn);
        const bool ip = TT > -1 || is_ip_face<MK>(TT);
#pragma unroll
        for (int m = 0; m < NE; m++) {
            g[m] += fvn[m];
            if (ip) g[m] -= G.cIP * (WL[m] - WR[m]);
        }
    }
    if constexpr (AV) {
        if (TT > -1) av_face_flux<NE>(G, av, e, TT, n, dL, WL, WR, dWL, dWR, g);
    }
}

// volume fluxes: convective in x and y, minus the viscous ones, minus the damping ones
// (DGpure.residuum's order)
template <int MK, int NE, bool VISC, bool AV, class T>
__device__ inline void quad_volume_flux(const ModelParams& P, const DGParams& G, const double* av, int e, T* Wq, T* dWq, T* fx,
                                        T* fy) {
    const double ex[2] = {1.0, 0.0}, ey[2] = {0.0, 1.0};
    convective_flux<MK>(P, Wq, ex, fx);
    convective_flux<MK>(P, Wq, ey, fy);
    if constexpr (VISC) {
        T gx[NE], gy[NE];
        diffusive_flux<MK>(P, Wq, dWq, ex, gx);
        diffusive_flux<MK>(P, Wq, dWq, ey, gy);
#pragma unroll
        for (int m = 0; m < NE; m++) { fx[m] -= gx[m]; fy[m] -= gy[m]; }
    }
    if constexpr (AV) av_volume_flux<NE>(G, av, e, dWq, fx, fy);
}

// state and physical gradient at a point from coefficients c(m*NB + j)
template <int NB, int NE, class T, class Acc>
__device__ inline void quad_interp(Acc c, int el, const double* phi, const double* gr, T* Wp, T* dWp) {
#pragma unroll
    for (int m = 0; m < NE; m++) {
        T s = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const T a = c(el, m * NB + j);
            s += a * phi[j];
            sx += a * gr[2 * j];
            sy += a * gr[2 * j + 1];
        }
        Wp[m] = s;
        dWp[m] = sx;
        dWp[NE + m] = sy;
    }
}

// Residual of element e: sink.face(k, p, phiL, phiR, jw, g) per face point (phiR null on boundary
// faces), sink.face_end(k), then sink.vol(jw, gradients, fx, fy) per volume point (order > 1).
template <int MK, int NB, int NQV, int NQF, bool VISC, bool AV, class T, class Own, class Nb, class Sink>
__device__ void quad_residual(const ModelParams& P, const DGParams& G, const QuadTables& tb, const QuadMesh& mq, int e, Own own,
                              Nb nbw, Sink& sk) {
    constexpr int NE = neq_of<MK>();
    constexpr int NQT = NQV + 4 * NQF;
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        const int TT = mq.nbr[e * 4 + k];
        const double* fg = mq.fgeo + ((size_t)e * 4 + k) * 4;
        const double n[2] = {fg[0], fg[1]};
        const double jf = fg[2], dL = fg[3];
#pragma unroll 1
        for (int p = 0; p < NQF; p++) {
            const double* phiL = tb.fb + (k * NQF + p) * NB;
            const double* gL = mq.qf + (((size_t)e * 4 + k) * NQF + p) * 2 * NB;
            T WL[NE], dWL[2 * NE], WR[NE], dWR[2 * NE];
            quad_interp<NB, NE, T>(own, e, phiL, gL, WL, dWL);
            const double* phiR = nullptr;
            if (TT > -1) {
                phiR = mq.qn + (((size_t)e * 4 + k) * NQF + p) * 3 * NB;
                quad_interp<NB, NE, T>(nbw, TT, phiR, phiR + NB, WR, dWR);
            }
            const double* X = mq.xq + ((size_t)e * NQT + NQV + k * NQF + p) * 2;
            T g[NE];
            quad_face_flux<MK, NE, VISC, AV>(P, G, mq.av, e, TT, n, dL, X, WL, dWL, WR, dWR, g);
            sk.face(k, p, phiL, phiR, jf * tb.fw[p], g);
        }
        sk.face_end(k);
    }
    if constexpr (NB > 1) {
#pragma unroll 1
        for (int q = 0; q < NQV; q++) {
            const double* qv = mq.qv + ((size_t)e * NQV + q) * (1 + 2 * NB);
            T Wq[NE], dWq[2 * NE], fx[NE], fy[NE];
            quad_interp<NB, NE, T>(own, e, tb.vb + q * NB, qv + 1, Wq, dWq);
            quad_volume_flux<MK, NE, VISC, AV>(P, G, mq.av, e, Wq, dWq, fx, fy);
            sk.vol(qv[0], qv + 1, fx, fy);
        }
    }
}

// K accumulated in value (residual) or tangent (J v, Jacobian columns) form
template <int NB, int NE>
struct QValueSink {
    double K[NB * NE];
    __device__ void face(int, int, const double* phiL, const double*, double jw, const double* g) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * g[m];
    }
    __device__ void face_end(int) {}
    __device__ void vol(double jw, const double* gr, const double* fx, const double* fy) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += jw * (fx[m] * gr[2 * j] + fy[m] * gr[2 * j + 1]);
    }
};
template <int NB, int NE>
struct QTangentSink {
    double K[NB * NE];
    __device__ void face(int, int, const double* phiL, const double*, double jw, const dual* g) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * g[m].d;
    }
    __device__ void face_end(int) {}
    __device__ void vol(double jw, const double* gr, const dual* fx, const dual* fy) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += jw * (fx[m].d * gr[2 * j] + fy[m].d * gr[2 * j + 1]);
    }
};

// R, RHS = R - M (a1 W + a2 Wo + a3 Wo2), per-element R_m^T M^-1 R_m and RHS_m^T M^-1 RHS_m
template <int MK, int NB, int NQV, int NQF, bool VISC, bool AV>
__global__ void __launch_bounds__(128) k_residual_quad(ModelParams P, DGParams G, QuadTables tb, QuadMesh mq, const double* __restrict__ W,
                                                       const double* __restrict__ Wo, const double* __restrict__ Wo2, double a1,
                                                       double a2, double a3, double* __restrict__ R, double* __restrict__ RHS,
                                                       double* __restrict__ enorm) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mq.nElem) return;
    QValueSink<NB, NE> sk;
#pragma unroll
    for (int i = 0; i < B; i++) sk.K[i] = 0.0;
    quad_residual<MK, NB, NQV, NQF, VISC, AV, double>(P, G, tb, mq, e, QPlain{W, B}, QPlain{W, B}, sk);
    const double* M = mq.mass + (size_t)e * NB * NB;
    double H[B];
#pragma unroll
    for (int m = 0; m < NE; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const size_t o = (size_t)e * B + m * NB + i;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const size_t oj = (size_t)e * B + m * NB + j;
                const double mij = M[i * NB + j];
                s += mij * a1 * W[oj] + mij * a2 * Wo[oj] + mij * a3 * Wo2[oj];
            }
            R[o] = sk.K[m * NB + i];
            H[m * NB + i] = sk.K[m * NB + i] - s;
            RHS[o] = H[m * NB + i];
        }
    if (!enorm) return;
    const double* Mi = mq.minv + (size_t)e * NB * NB;
#pragma unroll
    for (int which = 0; which < 2; which++)
#pragma unroll
        for (int m = 0; m < NE; m++) {
            const double* v = (which ? H : sk.K) + m * NB;
            double sum = 0.0;
#pragma unroll
            for (int a = 0; a < NB; a++) {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < NB; b++) t += Mi[a * NB + b] * v[b];
                sum += v[a] * t;
            }
            enorm[(size_t)e * 2 * NE + which * NE + m] = sum;
        }
}

// Thread (e, dir).  mode 0: column dir of the diagonal block (-dK/dW + a1 M) and of the
// neighbour blocks (dKR/dW, the neighbour-side integrand of e's face points); mode 1:
// out[e*B+dir] = ((a1 M - dR/dW)^T v)[e*B+dir].
template <int NB, int NE, typename BT>
struct QColumnSink : QTangentSink<NB, NE> {
    static constexpr int B = NB * NE;
    const QuadMesh* mq;
    BT* blocks;
    float* fblocks;
    const double* v;
    int e, dir, mode;
    double dot;
    __device__ void face(int k, int p, const double* phiL, const double* phiR, double jw, const dual* g) {
        QTangentSink<NB, NE>::face(k, p, phiL, phiR, jw, g);
        if (!phiR) return;
        if (mode == 0) {
            // the thread owns column dir of the neighbour block: the first point stores, the others add
            const size_t base = (size_t)mq->blkn[e * 4 + k] * B * B + (size_t)dir * B;
#pragma unroll
            for (int m = 0; m < NE; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const double a = (jw * phiR[j]) * g[m].d;
                    BT* c = blocks + base + m * NB + j;
                    *c = p == 0 ? (BT)a : (BT)(*c + a);
                }
        } else {
            const double* vn = v + (size_t)mq->nbr[e * 4 + k] * B;
#pragma unroll
            for (int m = 0; m < NE; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) dot += ((jw * phiR[j]) * g[m].d) * vn[m * NB + j];
        }
    }
    __device__ void face_end(int k) {
        if (mode != 0 || !fblocks) return;
        const int blk = mq->blkn[e * 4 + k];
        if (blk < 0) return;
        const BT* c = blocks + (size_t)blk * B * B + (size_t)dir * B;
        float* f = fblocks + (size_t)blk * B * B + (size_t)dir * B;
#pragma unroll
        for (int i = 0; i < B; i++) f[i] = (float)c[i];
    }
};

template <int MK, int NB, int NQV, int NQF, bool VISC, bool AV, typename BT>
__global__ void __launch_bounds__(64) k_jacobian_quad(ModelParams P, DGParams G, QuadTables tb, QuadMesh mq, const double* __restrict__ W,
                                                      double a1, BT* __restrict__ blocks, float* __restrict__ fblocks,
                                                      const double* __restrict__ v, double* __restrict__ out, int mode) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = (int)(t / B), dir = (int)(t % B);
    if (e >= mq.nElem) return;
    QColumnSink<NB, NE, BT> sk;
#pragma unroll
    for (int i = 0; i < B; i++) sk.K[i] = 0.0;
    sk.mq = &mq;
    sk.blocks = blocks;
    sk.fblocks = fblocks;
    sk.v = v;
    sk.e = e;
    sk.dir = dir;
    sk.mode = mode;
    sk.dot = 0.0;
    quad_residual<MK, NB, NQV, NQF, VISC, AV, dual>(P, G, tb, mq, e, QColumn{W, B, dir}, QConst{W, B}, sk);
    const double* M = mq.mass + (size_t)e * NB * NB;
    const int md = dir / NB, jd = dir % NB;
    if (mode == 0) {
        const size_t base = (size_t)mq.blkd[e] * B * B + (size_t)dir * B;
#pragma unroll
        for (int i = 0; i < B; i++) {
            double a = -sk.K[i];
            if (i / NB == md) a += a1 * M[jd * NB + (i % NB)];
            blocks[base + i] = (BT)a;
            if (fblocks) fblocks[base + i] = (float)a;
        }
    } else {
        const double* ve = v + (size_t)e * B;
        double s = sk.dot;
#pragma unroll
        for (int i = 0; i < B; i++) {
            double a = -sk.K[i];
            if (i / NB == md) a += a1 * M[jd * NB + (i % NB)];
            s += a * ve[i];
        }
        out[(size_t)e * B + dir] = s;
    }
}

// out = (a1 M - dR/dW) v by one forward tangent pass per element
template <int MK, int NB, int NQV, int NQF, bool VISC, bool AV>
__global__ void __launch_bounds__(64) k_jvp_quad(ModelParams P, DGParams G, QuadTables tb, QuadMesh mq, const double* __restrict__ W,
                                                 const double* __restrict__ v, double a1, double* __restrict__ out) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mq.nElem) return;
    QTangentSink<NB, NE> sk;
#pragma unroll
    for (int i = 0; i < B; i++) sk.K[i] = 0.0;
    quad_residual<MK, NB, NQV, NQF, VISC, AV, dual>(P, G, tb, mq, e, QTangent{W, v, B}, QTangent{W, v, B}, sk);
    const double* M = mq.mass + (size_t)e * NB * NB;
#pragma unroll
    for (int m = 0; m < NE; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) s += M[i * NB + j] * v[(size_t)e * B + m * NB + j];
            out[(size_t)e * B + m * NB + i] = a1 * s - sk.K[m * NB + i];
        }
}

// elemSize / lambda_max(mean W) per element (State.getTimeStep, Aerodynamics.maxEigenvalue)
template <int MK, int NB>
__global__ void k_size_over_lambda_quad(ModelParams P, QuadMesh mq, const double* __restrict__ W, const double* __restrict__ esize,
                                        double* __restrict__ out) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mq.nElem) return;
    const double* is = mq.is + (size_t)e * NB;
    double wa[NE];
#pragma unroll
    for (int m = 0; m < NE; m++) {
        wa[m] = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) wa[m] = wa[m] + W[(size_t)e * B + m * NB + j] * is[j];
    }
    const double lam = max_eigenvalue<MK>(P, wa);
    out[e] = lam > 0 ? esize[e] / lam : __builtin_huge_val();
}

// DGpure.limiter (FlowPro/src/main/java/flowpro/core/element/DGpure.java:358-400) on a
// quadrilateral: eps = lambda(mean W) elemSize / k g_shock, the inner damping with its decaying
// memory.  The bilinear map is not affine, so the truncation matrix is the element's own
// (trunc: [nElem][NB][NB], DGpure.computeOrderTruncationMatrix with the physical Taylor basis)
// and the sensor integrals take the element's |det J| w per point.  The Aerodynamics models give
// every equation the density sensor.
template <int MK, int NB, int NQV>
__global__ void k_av_limiter_quad(ModelParams P, QuadTables tb, QuadMesh mq, AvParams A, const double* __restrict__ W,
                                  const double* __restrict__ esize, const double* __restrict__ trunc, double* __restrict__ av,
                                  double* __restrict__ old) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mq.nElem) return;
    double* a = av + (size_t)e * (1 + NE);
    a[0] = 0.0;
    if (A.order <= 1) return;
    const double* is = mq.is + (size_t)e * NB;
    double w[NB], wa[NE];
#pragma unroll
    for (int m = 0; m < NE; m++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) s = s + W[(size_t)e * B + m * NB + j] * is[j];
        wa[m] = s;
    }
    const double mean0 = wa[0];
#pragma unroll
    for (int j = 0; j < NB; j++) w[j] = W[(size_t)e * B + j];
    const double lam = max_eigenvalue<MK>(P, wa);
    const double scale = lam * esize[e] / A.order;
    const double* qv = mq.qv + (size_t)e * NQV * (1 + 2 * NB);
    auto jw = [&](int q) { return qv[(size_t)q * (1 + 2 * NB)]; };
    auto phi = [&](int q) { return tb.vb + q * NB; };
    const double* tr = trunc ? trunc + (size_t)e * NB * NB : nullptr;
    if (A.dampTol > 0) a[0] = scale * shock_sensor_at<NB>(phi, jw, NQV, w, mean0, tr, A.order, A.dampTol);
    if (A.dampInnerTol > 0) {
        const double g = shock_sensor_at<NB>(phi, jw, NQV, w, mean0, tr, A.order, A.dampInnerTol);
#pragma unroll
        for (int m = 0; m < NE; m++) {
            double& o = old[(size_t)e * NE + m];
            const double ind = fmax(g, o);
            o = 0.1 * ind;
            a[1 + m] = A.nCoef > 0 ? (A.nCoef == 1 ? A.coef[0] : A.coef[m]) * ind : scale * ind;
        }
    }
}

// Element.positivityLimiter (Element.java:385-476) over the volume and the four faces' points
template <int MK, int NB, int NQV, int NQF>
__global__ void k_positivity_quad(QuadTables tb, QuadMesh mq, double* __restrict__ W, int* __restrict__ count) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mq.nElem || NB == 1) return;
    const double* is = mq.is + (size_t)e * NB;
    double w[B], wa[NE];
#pragma unroll
    for (int i = 0; i < B; i++) w[i] = W[(size_t)e * B + i];
#pragma unroll
    for (int m = 0; m < NE; m++) {
        wa[m] = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) wa[m] = wa[m] + w[m * NB + j] * is[j];
    }
    if (!admissible<NE>(wa)) return;
    double theta = 1.0;
    auto at = [&](const double* phi) {
        double wx[NE];
#pragma unroll
        for (int m = 0; m < NE; m++) {
            wx[m] = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) wx[m] += w[m * NB + j] * phi[j];
        }
        theta = fmin(theta, admissible_theta<NE>(wa, wx));
    };
    for (int q = 0; q < NQV; q++) at(tb.vb + q * NB);
    for (int k = 0; k < 4; k++)
        for (int p = 0; p < NQF; p++) at(tb.fb + (k * NQF + p) * NB);
    if (theta >= 1) return;
#pragma unroll
    for (int m = 0; m < NE; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) W[(size_t)e * B + m * NB + i] = wa[m] + theta * (w[m * NB + i] - wa[m]);
    atomicAdd(count, 1);
}

// ------------------------------------------------------------------ dispatch
// (model, NB, NQV, NQF, viscous) for element orders 1-5: NB = NQV = k^2, NQF = k
// DragCoefficient (which 0) / LiftCoefficient (which 1) contribution of every quadrilateral:
// the pressure over its wall faces times the normal component, and the exact gradient
template <int NB, int NQF>
__global__ void k_functional_quad(QuadTables tb, QuadMesh mq, const double* __restrict__ W, double kappa, int which,
                                  double* __restrict__ ival, double* __restrict__ grad) {
    constexpr int NEQ = 4, B = NB * NEQ;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mq.nElem) return;
    double g[B];
#pragma unroll
    for (int i = 0; i < B; i++) g[i] = 0.0;
    double I = 0.0;
    for (int k = 0; k < 4; k++) {
        if (mq.nbr[e * 4 + k] != BT_WALL) continue;
        const double* fg = mq.fgeo + ((size_t)e * 4 + k) * 4;
        const double nk = fg[which];
        for (int p = 0; p < NQF; p++) {
            const double* phi = tb.fb + (k * NQF + p) * NB;
            double w[NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) s += W[(size_t)e * B + m * NB + j] * phi[j];
                w[m] = s;
            }
            const double mom2 = w[1] * w[1] + w[2] * w[2];
            const double pr = (kappa - 1) * (w[3] - mom2 / (2 * w[0]));
            const double jw = fg[2] * tb.fw[p];
            I += jw * pr * nk;
            const double dp[NEQ] = {(kappa - 1) * mom2 / (2 * w[0] * w[0]), -(kappa - 1) * w[1] / w[0], -(kappa - 1) * w[2] / w[0],
                                    kappa - 1};
#pragma unroll
            for (int m = 0; m < NEQ; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) g[m * NB + j] += jw * nk * dp[m] * phi[j];
        }
    }
    ival[e] = I;
    if (grad) {
#pragma unroll
        for (int i = 0; i < B; i++) grad[(size_t)e * B + i] = g[i];
    }
}

// The file is compiled twice so that its halves build in parallel: as itself (FP_QUAD_PART 0: the
// kernels without artificial viscosity, and the entry points the host calls, which hand the damped
// cases on) and through fp_quad_av.hip (FP_QUAD_PART 1: the kernels with it, entry points named _av).
#ifndef FP_QUAD_PART
#define FP_QUAD_PART 0
#endif
// (model, NB, NQV, NQF, viscous, artificial viscosity)
#define FP_QUAD_SHAPES(X, A)                    \
    X(MODEL_NAVIER_STOKES, 1, 1, 1, false, A)   \
    X(MODEL_NAVIER_STOKES, 4, 4, 2, false, A)   \
    X(MODEL_NAVIER_STOKES, 9, 9, 3, false, A)   \
    X(MODEL_NAVIER_STOKES, 16, 16, 4, false, A) \
    X(MODEL_NAVIER_STOKES, 25, 25, 5, false, A) \
    X(MODEL_NAVIER_STOKES, 1, 1, 1, true, A)    \
    X(MODEL_NAVIER_STOKES, 4, 4, 2, true, A)    \
    X(MODEL_NAVIER_STOKES, 9, 9, 3, true, A)    \
    X(MODEL_NAVIER_STOKES, 16, 16, 4, true, A)  \
    X(MODEL_NAVIER_STOKES, 25, 25, 5, true, A)  \
    X(MODEL_ADIABATIC, 1, 1, 1, false, A)       \
    X(MODEL_ADIABATIC, 4, 4, 2, false, A)       \
    X(MODEL_ADIABATIC, 9, 9, 3, false, A)       \
    X(MODEL_ADIABATIC, 16, 16, 4, false, A)     \
    X(MODEL_ADIABATIC, 25, 25, 5, false, A)
#if FP_QUAD_PART == 0
#define FP_QUAD_CASES(X) FP_QUAD_SHAPES(X, false)
#define FP_QFN(name) name
#define FP_QAV(call) \
    if (L.av) return call;
bool quad_supported_av(const ElemLaunch& L);
hipError_t launch_residual_quad_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const QuadTables& tb,
                                   const QuadMesh& mq, const double* W, const double* Wo, const double* Wo2, double a1, double a2,
                                   double a3, double* R, double* RHS, double* enorm);
hipError_t launch_jacobian_quad_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const QuadTables& tb,
                                   const QuadMesh& mq, const double* W, double a1, double* blocks, float* fblocks, const double* v,
                                   double* out, int mode);
hipError_t launch_jvp_quad_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const QuadTables& tb,
                              const QuadMesh& mq, const double* W, const double* v, double a1, double* out);
hipError_t launch_av_limiter_quad_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const QuadTables& tb,
                                     const QuadMesh& mq, const AvParams& Ap, const double* W, const double* esize,
                                     const double* trunc, double* av, double* old);
#else
#define FP_QUAD_CASES(X) FP_QUAD_SHAPES(X, true)
#define FP_QFN(name) name##_av
#define FP_QAV(call)
#endif
#define FP_QMATCH(MK, NB, NQV, NQF, V, A) \
    (L.model == MK && L.nb == NB && L.nqv == NQV && L.nqf == NQF && (L.visc != 0) == V && (L.av != 0) == A)
// the kernels that do not depend on the damping: every shape, damped or not
#define FP_QSHAPE(MK, NB, NQV, NQF, V) (L.model == MK && L.nb == NB && L.nqv == NQV && L.nqf == NQF && (L.visc != 0) == V)

bool FP_QFN(quad_supported)(const ElemLaunch& L) {
    FP_QAV(quad_supported_av(L))
#define X(MK, NB, NQV, NQF, V, A) \
    if (FP_QMATCH(MK, NB, NQV, NQF, V, A)) return true;
    FP_QUAD_CASES(X)
#undef X
    return false;
}

hipError_t FP_QFN(launch_residual_quad)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G,
                                        const QuadTables& tb, const QuadMesh& mq, const double* W, const double* Wo,
                                        const double* Wo2, double a1, double a2, double a3, double* R, double* RHS, double* enorm) {
    FP_QAV(launch_residual_quad_av(L, s, P, G, tb, mq, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm))
    dim3 blk(128), grd((mq.nElem + 127) / 128);
#define X(MK, NB, NQV, NQF, V, A)                                                                                       \
    if (FP_QMATCH(MK, NB, NQV, NQF, V, A)) {                                                                            \
        hipLaunchKernelGGL((k_residual_quad<MK, NB, NQV, NQF, V, A>), grd, blk, 0, s, P, G, tb, mq, W, Wo, Wo2, a1, a2, a3, R, \
                           RHS, enorm);                                                                              \
        return hipGetLastError();                                                                                    \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t FP_QFN(launch_jacobian_quad)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G,
                                        const QuadTables& tb, const QuadMesh& mq, const double* W, double a1, double* blocks,
                                        float* fblocks, const double* v, double* out, int mode) {
    FP_QAV(launch_jacobian_quad_av(L, s, P, G, tb, mq, W, a1, blocks, fblocks, v, out, mode))
    const long long threads = (long long)mq.nElem * L.nb * L.neq;
    dim3 blk(64), grd((unsigned)((threads + 63) / 64));
#define X(MK, NB, NQV, NQF, V, A)                                                                                          \
    if (FP_QMATCH(MK, NB, NQV, NQF, V, A)) {                                                                               \
        hipLaunchKernelGGL((k_jacobian_quad<MK, NB, NQV, NQF, V, A, double>), grd, blk, 0, s, P, G, tb, mq, W, a1, blocks, fblocks, v, \
                           out, mode);                                                                                  \
        return hipGetLastError();                                                                                       \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t FP_QFN(launch_jvp_quad)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const QuadTables& tb,
                                   const QuadMesh& mq, const double* W, const double* v, double a1, double* out) {
    FP_QAV(launch_jvp_quad_av(L, s, P, G, tb, mq, W, v, a1, out))
    dim3 blk(64), grd((mq.nElem + 63) / 64);
#define X(MK, NB, NQV, NQF, V, A)                                                                                 \
    if (FP_QMATCH(MK, NB, NQV, NQF, V, A)) {                                                                      \
        hipLaunchKernelGGL((k_jvp_quad<MK, NB, NQV, NQF, V, A>), grd, blk, 0, s, P, G, tb, mq, W, v, a1, out);    \
        return hipGetLastError();                                                                              \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

#if FP_QUAD_PART == 0
hipError_t launch_size_over_lambda_quad(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const QuadMesh& mq, const double* W,
                                        const double* esize, double* out) {
    dim3 blk(256), grd((mq.nElem + 255) / 256);
#define X(MK, NB, NQV, NQF, V, A)                                                                          \
    if (FP_QSHAPE(MK, NB, NQV, NQF, V)) {                                                               \
        hipLaunchKernelGGL((k_size_over_lambda_quad<MK, NB>), grd, blk, 0, s, P, mq, W, esize, out);    \
        return hipGetLastError();                                                                       \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}
#endif

hipError_t FP_QFN(launch_av_limiter_quad)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const QuadTables& tb,
                                          const QuadMesh& mq, const AvParams& Ap, const double* W, const double* esize,
                                          const double* trunc, double* av, double* old) {
    FP_QAV(launch_av_limiter_quad_av(L, s, P, tb, mq, Ap, W, esize, trunc, av, old))
    dim3 blk(128), grd((mq.nElem + 127) / 128);
#define X(MK, NB, NQV, NQF, V, A)                                                                                     \
    if (A && FP_QMATCH(MK, NB, NQV, NQF, V, A)) {                                                                     \
        hipLaunchKernelGGL((k_av_limiter_quad<MK, NB, NQV>), grd, blk, 0, s, P, tb, mq, Ap, W, esize, trunc, av, old); \
        return hipGetLastError();                                                                                     \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

#if FP_QUAD_PART == 0
hipError_t launch_positivity_quad(const ElemLaunch& L, hipStream_t s, const QuadTables& tb, const QuadMesh& mq, double* W, int* count) {
    dim3 blk(128), grd((mq.nElem + 127) / 128);
#define X(MK, NB, NQV, NQF, V, A)                                                                          \
    if (FP_QSHAPE(MK, NB, NQV, NQF, V)) {                                                               \
        hipLaunchKernelGGL((k_positivity_quad<MK, NB, NQV, NQF>), grd, blk, 0, s, tb, mq, W, count);    \
        return hipGetLastError();                                                                       \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_functional_quad(const ElemLaunch& L, hipStream_t s, const QuadTables& tb, const QuadMesh& mq, const double* W,
                                  double kappa, int which, double* ival, double* grad) {
    if (L.model != MODEL_NAVIER_STOKES) return hipErrorInvalidValue;
    dim3 blk(128), grd((mq.nElem + 127) / 128);
#define X(MK, NB, NQV, NQF, V, A)                                                                                   \
    if (MK == MODEL_NAVIER_STOKES && FP_QSHAPE(MK, NB, NQV, NQF, V)) {                                            \
        hipLaunchKernelGGL((k_functional_quad<NB, NQF>), grd, blk, 0, s, tb, mq, W, kappa, which, ival, grad);   \
        return hipGetLastError();                                                                                 \
    }
    FP_QUAD_CASES(X)
#undef X
    return hipErrorInvalidValue;
}
#endif

}  // namespace fp
