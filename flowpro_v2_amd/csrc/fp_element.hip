// Element kernels of the implicit DG step on gfx950:
//   k_residual  - spatial residual R(W), BDF right-hand side, per-element norms
//   k_jacobian  - exact element Jacobian columns written into the BSR matrix
//                 (-dR/dW + a1 M, the reference's convention), or the matrix-free
//                 transposed product (J^T v) when a vector is given
//   k_jvp       - matrix-free product (a1 M - dR/dW) v
// All three run the same DG residual routine `dg_residual`, a restatement of
// DGpure.residuum / residuumWall (FlowPro/src/main/java/flowpro/core/element/
// DGpure.java:108-355) over a scalar type: double, or a forward-mode dual number
// whose tangent seeds one DOF (Jacobian column) or a direction vector.
#include <hip/hip_runtime.h>

#include "fp_common.hpp"

namespace fp {

// physical basis gradient from the reference one: d_x = iT^T d_xi
// (Transformation.transformBasis, FlowPro/src/main/java/flowpro/core/transformation/Transformation.java:146-163)
__device__ inline void phys_grad(const double* iT, double dxi, double deta, double& gx, double& gy) {
    gx = iT[0] * dxi + iT[2] * deta;
    gy = iT[1] * dxi + iT[3] * deta;
}

// Integrand policy: what is done with the face integrand seen from the neighbour (KR).
struct NoKR {
    __device__ void face_begin(int) {}
    template <class T, int NB, int NEQ>
    __device__ void add(int, const double*, double, const T*) {}
    __device__ void face_end(int) {}
};

// DG residual of element e with own coefficients W (type T) and neighbour
// coefficients from `nbw(elem, idx)` (type T).  K receives the element residual,
// `kr` the neighbour-side face integrands of interior faces.
template <int MK, int NB, int NEQ, int NQV, int NQF, bool VISC, class T, class NbW, class KR>
__device__ void dg_residual(const ModelParams& P, const DGParams& G, const Tables& tb, const Mesh& ms, int e,
                            const T* W, NbW nbw, T* K, KR& kr) {
    const double* ge = ms.geo + (size_t)e * GEO;
    const double iT[4] = {ge[0], ge[1], ge[2], ge[3]};
    const double jac = ge[4];
    // ---------------- faces (DGpure.residuumWall) ----------------
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        const int TT = ms.nbr[e * 3 + k];
        const double* fg = ms.fgeo + ((size_t)e * 3 + k) * 4;
        const double n[2] = {fg[0], fg[1]};
        const double jacF = fg[2], dL = fg[3];
        int kr_face = 0;
        double iTR[4] = {0, 0, 0, 0};
        if (TT > -1) {
            kr_face = ms.nbrf[e * 3 + k];
            const double* gn = ms.geo + (size_t)TT * GEO;
            iTR[0] = gn[0]; iTR[1] = gn[1]; iTR[2] = gn[2]; iTR[3] = gn[3];
        }
        kr.face_begin(k);
#pragma unroll 1
        for (int p = 0; p < NQF; p++) {
            const double* phiL = tb.fb + (k * NQF + p) * NB;
            const double* dphiL = tb.fd + (k * NQF + p) * NB * 2;
            T WL[NEQ], WR[NEQ], dWL[2 * NEQ], dWR[2 * NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                T s = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    double gx, gy;
                    phys_grad(iT, dphiL[2 * j], dphiL[2 * j + 1], gx, gy);
                    T w = W[m * NB + j];
                    s += w * phiL[j];
                    sx += w * gx;
                    sy += w * gy;
                }
                WL[m] = s;
                dWL[m] = sx;
                dWL[NEQ + m] = sy;
            }
            const int pr = NQF - 1 - p;   // the neighbour traverses the shared face backwards
            if (TT > -1) {
                const double* phiR = tb.fb + (kr_face * NQF + pr) * NB;
                const double* dphiR = tb.fd + (kr_face * NQF + pr) * NB * 2;
#pragma unroll
                for (int m = 0; m < NEQ; m++) {
                    T s = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        double gx, gy;
                        phys_grad(iTR, dphiR[2 * j], dphiR[2 * j + 1], gx, gy);
                        T w = nbw(TT, m * NB + j);
                        s += w * phiR[j];
                        sx += w * gx;
                        sy += w * gy;
                    }
                    WR[m] = s;
                    dWR[m] = sx;
                    dWR[NEQ + m] = sy;
                }
            } else {
                double X[2];
                const double* xi = tb.fx + (k * NQF + p) * 2;
                X[0] = ge[5] * xi[0] + ge[6] * xi[1] + ge[7];
                X[1] = ge[8] * xi[0] + ge[9] * xi[1] + ge[10];
                boundary_value<MK>(P, WL, n, TT, X, WR);
#pragma unroll
                for (int m = 0; m < NEQ; m++) { dWR[m] = dWL[m]; dWR[NEQ + m] = dWL[NEQ + m]; }
            }
            T fn[NEQ];
            numerical_convective_flux<MK>(P, WL, WR, n, TT, fn);
            T g[NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) g[m] = -fn[m];
            if constexpr (VISC) {
                T Wc[NEQ], dWc[2 * NEQ], fvn[NEQ];
#pragma unroll
                for (int m = 0; m < NEQ; m++) {
                    Wc[m] = TT > -1 ? (WL[m] + WR[m]) / 2.0 : WR[m];
                    T jump = G.beta0 * (WR[m] - WL[m]) / dL;
                    dWc[m] = (dWL[m] + dWR[m]) / 2.0 + jump * n[0];
                    dWc[NEQ + m] = (dWL[NEQ + m] + dWR[NEQ + m]) / 2.0 + jump * n[1];
                }
                double Df = 0.0;   // k-omega SST: the wall distance along the face (Element.java:299-312)
                if constexpr (MK == MODEL_KOMEGA) {
                    const double* wd = ms.wdist + (size_t)e * 3;
                    Df = tb.fint[2 * p] * wd[k] + tb.fint[2 * p + 1] * wd[k == 2 ? 0 : k + 1];
                }
                numerical_diffusive_flux<MK>(P, Wc, dWc, n, TT, fvn, Df);
                const bool ip = TT > -1 || is_ip_face<MK>(TT);
#pragma unroll
                for (int m = 0; m < NEQ; m++) {
                    g[m] += fvn[m];
                    if (ip) g[m] -= G.cIP * (WL[m] - WR[m]);
                }
            }
            if (G.av && TT > -1) av_face_flux<NEQ>(G, ms, e, TT, n, dL, WL, WR, dWL, dWR, g);
            const double jw = jacF * tb.fw[p];
#pragma unroll
            for (int m = 0; m < NEQ; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * g[m];
            if (TT > -1) kr.template add<T, NB, NEQ>(k, tb.fb + (kr_face * NQF + pr) * NB, jw, g);
        }
        kr.face_end(k);
    }
    // ---------------- volume (DGpure.residuum) ----------------
    if constexpr (NB > 1 || MK == MODEL_SCALAR || has_rans_source<MK>()) {
#pragma unroll 1
        for (int q = 0; q < NQV; q++) {
            const double* phi = tb.vb + q * NB;
            const double* dphi = tb.vd + q * NB * 2;
            T Wq[NEQ], dWq[2 * NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                T s = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    double gx, gy;
                    phys_grad(iT, dphi[2 * j], dphi[2 * j + 1], gx, gy);
                    T w = W[m * NB + j];
                    s += w * phi[j];
                    sx += w * gx;
                    sy += w * gy;
                }
                Wq[m] = s;
                dWq[m] = sx;
                dWq[NEQ + m] = sy;
            }
            const double jw = jac * tb.vw[q];
            double Dq = 0.0;   // wall distance at the point (Element.java:263-275)
            if constexpr (has_rans_source<MK>()) {
                const double* xi = tb.vx + q * 2;
                const double* wd = ms.wdist + (size_t)e * 3;
                Dq = (1.0 - xi[0] - xi[1]) * wd[0] + xi[0] * wd[1] + xi[1] * wd[2];
            }
            T fx[NEQ], fy[NEQ];
            if constexpr (NB > 1) {
                const double ex[2] = {1.0, 0.0}, ey[2] = {0.0, 1.0};
                convective_flux<MK>(P, Wq, ex, fx);
                convective_flux<MK>(P, Wq, ey, fy);
                if constexpr (VISC) {
                    T gx[NEQ], gy[NEQ];
                    diffusive_flux<MK>(P, Wq, dWq, ex, gx, Dq);
                    diffusive_flux<MK>(P, Wq, dWq, ey, gy, Dq);
#pragma unroll
                    for (int m = 0; m < NEQ; m++) { fx[m] -= gx[m]; fy[m] -= gy[m]; }
                }
                if (G.av) av_volume_flux<NEQ>(G, ms, e, dWq, fx, fy);
            } else {
#pragma unroll
                for (int m = 0; m < NEQ; m++) { fx[m] = 0.0; fy[m] = 0.0; }
            }
            T src[NEQ];
            if constexpr (MK == MODEL_SCALAR) {
                const double* xi = tb.vx + q * 2;
                double X[2] = {ge[5] * xi[0] + ge[6] * xi[1] + ge[7], ge[8] * xi[0] + ge[9] * xi[1] + ge[10]};
                source_term(P, X, src);
            }
            if constexpr (MK == MODEL_SA) sa_source(P, Wq, dWq, Dq, src);
            if constexpr (MK == MODEL_KOMEGA) kw_source(P, Wq, dWq, Dq, src);
#pragma unroll
            for (int j = 0; j < NB; j++) {
                double gx, gy;
                phys_grad(iT, dphi[2 * j], dphi[2 * j + 1], gx, gy);
#pragma unroll
                for (int m = 0; m < NEQ; m++) {
                    T v = fx[m] * gx + fy[m] * gy;
                    if constexpr (MK == MODEL_SCALAR || has_rans_source<MK>()) v += src[m] * phi[j];
                    K[m * NB + j] += jw * v;
                }
            }
        }
    }
}

struct PlainNb {
    const double* W;
    int B;
    __device__ double operator()(int el, int i) const { return W[(size_t)el * B + i]; }
};
struct TangentNb {
    const double* W;
    const double* V;
    int B;
    __device__ dual operator()(int el, int i) const {
        size_t o = (size_t)el * B + i;
        return dual(W[o], V[o]);
    }
};
struct ZeroTangentNb {
    const double* W;
    int B;
    __device__ dual operator()(int el, int i) const { return dual(W[(size_t)el * B + i], 0.0); }
};

// ------------------------------------------------------------------ residual
// R = R(W); RHS = R - M (a1 W + a2 Wo + a3 Wo2) (BDF1.assembleRHS / BDF2.assembleRHS,
// FlowPro/src/main/java/flowpro/core/element/BDF1.java:34-49, BDF2.java:32-51);
// enorm[e][0..NEQ) = R_m^T M^-1 R_m, enorm[e][NEQ..2NEQ) = RHS_m^T M^-1 RHS_m
// (Element.l2NormsSquared, FlowPro/src/main/java/flowpro/core/element/Element.java:357-373).
template <int MK, int NB, int NEQ, int NQV, int NQF, bool VISC>
__global__ void __launch_bounds__(128) k_residual(ModelParams P, DGParams G, Tables tb, Mesh ms, const double* __restrict__ W,
                                                  const double* __restrict__ Wo, const double* __restrict__ Wo2, double a1,
                                                  double a2, double a3, double* __restrict__ R, double* __restrict__ RHS,
                                                  double* __restrict__ enorm) {
    constexpr int B = NB * NEQ;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ms.nElem) return;
    double w[B], K[B];
#pragma unroll
    for (int i = 0; i < B; i++) { w[i] = W[(size_t)e * B + i]; K[i] = 0.0; }
    NoKR kr;
    dg_residual<MK, NB, NEQ, NQV, NQF, VISC, double>(P, G, tb, ms, e, w, PlainNb{W, B}, K, kr);
    const double jac = ms.geo[(size_t)e * GEO + 4];
    double rhs[B];
#pragma unroll
    for (int m = 0; m < NEQ; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                size_t o = (size_t)e * B + m * NB + j;
                double mij = jac * tb.mref[i * NB + j];
                s += mij * a1 * w[m * NB + j] + mij * a2 * Wo[o] + mij * a3 * Wo2[o];
            }
            rhs[m * NB + i] = K[m * NB + i] - s;
        }
#pragma unroll
    for (int i = 0; i < B; i++) {
        R[(size_t)e * B + i] = K[i];
        RHS[(size_t)e * B + i] = rhs[i];
    }
    if (enorm) {
#pragma unroll
        for (int m = 0; m < NEQ; m++) {
            double sr = 0.0, sh = 0.0;
#pragma unroll
            for (int i = 0; i < NB; i++) {
                double tr = 0.0, th = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    double mi = tb.mrefinv[i * NB + j];
                    tr += mi * K[m * NB + j];
                    th += mi * rhs[m * NB + j];
                }
                sr += K[m * NB + i] * tr;
                sh += rhs[m * NB + i] * th;
            }
            enorm[(size_t)e * 2 * NEQ + m] = sr / jac;
            enorm[(size_t)e * 2 * NEQ + NEQ + m] = sh / jac;
        }
    }
}

// ------------------------------------------------------------------ Jacobian / J^T v
// KR policy for the Jacobian: accumulate the tangent of the neighbour-side face
// integrand and write it as column `dir` of block (row neighbour, column e).
template <int NB, int NEQ>
struct KRColumn {
    double acc[NB * NEQ];
    __device__ void face_begin(int) {
#pragma unroll
        for (int i = 0; i < NB * NEQ; i++) acc[i] = 0.0;
    }
    template <class T, int NB_, int NEQ_>
    __device__ void add(int, const double* phiR, double jw, const T* g) {
#pragma unroll
        for (int m = 0; m < NEQ; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[m * NB + j] += (jw * phiR[j]) * tangent(g[m]);
    }
    __device__ void face_end(int) {}
};

// Thread (e, dir): tangent of the residual of element e in the direction of its own
// DOF `dir`.  mode 0: store -dK/dW (+ a1 M) as column dir of the diagonal block and
// dKR/dW as column dir of the neighbour blocks (the reference's ADiag / ANeighs,
// BDF1.java:89-118, in column-major BSR).  mode 1: out[e*B+dir] = (A^T v)[e*B+dir].
template <int MK, int NB, int NEQ, int NQV, int NQF, bool VISC, typename BT>
__global__ void __launch_bounds__(128) k_jacobian(ModelParams P, DGParams G, Tables tb, Mesh ms, const double* __restrict__ W,
                                                  double a1, BT* __restrict__ blocks, const double* __restrict__ v,
                                                  double* __restrict__ out, int mode) {
    constexpr int B = NB * NEQ;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int e = (int)(t / B);
    int dir = (int)(t % B);
    if (e >= ms.nElem) return;
    dual w[B], K[B];
#pragma unroll
    for (int i = 0; i < B; i++) { w[i] = dual(W[(size_t)e * B + i], i == dir ? 1.0 : 0.0); K[i] = dual(0.0, 0.0); }
    // faces are processed one at a time: the neighbour column is emitted at face end
    struct Emit : KRColumn<NB, NEQ> {
        const Mesh* ms; BT* blocks; int e, dir, mode; const double* v; double dot;
        __device__ void face_end(int k) {
            int blk = ms->blkn[e * 3 + k];
            if (blk < 0) return;
            if (mode == 0) {
                BT* col = blocks + (size_t)blk * B * B + (size_t)dir * B;
#pragma unroll
                for (int i = 0; i < B; i++) col[i] = (BT)this->acc[i];
            } else {
                const double* vn = v + (size_t)ms->nbr[e * 3 + k] * B;
#pragma unroll
                for (int i = 0; i < B; i++) dot += this->acc[i] * vn[i];
            }
        }
    } kr;
    kr.ms = &ms; kr.blocks = blocks; kr.e = e; kr.dir = dir; kr.mode = mode; kr.v = v; kr.dot = 0.0;
    dg_residual<MK, NB, NEQ, NQV, NQF, VISC, dual>(P, G, tb, ms, e, w, ZeroTangentNb{W, B}, K, kr);
    const double jac = ms.geo[(size_t)e * GEO + 4];
    const int md = dir / NB, jd = dir % NB;
    if (mode == 0) {
        BT* col = blocks + (size_t)ms.blkd[e] * B * B + (size_t)dir * B;
#pragma unroll
        for (int i = 0; i < B; i++) {
            double a = -K[i].d;
            if (i / NB == md) a += a1 * (jac * tb.mref[jd * NB + (i % NB)]);
            col[i] = (BT)a;
        }
    } else {
        const double* ve = v + (size_t)e * B;
        double s = kr.dot;
#pragma unroll
        for (int i = 0; i < B; i++) {
            double a = -K[i].d;
            if (i / NB == md) a += a1 * (jac * tb.mref[jd * NB + (i % NB)]);
            s += a * ve[i];
        }
        out[(size_t)e * B + dir] = s;
    }
}

// ------------------------------------------------------------------ J v
template <int MK, int NB, int NEQ, int NQV, int NQF, bool VISC>
__global__ void __launch_bounds__(128) k_jvp(ModelParams P, DGParams G, Tables tb, Mesh ms, const double* __restrict__ W,
                                             const double* __restrict__ v, double a1, double* __restrict__ out) {
    constexpr int B = NB * NEQ;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ms.nElem) return;
    dual w[B], K[B];
#pragma unroll
    for (int i = 0; i < B; i++) { w[i] = dual(W[(size_t)e * B + i], v[(size_t)e * B + i]); K[i] = dual(0.0, 0.0); }
    NoKR kr;
    dg_residual<MK, NB, NEQ, NQV, NQF, VISC, dual>(P, G, tb, ms, e, w, TangentNb{W, v, B}, K, kr);
    const double jac = ms.geo[(size_t)e * GEO + 4];
#pragma unroll
    for (int m = 0; m < NEQ; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) s += (jac * tb.mref[i * NB + j]) * v[(size_t)e * B + m * NB + j];
            out[(size_t)e * B + m * NB + i] = a1 * s - K[m * NB + i].d;
        }
}

// ------------------------------------------------------------------ dispatch

#define FP_CASES(X)                          \
    X(MODEL_NAVIER_STOKES, 3, 4, 6, 2, false) \
    X(MODEL_NAVIER_STOKES, 6, 4, 7, 3, false) \
    X(MODEL_NAVIER_STOKES, 10, 4, 16, 4, false) \
    X(MODEL_NAVIER_STOKES, 3, 4, 6, 2, true)  \
    X(MODEL_NAVIER_STOKES, 6, 4, 7, 3, true)  \
    X(MODEL_NAVIER_STOKES, 10, 4, 16, 4, true) \
    X(MODEL_SA, 3, 5, 6, 2, true)             \
    X(MODEL_SA, 6, 5, 7, 3, true)             \
    X(MODEL_KOMEGA, 3, 6, 6, 2, true)         \
    X(MODEL_KOMEGA, 6, 6, 7, 3, true)         \
    X(MODEL_ADIABATIC, 3, 3, 6, 2, false)     \
    X(MODEL_ADIABATIC, 6, 3, 7, 3, false)     \
    X(MODEL_ADIABATIC, 10, 3, 16, 4, false)   \
    X(MODEL_SCALAR, 1, 1, 1, 1, false)        \
    X(MODEL_SCALAR, 3, 1, 6, 2, false)        \
    X(MODEL_SCALAR, 6, 1, 7, 3, false)        \
    X(MODEL_SCALAR, 10, 1, 16, 4, false)      \
    X(MODEL_SCALAR, 3, 1, 6, 2, true)         \
    X(MODEL_SCALAR, 6, 1, 7, 3, true)         \
    X(MODEL_SCALAR, 10, 1, 16, 4, true)

#define FP_MATCH(MK, NB, NEQ, NQV, NQF, V) \
    (L.model == MK && L.nb == NB && L.neq == NEQ && L.nqv == NQV && L.nqf == NQF && (bool)L.visc == V)

bool element_supported(const ElemLaunch& L) {
#define X(MK, NB, NEQ, NQV, NQF, V) \
    if (FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) return true;
    FP_CASES(X)
#undef X
    return false;
}

hipError_t launch_residual(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                           const Mesh& ms, const double* W, const double* Wo, const double* Wo2, double a1, double a2,
                           double a3, double* R, double* RHS, double* enorm) {
    dim3 blk(128), grd((ms.nElem + 127) / 128);
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                                  \
    if (FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                                                         \
        hipLaunchKernelGGL((k_residual<MK, NB, NEQ, NQV, NQF, V>), grd, blk, 0, s, P, G, tb, ms, W, Wo, Wo2, a1, a2, a3, R, \
                           RHS, enorm);                                                                               \
        return hipGetLastError();                                                                                     \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

template <typename BT>
hipError_t launch_jacobian(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                           const Mesh& ms, const double* W, double a1, BT* blocks, const double* v, double* out, int mode) {
    long long nthreads = (long long)ms.nElem * L.nb * L.neq;
    dim3 blk(128), grd((unsigned)((nthreads + 127) / 128));
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                                     \
    if (FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                                                            \
        hipLaunchKernelGGL((k_jacobian<MK, NB, NEQ, NQV, NQF, V, BT>), grd, blk, 0, s, P, G, tb, ms, W, a1, blocks, v, out, \
                           mode);                                                                                        \
        return hipGetLastError();                                                                                        \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}
template hipError_t launch_jacobian<double>(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&,
                                            const Mesh&, const double*, double, double*, const double*, double*, int);

hipError_t launch_jvp(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                      const Mesh& ms, const double* W, const double* v, double a1, double* out) {
    dim3 blk(128), grd((ms.nElem + 127) / 128);
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                         \
    if (FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                                                \
        hipLaunchKernelGGL((k_jvp<MK, NB, NEQ, NQV, NQF, V>), grd, blk, 0, s, P, G, tb, ms, W, v, a1, out); \
        return hipGetLastError();                                                                            \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ time step
// elemSize / lambda_max(mean W) per element (State.getTimeStep, FlowPro/src/main/java/flowpro/core/State.java:177-192;
// Element.calculateMaxEigenvalue, Element.java:510-513; Aerodynamics.maxEigenvalue :323-337)
template <int MK, int NB, int NEQ>
__global__ void k_size_over_lambda(ModelParams P, Tables tb, Mesh ms, const double* __restrict__ W, const double* __restrict__ esize,
                                   const double* __restrict__ isref, double* __restrict__ out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ms.nElem) return;
    constexpr int B = NB * NEQ;
    double wa[NEQ];
#pragma unroll
    for (int m = 0; m < NEQ; m++) wa[m] = 0.0;
#pragma unroll
    for (int m = 0; m < NEQ; m++)
#pragma unroll
        for (int j = 0; j < NB; j++) wa[m] = wa[m] + W[(size_t)e * B + m * NB + j] * isref[j];
    double lam = max_eigenvalue<MK>(P, wa);
    out[e] = lam > 0 ? esize[e] / lam : __builtin_huge_val();
}

// ------------------------------------------------------------------ artificial viscosity
// DGpure.limiter (FlowPro/src/main/java/flowpro/core/element/DGpure.java:358-400) with the
// sensors shock_senzor / shock_senzor_all_eqn (:403-503): the L2 distance of an equation's
// solution from its truncation (projection onto linear functions at order > 2, the mean
// otherwise) relative to its L2 norm, on a log10 scale, through a sine ramp of half-width kap
// around S0 = log10(1 / (k-1)^4).  The Aerodynamics models give every equation the density
// sensor (Aerodynamics.combineShockSensors); the scalar model has one equation.  trunc is the
// projection matrix in the nodal basis, the same for every straight triangle (the map is
// affine, so linear functions and the L2 inner product up to |det J| are preserved), and the
// reference weights stand for |det J| w.
template <int NB>
__device__ inline double shock_sensor(const Tables& tb, int nqv, const double* w, double mean, const double* trunc, int order,
                                      double kap) {
    return shock_sensor_at<NB>([&](int q) { return tb.vb + q * NB; }, [&](int q) { return tb.vw[q]; }, nqv, w, mean, trunc, order,
                               kap);
}

template <int MK, int NB, int NEQ>
__global__ void k_av_limiter(ModelParams P, Tables tb, int nqv, Mesh ms, AvParams A, const double* __restrict__ W,
                             const double* __restrict__ esize, const double* __restrict__ isref, const double* __restrict__ trunc,
                             double* __restrict__ av, double* __restrict__ old) {
    constexpr int B = NB * NEQ;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ms.nElem) return;
    double* a = av + (size_t)e * (1 + NEQ);
    a[0] = 0.0;
    if (A.order <= 1) return;
    double w[B], mean[NEQ], wa[NEQ];
#pragma unroll
    for (int i = 0; i < B; i++) w[i] = W[(size_t)e * B + i];
#pragma unroll
    for (int m = 0; m < NEQ; m++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) s = s + w[m * NB + j] * isref[j];
        mean[m] = s;
        wa[m] = s;
    }
    const double lam = max_eigenvalue<MK>(P, wa);
    const double scale = lam * esize[e] / A.order;
    if (A.dampTol > 0) a[0] = scale * shock_sensor<NB>(tb, nqv, w, mean[0], trunc, A.order, A.dampTol);
    if (A.dampInnerTol > 0) {
        const double g0 = shock_sensor<NB>(tb, nqv, w, mean[0], trunc, A.order, A.dampInnerTol);
#pragma unroll
        for (int m = 0; m < NEQ; m++) {
            const double g = m == 0 || MK != MODEL_SCALAR ? g0 : shock_sensor<NB>(tb, nqv, w + m * NB, mean[m], trunc, A.order, A.dampInnerTol);
            double& o = old[(size_t)e * NEQ + m];
            const double ind = fmax(g, o);
            o = 0.1 * ind;
            a[1 + m] = A.nCoef > 0 ? (A.nCoef == 1 ? A.coef[0] : A.coef[m]) * ind : scale * ind;
        }
    }
}

hipError_t launch_av_limiter(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const Tables& tb, const Mesh& ms,
                             const AvParams& A, const double* W, const double* esize, const double* isref, const double* trunc,
                             double* av, double* old) {
    dim3 blk(128), grd((ms.nElem + 127) / 128);
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                                   \
    if (FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                                                          \
        hipLaunchKernelGGL((k_av_limiter<MK, NB, NEQ>), grd, blk, 0, s, P, tb, NQV, ms, A, W, esize, isref, trunc, av, old); \
        return hipGetLastError();                                                                                      \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ positivity limiter
template <int NB, int NEQ, int NQV, int NQF>
__global__ void k_positivity(Tables tb, Mesh ms, double* __restrict__ W, const double* __restrict__ isref, int* __restrict__ count) {
    constexpr int B = NB * NEQ;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ms.nElem || NB == 1) return;
    double w[B];
#pragma unroll
    for (int i = 0; i < B; i++) w[i] = W[(size_t)e * B + i];
    double wa[NEQ];
#pragma unroll
    for (int m = 0; m < NEQ; m++) wa[m] = 0.0;
#pragma unroll
    for (int m = 0; m < NEQ; m++)
#pragma unroll
        for (int j = 0; j < NB; j++) wa[m] = wa[m] + w[m * NB + j] * isref[j];
    if (!admissible<NEQ>(wa)) return;
    double theta = 1.0;
    auto at = [&](const double* phi) {
        double wx[NEQ];
        for (int m = 0; m < NEQ; m++) wx[m] = 0.0;
        for (int m = 0; m < NEQ; m++)
            for (int j = 0; j < NB; j++) wx[m] += w[m * NB + j] * phi[j];
        theta = fmin(theta, admissible_theta<NEQ>(wa, wx));
    };
    for (int q = 0; q < NQV; q++) at(tb.vb + q * NB);
    for (int k = 0; k < 3; k++)
        for (int p = 0; p < NQF; p++) at(tb.fb + (k * NQF + p) * NB);
    if (theta >= 1) return;
#pragma unroll
    for (int m = 0; m < NEQ; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double mean = wa[m];
            W[(size_t)e * B + m * NB + i] = mean + theta * (w[m * NB + i] - mean);
        }
    atomicAdd(count, 1);
}

hipError_t launch_size_over_lambda(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const Tables& tb, const Mesh& ms,
                                   const double* W, const double* esize, const double* isref, double* out) {
    dim3 blk(256), grd((ms.nElem + 255) / 256);
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                                 \
    if (FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                                                        \
        hipLaunchKernelGGL((k_size_over_lambda<MK, NB, NEQ>), grd, blk, 0, s, P, tb, ms, W, esize, isref, out);     \
        return hipGetLastError();                                                                                    \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_positivity(const ElemLaunch& L, hipStream_t s, const Tables& tb, const Mesh& ms, double* W,
                             const double* isref, int* count) {
    if (L.model == MODEL_SCALAR) return hipSuccess;
    dim3 blk(128), grd((ms.nElem + 127) / 128);
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                        \
    if (MK != MODEL_SCALAR && FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                         \
        hipLaunchKernelGGL((k_positivity<NB, NEQ, NQV, NQF>), grd, blk, 0, s, tb, ms, W, isref, count);    \
        return hipGetLastError();                                                                      \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ functionals
// DragCoefficient / LiftCoefficient (ModuleProjects/FunctionalsOfInterest/src/main/java/flowpro/
// user/functional/DragCoefficient.java:31-48, LiftCoefficient.java:32-34): I = sum over WALL
// faces of int p(W_L) n_which, p = (kappa - 1)(E - |m|^2 / (2 rho)) without density clipping,
// as Element.computeFunctional integrates it (Element.java:789-862).  Also the exact gradient
// dI/dW in the element layout.
template <int NB, int NQF>
__global__ void k_functional(Tables tb, Mesh ms, const double* __restrict__ W, double kappa, int which,
                             double* __restrict__ ival, double* __restrict__ grad) {
    constexpr int NEQ = 4, B = NB * NEQ;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ms.nElem) return;
    double g[B];
#pragma unroll
    for (int i = 0; i < B; i++) g[i] = 0.0;
    double I = 0.0;
    for (int k = 0; k < 3; k++) {
        if (ms.nbr[e * 3 + k] != BT_WALL) continue;
        const double* fg = ms.fgeo + ((size_t)e * 3 + k) * 4;
        const double nk = fg[which];
        for (int p = 0; p < NQF; p++) {
            const double* phi = tb.fb + (k * NQF + p) * NB;
            double w[4];
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
            const double dp[4] = {(kappa - 1) * mom2 / (2 * w[0] * w[0]), -(kappa - 1) * w[1] / w[0], -(kappa - 1) * w[2] / w[0],
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

hipError_t launch_functional(const ElemLaunch& L, hipStream_t s, const Tables& tb, const Mesh& ms, const double* W, double kappa,
                             int which, double* ival, double* grad) {
    if (L.model != MODEL_NAVIER_STOKES) return hipErrorInvalidValue;
    dim3 blk(128), grd((ms.nElem + 127) / 128);
#define X(MK, NB, NEQ, NQV, NQF, V)                                                                          \
    if (MK == MODEL_NAVIER_STOKES && FP_MATCH(MK, NB, NEQ, NQV, NQF, V)) {                                    \
        hipLaunchKernelGGL((k_functional<NB, NQF>), grd, blk, 0, s, tb, ms, W, kappa, which, ival, grad);   \
        return hipGetLastError();                                                                             \
    }
    FP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace fp
