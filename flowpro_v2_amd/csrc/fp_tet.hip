// DG kernels on straight tetrahedra (tetrahedralElement / basis3Dtetra orders 1 and 2,
// transformation3Dtetra, faceTransformation3Dtriangle) for the inviscid 3D model.
//
// One thread per element (residual, J v) or per (element, own DOF) (exact Jacobian columns
// and J^T v); P1 tetrahedra have 20 unknowns per element.  The arithmetic is the one of
// DGpure.residuum / residuumWall (FlowPro/src/main/java/flowpro/core/element/DGpure.java:
// 108-355): faces first, then the volume; the neighbour-side values at a face point are the
// neighbour's basis at the point mapped through the neighbour's inverse affine map, as
// Integration.Face does (FlowPro/src/main/java/flowpro/core/Integration.java:140-151).
#include <hip/hip_runtime.h>

#include "fp_common.hpp"
#include "fp_model3d.hpp"

namespace fp {

// P0: 1; P1: the Lagrange basis on the vertices (basis3Dtetra with the 4 vertex nodes) =
// barycentric coordinates (1 - xi - eta - zeta, xi, eta, zeta)
template <int NB> __device__ inline void tet_basis(const double* xi, double* phi) {
    if constexpr (NB == 1) {
        phi[0] = 1.0;
    } else {
        phi[0] = 1.0 - xi[0] - xi[1] - xi[2];
        phi[1] = xi[0];
        phi[2] = xi[1];
        phi[3] = xi[2];
    }
}

struct TetNoKR {
    __device__ void face_begin(int) {}
    template <class T, int NB>
    __device__ void add(int, const double*, double, const T*) {}
    __device__ void face_end(int) {}
};

// Residual of element e.  Faces first (k = 0..3), then the volume fluxes (order 2).
template <int NB, int NQV, int NQF, class T, class NbW, class KR>
__device__ void tet_residual(const ModelParams& P, const TetTables& tb, const TetMesh& mt, int e, const T* W, NbW nbw, T* K,
                             KR& kr) {
    constexpr int NE = 5;
    const double* ge = mt.geo + (size_t)e * TGEO;
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        const int TT = mt.nbr[e * 4 + k];
        const double* fg = mt.fgeo + ((size_t)e * 4 + k) * TFGEO;
        const double n[3] = {fg[0], fg[1], fg[2]};
        const double area = fg[3];
        const double* gn = TT > -1 ? mt.geo + (size_t)TT * TGEO : nullptr;
        kr.face_begin(k);
#pragma unroll 1
        for (int p = 0; p < NQF; p++) {
            const double* phiL = tb.fb + (k * NQF + p) * NB;
            T WL[NE], WR[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) {
                T s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) s += W[m * NB + j] * phiL[j];
                WL[m] = s;
            }
            double phiR[NB];
            if (TT > -1) {
                // physical point from the own map, reference point of the neighbour from its inverse map
                const double* xr = tb.fxr + (k * NQF + p) * 3;
                double X[3], xn[3];
#pragma unroll
                for (int a = 0; a < 3; a++) X[a] = ge[10 + 3 * a] * xr[0] + ge[11 + 3 * a] * xr[1] + ge[12 + 3 * a] * xr[2] + ge[19 + a];
#pragma unroll
                for (int a = 0; a < 3; a++)
                    xn[a] = gn[3 * a] * (X[0] - gn[19]) + gn[3 * a + 1] * (X[1] - gn[20]) + gn[3 * a + 2] * (X[2] - gn[21]);
                tet_basis<NB>(xn, phiR);
#pragma unroll
                for (int m = 0; m < NE; m++) {
                    T s = 0.0;
#pragma unroll
                    for (int j = 0; j < NB; j++) s += nbw(TT, m * NB + j) * phiR[j];
                    WR[m] = s;
                }
            } else {
                boundary_value3(P, WL, n, TT, WR);
            }
            T fn[NE];
            numerical_convective_flux3(P, WL, WR, n, TT, fn);
            T g[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) g[m] = -fn[m];
            const double jw = area * tb.fw[p];
#pragma unroll
            for (int m = 0; m < NE; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * g[m];
            if (TT > -1) kr.template add<T, NB>(k, phiR, jw, g);
        }
        kr.face_end(k);
    }
    if constexpr (NB > 1) {
        const double jac = ge[9];
#pragma unroll 1
        for (int q = 0; q < NQV; q++) {
            const double* phi = tb.vb + q * NB;
            const double* dphi = tb.vd + q * NB * 3;
            T Wq[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) {
                T s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) s += W[m * NB + j] * phi[j];
                Wq[m] = s;
            }
            T f[3][NE];
            const double ex[3] = {1.0, 0.0, 0.0}, ey[3] = {0.0, 1.0, 0.0}, ez[3] = {0.0, 0.0, 1.0};
            convective_flux3(P, Wq, ex, f[0]);
            convective_flux3(P, Wq, ey, f[1]);
            convective_flux3(P, Wq, ez, f[2]);
            const double jw = jac * tb.vw[q];
#pragma unroll
            for (int j = 0; j < NB; j++) {
                // physical gradient d phi / d x_a = sum_b iT[b][a] d phi / d xi_b
                double gr[3];
#pragma unroll
                for (int a = 0; a < 3; a++) gr[a] = ge[a] * dphi[3 * j] + ge[3 + a] * dphi[3 * j + 1] + ge[6 + a] * dphi[3 * j + 2];
#pragma unroll
                for (int m = 0; m < NE; m++) K[m * NB + j] += jw * (f[0][m] * gr[0] + f[1][m] * gr[1] + f[2][m] * gr[2]);
            }
        }
    }
}

struct TetPlainNb {
    const double* W;
    int B;
    __device__ double operator()(int el, int i) const { return W[(size_t)el * B + i]; }
};
struct TetTangentNb {
    const double* W;
    const double* V;
    int B;
    __device__ dual operator()(int el, int i) const {
        size_t o = (size_t)el * B + i;
        return dual(W[o], V[o]);
    }
};
struct TetZeroTangentNb {
    const double* W;
    int B;
    __device__ dual operator()(int el, int i) const { return dual(W[(size_t)el * B + i], 0.0); }
};

// R, RHS = R - M (a1 W + a2 Wo + a3 Wo2), per-element R_m^T M^-1 R_m and RHS_m^T M^-1 RHS_m
template <int NB, int NQV, int NQF>
__global__ void __launch_bounds__(128) k_residual_tet(ModelParams P, TetTables tb, TetMesh mt, const double* __restrict__ W,
                                                      const double* __restrict__ Wo, const double* __restrict__ Wo2, double a1,
                                                      double a2, double a3, double* __restrict__ R, double* __restrict__ RHS,
                                                      double* __restrict__ enorm) {
    constexpr int NE = 5, B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mt.nElem) return;
    double w[B], K[B];
#pragma unroll
    for (int i = 0; i < B; i++) { w[i] = W[(size_t)e * B + i]; K[i] = 0.0; }
    TetNoKR kr;
    tet_residual<NB, NQV, NQF, double>(P, tb, mt, e, w, TetPlainNb{W, B}, K, kr);
    const double jac = mt.geo[(size_t)e * TGEO + 9];
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
                const double mij = jac * tb.mref[i * NB + j];
                s += mij * a1 * w[m * NB + j] + mij * a2 * Wo[oj] + mij * a3 * Wo2[oj];
            }
            R[o] = K[m * NB + i];
            H[m * NB + i] = K[m * NB + i] - s;
            RHS[o] = H[m * NB + i];
        }
    if (!enorm) return;
#pragma unroll
    for (int which = 0; which < 2; which++)
#pragma unroll
        for (int m = 0; m < NE; m++) {
            const double* v = (which ? H : K) + m * NB;
            double sum = 0.0;
#pragma unroll
            for (int a = 0; a < NB; a++) {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < NB; b++) t += tb.mrefinv[a * NB + b] * v[b];
                sum += v[a] * t;
            }
            enorm[(size_t)e * 2 * NE + which * NE + m] = sum / jac;
        }
}

template <int NB>
struct TetKRColumn {
    double acc[NB * 5];
    __device__ void face_begin(int) {
#pragma unroll
        for (int i = 0; i < NB * 5; i++) acc[i] = 0.0;
    }
    template <class T, int NB_>
    __device__ void add(int, const double* phiR, double jw, const T* g) {
#pragma unroll
        for (int m = 0; m < 5; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[m * NB + j] += (jw * phiR[j]) * tangent(g[m]);
    }
};

// Thread (e, dir).  mode 0: column dir of the diagonal block (-dK/dW + a1 M) and of the
// neighbour blocks (dKR/dW); mode 1: out[e*B+dir] = ((a1 M - dR/dW)^T v)[e*B+dir].
template <int NB, int NQV, int NQF, typename BT>
__global__ void __launch_bounds__(128) k_jacobian_tet(ModelParams P, TetTables tb, TetMesh mt, const double* __restrict__ W,
                                                      double a1, BT* __restrict__ blocks, const double* __restrict__ v,
                                                      double* __restrict__ out, int mode) {
    constexpr int NE = 5, B = NB * NE;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = (int)(t / B), dir = (int)(t % B);
    if (e >= mt.nElem) return;
    dual w[B], K[B];
#pragma unroll
    for (int i = 0; i < B; i++) { w[i] = dual(W[(size_t)e * B + i], i == dir ? 1.0 : 0.0); K[i] = dual(0.0, 0.0); }
    struct Emit : TetKRColumn<NB> {
        const TetMesh* mt; BT* blocks; int e, dir, mode; const double* v; double dot;
        __device__ void face_end(int k) {
            const int blk = mt->blkn[e * 4 + k];
            if (blk < 0) return;
            if (mode == 0) {
                BT* col = blocks + (size_t)blk * B * B + (size_t)dir * B;
#pragma unroll
                for (int i = 0; i < B; i++) col[i] = (BT)this->acc[i];
            } else {
                const double* vn = v + (size_t)mt->nbr[e * 4 + k] * B;
#pragma unroll
                for (int i = 0; i < B; i++) dot += this->acc[i] * vn[i];
            }
        }
    } kr;
    kr.mt = &mt; kr.blocks = blocks; kr.e = e; kr.dir = dir; kr.mode = mode; kr.v = v; kr.dot = 0.0;
    tet_residual<NB, NQV, NQF, dual>(P, tb, mt, e, w, TetZeroTangentNb{W, B}, K, kr);
    const double jac = mt.geo[(size_t)e * TGEO + 9];
    const int md = dir / NB, jd = dir % NB;
    if (mode == 0) {
        BT* col = blocks + (size_t)mt.blkd[e] * B * B + (size_t)dir * B;
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

// Team form of the exact Jacobian (assembly mode): a workgroup owns EPB elements.  Phase 1
// interpolates the own and the neighbour state at every face and volume point; phase 2 takes the
// derivative of each point's integrand (face: g = -fn; volume: the fluxes in x, y, z) with respect
// to each of the five own-state values with one dual pass per (point, value); phase 3 thread
// (element, column dir) contracts those with the test functions into column dir of the diagonal
// block (-dK/dW + a1 M) and of the four neighbour blocks (the neighbour-side integrand, KR).
template <int NB, int NQV, int NQF>
struct TetTeam {
    static constexpr int NE = 5, B = NB * NE, NQFT = 4 * NQF, NQ = NQV + NQFT;
    // per element (doubles): own coefficients, point states, neighbour basis at the face points,
    // integrand derivatives
    static constexpr int E_W = 0;
    static constexpr int E_SV = E_W + B;                       // [NQV][NE]
    static constexpr int E_SF = E_SV + NQV * NE;               // [NQFT][2 NE] WL, WR
    static constexpr int E_PR = E_SF + NQFT * 2 * NE;          // [NQFT][NB] neighbour basis
    static constexpr int E_DV = E_PR + NQFT * NB;              // [NQV][NE inputs][3 NE]
    static constexpr int E_DF = E_DV + NQV * NE * 3 * NE;      // [NQFT][NE inputs][NE]
    static constexpr int PER = E_DF + NQFT * NE * NE;
    static constexpr int EPB0 = (56 * 1024 / 8) / PER;
    static constexpr int EPB = EPB0 < 1 ? 1 : (EPB0 > 8 ? 8 : EPB0);
};

template <int NB, int NQV, int NQF>
__global__ void __launch_bounds__(256) k_jacobian_tet_team(ModelParams P, TetTables tb, TetMesh mt, const double* __restrict__ W,
                                                           double a1, double* __restrict__ blocks, float* __restrict__ fblocks) {
    using LY = TetTeam<NB, NQV, NQF>;
    constexpr int NE = LY::NE, B = LY::B, NQ = LY::NQ, EPB = LY::EPB;
    __shared__ double eb[EPB * LY::PER];
    const int e0 = blockIdx.x * EPB;
    const int nel = min(EPB, mt.nElem - e0);
    if (nel <= 0) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int q = tid; q < nel * B; q += nt) {
        const int el = q / B, i = q % B;
        eb[el * LY::PER + LY::E_W + i] = W[(size_t)(e0 + el) * B + i];
    }
    __syncthreads();
    // phase 1: states at the points
    for (int q = tid; q < nel * NQ; q += nt) {
        const int el = q / NQ, qp = q % NQ, e = e0 + el;
        double* E = eb + el * LY::PER;
        const double* w = E + LY::E_W;
        if (qp < NQV) {
            const double* phi = tb.vb + qp * NB;
#pragma unroll
            for (int m = 0; m < NE; m++) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) s += w[m * NB + j] * phi[j];
                E[LY::E_SV + qp * NE + m] = s;
            }
        } else {
            const int f = qp - NQV, k = f / NQF;
            const double* phiL = tb.fb + f * NB;
            double* st = E + LY::E_SF + f * 2 * NE;
#pragma unroll
            for (int m = 0; m < NE; m++) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) s += w[m * NB + j] * phiL[j];
                st[m] = s;
            }
            const int TT = mt.nbr[e * 4 + k];
            double* pr = E + LY::E_PR + f * NB;
            if (TT > -1) {
                const double* ge = mt.geo + (size_t)e * TGEO;
                const double* gn = mt.geo + (size_t)TT * TGEO;
                const double* xr = tb.fxr + f * 3;
                double X[3], xn[3], phiR[NB];
#pragma unroll
                for (int a = 0; a < 3; a++) X[a] = ge[10 + 3 * a] * xr[0] + ge[11 + 3 * a] * xr[1] + ge[12 + 3 * a] * xr[2] + ge[19 + a];
#pragma unroll
                for (int a = 0; a < 3; a++)
                    xn[a] = gn[3 * a] * (X[0] - gn[19]) + gn[3 * a + 1] * (X[1] - gn[20]) + gn[3 * a + 2] * (X[2] - gn[21]);
                tet_basis<NB>(xn, phiR);
                const double* wn = W + (size_t)TT * B;
#pragma unroll
                for (int m = 0; m < NE; m++) {
                    double s = 0.0;
#pragma unroll
                    for (int j = 0; j < NB; j++) s += wn[m * NB + j] * phiR[j];
                    st[NE + m] = s;
                }
#pragma unroll
                for (int j = 0; j < NB; j++) pr[j] = phiR[j];
            } else {
#pragma unroll
                for (int j = 0; j < NB; j++) pr[j] = 0.0;
            }
        }
    }
    __syncthreads();
    // phase 2: derivatives of the weighted integrands, one (point, own value) pair per task
    for (int q = tid; q < nel * NQ * NE; q += nt) {
        const int el = q / (NQ * NE), r = q % (NQ * NE), qp = r / NE, c = r % NE, e = e0 + el;
        double* E = eb + el * LY::PER;
        if (qp < NQV) {
            if constexpr (NB > 1) {
                dual Wq[NE];
#pragma unroll
                for (int m = 0; m < NE; m++) Wq[m] = dual(E[LY::E_SV + qp * NE + m], m == c ? 1.0 : 0.0);
                dual f[3][NE];
                const double ex[3] = {1.0, 0.0, 0.0}, ey[3] = {0.0, 1.0, 0.0}, ez[3] = {0.0, 0.0, 1.0};
                convective_flux3(P, Wq, ex, f[0]);
                convective_flux3(P, Wq, ey, f[1]);
                convective_flux3(P, Wq, ez, f[2]);
                const double jw = mt.geo[(size_t)e * TGEO + 9] * tb.vw[qp];
                double* out = E + LY::E_DV + (qp * NE + c) * 3 * NE;
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int m = 0; m < NE; m++) out[d * NE + m] = jw * f[d][m].d;
            }
        } else {
            const int f = qp - NQV, k = f / NQF, p = f % NQF;
            const int TT = mt.nbr[e * 4 + k];
            const double* fg = mt.fgeo + ((size_t)e * 4 + k) * TFGEO;
            const double n[3] = {fg[0], fg[1], fg[2]};
            const double* st = E + LY::E_SF + f * 2 * NE;
            dual WL[NE], WR[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) WL[m] = dual(st[m], m == c ? 1.0 : 0.0);
            if (TT > -1) {
#pragma unroll
                for (int m = 0; m < NE; m++) WR[m] = dual(st[NE + m], 0.0);
            } else {
                boundary_value3(P, WL, n, TT, WR);
            }
            dual fn[NE];
            numerical_convective_flux3(P, WL, WR, n, TT, fn);
            const double jw = fg[3] * tb.fw[p];
            double* out = E + LY::E_DF + (f * NE + c) * NE;
#pragma unroll
            for (int m = 0; m < NE; m++) out[m] = -jw * fn[m].d;
        }
    }
    __syncthreads();
    // phase 3: thread (element, column dir = (mp, jp))
    const size_t BB = (size_t)B * B;
    for (int q = tid; q < nel * B; q += nt) {
        const int el = q / B, dir = q % B, mp = dir / NB, jp = dir % NB, e = e0 + el;
        const double* E = eb + el * LY::PER;
        double acc[B];
#pragma unroll
        for (int o = 0; o < B; o++) acc[o] = 0.0;
        for (int k = 0; k < 4; k++) {
            const int blk = mt.blkn[e * 4 + k];
            double accR[B];
#pragma unroll
            for (int o = 0; o < B; o++) accR[o] = 0.0;
            for (int p = 0; p < NQF; p++) {
                const int f = k * NQF + p;
                const double psi = tb.fb[f * NB + jp];
                const double* dg = E + LY::E_DF + (f * NE + mp) * NE;
                const double* phiL = tb.fb + f * NB;
                const double* phiR = E + LY::E_PR + f * NB;
#pragma unroll
                for (int m = 0; m < NE; m++) {
                    const double t = dg[m] * psi;
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        acc[m * NB + j] += phiL[j] * t;
                        accR[m * NB + j] += phiR[j] * t;
                    }
                }
            }
            if (blk >= 0) {
                double* col = blocks + (size_t)blk * BB + (size_t)dir * B;
#pragma unroll
                for (int o = 0; o < B; o++) col[o] = accR[o];
                if (fblocks) {
                    float* fc = fblocks + (size_t)blk * BB + (size_t)dir * B;
#pragma unroll
                    for (int o = 0; o < B; o++) fc[o] = (float)accR[o];
                }
            }
        }
        if constexpr (NB > 1) {
            const double* ge = mt.geo + (size_t)e * TGEO;
            for (int qv = 0; qv < NQV; qv++) {
                const double psi = tb.vb[qv * NB + jp];
                const double* df = E + LY::E_DV + (qv * NE + mp) * 3 * NE;
                const double* dphi = tb.vd + qv * NB * 3;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    double gr[3];
#pragma unroll
                    for (int a = 0; a < 3; a++) gr[a] = ge[a] * dphi[3 * j] + ge[3 + a] * dphi[3 * j + 1] + ge[6 + a] * dphi[3 * j + 2];
#pragma unroll
                    for (int m = 0; m < NE; m++) acc[m * NB + j] += (df[m] * gr[0] + df[NE + m] * gr[1] + df[2 * NE + m] * gr[2]) * psi;
                }
            }
        }
        const double jac = mt.geo[(size_t)e * TGEO + 9];
        double* col = blocks + (size_t)mt.blkd[e] * BB + (size_t)dir * B;
#pragma unroll
        for (int o = 0; o < B; o++) {
            double a = -acc[o];
            if (o / NB == mp) a += a1 * (jac * tb.mref[jp * NB + (o % NB)]);
            col[o] = a;
            if (fblocks) fblocks[(size_t)mt.blkd[e] * BB + (size_t)dir * B + o] = (float)a;
        }
    }
}

// out = (a1 M - dR/dW) v by one forward tangent pass per element
template <int NB, int NQV, int NQF>
__global__ void __launch_bounds__(128) k_jvp_tet(ModelParams P, TetTables tb, TetMesh mt, const double* __restrict__ W,
                                                 const double* __restrict__ v, double a1, double* __restrict__ out) {
    constexpr int NE = 5, B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mt.nElem) return;
    dual w[B], K[B];
#pragma unroll
    for (int i = 0; i < B; i++) { w[i] = dual(W[(size_t)e * B + i], v[(size_t)e * B + i]); K[i] = dual(0.0, 0.0); }
    TetNoKR kr;
    tet_residual<NB, NQV, NQF, dual>(P, tb, mt, e, w, TetTangentNb{W, v, B}, K, kr);
    const double jac = mt.geo[(size_t)e * TGEO + 9];
#pragma unroll
    for (int m = 0; m < NE; m++)
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) s += (jac * tb.mref[i * NB + j]) * v[(size_t)e * B + m * NB + j];
            out[(size_t)e * B + m * NB + i] = a1 * s - K[m * NB + i].d;
        }
}

// elemSize / lambda_max(mean W) per element (State.getTimeStep, Aerodynamics.maxEigenvalue)
template <int NB>
__global__ void k_size_over_lambda_tet(ModelParams P, TetMesh mt, const double* __restrict__ W, const double* __restrict__ esize,
                                       const double* __restrict__ isref, double* __restrict__ out) {
    constexpr int B = NB * 5;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mt.nElem) return;
    double wa[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        wa[m] = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) wa[m] = wa[m] + W[(size_t)e * B + m * NB + j] * isref[j];
    }
    const double lam = max_eigenvalue3(P, wa);
    out[e] = lam > 0 ? esize[e] / lam : __builtin_huge_val();
}

// Element.positivityLimiter with Aerodynamics.isAdmissible in 3D (rho, rho E - |m|^2 / (2 rho))
__device__ inline bool admissible3(const double* W) {
    if (!(W[0] > 1e-10)) return false;
    double m2 = W[1] * W[1];
    m2 += W[2] * W[2];
    m2 += W[3] * W[3];
    return W[4] - m2 / (2 * W[0]) > 1e-10;
}
template <int NB, int NQV, int NQF>
__global__ void k_positivity_tet(TetTables tb, TetMesh mt, double* __restrict__ W, const double* __restrict__ isref,
                                 int* __restrict__ count) {
    constexpr int NE = 5, B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mt.nElem || NB == 1) return;
    double w[B], wa[NE];
#pragma unroll
    for (int i = 0; i < B; i++) w[i] = W[(size_t)e * B + i];
#pragma unroll
    for (int m = 0; m < NE; m++) {
        wa[m] = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) wa[m] = wa[m] + w[m * NB + j] * isref[j];
    }
    if (!admissible3(wa)) return;
    double theta = 1.0;
    auto at = [&](const double* phi) {
        double wx[NE];
#pragma unroll
        for (int m = 0; m < NE; m++) {
            wx[m] = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) wx[m] += w[m * NB + j] * phi[j];
        }
        if (admissible3(wx)) return;
        double lo = 0, hi = 1, wt[NE];
        for (int it = 0; it < 30; it++) {
            const double mid = 0.5 * (lo + hi);
#pragma unroll
            for (int m = 0; m < NE; m++) wt[m] = wa[m] + mid * (wx[m] - wa[m]);
            if (admissible3(wt)) lo = mid;
            else hi = mid;
        }
        theta = fmin(theta, lo);
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
// (NB, NQV, NQF): order 1 with the 1-point rules, order 2 with gaussTetra/3 (6 points) and
// the positive triangle rule of degree 4 (6 points)
#define FP_TET_CASES(X) X(1, 1, 1) X(4, 6, 6)
#define FP_TMATCH3(NB, NQV, NQF) (L.nb == NB && L.nqv == NQV && L.nqf == NQF)

bool tet_supported(const ElemLaunch& L) {
#define X(NB, NQV, NQF) \
    if (FP_TMATCH3(NB, NQV, NQF)) return true;
    FP_TET_CASES(X)
#undef X
    return false;
}

hipError_t launch_residual_tet(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const TetTables& tb, const TetMesh& mt,
                               const double* W, const double* Wo, const double* Wo2, double a1, double a2, double a3, double* R,
                               double* RHS, double* enorm) {
    dim3 blk(128), grd((mt.nElem + 127) / 128);
#define X(NB, NQV, NQF)                                                                                               \
    if (FP_TMATCH3(NB, NQV, NQF)) {                                                                                   \
        hipLaunchKernelGGL((k_residual_tet<NB, NQV, NQF>), grd, blk, 0, s, P, tb, mt, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm); \
        return hipGetLastError();                                                                                     \
    }
    FP_TET_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_jacobian_tet(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const TetTables& tb, const TetMesh& mt,
                               const double* W, double a1, double* blocks, const double* v, double* out, int mode) {
    const long long threads = (long long)mt.nElem * L.nb * 5;
    dim3 blk(128), grd((unsigned)((threads + 127) / 128));
#define X(NB, NQV, NQF)                                                                                                     \
    if (FP_TMATCH3(NB, NQV, NQF)) {                                                                                         \
        hipLaunchKernelGGL((k_jacobian_tet<NB, NQV, NQF, double>), grd, blk, 0, s, P, tb, mt, W, a1, blocks, v, out, mode); \
        return hipGetLastError();                                                                                           \
    }
    FP_TET_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_jacobian_tet_team(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const TetTables& tb, const TetMesh& mt,
                                    const double* W, double a1, double* blocks, float* fblocks) {
#define X(NB, NQV, NQF)                                                                                          \
    if (FP_TMATCH3(NB, NQV, NQF)) {                                                                              \
        constexpr int EPB = TetTeam<NB, NQV, NQF>::EPB;                                                          \
        hipLaunchKernelGGL((k_jacobian_tet_team<NB, NQV, NQF>), dim3((mt.nElem + EPB - 1) / EPB), dim3(256), 0, s, P, tb, mt, W, \
                           a1, blocks, fblocks);                                                                 \
        return hipGetLastError();                                                                                \
    }
    FP_TET_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_jvp_tet(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const TetTables& tb, const TetMesh& mt,
                          const double* W, const double* v, double a1, double* out) {
    dim3 blk(128), grd((mt.nElem + 127) / 128);
#define X(NB, NQV, NQF)                                                                          \
    if (FP_TMATCH3(NB, NQV, NQF)) {                                                              \
        hipLaunchKernelGGL((k_jvp_tet<NB, NQV, NQF>), grd, blk, 0, s, P, tb, mt, W, v, a1, out); \
        return hipGetLastError();                                                                \
    }
    FP_TET_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_size_over_lambda_tet(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const TetMesh& mt, const double* W,
                                       const double* esize, const double* isref, double* out) {
    dim3 blk(256), grd((mt.nElem + 255) / 256);
    if (L.nb == 1) hipLaunchKernelGGL((k_size_over_lambda_tet<1>), grd, blk, 0, s, P, mt, W, esize, isref, out);
    else if (L.nb == 4) hipLaunchKernelGGL((k_size_over_lambda_tet<4>), grd, blk, 0, s, P, mt, W, esize, isref, out);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_positivity_tet(const ElemLaunch& L, hipStream_t s, const TetTables& tb, const TetMesh& mt, double* W,
                                 const double* isref, int* count) {
    dim3 blk(128), grd((mt.nElem + 127) / 128);
#define X(NB, NQV, NQF)                                                                           \
    if (FP_TMATCH3(NB, NQV, NQF)) {                                                               \
        hipLaunchKernelGGL((k_positivity_tet<NB, NQV, NQF>), grd, blk, 0, s, tb, mt, W, isref, count); \
        return hipGetLastError();                                                                 \
    }
    FP_TET_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace fp
