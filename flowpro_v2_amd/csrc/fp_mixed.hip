// DG kernels on meshes that mix element types (triangleElement, squareElement) and element
// orders (order.txt), for the NavierStokes (Euler when inviscid) and AdiabaticEulers models.
//
// Every element is padded to NBM basis functions with zero tables, so one compile-time shape
// serves all elements; the point counts (volume rule of the element, face rule of the larger order
// of the two sides, Integration.Face, FlowPro/src/main/java/flowpro/core/Integration.java:114-117)
// are runtime loop bounds.  The neighbour-side basis and gradients at a face point are tabulated at
// setup from the neighbour's own basis at the point its inverse map returns.  One thread per
// element (residual, J v) or per (element, own padded DOF) (exact Jacobian columns, J^T v); the
// arithmetic is the one of DGpure.residuum / residuumWall (DGpure.java:108-355): faces first, then
// the volume.
#include <hip/hip_runtime.h>

#include "fp_common.hpp"

namespace fp {

namespace {
struct MPlain {
    const double* W;
    int B;
    __device__ double operator()(int el, int i) const { return W[(size_t)el * B + i]; }
};
struct MTangent {
    const double* W;
    const double* V;
    int B;
    __device__ dual operator()(int el, int i) const {
        const size_t o = (size_t)el * B + i;
        return dual(W[o], V[o]);
    }
};
struct MConst {
    const double* W;
    int B;
    __device__ dual operator()(int el, int i) const { return dual(W[(size_t)el * B + i], 0.0); }
};
struct MColumn {
    const double* W;
    int B, dir;
    __device__ dual operator()(int el, int i) const { return dual(W[(size_t)el * B + i], i == dir ? 1.0 : 0.0); }
};
}  // namespace

// face integrand: boundary value, convective flux, diffusive flux with the element's c_IP, then
// (AV) the damping flux of element e and its neighbour on interior faces
template <int MK, int NE, bool VISC, bool AV, class T>
__device__ inline void mixed_face_flux(const ModelParams& P, const DGParams& G, const MixedMesh& mm, int e, double cIP, int TT,
                                       const double* n, double dL, const double* X, T* WL, T* dWL, T* WR, T* dWR, T* g) {
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
            if (ip) g[m] -= cIP * (WL[m] - WR[m]);
        }
    }
    if constexpr (AV) {
        if (TT > -1) av_face_flux<NE>(G, mm.av, e, TT, n, dL, WL, WR, dWL, dWR, g);
    }
}

// volume fluxes at a point: convective in x and y, minus the viscous ones, minus (AV) the damping
// ones (DGpure.residuum's order)
template <int MK, int NE, bool VISC, bool AV, class T>
__device__ inline void mixed_volume_flux(const ModelParams& P, const DGParams& G, const MixedMesh& mm, int e, T* Wq, T* dWq, T* fx,
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
    if constexpr (AV) av_volume_flux<NE>(G, mm.av, e, dWq, fx, fy);
}

template <int NB, int NE, class T, class Acc>
__device__ inline void mixed_interp(Acc c, int el, const double* phi, const double* gr, T* Wp, T* dWp) {
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

// Diagonal entry of a padded unknown (the identity mass times a1): 1 at a1 = 0 (the steady adjoint's
// -dR/dW), so that the padded rows stay regular; nothing couples to them either way.
__device__ inline double padded_diagonal(double a1) { return a1 != 0.0 ? a1 : 1.0; }

// DGpure.limiter's c_IP: the penalty on elements of order > 1 of a diffusive model
__device__ inline double mixed_cip(const DGParams& G, const MixedMesh& mm, int e) { return mm.nbe[e] > 1 ? G.cIP : 0.0; }

template <int MK, int NB, bool VISC, bool AV, class T, class Own, class Nb, class Sink>
__device__ void mixed_residual(const ModelParams& P, const DGParams& G, const MixedMesh& mm, int e, Own own, Nb nbw, Sink& sk) {
    constexpr int NE = neq_of<MK>();
    constexpr int FS = 3 + 6 * NB;
    const double cIP = mixed_cip(G, mm, e);
    const int nv = mm.nv[e];
#pragma unroll 1
    for (int k = 0; k < nv; k++) {
        const int TT = mm.nbr[e * 4 + k];
        const double* fg = mm.fgeo + ((size_t)e * 4 + k) * 4;
        const double n[2] = {fg[0], fg[1]};
        const double dL = fg[3];
        const int nq = mm.nqf[e * 4 + k];
#pragma unroll 1
        for (int p = 0; p < nq; p++) {
            const double* fp = mm.face + (((size_t)e * 4 + k) * mm.nqfm + p) * FS;
            const double* phiL = fp + 3;
            T WL[NE], dWL[2 * NE], WR[NE], dWR[2 * NE];
            mixed_interp<NB, NE, T>(own, e, phiL, phiL + NB, WL, dWL);
            const double* phiR = nullptr;
            if (TT > -1) {
                phiR = fp + 3 + 3 * NB;
                mixed_interp<NB, NE, T>(nbw, TT, phiR, phiR + NB, WR, dWR);
            }
            T g[NE];
            mixed_face_flux<MK, NE, VISC, AV>(P, G, mm, e, cIP, TT, n, dL, fp + 1, WL, dWL, WR, dWR, g);
            sk.face(k, p, phiL, phiR, fp[0], g);
        }
    }
    if (mm.nbe[e] > 1) {   // the volume fluxes vanish at order 1 (DGpure.residuum)
        const int nq = mm.nqv[e];
#pragma unroll 1
        for (int q = 0; q < nq; q++) {
            const double* vp = mm.vol + ((size_t)e * mm.nqvm + q) * (1 + 3 * NB);
            T Wq[NE], dWq[2 * NE], fx[NE], fy[NE];
            mixed_interp<NB, NE, T>(own, e, vp + 1, vp + 1 + NB, Wq, dWq);
            mixed_volume_flux<MK, NE, VISC, AV>(P, G, mm, e, Wq, dWq, fx, fy);
            sk.vol(vp[0], vp + 1 + NB, fx, fy);
        }
    }
}

template <int NB, int NE>
struct MValueSink {
    double K[NB * NE];
    __device__ void face(int, int, const double* phiL, const double*, double jw, const double* g) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * g[m];
    }
    __device__ void vol(double jw, const double* gr, const double* fx, const double* fy) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += jw * (fx[m] * gr[2 * j] + fy[m] * gr[2 * j + 1]);
    }
};
template <int NB, int NE>
struct MTangentSink {
    double K[NB * NE];
    __device__ void face(int, int, const double* phiL, const double*, double jw, const dual* g) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * g[m].d;
    }
    __device__ void vol(double jw, const double* gr, const dual* fx, const dual* fy) {
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += jw * (fx[m].d * gr[2 * j] + fy[m].d * gr[2 * j + 1]);
    }
};

template <int MK, int NB, bool VISC, bool AV>
__global__ void __launch_bounds__(128) k_residual_mixed(ModelParams P, DGParams G, MixedMesh mm, const double* __restrict__ W,
                                                        const double* __restrict__ Wo, const double* __restrict__ Wo2, double a1,
                                                        double a2, double a3, double* __restrict__ R, double* __restrict__ RHS,
                                                        double* __restrict__ enorm) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm.nElem) return;
    MValueSink<NB, NE> sk;
#pragma unroll
    for (int i = 0; i < B; i++) sk.K[i] = 0.0;
    mixed_residual<MK, NB, VISC, AV, double>(P, G, mm, e, MPlain{W, B}, MPlain{W, B}, sk);
    const double* M = mm.mass + (size_t)e * NB * NB;
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
    const double* Mi = mm.minv + (size_t)e * NB * NB;
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

// face and volume points per element: 4 faces of up to 5 points (a face next to an element of
// order 5), 25 volume points (the 5 x 5 Gauss rule of an order-5 square; the triangle rule of
// order 5 has 19)
constexpr int MJ_MAXF = 20, MJ_MAXV = 25;
// threads of the element workgroups: one wavefront, two for the padded blocks of order-5 squares
// (B = 100 > 64 columns)
template <int B> constexpr int mixed_team_threads() { return B > 64 ? 128 : 64; }

// The residual, one workgroup per element (64 lanes, 128 above B = 64): lane = point evaluates the face integrand
// or the volume fluxes there (into LDS), then lane = own DOF (m, j) sums its test-function
// contributions in the same order as k_residual_mixed (faces, then volume), adds the BDF mass
// terms, and lanes (which, m) form the equation norms.  Same arithmetic as k_residual_mixed.
template <int MK, int NB, bool VISC, bool AV>
__global__ void __launch_bounds__(128) k_residual_mixed_elem(ModelParams P, DGParams G, MixedMesh mm, const double* __restrict__ W,
                                                            const double* __restrict__ Wo, const double* __restrict__ Wo2, double a1,
                                                            double a2, double a3, double* __restrict__ R, double* __restrict__ RHS,
                                                            double* __restrict__ enorm) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    __shared__ double GF[MJ_MAXF][NE];
    __shared__ double GV[MJ_MAXV][2 * NE];
    __shared__ double KS[B], HS[B];
    const int e = blockIdx.x, lane = threadIdx.x;
    if (e >= mm.nElem) return;
    const int nv = mm.nv[e], nqv = mm.nbe[e] > 1 ? mm.nqv[e] : 0;
    const double cIP = mixed_cip(G, mm, e);
    int fstart[5];
    fstart[0] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) fstart[k + 1] = fstart[k] + (k < nv ? mm.nqf[e * 4 + k] : 0);
    const int nfp = fstart[4];
    const MPlain own{W, B};
    for (int pt = lane; pt < nfp + nqv; pt += mixed_team_threads<B>()) {
        if (pt < nfp) {
            int k = 0;
            while (pt >= fstart[k + 1]) k++;
            const int q = pt - fstart[k];
            const int TT = mm.nbr[e * 4 + k];
            const double* fg = mm.fgeo + ((size_t)e * 4 + k) * 4;
            const double n[2] = {fg[0], fg[1]};
            const double* fp = mm.face + (((size_t)e * 4 + k) * mm.nqfm + q) * (3 + 6 * NB);
            double WL[NE], dWL[2 * NE], WR[NE], dWR[2 * NE], g[NE];
            mixed_interp<NB, NE, double>(own, e, fp + 3, fp + 3 + NB, WL, dWL);
            if (TT > -1) mixed_interp<NB, NE, double>(own, TT, fp + 3 + 3 * NB, fp + 3 + 4 * NB, WR, dWR);
            mixed_face_flux<MK, NE, VISC, AV>(P, G, mm, e, cIP, TT, n, fg[3], fp + 1, WL, dWL, WR, dWR, g);
#pragma unroll
            for (int m = 0; m < NE; m++) GF[pt][m] = g[m];
        } else {
            const int q = pt - nfp;
            const double* vp = mm.vol + ((size_t)e * mm.nqvm + q) * (1 + 3 * NB);
            double Wq[NE], dWq[2 * NE], fx[NE], fy[NE];
            mixed_interp<NB, NE, double>(own, e, vp + 1, vp + 1 + NB, Wq, dWq);
            mixed_volume_flux<MK, NE, VISC, AV>(P, G, mm, e, Wq, dWq, fx, fy);
#pragma unroll
            for (int m = 0; m < NE; m++) {
                GV[q][m] = fx[m];
                GV[q][NE + m] = fy[m];
            }
        }
    }
    __syncthreads();
    if (lane < B) {
        const int m = lane / NB, j = lane % NB;
        double K = 0.0;
        for (int k = 0; k < nv; k++)
            for (int q = 0; q < mm.nqf[e * 4 + k]; q++) {
                const double* fp = mm.face + (((size_t)e * 4 + k) * mm.nqfm + q) * (3 + 6 * NB);
                K += (fp[0] * fp[3 + j]) * GF[fstart[k] + q][m];
            }
        for (int q = 0; q < nqv; q++) {
            const double* vp = mm.vol + ((size_t)e * mm.nqvm + q) * (1 + 3 * NB);
            const double* gr = vp + 1 + NB;
            K += vp[0] * (GV[q][m] * gr[2 * j] + GV[q][NE + m] * gr[2 * j + 1]);
        }
        const double* M = mm.mass + (size_t)e * NB * NB;
        double sm = 0.0;
#pragma unroll
        for (int jj = 0; jj < NB; jj++) {
            const size_t oj = (size_t)e * B + m * NB + jj;
            const double mij = M[j * NB + jj];
            sm += mij * a1 * W[oj] + mij * a2 * Wo[oj] + mij * a3 * Wo2[oj];
        }
        const size_t o = (size_t)e * B + lane;
        R[o] = K;
        RHS[o] = K - sm;
        KS[lane] = K;
        HS[lane] = K - sm;
    }
    if (!enorm) return;
    __syncthreads();
    if (lane < 2 * NE) {
        const int which = lane / NE, m = lane % NE;
        const double* Mi = mm.minv + (size_t)e * NB * NB;
        const double* v = (which ? HS : KS) + m * NB;
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

// Thread (e, dir) over the padded DOFs.  mode 0: column dir of the diagonal block (-dK/dW + a1 M)
// and of the neighbour blocks (the neighbour-side integrand, KR); mode 1: ((a1 M - dR/dW)^T v)[e, dir].
template <int NB, int NE, typename BT>
struct MColumnSink : MTangentSink<NB, NE> {
    static constexpr int B = NB * NE;
    const MixedMesh* mm;
    BT* blocks;
    float* fblocks;
    const double* v;
    int e, dir, mode;
    double dot;
    __device__ void face(int k, int p, const double* phiL, const double* phiR, double jw, const dual* g) {
        MTangentSink<NB, NE>::face(k, p, phiL, phiR, jw, g);
        if (!phiR) return;
        if (mode == 0) {
            const size_t base = (size_t)mm->blkn[e * 4 + k] * B * B + (size_t)dir * B;
#pragma unroll
            for (int m = 0; m < NE; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const double a = (jw * phiR[j]) * g[m].d;
                    BT* c = blocks + base + m * NB + j;
                    *c = p == 0 ? (BT)a : (BT)(*c + a);
                    if (fblocks && p + 1 == mm->nqf[e * 4 + k]) fblocks[base + m * NB + j] = (float)*c;
                }
        } else {
            const double* vn = v + (size_t)mm->nbr[e * 4 + k] * B;
#pragma unroll
            for (int m = 0; m < NE; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) dot += ((jw * phiR[j]) * g[m].d) * vn[m * NB + j];
        }
    }
};

// Thread t takes column t % B of element t / B.  Assembly (mode 0) runs k_jacobian_mixed_elem;
// this per-column form serves J^T v (mode 1).
template <int MK, int NB, bool VISC, bool AV, typename BT>
__global__ void __launch_bounds__(64) k_jacobian_mixed(ModelParams P, DGParams G, MixedMesh mm, const double* __restrict__ W, double a1,
                                                       BT* __restrict__ blocks, float* __restrict__ fblocks,
                                                       const double* __restrict__ v, double* __restrict__ out, int mode) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = (int)(t / B), dir = (int)(t % B);
    if (e >= mm.nElem) return;
    if (dir % NB >= mm.nbe[e]) {
        // a padded unknown: nothing depends on it, so its column is a1 e_dir in the diagonal
        // block (the identity mass; e_dir at a1 = 0) and zero in the neighbour blocks
        const double pd = padded_diagonal(a1);
        if (mode == 0) {
            const size_t base = (size_t)mm.blkd[e] * B * B + (size_t)dir * B;
            for (int i = 0; i < B; i++) {
                blocks[base + i] = (BT)(i == dir ? pd : 0.0);
                if (fblocks) fblocks[base + i] = (float)(i == dir ? pd : 0.0);
            }
            for (int k = 0; k < mm.nv[e]; k++)
                if (mm.nbr[e * 4 + k] > -1) {
                    const size_t nb = (size_t)mm.blkn[e * 4 + k] * B * B + (size_t)dir * B;
                    for (int i = 0; i < B; i++) {
                        blocks[nb + i] = (BT)0;
                        if (fblocks) fblocks[nb + i] = 0.0f;
                    }
                }
        } else {
            out[(size_t)e * B + dir] = pd * v[(size_t)e * B + dir];
        }
        return;
    }
    MColumnSink<NB, NE, BT> sk;
#pragma unroll
    for (int i = 0; i < B; i++) sk.K[i] = 0.0;
    sk.mm = &mm;
    sk.blocks = blocks;
    sk.fblocks = fblocks;
    sk.v = v;
    sk.e = e;
    sk.dir = dir;
    sk.mode = mode;
    sk.dot = 0.0;
    mixed_residual<MK, NB, VISC, AV, dual>(P, G, mm, e, MColumn{W, B, dir}, MConst{W, B}, sk);
    const double* M = mm.mass + (size_t)e * NB * NB;
    const int md = dir / NB, jd = dir % NB;
    if (mode == 0) {
        const size_t base = (size_t)mm.blkd[e] * B * B + (size_t)dir * B;
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

// Exact Jacobian, one workgroup per element (64 lanes, 128 above B = 64; assembly, mode 0 of k_jacobian_mixed).
// Phase 1: the derivatives of the face integrand g(WL, dWL; WR, dWR) and of the volume fluxes
// (fx, fy)(W, dW) at every point with respect to the 3 nEqs own inputs (W_m, dW_m/dx, dW_m/dy;
// nEqs when inviscid), one dual evaluation per (point, input), spread over the lanes and kept
// in LDS.  Phase 2: lane = own column (m, j); the column's tangent at a point is the chain rule
// over those inputs (phi_j, dphi_j/dx, dphi_j/dy in equation m), contracted with the own test
// functions into the diagonal block and with the neighbour's into the neighbour blocks.  The
// dual pass per (column, point) of k_jacobian_mixed becomes one per (input, point).

template <int MK, int NB, bool VISC, bool AV>
__global__ void __launch_bounds__(128) k_jacobian_mixed_elem(ModelParams P, DGParams G, MixedMesh mm, const double* __restrict__ W,
                                                            double a1, double* __restrict__ blocks, float* __restrict__ fblocks) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    constexpr bool GRAD = VISC || AV;   // the diffusive and the damping terms see the gradients
    constexpr int NIN = GRAD ? 3 * NE : NE;
    __shared__ double JF[MJ_MAXF][NIN][NE];
    __shared__ double JV[MJ_MAXV][NIN][2 * NE];
    const int e = blockIdx.x, lane = threadIdx.x;
    if (e >= mm.nElem) return;
    const int nv = mm.nv[e], nbe = mm.nbe[e], nqv = nbe > 1 ? mm.nqv[e] : 0;
    const double cIP = mixed_cip(G, mm, e);
    int fstart[5];
    fstart[0] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) fstart[k + 1] = fstart[k] + (k < nv ? mm.nqf[e * 4 + k] : 0);
    const int nfp = fstart[4];
    const MPlain own{W, B};
    // phase 1
    for (int it = lane; it < (nfp + nqv) * NIN; it += mixed_team_threads<B>()) {
        const int pt = it / NIN, d = it % NIN;
        if (pt < nfp) {
            int k = 0;
            while (pt >= fstart[k + 1]) k++;
            const int q = pt - fstart[k];
            const int TT = mm.nbr[e * 4 + k];
            const double* fg = mm.fgeo + ((size_t)e * 4 + k) * 4;
            const double n[2] = {fg[0], fg[1]};
            const double* fp = mm.face + (((size_t)e * 4 + k) * mm.nqfm + q) * (3 + 6 * NB);
            double wl[NE], dwl[2 * NE];
            mixed_interp<NB, NE, double>(own, e, fp + 3, fp + 3 + NB, wl, dwl);
            dual WL[NE], dWL[2 * NE], WR[NE], dWR[2 * NE], g[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) WL[m] = dual(wl[m], d == m ? 1.0 : 0.0);
#pragma unroll
            for (int m = 0; m < 2 * NE; m++) dWL[m] = dual(dwl[m], d == NE + m ? 1.0 : 0.0);
            if (TT > -1) {
                double wr[NE], dwr[2 * NE];
                mixed_interp<NB, NE, double>(own, TT, fp + 3 + 3 * NB, fp + 3 + 4 * NB, wr, dwr);
#pragma unroll
                for (int m = 0; m < NE; m++) WR[m] = dual(wr[m], 0.0);
#pragma unroll
                for (int m = 0; m < 2 * NE; m++) dWR[m] = dual(dwr[m], 0.0);
            }
            mixed_face_flux<MK, NE, VISC, AV>(P, G, mm, e, cIP, TT, n, fg[3], fp + 1, WL, dWL, WR, dWR, g);
#pragma unroll
            for (int m = 0; m < NE; m++) JF[pt][d][m] = g[m].d;
        } else {
            const int q = pt - nfp;
            const double* vp = mm.vol + ((size_t)e * mm.nqvm + q) * (1 + 3 * NB);
            double wq[NE], dwq[2 * NE];
            mixed_interp<NB, NE, double>(own, e, vp + 1, vp + 1 + NB, wq, dwq);
            dual Wq[NE], dWq[2 * NE], fx[NE], fy[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) Wq[m] = dual(wq[m], d == m ? 1.0 : 0.0);
#pragma unroll
            for (int m = 0; m < 2 * NE; m++) dWq[m] = dual(dwq[m], d == NE + m ? 1.0 : 0.0);
            mixed_volume_flux<MK, NE, VISC, AV>(P, G, mm, e, Wq, dWq, fx, fy);
#pragma unroll
            for (int m = 0; m < NE; m++) {
                JV[q][d][m] = fx[m].d;
                JV[q][d][NE + m] = fy[m].d;
            }
        }
    }
    __syncthreads();
    // phase 2
    if (lane >= B) return;
    const int dir = lane, md = dir / NB, jd = dir % NB;
    if (jd >= nbe) {   // a padded unknown: a1 e_dir (e_dir at a1 = 0) in the diagonal block, zero in the neighbour blocks
        const double pd = padded_diagonal(a1);
        const size_t base = (size_t)mm.blkd[e] * B * B + (size_t)dir * B;
        for (int i = 0; i < B; i++) {
            blocks[base + i] = i == dir ? pd : 0.0;
            if (fblocks) fblocks[base + i] = (float)(i == dir ? pd : 0.0);
        }
        for (int k = 0; k < nv; k++)
            if (mm.nbr[e * 4 + k] > -1) {
                const size_t nb = (size_t)mm.blkn[e * 4 + k] * B * B + (size_t)dir * B;
                for (int i = 0; i < B; i++) {
                    blocks[nb + i] = 0.0;
                    if (fblocks) fblocks[nb + i] = 0.0f;
                }
            }
        return;
    }
    double K[B];
#pragma unroll
    for (int i = 0; i < B; i++) K[i] = 0.0;
#pragma unroll 1
    for (int k = 0; k < nv; k++) {
        const int TT = mm.nbr[e * 4 + k];
        const int nq = mm.nqf[e * 4 + k];
        const size_t nbase = TT > -1 ? (size_t)mm.blkn[e * 4 + k] * B * B + (size_t)dir * B : 0;
#pragma unroll 1
        for (int q = 0; q < nq; q++) {
            const double* fp = mm.face + (((size_t)e * 4 + k) * mm.nqfm + q) * (3 + 6 * NB);
            const double jw = fp[0];
            const double* phiL = fp + 3;
            const double* grL = fp + 3 + NB;
            const double tw = phiL[jd], tx = grL[2 * jd], ty = grL[2 * jd + 1];
            const int pt = fstart[k] + q;
            double T[NE];
#pragma unroll
            for (int m = 0; m < NE; m++) {
                T[m] = JF[pt][md][m] * tw;
                if constexpr (GRAD) T[m] += JF[pt][NE + md][m] * tx + JF[pt][2 * NE + md][m] * ty;
            }
#pragma unroll
            for (int m = 0; m < NE; m++)
#pragma unroll
                for (int j = 0; j < NB; j++) K[m * NB + j] += (jw * phiL[j]) * T[m];
            if (TT > -1) {
                const double* phiR = fp + 3 + 3 * NB;
#pragma unroll
                for (int m = 0; m < NE; m++)
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        const double a = (jw * phiR[j]) * T[m];
                        double* c = blocks + nbase + m * NB + j;
                        *c = q == 0 ? a : *c + a;
                        if (fblocks && q + 1 == nq) fblocks[nbase + m * NB + j] = (float)*c;
                    }
            }
        }
    }
#pragma unroll 1
    for (int q = 0; q < nqv; q++) {
        const double* vp = mm.vol + ((size_t)e * mm.nqvm + q) * (1 + 3 * NB);
        const double jw = vp[0];
        const double* gr = vp + 1 + NB;
        const double tw = vp[1 + jd], tx = gr[2 * jd], ty = gr[2 * jd + 1];
        double Tx[NE], Ty[NE];
#pragma unroll
        for (int m = 0; m < NE; m++) {
            Tx[m] = JV[q][md][m] * tw;
            Ty[m] = JV[q][md][NE + m] * tw;
            if constexpr (GRAD) {
                Tx[m] += JV[q][NE + md][m] * tx + JV[q][2 * NE + md][m] * ty;
                Ty[m] += JV[q][NE + md][NE + m] * tx + JV[q][2 * NE + md][NE + m] * ty;
            }
        }
#pragma unroll
        for (int m = 0; m < NE; m++)
#pragma unroll
            for (int j = 0; j < NB; j++) K[m * NB + j] += jw * (Tx[m] * gr[2 * j] + Ty[m] * gr[2 * j + 1]);
    }
    const double* M = mm.mass + (size_t)e * NB * NB;
    const size_t base = (size_t)mm.blkd[e] * B * B + (size_t)dir * B;
#pragma unroll
    for (int i = 0; i < B; i++) {
        double a = -K[i];
        if (i / NB == md) a += a1 * M[jd * NB + (i % NB)];
        blocks[base + i] = a;
        if (fblocks) fblocks[base + i] = (float)a;
    }
}

template <int MK, int NB, bool VISC, bool AV>
__global__ void __launch_bounds__(64) k_jvp_mixed(ModelParams P, DGParams G, MixedMesh mm, const double* __restrict__ W,
                                                  const double* __restrict__ v, double a1, double* __restrict__ out) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm.nElem) return;
    MTangentSink<NB, NE> sk;
#pragma unroll
    for (int i = 0; i < B; i++) sk.K[i] = 0.0;
    mixed_residual<MK, NB, VISC, AV, dual>(P, G, mm, e, MTangent{W, v, B}, MTangent{W, v, B}, sk);
    const double* M = mm.mass + (size_t)e * NB * NB;
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

template <int MK, int NB>
__global__ void k_size_over_lambda_mixed(ModelParams P, MixedMesh mm, const double* __restrict__ W, const double* __restrict__ esize,
                                         double* __restrict__ out) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm.nElem) return;
    const double* is = mm.is + (size_t)e * NB;
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

// DGpure.limiter (FlowPro/src/main/java/flowpro/core/element/DGpure.java:358-400) per element of a
// mixed mesh, with the element's own order k (from its type and basis size): nothing at k = 1; eps
// = lambda(mean W) elemSize / k g_shock with S0 = log10(1 / (k-1)^4); the inner damping with its
// decaying memory.  trunc: [nElem][NB][NB], each element's TrunOrd (k > 2) padded with zeros; the
// sensor integrals take the element's |det J| w and basis values (zero on the padded functions) per
// point.  The Aerodynamics models give every equation the density sensor.
template <int MK, int NB>
__global__ void k_av_limiter_mixed(ModelParams P, MixedMesh mm, AvParams A, const double* __restrict__ W,
                                   const double* __restrict__ esize, const double* __restrict__ trunc, double* __restrict__ av,
                                   double* __restrict__ old) {
    constexpr int NE = neq_of<MK>(), B = NB * NE, VS = 1 + 3 * NB;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm.nElem) return;
    double* a = av + (size_t)e * (1 + NE);
    a[0] = 0.0;
    const int nbe = mm.nbe[e];
    int k = 1;
    if (mm.nv[e] == 4)
        while (k * k < nbe) k++;
    else
        while (k * (k + 1) / 2 < nbe) k++;
    if (k <= 1) return;
    const double* is = mm.is + (size_t)e * NB;
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
    const double scale = lam * esize[e] / k;
    const double* vp = mm.vol + (size_t)e * mm.nqvm * VS;
    auto jw = [&](int q) { return vp[(size_t)q * VS]; };
    auto phi = [&](int q) { return vp + (size_t)q * VS + 1; };
    const double* tr = trunc + (size_t)e * NB * NB;
    const int nq = mm.nqv[e];
    if (A.dampTol > 0) a[0] = scale * shock_sensor_at<NB>(phi, jw, nq, w, mean0, tr, k, A.dampTol);
    if (A.dampInnerTol > 0) {
        const double g = shock_sensor_at<NB>(phi, jw, nq, w, mean0, tr, k, A.dampInnerTol);
#pragma unroll
        for (int m = 0; m < NE; m++) {
            double& o = old[(size_t)e * NE + m];
            const double ind = fmax(g, o);
            o = 0.1 * ind;
            a[1 + m] = A.nCoef > 0 ? (A.nCoef == 1 ? A.coef[0] : A.coef[m]) * ind : scale * ind;
        }
    }
}

// Element.positivityLimiter (Element.java:385-476) over the element's volume and face points; only
// the element's own coefficients change (the padded ones stay zero)
template <int MK, int NB>
__global__ void k_positivity_mixed(MixedMesh mm, double* __restrict__ W, int* __restrict__ count) {
    constexpr int NE = neq_of<MK>(), B = NB * NE;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm.nElem || mm.nbe[e] == 1) return;
    const int nb = mm.nbe[e];
    const double* is = mm.is + (size_t)e * NB;
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
    for (int q = 0; q < mm.nqv[e]; q++) at(mm.vol + ((size_t)e * mm.nqvm + q) * (1 + 3 * NB) + 1);
    for (int k = 0; k < mm.nv[e]; k++)
        for (int p = 0; p < mm.nqf[e * 4 + k]; p++) at(mm.face + (((size_t)e * 4 + k) * mm.nqfm + p) * (3 + 6 * NB) + 3);
    if (theta >= 1) return;
#pragma unroll
    for (int m = 0; m < NE; m++)
#pragma unroll
        for (int i = 0; i < NB; i++)
            if (i < nb) W[(size_t)e * B + m * NB + i] = wa[m] + theta * (w[m * NB + i] - wa[m]);
    atomicAdd(count, 1);
}

// DragCoefficient (which 0) / LiftCoefficient (which 1) contribution of every element: the
// pressure over its wall faces times the normal component, at the face rule of the element's own
// order (a wall face has no second side), and the exact gradient with respect to the element's
// coefficients (zero on the padded ones)
template <int NB>
__global__ void k_functional_mixed(MixedMesh mm, const double* __restrict__ W, double kappa, int which, double* __restrict__ ival,
                                   double* __restrict__ grad) {
    constexpr int NEQ = 4, B = NB * NEQ;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm.nElem) return;
    double g[B];
#pragma unroll
    for (int i = 0; i < B; i++) g[i] = 0.0;
    double I = 0.0;
    for (int k = 0; k < mm.nv[e]; k++) {
        if (mm.nbr[e * 4 + k] != BT_WALL) continue;
        const double nk = mm.fgeo[((size_t)e * 4 + k) * 4 + which];
        for (int p = 0; p < mm.nqf[e * 4 + k]; p++) {
            const double* fp = mm.face + (((size_t)e * 4 + k) * mm.nqfm + p) * (3 + 6 * NB);
            const double* phi = fp + 3;
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
            const double jw = fp[0];
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

// ------------------------------------------------------------------ dispatch
// (model, padded basis size, viscous): the padded size is the largest basis of the mesh rounded up
// to 4, 9, 16 or 25 (element orders up to 5 on triangles and squares)
// The file is compiled twice so that its halves build in parallel: as itself (FP_MIXED_PART 0: the
// kernels without artificial viscosity, and the entry points the host calls, which hand the damped
// cases on) and through fp_mixed_av.hip (FP_MIXED_PART 1: the kernels with it, entry points named _av).
#ifndef FP_MIXED_PART
#define FP_MIXED_PART 0
#endif
#define FP_MIXED_SHAPES(X, A)                                                                                        \
    X(MODEL_NAVIER_STOKES, 4, false, A) X(MODEL_NAVIER_STOKES, 9, false, A) X(MODEL_NAVIER_STOKES, 16, false, A)     \
    X(MODEL_NAVIER_STOKES, 25, false, A)                                                                             \
    X(MODEL_NAVIER_STOKES, 4, true, A) X(MODEL_NAVIER_STOKES, 9, true, A) X(MODEL_NAVIER_STOKES, 16, true, A)        \
    X(MODEL_NAVIER_STOKES, 25, true, A)                                                                              \
    X(MODEL_ADIABATIC, 4, false, A) X(MODEL_ADIABATIC, 9, false, A) X(MODEL_ADIABATIC, 16, false, A)                 \
    X(MODEL_ADIABATIC, 25, false, A)
#if FP_MIXED_PART == 0
#define FP_MIXED_CASES(X) FP_MIXED_SHAPES(X, false)
#define FP_MFN(name) name
#define FP_MAV(call) \
    if (L.av) return call;
bool mixed_supported_av(const ElemLaunch& L);
hipError_t launch_residual_mixed_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const MixedMesh& mm,
                                    const double* W, const double* Wo, const double* Wo2, double a1, double a2, double a3,
                                    double* R, double* RHS, double* enorm);
hipError_t launch_jacobian_mixed_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const MixedMesh& mm,
                                    const double* W, double a1, double* blocks, float* fblocks, const double* v, double* out,
                                    int mode);
hipError_t launch_jvp_mixed_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const MixedMesh& mm,
                               const double* W, const double* v, double a1, double* out);
hipError_t launch_av_limiter_mixed_av(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const MixedMesh& mm,
                                      const AvParams& Ap, const double* W, const double* esize, const double* trunc, double* av,
                                      double* old);
#else
#define FP_MIXED_CASES(X) FP_MIXED_SHAPES(X, true)
#define FP_MFN(name) name##_av
#define FP_MAV(call)
#endif
#define FP_MMATCH(MK, NB, V, A) (L.model == MK && L.nb == NB && (L.visc != 0) == V && (L.av != 0) == A)
// the kernels that do not depend on the damping: every shape, damped or not
#define FP_MSHAPE(MK, NB, V) (L.model == MK && L.nb == NB && (L.visc != 0) == V)

bool FP_MFN(mixed_supported)(const ElemLaunch& L) {
    FP_MAV(mixed_supported_av(L))
#define X(MK, NB, V, A) \
    if (FP_MMATCH(MK, NB, V, A)) return true;
    FP_MIXED_CASES(X)
#undef X
    return false;
}

hipError_t FP_MFN(launch_residual_mixed)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G,
                                         const MixedMesh& mm, const double* W, const double* Wo, const double* Wo2, double a1,
                                         double a2, double a3, double* R, double* RHS, double* enorm) {
    FP_MAV(launch_residual_mixed_av(L, s, P, G, mm, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm))
    dim3 blk(128), grd((mm.nElem + 127) / 128);
    const bool elem = mm.nqfm * 4 <= MJ_MAXF && mm.nqvm <= MJ_MAXV && mm.nElem > 0;
#define X(MK, NB, V, A)                                                                                                      \
    if (FP_MMATCH(MK, NB, V, A) && elem) {                                                                                   \
        hipLaunchKernelGGL((k_residual_mixed_elem<MK, NB, V, A>), dim3(mm.nElem), dim3(mixed_team_threads<NB * neq_of<MK>()>()), 0, s, \
                           P, G, mm, W, Wo, Wo2, a1, a2, \
                           a3, R, RHS, enorm);                                                                            \
        return hipGetLastError();                                                                                         \
    }                                                                                                                     \
    if (FP_MMATCH(MK, NB, V, A)) {                                                                                           \
        hipLaunchKernelGGL((k_residual_mixed<MK, NB, V, A>), grd, blk, 0, s, P, G, mm, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm); \
        return hipGetLastError();                                                                                         \
    }
    FP_MIXED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

// mode 0 (assembly): one workgroup per element (k_jacobian_mixed_elem, every column including the
// padded ones); mode 1 (J^T v): one thread per padded column
hipError_t FP_MFN(launch_jacobian_mixed)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G,
                                         const MixedMesh& mm, const double* W, double a1, double* blocks, float* fblocks,
                                         const double* v, double* out, int mode) {
    FP_MAV(launch_jacobian_mixed_av(L, s, P, G, mm, W, a1, blocks, fblocks, v, out, mode))
    if (mm.nElem == 0) return hipSuccess;
    if (mode == 0 && (mm.nqfm * 4 > MJ_MAXF || mm.nqvm > MJ_MAXV)) return hipErrorInvalidValue;   // orders <= 5 only
    const long long threads = (long long)mm.nElem * L.nb * L.neq;
    dim3 blk(64), grd((unsigned)((threads + 63) / 64));
#define X(MK, NB, V, A)                                                                                                   \
    if (FP_MMATCH(MK, NB, V, A)) {                                                                                        \
        if (mode == 0)                                                                                                 \
            hipLaunchKernelGGL((k_jacobian_mixed_elem<MK, NB, V, A>), dim3(mm.nElem), dim3(mixed_team_threads<NB * neq_of<MK>()>()), \
                               0, s, P, G, mm, W, a1,                                                                  \
                               blocks, fblocks);                                                                       \
        else                                                                                                           \
            hipLaunchKernelGGL((k_jacobian_mixed<MK, NB, V, A, double>), grd, blk, 0, s, P, G, mm, W, a1, blocks, fblocks, \
                               v, out, mode);                                                                          \
        return hipGetLastError();                                                                                      \
    }
    FP_MIXED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t FP_MFN(launch_jvp_mixed)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const MixedMesh& mm,
                                    const double* W, const double* v, double a1, double* out) {
    FP_MAV(launch_jvp_mixed_av(L, s, P, G, mm, W, v, a1, out))
    dim3 blk(64), grd((mm.nElem + 63) / 64);
#define X(MK, NB, V, A)                                                                              \
    if (FP_MMATCH(MK, NB, V, A)) {                                                                   \
        hipLaunchKernelGGL((k_jvp_mixed<MK, NB, V, A>), grd, blk, 0, s, P, G, mm, W, v, a1, out);   \
        return hipGetLastError();                                                                 \
    }
    FP_MIXED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

#if FP_MIXED_PART == 0
hipError_t launch_size_over_lambda_mixed(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const MixedMesh& mm, const double* W,
                                         const double* esize, double* out) {
    dim3 blk(256), grd((mm.nElem + 255) / 256);
#define X(MK, NB, V, A)                                                                                   \
    if (FP_MSHAPE(MK, NB, V)) {                                                                        \
        hipLaunchKernelGGL((k_size_over_lambda_mixed<MK, NB>), grd, blk, 0, s, P, mm, W, esize, out);  \
        return hipGetLastError();                                                                      \
    }
    FP_MIXED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}
#endif

hipError_t FP_MFN(launch_av_limiter_mixed)(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const MixedMesh& mm,
                                           const AvParams& Ap, const double* W, const double* esize, const double* trunc, double* av,
                                           double* old) {
    FP_MAV(launch_av_limiter_mixed_av(L, s, P, mm, Ap, W, esize, trunc, av, old))
    dim3 blk(128), grd((mm.nElem + 127) / 128);
#define X(MK, NB, V, A)                                                                                          \
    if (A && FP_MMATCH(MK, NB, V, A)) {                                                                          \
        hipLaunchKernelGGL((k_av_limiter_mixed<MK, NB>), grd, blk, 0, s, P, mm, Ap, W, esize, trunc, av, old);   \
        return hipGetLastError();                                                                                \
    }
    FP_MIXED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

#if FP_MIXED_PART == 0
hipError_t launch_positivity_mixed(const ElemLaunch& L, hipStream_t s, const MixedMesh& mm, double* W, int* count) {
    dim3 blk(128), grd((mm.nElem + 127) / 128);
#define X(MK, NB, V, A)                                                                      \
    if (FP_MSHAPE(MK, NB, V)) {                                                           \
        hipLaunchKernelGGL((k_positivity_mixed<MK, NB>), grd, blk, 0, s, mm, W, count);   \
        return hipGetLastError();                                                         \
    }
    FP_MIXED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_functional_mixed(const ElemLaunch& L, hipStream_t s, const MixedMesh& mm, const double* W, double kappa, int which,
                                   double* ival, double* grad) {
    if (L.model != MODEL_NAVIER_STOKES) return hipErrorInvalidValue;
    dim3 blk(128), grd((mm.nElem + 127) / 128);
    if (L.nb == 4) hipLaunchKernelGGL((k_functional_mixed<4>), grd, blk, 0, s, mm, W, kappa, which, ival, grad);
    else if (L.nb == 9) hipLaunchKernelGGL((k_functional_mixed<9>), grd, blk, 0, s, mm, W, kappa, which, ival, grad);
    else if (L.nb == 16) hipLaunchKernelGGL((k_functional_mixed<16>), grd, blk, 0, s, mm, W, kappa, which, ival, grad);
    else if (L.nb == 25) hipLaunchKernelGGL((k_functional_mixed<25>), grd, blk, 0, s, mm, W, kappa, which, ival, grad);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}
#endif

}  // namespace fp
