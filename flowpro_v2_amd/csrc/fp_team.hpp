// Element-team kernels of the implicit DG step (residual and exact Jacobian assembly).
//
// A workgroup owns EPB elements.  Phase 0 stages the reference tables, the elements'
// coefficients and their physical basis gradients in LDS; phase 1a interpolates the
// state (own and neighbour side) at every volume and face quadrature point; phase 1b
// evaluates the pointwise integrands (residual) or their derivatives with respect to
// the 3*nEqs local inputs (W, dW/dx, dW/dy) with forward-mode dual numbers (Jacobian);
// phase 2 contracts them with the test functions.  The arithmetic is the one of
// DGpure.residuum / residuumWall (FlowPro/src/main/java/flowpro/core/element/
// DGpure.java:108-355): faces first, then the volume; interior faces evaluated from
// both sides; the neighbour-side face integrand (KR) gives the off-diagonal blocks
// as BDF1.assembleJacobiMatrix does (BDF1.java:89-118), here exactly.
#pragma once
#include <hip/hip_runtime.h>

#include "fp_common.hpp"

namespace fp {

constexpr int TEAM_THREADS = 256;   // workgroup size of the team kernels

// The physical gradients of the basis are not staged: the teams contract on the reference gradients
// and apply the element map once per sum (d_x = iT^T d_xi), which leaves the LDS to more elements.
template <int NB, int NEQ, int NQV, int NQF>
struct TeamLayout {
    static constexpr int B = NB * NEQ;
    static constexpr int NQFT = 3 * NQF;                 // face points of an element
    static constexpr int NQ = NQV + NQFT;
    // shared tables
    static constexpr int T_VB = 0;                       // [NQV][NB]
    static constexpr int T_VW = T_VB + NQV * NB;         // [NQV]
    static constexpr int T_FB = T_VW + NQV;              // [3*NQF][NB]
    static constexpr int T_FW = T_FB + NQFT * NB;        // [NQF]
    static constexpr int T_VD = T_FW + NQF;              // [NQV][NB][2] reference gradients
    static constexpr int T_FD = T_VD + NQV * NB * 2;     // [3*NQF][NB][2]
    static constexpr int T_M = T_FD + NQFT * NB * 2;     // [NB][NB]
    static constexpr int T_MI = T_M + NB * NB;
    static constexpr int T_FX = T_MI + NB * NB;          // [3*NQF][2]
    static constexpr int T_VX = T_FX + NQFT * 2;         // [NQV][2]
    static constexpr int TAB = T_VX + NQV * 2;
    // per element
    static constexpr int E_W = 0;                        // [B]
    static constexpr int SV = 3 * NEQ;                   // W, dW/dx, dW/dy at a volume point
    static constexpr int SF = 6 * NEQ;                   // WL, dWL/dx, dWL/dy, WR, dWR/dx, dWR/dy at a face point
    static constexpr int E_SV = E_W + B;                 // [NQV][SV]
    static constexpr int E_SF = E_SV + NQV * SV;         // [3*NQF][SF]
    static constexpr int ELEM0 = E_SF + NQFT * SF;
};

// Task q of a loop over the (point, input) pairs of a team's elements: the volume points of every
// element come first, then the face points, so that a wave runs the volume or the face code, not
// both (element by element the two kinds of point would share every wave).  Each task still
// writes its own LDS slot: the order changes nothing in the results.
template <int NQV, int NQFT, int NIN>
__device__ inline void team_task(int q, int nel, int& el, int& qp, int& c) {
    if (q < nel * NQV * NIN) {
        el = q / (NQV * NIN);
        const int r = q % (NQV * NIN);
        qp = r / NIN;
        c = r % NIN;
    } else {
        const int q2 = q - nel * NQV * NIN;
        el = q2 / (NQFT * NIN);
        const int r = q2 % (NQFT * NIN);
        qp = NQV + r / NIN;
        c = r % NIN;
    }
}

// Task q of the Jacobian team's (point, input) loop: the value inputs (W_m) of the volume points, their
// gradient inputs (dW_m/dx, dW_m/dy), then the same two for the face points, so that a wave holds one
// kind of input at one kind of point: the waves of gradient inputs skip the convective fluxes.
template <int NQV, int NQFT, int NEQ, int NIN>
__device__ inline void jacobian_task(int q, int nel, int& el, int& qp, int& c) {
    constexpr int NGI = NIN - NEQ;
    const int a0 = nel * NQV * NEQ, a1 = a0 + nel * NQV * NGI, a2 = a1 + nel * NQFT * NEQ;
    int r, np, base, nin, c0;
    if (q < a0) { r = q; np = NQV; base = 0; nin = NEQ; c0 = 0; }
    else if (q < a1) { r = q - a0; np = NQV; base = 0; nin = NGI; c0 = NEQ; }
    else if (q < a2) { r = q - a1; np = NQFT; base = NQV; nin = NEQ; c0 = 0; }
    else { r = q - a2; np = NQFT; base = NQV; nin = NGI; c0 = NEQ; }
    el = r / (np * nin);
    const int rr = r % (np * nin);
    qp = base + rr / nin;
    c = c0 + rr % nin;
}

// ------------------------------------------------------------------ phase 0 / 1a helpers
template <int NB, int NEQ, int NQV, int NQF>
__device__ inline void team_stage(const Tables& tb, const Mesh& ms, const double* __restrict__ W, int e0, int nel, double* tab,
                                  double* eb, int elemStride) {
    using LY = TeamLayout<NB, NEQ, NQV, NQF>;
    constexpr int B = LY::B;
    const int tid = threadIdx.x, nt = blockDim.x;
    // one flat copy of the tables; the trip count is a compile-time constant so that all loads
    // of a thread are in flight before its LDS stores
    constexpr int NTAB = LY::TAB;
    constexpr int PER = (NTAB + TEAM_THREADS - 1) / TEAM_THREADS;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int q = tid + TEAM_THREADS * u;
        if (q < NTAB) {
            double v;
            if (q < LY::T_VW) v = tb.vb[q - LY::T_VB];
            else if (q < LY::T_FB) v = tb.vw[q - LY::T_VW];
            else if (q < LY::T_FW) v = tb.fb[q - LY::T_FB];
            else if (q < LY::T_VD) v = tb.fw[q - LY::T_FW];
            else if (q < LY::T_FD) v = tb.vd[q - LY::T_VD];
            else if (q < LY::T_M) v = tb.fd[q - LY::T_FD];
            else if (q < LY::T_MI) v = tb.mref[q - LY::T_M];
            else if (q < LY::T_FX) v = tb.mrefinv[q - LY::T_MI];
            else if (q < LY::T_VX) v = tb.fx[q - LY::T_FX];
            else v = tb.vx[q - LY::T_VX];
            tab[q] = v;
        }
    }
    for (int q = tid; q < nel * B; q += nt) {
        const int el = q / B, i = q % B;
        eb[el * elemStride + LY::E_W + i] = W[(size_t)(e0 + el) * B + i];
    }
    __syncthreads();
    // states at the quadrature points (own side from LDS, neighbour side from HBM)
    for (int q = tid; q < nel * LY::NQ; q += nt) {
        int el, qp, c0;
        team_task<NQV, LY::NQFT, 1>(q, nel, el, qp, c0);
        double* E = eb + el * elemStride;
        const double* w = E + LY::E_W;
        // own-side value and gradients at a point: d_x W = iT^T (sum_j W_j d_xi phi_j)
        auto own = [&](const double* phi, const double* dref, double* st) {
            const double* g = ms.geo + (size_t)(e0 + el) * GEO;
            const double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                double s = 0, sa = 0, sb = 0;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const double c = w[m * NB + j];
                    s += c * phi[j];
                    sa += c * dref[2 * j];
                    sb += c * dref[2 * j + 1];
                }
                st[m] = s;
                st[NEQ + m] = g0 * sa + g2 * sb;
                st[2 * NEQ + m] = g1 * sa + g3 * sb;
            }
        };
        if (qp < NQV) {
            own(&tab[LY::T_VB + qp * NB], &tab[LY::T_VD + qp * NB * 2], &E[LY::E_SV + qp * LY::SV]);
        } else {
            const int f = qp - NQV, k = f / NQF, p = f % NQF;
            const int e = e0 + el;
            double* st = &E[LY::E_SF + f * LY::SF];
            own(&tab[LY::T_FB + f * NB], &tab[LY::T_FD + f * NB * 2], st);
            const int TT = ms.nbr[e * 3 + k];
            if (TT > -1) {
                const int kr = ms.nbrf[e * 3 + k];
                const int pr = NQF - 1 - p;
                const double* gn = ms.geo + (size_t)TT * GEO;
                const double* phR = &tab[LY::T_FB + (kr * NQF + pr) * NB];
                const double* dR = &tab[LY::T_FD + (kr * NQF + pr) * NB * 2];
                const double* wn = W + (size_t)TT * B;
                double wv[B];
#pragma unroll
                for (int q2 = 0; q2 < B; q2++) wv[q2] = wn[q2];
#pragma unroll
                for (int m = 0; m < NEQ; m++) {
                    double s = 0, sa = 0, sb = 0;
#pragma unroll
                    for (int j = 0; j < NB; j++) {
                        const double c = wv[m * NB + j];
                        s += c * phR[j];
                        sa += c * dR[2 * j];
                        sb += c * dR[2 * j + 1];
                    }
                    const double sx = gn[0] * sa + gn[2] * sb, sy = gn[1] * sa + gn[3] * sb;
                    st[3 * NEQ + m] = s;
                    st[4 * NEQ + m] = sx;
                    st[5 * NEQ + m] = sy;
                }
            }
        }
    }
    __syncthreads();
}

// pointwise face integrand g (the reference's K update per face point divided by the
// test function): -fn + fvn - cIP [W] for the given own-side state.  GRADONLY (the tangent of a
// gradient input, the states' tangents zero): the convective flux and the penalty have zero tangents
// and are not evaluated; their only effect on the rest, the density clipping of limiteRho, is the
// identity above RHO_TOL, and below it they are evaluated as usual.
template <int MK, int NEQ, bool VISC, bool AV, class T, bool GRADONLY = false>
__device__ inline void face_integrand(const ModelParams& P, const DGParams& G, const Mesh& ms, int e, int TT, const double* n, double dL,
                                      const double* X, T* WL, T* dWL, const double* st, T* g, double D) {
    T WR[NEQ], dWR[2 * NEQ];
    if (TT > -1) {
#pragma unroll
        for (int m = 0; m < NEQ; m++) { WR[m] = st[3 * NEQ + m]; dWR[m] = st[4 * NEQ + m]; dWR[NEQ + m] = st[5 * NEQ + m]; }
    } else {
        boundary_value<MK>(P, WL, n, TT, X, WR);
#pragma unroll
        for (int m = 0; m < NEQ; m++) { dWR[m] = dWL[m]; dWR[NEQ + m] = dWL[NEQ + m]; }
    }
    const bool conv = !GRADONLY || value(WL[0]) < RHO_TOL || value(WR[0]) < RHO_TOL;
    if (conv) {
        T fn[NEQ];
        numerical_convective_flux<MK>(P, WL, WR, n, TT, fn);
#pragma unroll
        for (int m = 0; m < NEQ; m++) g[m] = -fn[m];
    } else {
#pragma unroll
        for (int m = 0; m < NEQ; m++) g[m] = 0.0;
    }
    if constexpr (VISC) {
        T Wc[NEQ], dWc[2 * NEQ], fvn[NEQ];
#pragma unroll
        for (int m = 0; m < NEQ; m++) {
            Wc[m] = TT > -1 ? (WL[m] + WR[m]) / 2.0 : WR[m];
            T jump = G.beta0 * (WR[m] - WL[m]) / dL;
            dWc[m] = (dWL[m] + dWR[m]) / 2.0 + jump * n[0];
            dWc[NEQ + m] = (dWL[NEQ + m] + dWR[NEQ + m]) / 2.0 + jump * n[1];
        }
        numerical_diffusive_flux<MK>(P, Wc, dWc, n, TT, fvn, D);
        const bool ip = conv && (TT > -1 || is_ip_face<MK>(TT));
#pragma unroll
        for (int m = 0; m < NEQ; m++) {
            g[m] += fvn[m];
            if (ip) g[m] -= G.cIP * (WL[m] - WR[m]);
        }
    }
    if constexpr (AV) {
        if (TT > -1) av_face_flux<NEQ>(G, ms, e, TT, n, dL, WL, WR, dWL, dWR, g);
    }
}

// number of volume integrand slots: fluxes in x and y, plus the state-dependent source of
// the Spalart-Allmaras and k-omega SST models (contracted with the test function itself)
template <int MK> constexpr int volume_slots() { return has_rans_source<MK>() ? 3 : 2; }

// pointwise volume fluxes (convective minus viscous) in x and y and the model source, called
// in the order of DGpure.residuum (convection x, y, diffusion x, y, source): each call clips
// the density of the shared state as the reference does.  D: wall distance at the point.
template <int MK, int NB, int NEQ, bool VISC, bool AV, class T, bool GRADONLY = false>
__device__ inline void volume_integrand(const ModelParams& P, const DGParams& G, const Mesh& ms, int e, T* Wq, T* dWq, double D, T* fx,
                                        T* fy, T* sv) {
    if constexpr (NB > 1) {
        const double ex[2] = {1.0, 0.0}, ey[2] = {0.0, 1.0};
        // GRADONLY: the convective fluxes have zero tangents (see face_integrand)
        if (!GRADONLY || value(Wq[0]) < RHO_TOL) {
            convective_flux<MK>(P, Wq, ex, fx);
            convective_flux<MK>(P, Wq, ey, fy);
        } else {
#pragma unroll
            for (int m = 0; m < NEQ; m++) { fx[m] = 0.0; fy[m] = 0.0; }
        }
        if constexpr (VISC) {
            T gx[NEQ], gy[NEQ];
            diffusive_flux<MK>(P, Wq, dWq, ex, gx, D);
            diffusive_flux<MK>(P, Wq, dWq, ey, gy, D);
#pragma unroll
            for (int m = 0; m < NEQ; m++) { fx[m] -= gx[m]; fy[m] -= gy[m]; }
        }
        if constexpr (AV) av_volume_flux<NEQ>(G, ms, e, dWq, fx, fy);
    } else {
#pragma unroll
        for (int m = 0; m < NEQ; m++) { fx[m] = 0.0; fy[m] = 0.0; }
    }
    if constexpr (MK == MODEL_SA) sa_source(P, Wq, dWq, D, sv);
    if constexpr (MK == MODEL_KOMEGA) kw_source(P, Wq, dWq, D, sv);
}

// wall distance at reference point xi: the vertex distances interpolated linearly
// (Element.interpolateVelocityAndFillElementDataObjectOnVolume, Element.java:263-275)
__device__ inline double wall_distance_at(const Mesh& ms, int e, const double* xi) {
    const double* w = ms.wdist + (size_t)e * 3;
    return (1.0 - xi[0] - xi[1]) * w[0] + xi[0] * w[1] + xi[1] * w[2];
}
// wall distance at point p of face k: the distances of the face's two vertices interpolated with the
// face rule's 1D interpolants (Element.interpolateVelocityAndFillElementDataObjectOnFace, Element.java:299-312)
__device__ inline double face_wall_distance(const Tables& tb, const Mesh& ms, int e, int k, int p) {
    const double* w = ms.wdist + (size_t)e * 3;
    return tb.fint[2 * p] * w[k] + tb.fint[2 * p + 1] * w[k == 2 ? 0 : k + 1];
}

// elements per team whose staged states and EXTRA doubles each fit in KB KiB of LDS beside the
// tables (1 to 16)
template <int NB, int NEQ, int NQV, int NQF, int EXTRA, int KB>
struct TeamEPB {
    using LY = TeamLayout<NB, NEQ, NQV, NQF>;
    static constexpr int PER = LY::ELEM0 + EXTRA;
    static constexpr int BUDGET = KB * 1024 / 8;   // doubles of LDS per workgroup
    static constexpr int E0 = (BUDGET - LY::TAB) / PER;
    static constexpr int value = E0 < 1 ? 1 : (E0 > 16 ? 16 : E0);
};

// ------------------------------------------------------------------ residual
// LDS per element of the residual team beside the staged states: the integrands (fluxes and source
// at the volume points, the face integrand at the face points), K and the right-hand side
template <int MK, int NB, int NEQ, int NQV, int NQF>
constexpr int residual_team_extra() {
    return NQV * volume_slots<MK>() * NEQ + 3 * NQF * NEQ + 2 * NB * NEQ;
}
// Elements per residual team: one pass of the team's threads over the quadrature points of its
// elements (phase 1b, the flux evaluations) and at most 76 KiB of LDS, so that two teams share a CU;
// the registers (up to 255 per lane, no spills) hold the CU to two teams anyway.  Measured on MI355X at
// NS p=3, 200,000 triangles (tasks in element order): 4 elements per team with the physical gradients
// staged and 3 teams per CU (88 VGPRs spilled) 1.23 ms, 8 elements 0.77 ms, 9 elements 0.71 ms; with
// the volume tasks first 0.68 ms.
constexpr int RESIDUAL_LDS_KB = 76;
template <int MK, int NB, int NEQ, int NQV, int NQF>
constexpr int residual_team_epb() {
    constexpr int fit = TeamEPB<NB, NEQ, NQV, NQF, residual_team_extra<MK, NB, NEQ, NQV, NQF>(), RESIDUAL_LDS_KB>::value;
    constexpr int pass = TEAM_THREADS / (NQV + 3 * NQF);
    return fit < pass ? fit : (pass < 1 ? 1 : pass);
}

// LDS per element of the Jacobian team beside the staged states: the derivative of every integrand
// at every point with respect to each of the NIN inputs of the point; elements per team within
// 76 KiB of LDS (two teams per CU, the registers allow no more).  NS p=3 at 200,000 triangles on
// MI355X: 2 elements per team with the physical gradients staged 7.8 ms (J^T v 6.2 ms), 3 elements
// without them 7.0 ms (5.6 ms).
template <int MK, int NB, int NEQ, int NQV, int NQF, bool GRAD>
constexpr int jacobian_team_extra() {
    return (NQV * volume_slots<MK>() + 3 * NQF) * (GRAD ? 3 * NEQ : NEQ) * NEQ;
}
template <int MK, int NB, int NEQ, int NQV, int NQF, bool GRAD>
constexpr int jacobian_team_epb() {
    // the transposed product adds its per-(column, equation) partial sums
    return TeamEPB<NB, NEQ, NQV, NQF, jacobian_team_extra<MK, NB, NEQ, NQV, NQF, GRAD>() + NB * NEQ * NEQ, 76>::value;
}

// Launch bounds: two teams per CU for the residual (255 registers, no spills) and for the Jacobian;
// held to 3 waves per SIMD the residual spilled, and the Jacobian's doubled occupancy over spill-free
// single-wave code hides the LDS and HBM latency better (measured on MI355X: NS p=3 Jacobian 12.7 ->
// 10.7 ms, SA p=2 23.2 -> 4.0 ms at 96k elements).
template <int MK, int NB, int NEQ, int NQV, int NQF, bool VISC, bool AV>
__global__ void __launch_bounds__(256, 2) k_residual_team(ModelParams P, DGParams G, Tables tb, Mesh ms, const double* __restrict__ W,
                                                       const double* __restrict__ Wo, const double* __restrict__ Wo2, double a1,
                                                       double a2, double a3, double* __restrict__ R, double* __restrict__ RHS,
                                                       double* __restrict__ enorm) {
    using LY = TeamLayout<NB, NEQ, NQV, NQF>;
    constexpr int B = LY::B;
    constexpr int NVS = volume_slots<MK>();
    constexpr int EXTRA = residual_team_extra<MK, NB, NEQ, NQV, NQF>();
    constexpr int EPB = residual_team_epb<MK, NB, NEQ, NQV, NQF>();
    constexpr int STRIDE = LY::ELEM0 + EXTRA;
    __shared__ double tab[LY::TAB];
    __shared__ double eb[EPB * STRIDE];
    const int e0 = blockIdx.x * EPB;
    const int nel = min(EPB, ms.nElem - e0);
    if (nel <= 0) return;
    team_stage<NB, NEQ, NQV, NQF>(tb, ms, W, e0, nel, tab, eb, STRIDE);
    const int tid = threadIdx.x, nt = blockDim.x;
    constexpr int O_IV = LY::ELEM0, O_IF = O_IV + NQV * NVS * NEQ, O_K = O_IF + LY::NQFT * NEQ, O_H = O_K + B;
    // phase 1b: integrands at every quadrature point, weighted by J w
    for (int q = tid; q < nel * LY::NQ; q += nt) {
        int el, qp, c0;
        team_task<NQV, LY::NQFT, 1>(q, nel, el, qp, c0);
        const int e = e0 + el;
        double* E = eb + el * STRIDE;
        const double* ge = ms.geo + (size_t)e * GEO;
        if (qp < NQV) {
            const double* st = &E[LY::E_SV + qp * LY::SV];
            double Wq[NEQ], dWq[2 * NEQ], fx[NEQ], fy[NEQ], sv[NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) { Wq[m] = st[m]; dWq[m] = st[NEQ + m]; dWq[NEQ + m] = st[2 * NEQ + m]; }
            const double D = has_rans_source<MK>() ? wall_distance_at(ms, e, &tab[LY::T_VX + 2 * qp]) : 0.0;
            volume_integrand<MK, NB, NEQ, VISC, AV>(P, G, ms, e, Wq, dWq, D, fx, fy, sv);
            const double jw = ge[4] * tab[LY::T_VW + qp];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                E[O_IV + (qp * NVS) * NEQ + m] = jw * fx[m];
                E[O_IV + (qp * NVS + 1) * NEQ + m] = jw * fy[m];
                if constexpr (NVS == 3) E[O_IV + (qp * NVS + 2) * NEQ + m] = jw * sv[m];
            }
        } else {
            const int f = qp - NQV, k = f / NQF, p = f % NQF;
            const int TT = ms.nbr[e * 3 + k];
            const double* fg = ms.fgeo + ((size_t)e * 3 + k) * 4;
            const double n[2] = {fg[0], fg[1]};
            const double* st = &E[LY::E_SF + f * LY::SF];
            double WL[NEQ], dWL[2 * NEQ], g[NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) { WL[m] = st[m]; dWL[m] = st[NEQ + m]; dWL[NEQ + m] = st[2 * NEQ + m]; }
            const double* xi = &tab[LY::T_FX + 2 * f];
            double X[2] = {ge[5] * xi[0] + ge[6] * xi[1] + ge[7], ge[8] * xi[0] + ge[9] * xi[1] + ge[10]};
            const double Df = MK == MODEL_KOMEGA ? face_wall_distance(tb, ms, e, k, p) : 0.0;
            face_integrand<MK, NEQ, VISC, AV>(P, G, ms, e, TT, n, fg[3], X, WL, dWL, st, g, Df);
            const double jw = fg[2] * tab[LY::T_FW + p];
#pragma unroll
            for (int m = 0; m < NEQ; m++) E[O_IF + f * NEQ + m] = jw * g[m];
        }
    }
    __syncthreads();
    // phase 2: test-function contraction, faces first (DGpure.residuum order), then RHS
    for (int q = tid; q < nel * B; q += nt) {
        const int el = q / B, i = q % B, m = i / NB, j = i % NB, e = e0 + el;
        double* E = eb + el * STRIDE;
        double K = 0.0;
        for (int f = 0; f < LY::NQFT; f++) K += tab[LY::T_FB + f * NB + j] * E[O_IF + f * NEQ + m];
        if constexpr (NB > 1) {
            // sum_q fx d_x phi_j + fy d_y phi_j with d_x phi = iT^T d_xi phi: contracted on the reference
            // gradients, mapped once
            double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
            for (int qv = 0; qv < NQV; qv++) {
                const double fx = E[O_IV + (qv * NVS) * NEQ + m], fy = E[O_IV + (qv * NVS + 1) * NEQ + m];
                const double dx = tab[LY::T_VD + (qv * NB + j) * 2], dy = tab[LY::T_VD + (qv * NB + j) * 2 + 1];
                a += fx * dx;
                b += fx * dy;
                c += fy * dx;
                d += fy * dy;
            }
            const double* g = ms.geo + (size_t)e * GEO;
            K += g[0] * a + g[2] * b + g[1] * c + g[3] * d;
        }
        if constexpr (NVS == 3) {
            for (int qv = 0; qv < NQV; qv++) K += E[O_IV + (qv * NVS + 2) * NEQ + m] * tab[LY::T_VB + qv * NB + j];
        }
        if constexpr (MK == MODEL_SCALAR) {
            const double* ge = ms.geo + (size_t)e * GEO;
            for (int qv = 0; qv < NQV; qv++) {
                const double* xi = &tab[LY::T_VX + 2 * qv];
                double X[2] = {ge[5] * xi[0] + ge[6] * xi[1] + ge[7], ge[8] * xi[0] + ge[9] * xi[1] + ge[10]};
                double src[1];
                source_term(P, X, src);
                K += ge[4] * tab[LY::T_VW + qv] * src[0] * tab[LY::T_VB + qv * NB + j];
            }
        }
        const double jac = ms.geo[(size_t)e * GEO + 4];
        double s = 0.0;
        for (int jj = 0; jj < NB; jj++) {
            const size_t o = (size_t)e * B + m * NB + jj;
            const double mij = jac * tab[LY::T_M + j * NB + jj];
            s += mij * a1 * E[LY::E_W + m * NB + jj] + mij * a2 * Wo[o] + mij * a3 * Wo2[o];
        }
        R[(size_t)e * B + i] = K;
        RHS[(size_t)e * B + i] = K - s;
        E[O_K + i] = K;
        E[O_H + i] = K - s;
    }
    if (!enorm) return;
    __syncthreads();
    // per-element R_m^T M^-1 R_m and RHS_m^T M^-1 RHS_m (Element.l2NormsSquared)
    for (int q = tid; q < nel * NEQ * 2; q += nt) {
        const int el = q / (NEQ * 2), c = q % (NEQ * 2), m = c % NEQ, which = c / NEQ, e = e0 + el;
        const double* v = eb + el * STRIDE + (which ? O_H : O_K) + m * NB;
        double sum = 0.0;
        for (int a = 0; a < NB; a++) {
            double t = 0.0;
            for (int b = 0; b < NB; b++) t += tab[LY::T_MI + a * NB + b] * v[b];
            sum += v[a] * t;
        }
        enorm[(size_t)e * 2 * NEQ + c] = sum / ms.geo[(size_t)e * GEO + 4];
    }
}

// ------------------------------------------------------------------ Jacobian
// NIN local inputs per point: the value (and with viscosity the x and y derivatives) of
// every equation.  Phase 1b stores d(J w * integrand)/d(input) for every point and input;
// phase 2 thread (element, column i = (m', j')) builds column i of the diagonal block and
// of the three neighbour blocks.  JTV: the columns are not stored but contracted with v, which
// gives the transposed product (a1 M - dR/dW)^T v without any write to another element's rows:
// column i of element e meets the rows of e (diagonal block) and of its face neighbours (the
// neighbour-side blocks its own face integrand yields), so out[e] is complete in e's team.
template <int MK, int NB, int NEQ, int NQV, int NQF, bool VISC, bool AV, bool JTV>
__global__ void __launch_bounds__(256, 2) k_jacobian_team(ModelParams P, DGParams G, Tables tb, Mesh ms, const double* __restrict__ W,
                                                       double a1, double* __restrict__ blocks, float* __restrict__ fblocks,
                                                       const double* __restrict__ v, double* __restrict__ out) {
    using LY = TeamLayout<NB, NEQ, NQV, NQF>;
    constexpr int B = LY::B;
    constexpr int NIN = (VISC || AV) ? 3 * NEQ : NEQ;   // the damping terms see the gradients too
    constexpr int NVS = volume_slots<MK>();
    constexpr int EXTRA = jacobian_team_extra<MK, NB, NEQ, NQV, NQF, VISC || AV>();
    constexpr int EPB = jacobian_team_epb<MK, NB, NEQ, NQV, NQF, VISC || AV>();
    constexpr int STRIDE = LY::ELEM0 + EXTRA;
    constexpr int O_DV = LY::ELEM0, O_DF = O_DV + NQV * NIN * NVS * NEQ;
    __shared__ double tab[LY::TAB];
    __shared__ double eb[EPB * STRIDE];
    __shared__ double red[JTV ? EPB * B * NEQ : 1];   // JTV: per (element, column, equation) partial sums
    const int e0 = blockIdx.x * EPB;
    const int nel = min(EPB, ms.nElem - e0);
    if (nel <= 0) return;
    team_stage<NB, NEQ, NQV, NQF>(tb, ms, W, e0, nel, tab, eb, STRIDE);
    const int tid = threadIdx.x, nt = blockDim.x;
    // phase 1b: tangents of the integrands, one (point, input) pair per task
    for (int q = tid; q < nel * LY::NQ * NIN; q += nt) {
        int el, qp, c;
        jacobian_task<NQV, LY::NQFT, NEQ, NIN>(q, nel, el, qp, c);
        const int e = e0 + el;
        // uniform over a wave but at section boundaries; NavierStokes only: with the Spalart-Allmaras
        // integrands the second copy of the code cost more than the skipped fluxes saved (SA p=2 on 500k
        // triangles 19.0 ms against 17.3 ms without the skip)
        const bool gradin = MK == MODEL_NAVIER_STOKES && NIN > NEQ && c >= NEQ;
        const int s = c / NEQ, mp = c % NEQ;
        double* E = eb + el * STRIDE;
        const double* ge = ms.geo + (size_t)e * GEO;
        if (qp < NQV) {
            const double* st = &E[LY::E_SV + qp * LY::SV];
            dual Wq[NEQ], dWq[2 * NEQ], fx[NEQ], fy[NEQ], sv[NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                Wq[m] = dual(st[m], (s == 0 && m == mp) ? 1.0 : 0.0);
                dWq[m] = dual(st[NEQ + m], (s == 1 && m == mp) ? 1.0 : 0.0);
                dWq[NEQ + m] = dual(st[2 * NEQ + m], (s == 2 && m == mp) ? 1.0 : 0.0);
            }
            const double D = has_rans_source<MK>() ? wall_distance_at(ms, e, &tab[LY::T_VX + 2 * qp]) : 0.0;
            if (gradin) volume_integrand<MK, NB, NEQ, VISC, AV, dual, true>(P, G, ms, e, Wq, dWq, D, fx, fy, sv);
            else volume_integrand<MK, NB, NEQ, VISC, AV, dual, false>(P, G, ms, e, Wq, dWq, D, fx, fy, sv);
            const double jw = ge[4] * tab[LY::T_VW + qp];
            double* out = &E[O_DV + ((qp * NIN + c) * NVS) * NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                out[m] = jw * fx[m].d;
                out[NEQ + m] = jw * fy[m].d;
                if constexpr (NVS == 3) out[2 * NEQ + m] = jw * sv[m].d;
            }
        } else {
            const int f = qp - NQV, k = f / NQF, p = f % NQF;
            const int TT = ms.nbr[e * 3 + k];
            const double* fg = ms.fgeo + ((size_t)e * 3 + k) * 4;
            const double n[2] = {fg[0], fg[1]};
            const double* st = &E[LY::E_SF + f * LY::SF];
            dual WL[NEQ], dWL[2 * NEQ], g[NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) {
                WL[m] = dual(st[m], (s == 0 && m == mp) ? 1.0 : 0.0);
                dWL[m] = dual(st[NEQ + m], (s == 1 && m == mp) ? 1.0 : 0.0);
                dWL[NEQ + m] = dual(st[2 * NEQ + m], (s == 2 && m == mp) ? 1.0 : 0.0);
            }
            const double* xi = &tab[LY::T_FX + 2 * f];
            double X[2] = {ge[5] * xi[0] + ge[6] * xi[1] + ge[7], ge[8] * xi[0] + ge[9] * xi[1] + ge[10]};
            const double Df = MK == MODEL_KOMEGA ? face_wall_distance(tb, ms, e, k, p) : 0.0;
            if (gradin) face_integrand<MK, NEQ, VISC, AV, dual, true>(P, G, ms, e, TT, n, fg[3], X, WL, dWL, st, g, Df);
            else face_integrand<MK, NEQ, VISC, AV, dual, false>(P, G, ms, e, TT, n, fg[3], X, WL, dWL, st, g, Df);
            const double jw = fg[2] * tab[LY::T_FW + p];
            double* out = &E[O_DF + (f * NIN + c) * NEQ];
#pragma unroll
            for (int m = 0; m < NEQ; m++) out[m] = jw * g[m].d;
        }
    }
    __syncthreads();
    // phase 2: thread (element, column i = (mp, jp), equation m): the NB rows of equation m in
    // column i of the diagonal block and of the three neighbour blocks.  One equation per thread
    // keeps the accumulators small and gives the contraction NEQ times the threads of a
    // column-per-thread split (the element count per workgroup is small: the LDS holds the
    // point derivatives); consecutive threads write consecutive rows of a column.
    const size_t BB = (size_t)B * B;
    for (int q = tid; q < nel * B * NEQ; q += nt) {
        const int el = q / (B * NEQ), r2 = q % (B * NEQ), i = r2 / NEQ, m = r2 % NEQ;
        const int mp = i / NB, jp = i % NB, e = e0 + el;
        const double* E = eb + el * STRIDE;
        // the element map: d_x psi = g0 d_xi psi + g2 d_eta psi, d_y psi = g1 d_xi psi + g3 d_eta psi
        const double* ge = ms.geo + (size_t)e * GEO;
        const double g0 = ge[0], g1 = ge[1], g2 = ge[2], g3 = ge[3];
        double acc[NB];
        double part = 0.0;   // JTV
#pragma unroll
        for (int j = 0; j < NB; j++) acc[j] = 0.0;
#pragma unroll 1
        for (int k = 0; k < 3; k++) {
            const int blkn = ms.blkn[e * 3 + k];
            const int kr = ms.nbrf[e * 3 + k];
            double accR[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) accR[j] = 0.0;
#pragma unroll 1
            for (int p = 0; p < NQF; p++) {
                const int f = k * NQF + p;
                // t = sum_s d g_m / d u_(s, mp) * psi_s(jp)
                double t = E[O_DF + (f * NIN + mp) * NEQ + m] * tab[LY::T_FB + f * NB + jp];
                if constexpr (VISC || AV) {
                    const double rx = tab[LY::T_FD + (f * NB + jp) * 2], ry = tab[LY::T_FD + (f * NB + jp) * 2 + 1];
                    const double psx = g0 * rx + g2 * ry, psy = g1 * rx + g3 * ry;
                    t += E[O_DF + (f * NIN + NEQ + mp) * NEQ + m] * psx + E[O_DF + (f * NIN + 2 * NEQ + mp) * NEQ + m] * psy;
                }
#pragma unroll
                for (int j = 0; j < NB; j++) acc[j] += tab[LY::T_FB + f * NB + j] * t;
                if (blkn >= 0) {
                    const double* phR = &tab[LY::T_FB + (kr * NQF + (NQF - 1 - p)) * NB];
#pragma unroll
                    for (int j = 0; j < NB; j++) accR[j] += phR[j] * t;
                }
            }
            if (blkn >= 0) {
                if constexpr (JTV) {
                    const double* vt = v + (size_t)ms.nbr[e * 3 + k] * B + (size_t)m * NB;
#pragma unroll
                    for (int j = 0; j < NB; j++) part += accR[j] * vt[j];
                } else {
                    const size_t o0 = (size_t)blkn * BB + (size_t)i * B + (size_t)m * NB;
#pragma unroll
                    for (int j = 0; j < NB; j++) blocks[o0 + j] = accR[j];
                    if (fblocks) {   // the FP32 copy the block-ILU0 factorisation starts from
#pragma unroll
                        for (int j = 0; j < NB; j++) fblocks[o0 + j] = (float)accR[j];
                    }
                }
            }
        }
        if constexpr (NB > 1) {
#pragma unroll 1
            for (int qv = 0; qv < NQV; qv++) {
                const double psi0 = tab[LY::T_VB + qv * NB + jp];
                const double* d0 = &E[O_DV + ((qv * NIN + mp) * NVS) * NEQ];
                double tx = d0[m] * psi0, ty = d0[NEQ + m] * psi0, ts = 0.0;
                if constexpr (NVS == 3) ts = d0[2 * NEQ + m] * psi0;
                if constexpr (VISC || AV) {
                    const double rx = tab[LY::T_VD + (qv * NB + jp) * 2], ry = tab[LY::T_VD + (qv * NB + jp) * 2 + 1];
                    const double psx = g0 * rx + g2 * ry, psy = g1 * rx + g3 * ry;
                    const double* d1 = &E[O_DV + ((qv * NIN + NEQ + mp) * NVS) * NEQ];
                    const double* d2 = &E[O_DV + ((qv * NIN + 2 * NEQ + mp) * NVS) * NEQ];
                    tx += d1[m] * psx + d2[m] * psy;
                    ty += d1[NEQ + m] * psx + d2[NEQ + m] * psy;
                    if constexpr (NVS == 3) ts += d1[2 * NEQ + m] * psx + d2[2 * NEQ + m] * psy;
                }
                // tx d_x psi_j + ty d_y psi_j on the reference gradients
                const double ta = g0 * tx + g1 * ty, tb2 = g2 * tx + g3 * ty;
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    acc[j] += ta * tab[LY::T_VD + (qv * NB + j) * 2] + tb2 * tab[LY::T_VD + (qv * NB + j) * 2 + 1];
                    if constexpr (NVS == 3) acc[j] += ts * tab[LY::T_VB + qv * NB + j];
                }
            }
        }
        const double jac = ms.geo[(size_t)e * GEO + 4];
        if constexpr (JTV) {
            const double* ve = v + (size_t)e * B + (size_t)m * NB;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                double a = -acc[j];
                if (m == mp) a += a1 * (jac * tab[LY::T_M + jp * NB + j]);
                part += a * ve[j];
            }
            red[(el * B + i) * NEQ + m] = part;
        } else {
            const size_t o0 = (size_t)ms.blkd[e] * BB + (size_t)i * B + (size_t)m * NB;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                double a = -acc[j];
                if (m == mp) a += a1 * (jac * tab[LY::T_M + jp * NB + j]);
                blocks[o0 + j] = a;
                if (fblocks) fblocks[o0 + j] = (float)a;
            }
        }
    }
    if constexpr (JTV) {
        __syncthreads();
        for (int q = tid; q < nel * B; q += nt) {
            double sum = 0.0;
#pragma unroll
            for (int m = 0; m < NEQ; m++) sum += red[q * NEQ + m];
            out[(size_t)e0 * B + q] = sum;
        }
    }
}

// the k-omega SST cases (model, Nb, nEqs, volume points, face points, viscous, artificial viscosity)
#define FP_TEAM_KW_CASES(X)                  \
    X(MODEL_KOMEGA, 3, 6, 6, 2, true, false) \
    X(MODEL_KOMEGA, 6, 6, 7, 3, true, false) \
    X(MODEL_KOMEGA, 3, 6, 6, 2, true, true)  \
    X(MODEL_KOMEGA, 6, 6, 7, 3, true, true)

// the k-omega SST team kernels, compiled in a translation unit of their own (fp_team_kw.hip)
hipError_t launch_residual_team_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                   const Mesh& ms, const double* W, const double* Wo, const double* Wo2, double a1, double a2,
                                   double a3, double* R, double* RHS, double* enorm);
hipError_t launch_jacobian_team_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                   const Mesh& ms, const double* W, double a1, double* blocks, float* fblocks);
hipError_t launch_jtv_team_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                              const Mesh& ms, const double* W, const double* v, double a1, double* out);

}  // namespace fp
