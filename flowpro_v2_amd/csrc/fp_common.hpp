// Shared device-side data structures of the implicit-DG path.
#pragma once
#include <hip/hip_runtime.h>

#include "fp_model.hpp"

namespace fp {

// Per-case constants handed to every element kernel by value.
//
// Element data live in HBM in the element order of the degrees of freedom
// (reverse Cuthill-McKee by default), element e owning DOFs [e*B, (e+1)*B) with
// the reference's local numbering m*Nb + j (equation-major inside an element):
//   geo  [e*GEO + ...]   inverse Jacobian iT00 iT01 iT10 iT11, |det J|/2, affine A0 A1 A2 B0 B1 B2
//   fgeo [(e*3+k)*4 + .] face normal nx ny, face Jacobian (edge length / 2), dL
//   nbr  [e*3+k]         neighbour element or boundary code (< 0)
//   nbrf [e*3+k]         neighbour's local face index of the shared face
//   blkn [e*3+k]         BSR block index of (row nbr, column e), -1 on boundary faces
//   blkd [e]             BSR block index of the diagonal block
//   wdist[e*3+v]         wall distance of vertex v (generateWallDistanceFunction), SA only
//   av   [e*(1+NEQ) + .] artificial viscosity eps and the equation-wise dampInner[m] of the
//                        last limiter call (DGpure.limiter), null when the case has no damping
// Reference-element tables (tab): basis and reference gradients at the volume
// points, per local face at the face points, weights, point coordinates and the
// reference mass matrix and its inverse.
constexpr int GEO = 12;

struct Tables {
    const double* vb;    // [NQV][NB]
    const double* vd;    // [NQV][NB][2]
    const double* vw;    // [NQV]
    const double* vx;    // [NQV][2]
    const double* fb;    // [3][NQF][NB]
    const double* fd;    // [3][NQF][NB][2]
    const double* fw;    // [NQF]
    const double* fx;    // [3][NQF][2]
    const double* mref;  // [NB][NB]
    const double* mrefinv;
    const double* fint;  // [NQF][2]: the face rule's interpolants (1 - t)/2, (1 + t)/2 of a face's two vertices
};

struct Mesh {
    int nElem;
    const double* geo;
    const double* fgeo;
    const int* nbr;
    const int* nbrf;
    const int* blkn;
    const int* blkd;
    const double* wdist;   // [e*3 + v] wall distance of the element's vertices (Spalart-Allmaras, k-omega SST)
    const double* av;      // [e*(1+NEQ) + .] eps, dampInner[0..NEQ) (null without damping)
};

// Straight tetrahedra (transformation3Dtetra: X = J xi + X0), element data in DOF order:
//   geo  [e*TGEO + ...]   iT (3x3 row-major, d xi_b / d x_a = iT[b][a]), |det J|, J (3x3 row-major), X0
//   fgeo [(e*4+k)*TFGEO]  outward unit normal (3), face Jacobian = triangle area, dL
//   nbr  [e*4+k]          neighbour element or boundary code (< 0)
//   blkn [e*4+k]          BSR block index of (row nbr, column e), -1 on boundary faces
// Reference tables: basis and reference gradients at the volume points, basis and reference
// coordinates at the face points of each local face, weights, reference mass matrix.
constexpr int TGEO = 22, TFGEO = 5;

struct TetTables {
    const double* vb;    // [NQV][NB]
    const double* vd;    // [NQV][NB][3]
    const double* vw;    // [NQV]
    const double* fb;    // [4][NQF][NB]
    const double* fw;    // [NQF]
    const double* fxr;   // [4][NQF][3]
    const double* mref;  // [NB][NB]
    const double* mrefinv;
};

struct TetMesh {
    int nElem;
    const double* geo;
    const double* fgeo;
    const int* nbr;
    const int* blkn;
    const int* blkd;
};

// Quadrilaterals (squareElement: basis2Dsquare, QuadratureSquare, transformation2Dsquare).  The
// bilinear map makes the Jacobian vary over the element, so the geometry is stored per point,
// element data in DOF order:
//   qv   [(e*NQV + q)*(1+2NB)]        |det J| w at volume point q, then the physical basis gradients
//                                     (d phi_j/dx, d phi_j/dy) pairs
//   qf   [((e*4 + k)*NQF + p)*2NB]    physical basis gradients at face point p of face k
//   qn   [((e*4 + k)*NQF + p)*3NB]    the neighbour's basis (NB) and physical gradients (2NB) at that
//                                     point, from the neighbour's inverse map (transformation2Dsquare.getXi)
//   xq   [(e*(NQV+4NQF) + i)*2]       physical coordinates of the volume points, then the face points
//   fgeo [(e*4 + k)*4]                face normal nx ny, face Jacobian (edge length / 2), dL
//   mass, minv [e*NB*NB]              element mass matrix and its inverse
//   is   [e*NB]                       Is_j = int phi_j / |K| (element mean = sum_j W_j Is_j)
//   nbr, nbrf, blkn [e*4 + k], blkd [e]  as for triangles
struct QuadTables {
    const double* vb;    // [NQV][NB]
    const double* vw;    // [NQV]
    const double* fb;    // [4][NQF][NB]
    const double* fw;    // [NQF]
};

struct QuadMesh {
    int nElem;
    const double* qv;
    const double* qf;
    const double* qn;
    const double* xq;
    const double* fgeo;
    const double* mass;
    const double* minv;
    const double* is;
    const int* nbr;
    const int* nbrf;
    const int* blkn;
    const int* blkd;
    const double* av;   // eps, dampInner[NEQ] per element (the artificial viscosity), or null
};

// Meshes that mix element types (triangles, quadrilaterals) and orders.  Every element is padded
// to the largest basis of the mesh, NBM functions: coefficient m*NBM + j of element e is the
// reference's m*nb_e + j for j < nb_e and zero otherwise; the padded unknowns are decoupled by an
// identity block in the mass matrix, so the padded system has the same solution and the
// factorisation the same real entries.  Per element (DOF order), all tables zero padded:
//   vol  [(e*NQVM + q)*(1 + 3NBM)]            |det J| w, basis (NBM), physical gradients (2NBM)
//   face [((e*4 + k)*NQFM + p)*(3 + 6NBM)]   JacobianFace w, X (2), own basis, own gradients,
//                                             the neighbour's basis and gradients at the point
//   fgeo [(e*4 + k)*4]                        nx, ny, 0, dL
//   nv, nqv, nbe [e]; nqf [e*4 + k]           faces, volume points, basis functions; face points
struct MixedMesh {
    int nElem, nqvm, nqfm;
    const int* nv;
    const int* nqv;
    const int* nbe;
    const int* nqf;
    const double* vol;
    const double* face;
    const double* fgeo;
    const double* mass;
    const double* minv;
    const double* is;
    const int* nbr;
    const int* blkn;
    const int* blkd;
    const double* av;   // eps, dampInner[NEQ] per element (the artificial viscosity), or null
};

// compile-time shape of the element kernels selected at run time
// Device-resident GMRES control block (Gmres.solve bookkeeping, Gmres.java:42-122): the Givens
// rotations, the convergence test and the triangular solve run on the device, and every kernel of
// an Arnoldi step returns at once when `stop` is set, so the host enqueues several steps without
// waiting for their Hessenberg columns.
struct GmresCtl {
    int conv;     // the relative residual reached tol (the solve is done)
    int stop;     // the current cycle is over (converged, or its last step done)
    int nused;    // Arnoldi vectors of the cycle's update
    int iters;    // Arnoldi steps performed in this solve
    double bnrm2; // |M^-1 b| (1 if zero)
    double error; // the last relative residual estimate
    double nr;    // |r| at the start of the cycle
    double tol;
};

struct ElemLaunch {
    int model, nb, neq, nqv, nqf, visc, av;
};

// vectors of one preconditioner application, read by the sweep kernels through a device
// pointer so that the launch sequence can be replayed from a graph for any vectors
struct IluIO {
    const double* b;
    double* x;
    // dataflow sweeps: the application's number (tags the published rows), the row tickets of
    // the forward and the backward sweep, and a flag set when a wait exceeded its spin limit
    unsigned epoch;
    unsigned ticket[2];
    int err;
};

struct DGParams {
    double beta0;   // DDG constant (0 when the model is diffusive and the global order is 1)
    double cIP;     // interior penalty (penalty when diffusive and order > 1, else 0)
    double dampConst;   // constant artificial viscosity
    int av;             // the damping terms are on (dampTol, dampConst or dampInnerTol set)
};

// DGpure.limiter parameters of the artificial viscosity (DGpure.initMethod)
struct AvParams {
    double dampTol, dampInnerTol;
    int nCoef;          // dampInnerCoef entries (0: lambda * elemSize / order is used)
    double coef[5];
    int order;          // element order k (the sensor's S0 = log10(1 / (k-1)^4))
};

// Artificial viscosity of DGpure (FlowPro/src/main/java/flowpro/core/element/DGpure.java:
// 176, 303-320), with eps and dampInner fixed by the last limiter call.  Interior faces:
// g += epsFace (avg(grad W) . n + beta0 [W] / dL), epsFace = mean eps + dampConst + mean
// dampInner[m]; volume: the flux loses (eps + dampConst + dampInner[m]) grad W_m.
template <int NEQ, class T>
__device__ inline void av_face_flux(const DGParams& G, const double* av, int e, int TT, const double* n, double dL, const T* WL,
                                    const T* WR, const T* dWL, const T* dWR, T* g) {
    const double* ae = av + (size_t)e * (1 + NEQ);
    const double* an = av + (size_t)TT * (1 + NEQ);
#pragma unroll
    for (int m = 0; m < NEQ; m++) {
        const double ef = 0.5 * (ae[0] + an[0]) + G.dampConst + 0.5 * (ae[1 + m] + an[1 + m]);
        T dWs = (dWL[m] + dWR[m]) / 2.0 * n[0];
        dWs += (dWL[NEQ + m] + dWR[NEQ + m]) / 2.0 * n[1];
        g[m] += ef * (dWs + G.beta0 * (WR[m] - WL[m]) / dL);
    }
}
template <int NEQ, class T>
__device__ inline void av_volume_flux(const DGParams& G, const double* av, int e, const T* dWq, T* fx, T* fy) {
    const double* ae = av + (size_t)e * (1 + NEQ);
#pragma unroll
    for (int m = 0; m < NEQ; m++) {
        const double c = ae[0] + G.dampConst + ae[1 + m];
        fx[m] -= c * dWq[m];
        fy[m] -= c * dWq[NEQ + m];
    }
}
template <int NEQ, class T>
__device__ inline void av_face_flux(const DGParams& G, const Mesh& ms, int e, int TT, const double* n, double dL, const T* WL,
                                    const T* WR, const T* dWL, const T* dWR, T* g) {
    av_face_flux<NEQ>(G, ms.av, e, TT, n, dL, WL, WR, dWL, dWR, g);
}
template <int NEQ, class T>
__device__ inline void av_volume_flux(const DGParams& G, const Mesh& ms, int e, const T* dWq, T* fx, T* fy) {
    av_volume_flux<NEQ>(G, ms.av, e, dWq, fx, fy);
}

// DGpure.shock_senzor / shock_senzor_all_eqn (FlowPro/src/main/java/flowpro/core/element/
// DGpure.java:403-503) for one equation: the L2 distance of the solution from its truncation
// (trunc = TrunOrd, the projection onto linear functions in the nodal basis, at order > 2; the
// mean otherwise) relative to its L2 norm, on a log10 scale, through a sine ramp of half-width kap
// around S0 = log10(1 / (k-1)^4).  w: the element's NB coefficients of the equation, phi(q): the
// NB basis values at volume point q of the nqv, jw(q): the volume weight of point q.
template <int NB, class PHI, class JW>
__device__ inline double shock_sensor_at(PHI phi, JW jw, int nqv, const double* w, double mean, const double* trunc, int order,
                                         double kap) {
    double tc[NB];
    if (order > 2) {
#pragma unroll
        for (int i = 0; i < NB; i++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; j++) s += trunc[i * NB + j] * w[j];
            tc[i] = s;
        }
    } else {
#pragma unroll
        for (int i = 0; i < NB; i++) tc[i] = mean;
    }
    double Se = 0.0, pod = 0.0;
    for (int q = 0; q < nqv; q++) {
        const double* b = phi(q);
        double vi = 0.0, vt = 0.0;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            vi += w[i] * b[i];
            vt += tc[i] * b[i];
        }
        const double c = jw(q);
        Se += c * (vi - vt) * (vi - vt);
        pod += c * vi * vi;
    }
    Se = log10(fabs(Se / pod));
    const double S0 = log10(1.0 / pow((double)(order - 1), 4.0));
    double s = 0.5 * (1.0 + sin(M_PI * (Se - S0) / (2.0 * kap)));
    if (Se < S0 - kap) s = 0.0;
    if (Se > S0 + kap) s = 1.0;
    return isnan(s) ? 0.0 : s;
}

// Element.positivityLimiter (FlowPro/src/main/java/flowpro/core/element/Element.java:385-476) with
// Aerodynamics.isAdmissible (:348-357); the L2 projection of the constant 1 onto the nodal basis is 1.
// AdiabaticEulers (three equations) admits every positive density (AdiabaticEulers.java:19-22)
template <int NEQ> __device__ inline bool admissible(const double* W) {
    if (!(W[0] > 1e-10)) return false;
    if constexpr (NEQ == 3) return true;
    double m2 = W[1] * W[1];
    m2 += W[2] * W[2];
    return W[3] - m2 / (2 * W[0]) > 1e-10;
}
template <int NEQ> __device__ inline double admissible_theta(const double* Wavg, const double* Wx) {
    if (admissible<NEQ>(Wx)) return 1.0;
    constexpr int NA = NEQ < 4 ? NEQ : 4;   // admissibility depends on rho, rho u, rho v (, rho E) only
    double lo = 0, hi = 1, Wt[NA];
    for (int it = 0; it < 30; it++) {
        double mid = 0.5 * (lo + hi);
        for (int m = 0; m < NA; m++) Wt[m] = Wavg[m] + mid * (Wx[m] - Wavg[m]);
        if (admissible<NEQ>(Wt)) lo = mid;
        else hi = mid;
    }
    return lo;
}
}  // namespace fp
