// Oracle (test infrastructure only): models, DGpure residual, BDF assembly
// (FD as in the reference, or exact with dual numbers), J*v and J^T*v.
#include <array>
#include <cstring>

#include "oracle_core.hpp"
#include "oracle_model.hpp"

namespace oracle {

// ------------------------------------------------------------- DGpure residual
// C/element/DGpure.java:195-355 residuumWall.  Wown = W + V of this element,
// Wnb = neighbour state (nullptr on boundary faces).  KR may be nullptr.
template <class T>
static void residuumWall(const Case& c, const Elem& e, int k, const T* Wown, const T* Wnb, T* K, T* KR) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs, dim = 2, nBasis = e.nBasis;
    const int TT = e.TT[k];
    const int nR = TT > -1 ? c.E[TT].nBasis : nBasis;   // the neighbour's basis size
    if (KR) for (int j = 0; j < nR * nEqs; j++) KR[j] = 0.0;
    const bool convective = true;
    const bool diffusive = M.diffusive;
    double beta0 = c.P.beta0;
    if (diffusive && c.P.order == 1) beta0 = 0;   // DGpure.java:287-289, inside the isDiffusive branch
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
            int ei[2] = {k, (k + 1) % e.nF};
            double cXi[2] = {0, 0};
            q.meshVelocity[0] = q.meshVelocity[1] = 0;
            q.wallDistance = 0;
            for (int j = 0; j < 2; j++) {
                double ip = e.interpF[k][2 * p + j];
                for (int d = 0; d < dim; d++) cXi[d] += ip * elem_ref_vertex(e, ei[j])[d];
                q.wallDistance += ip * e.wallDist[ei[j]];
            }
            elem_getX(e, cXi, q.currentX);
        }
        T WL[MAXEQ], WR[MAXEQ], dWL[2 * MAXEQ], dWR[2 * MAXEQ];
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
        T Wale[MAXEQ];
        T fn[MAXEQ];
        if (convective) {
            for (int d = 0; d < dim; d++) vn = vn + q.meshVelocity[d] * nn[d];
            if (TT > -1)
                for (int j = 0; j < nEqs; j++) Wale[j] = (WL[j] + WR[j]) / 2;
            else
                for (int j = 0; j < nEqs; j++) Wale[j] = WR[j];
            numericalConvectiveFlux(M, WL, WR, nn, TT, q, fn);
        }
        T fvn[MAXEQ];
        if (diffusive) {
            T Wc[MAXEQ], dWc[2 * MAXEQ];
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
                double epsFace = 0.5 * (e.eps + er.eps) + c.P.damp_const + 0.5 * (e.dampInner[m] + er.dampInner[m]);
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
    for (int k = 0; k < e.nF; k++) residuumWall(c, e, k, Wown, Wnb[k], K, KR[k]);
    const bool volumeFluxes = e.order > 1;
    if (!(volumeFluxes || isSourcePresent(M))) return;
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
            for (int j = 0; j < e.nF; j++) {
                double ip = e.interpV[e.nF * p + j];
                for (int d = 0; d < dim; d++) cXi[d] += ip * elem_ref_vertex(e, j)[d];
                q.wallDistance += ip * e.wallDist[j];
            }
            elem_getX(e, cXi, q.currentX);
        }
        const double u[2] = {0, 0};
        T WInt[MAXEQ], dWInt[2 * MAXEQ];
        for (int m = 0; m < nEqs; m++) WInt[m] = 0.0;
        for (int m = 0; m < dim * nEqs; m++) dWInt[m] = 0.0;
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nBasis; j++) {
                WInt[m] += Wown[m * nBasis + j] * base[j];
                for (int d = 0; d < dim; d++) dWInt[nEqs * d + m] += Wown[m * nBasis + j] * dBase[j * 2 + d];
            }
        T f[2][MAXEQ], fv[2][MAXEQ], product[MAXEQ];
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
                diffusiveFlux(M, WInt, dWInt, nor, q, fv[d]);
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
                    K[nBasis * m + j] += Jac * weight * fsum - (e.eps + c.P.damp_const + e.dampInner[m]) * Jac * weight * dWsum;
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
// Element.calculateAvgW: W_avg[m] = sum_j W[m Nb + j] Is[j]
static void average_w(const Case& c, const Elem& e, double* Wavg) {
    for (int m = 0; m < c.Mc.nEqs; m++) {
        Wavg[m] = 0;
        for (int j = 0; j < e.nBasis; j++) Wavg[m] = Wavg[m] + e.W[m * e.nBasis + j] * e.Is[j];
    }
}

// DGpure.shock_senzor_all_eqn (C/element/DGpure.java:454-503): per equation, the L2 distance of
// the solution from its truncation to linear functions (order > 2) or to the mean, relative to
// its L2 norm, on a log10 scale, through a sine ramp of half-width kap around S0 = log10(1/(k-1)^4).
// shock_senzor (:403-451) is the m = 0 entry.
static void shock_sensors(const Case& c, const Elem& e, double kap, double* shock) {
    const int nb = e.nBasis;
    double Ws[MAXEQ];
    average_w(c, e, Ws);
    for (int m = 0; m < c.Mc.nEqs; m++) {
        std::vector<double> trunc(nb, 0.0);
        if (e.order > 2) {
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) trunc[i] += e.TrunOrd[i * nb + j] * e.W[m * nb + j];
        } else {
            for (int i = 0; i < nb; i++) trunc[i] = Ws[m];   // Lagrange basis: the mean in every coefficient
        }
        double Se = 0, pod = 0;
        for (int p = 0; p < e.nq; p++) {
            double vi = 0, vt = 0;
            for (int i = 0; i < nb; i++) {
                vi += e.W[m * nb + i] * e.baseV[p * nb + i];
                vt += trunc[i] * e.baseV[p * nb + i];
            }
            Se += e.JacV[p] * e.wV[p] * (vi - vt) * (vi - vt);
            pod += e.JacV[p] * e.wV[p] * vi * vi;
        }
        Se = std::log10(std::fabs(Se / pod));
        const double S0 = std::log10(1. / std::pow(e.order - 1, 4.0));
        double s = 0.5 * (1 + std::sin(M_PI * (Se - S0) / (2 * kap)));
        if (Se < S0 - kap) s = 0;
        if (Se > S0 + kap) s = 1;
        shock[m] = std::isnan(s) ? 0 : s;
    }
}

// DGpure.limiter (C/element/DGpure.java:358-400): eps from the density sensor (dampTol), the
// equation-wise damping from the sensors combined by the model (Aerodynamics.combineShockSensors:
// every equation takes the density sensor; the scalar model keeps its own) with the decaying
// memory innerIndicatorOld, and the interior penalty c_IP.
void limiter_all(Case& c, bool isFirstIter) {
    (void)isFirstIter;   // the first-iteration branch of the reference has an empty body
    const int nEqs = c.Mc.nEqs;
    for (auto& e : c.E) {
        e.eps = 0;
        e.c_IP = 0;
        if (e.order <= 1) continue;
        double Wavg[MAXEQ];
        average_w(c, e, Wavg);
        QPData q;
        q.meshVelocity[0] = q.meshVelocity[1] = q.meshVelocity[2] = 0;
        const double lam = maxEigenvalue(c.Mc, Wavg, q);
        double shock[MAXEQ];
        if (c.P.damp_tol > 0) {
            shock_sensors(c, e, c.P.damp_tol, shock);
            e.eps = lam * e.elemSize / e.order * shock[0];
        }
        if (c.P.damp_inner_tol > 0) {
            shock_sensors(c, e, c.P.damp_inner_tol, shock);
            if (c.Mc.kind != 2)
                for (int m = 1; m < nEqs; m++) shock[m] = shock[0];
            for (int m = 0; m < nEqs; m++) {
                const double indicator = std::max(shock[m], e.innerIndicatorOld[m]);
                e.innerIndicatorOld[m] = 0.1 * indicator;
                if (c.P.n_damp_inner_coef > 0)
                    e.dampInner[m] = (c.P.n_damp_inner_coef == 1 ? c.P.damp_inner_coef[0] : c.P.damp_inner_coef[m]) * indicator;
                else
                    e.dampInner[m] = lam * e.elemSize / e.order * indicator;
            }
        }
        if (c.Mc.diffusive) e.c_IP = c.P.penalty;
    }
}

// ------------------------------------------------------------- residual
void residual_all(Case& c, double* R) {
    const int nEqs = c.Mc.nEqs;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        const int B = e.nBasis * nEqs;
        const double* Wnb[MAXF];
        double* KR[MAXF] = {nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < e.nF; k++) Wnb[k] = e.TT[k] > -1 ? c.E[e.TT[k]].W.data() : nullptr;
        double* K = R + c.off[i];
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
void bdf_coefficients(const Case& c, double dt, double dto, double* a) { bdf_coeffs(c, dt, dto, a); }

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
        const double* Wnb[MAXF];
        double* KR[MAXF] = {nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < e.nF; k++) Wnb[k] = e.TT[k] > -1 ? c.E[e.TT[k]].W.data() : nullptr;
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
        std::vector<double> RwNeigh[MAXF], RwNeighH[MAXF];
        double* KR[MAXF] = {nullptr, nullptr, nullptr, nullptr};
        double* KRh[MAXF] = {nullptr, nullptr, nullptr, nullptr};
        const double* Wnb[MAXF];
        for (int k = 0; k < e.nF; k++) {
            Wnb[k] = e.TT[k] > -1 ? c.E[e.TT[k]].W.data() : nullptr;
            if (e.TT[k] > -1) {
                const int Bn = c.E[e.TT[k]].nBasis * nEqs;
                RwNeigh[k].assign(Bn, 0.0);
                RwNeighH[k].assign(Bn, 0.0);
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
                for (int k = 0; k < e.nF; k++) {
                    if (e.TT[k] > -1) {
                        Elem& nb = c.E[e.TT[k]];
                        const int Bn = nb.nBasis * nEqs;
                        double* Aaux = &nb.ANeigh[e.faceIndexReverse[k]][(size_t)ii * Bn];   // [B][B_nbr]
                        for (int j = 0; j < Bn; j++) Aaux[j] = (RwNeighH[k][j] - RwNeigh[k][j]) / h;
                    }
                }
            }
            for (int j = 0; j < B * B; j++) e.ADiag[j] /= -h;
        } else {
            // exact derivative: forward mode, one direction per perturbed DOF
            std::vector<Dual> Wd(B), K(B), KRd[MAXF];
            std::vector<Dual> Wnbd[MAXF];
            const Dual* Wnbp[MAXF];
            Dual* KRp[MAXF];
            for (int k = 0; k < e.nF; k++) {
                Wnbp[k] = nullptr;
                KRp[k] = nullptr;
                if (e.TT[k] > -1) {
                    const auto& Wn = c.E[e.TT[k]].W;
                    const int Bn = (int)Wn.size();
                    Wnbd[k].resize(Bn);
                    for (int j = 0; j < Bn; j++) Wnbd[k][j] = Dual(Wn[j], 0.0);
                    KRd[k].resize(Bn);
                    Wnbp[k] = Wnbd[k].data();
                    KRp[k] = KRd[k].data();
                }
            }
            for (int ii = 0; ii < B; ii++) {
                for (int j = 0; j < B; j++) { Wd[j] = Dual(e.W[j], j == ii ? 1.0 : 0.0); K[j] = Dual(0.0, 0.0); }
                residuum<Dual>(c, e, Wd.data(), Wnbp, K.data(), KRp);
                double* Ai = &e.ADiag[(size_t)ii * B];
                for (int j = 0; j < B; j++) Ai[j] = -K[j].d;
                for (int k = 0; k < e.nF; k++) {
                    if (e.TT[k] > -1) {
                        Elem& nb = c.E[e.TT[k]];
                        const int Bn = nb.nBasis * nEqs;
                        double* Aaux = &nb.ANeigh[e.faceIndexReverse[k]][(size_t)ii * Bn];
                        for (int j = 0; j < Bn; j++) Aaux[j] = KRd[k][j].d;
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
        std::vector<Dual> Wnbd[MAXF];
        const Dual* Wnbp[MAXF];
        Dual* KRp[MAXF] = {nullptr, nullptr, nullptr, nullptr};
        const size_t oi = c.off[i];
        for (int j = 0; j < B; j++) { Wd[j] = Dual(e.W[j], v[oi + j]); K[j] = Dual(0.0, 0.0); }
        for (int k = 0; k < e.nF; k++) {
            Wnbp[k] = nullptr;
            int t = e.TT[k];
            if (t > -1) {
                const int Bn = (int)c.E[t].W.size();
                Wnbd[k].resize(Bn);
                for (int j = 0; j < Bn; j++) Wnbd[k][j] = Dual(c.E[t].W[j], v[(size_t)c.off[t] + j]);
                Wnbp[k] = Wnbd[k].data();
            }
        }
        residuum<Dual>(c, e, Wd.data(), Wnbp, K.data(), KRp);
        for (int m = 0; m < nEqs; m++)
            for (int r = 0; r < Nb; r++) {
                double mv = 0;
                for (int s = 0; s < Nb; s++) mv += e.M[r * Nb + s] * v[oi + m * Nb + s];
                out[oi + m * Nb + r] = a1 * mv - K[m * Nb + r].d;
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
        std::vector<Dual> Wd(B), K(B), KRd[MAXF];
        std::vector<Dual> Wnbd[MAXF];
        const Dual* Wnbp[MAXF];
        Dual* KRp[MAXF];
        for (int k = 0; k < e.nF; k++) {
            Wnbp[k] = nullptr;
            KRp[k] = nullptr;
            int t = e.TT[k];
            if (t > -1) {
                const int Bn = (int)c.E[t].W.size();
                Wnbd[k].resize(Bn);
                for (int j = 0; j < Bn; j++) Wnbd[k][j] = Dual(c.E[t].W[j], 0.0);
                KRd[k].resize(Bn);
                Wnbp[k] = Wnbd[k].data();
                KRp[k] = KRd[k].data();
            }
        }
        for (int ii = 0; ii < B; ii++) {
            for (int j = 0; j < B; j++) { Wd[j] = Dual(e.W[j], j == ii ? 1.0 : 0.0); K[j] = Dual(0.0, 0.0); }
            residuum<Dual>(c, e, Wd.data(), Wnbp, K.data(), KRp);
            double s = 0;
            for (int j = 0; j < B; j++) s += -K[j].d * v[(size_t)c.off[i] + j];
            for (int k = 0; k < e.nF; k++) {
                int t = e.TT[k];
                if (t > -1)
                    for (int j = 0; j < (int)KRd[k].size(); j++) s += KRd[k][j].d * v[(size_t)c.off[t] + j];
            }
            int m = ii / Nb, r = ii % Nb;
            double mv = 0;
            for (int q = 0; q < Nb; q++) mv += e.M[q * Nb + r] * v[(size_t)c.off[i] + m * Nb + q];
            out[(size_t)c.off[i] + ii] = s + a1 * mv;
        }
    }
}

// ------------------------------------------------------------- functionals
// Element.computeFunctional / computeFunctionalsWall (C/element/Element.java:789-862) with
// FOI/DragCoefficient.java:31-48 (which = 0, p n_x) and FOI/LiftCoefficient.java:32-34 (which = 1, p n_y).
static double functional_elem(const Case& c, const Elem& e, int which, const double* W) {
    const int nEqs = c.Mc.nEqs, Nb = e.nBasis;
    double f = 0;
    for (int k = 0; k < e.nF; k++) {
        double fk = 0;
        for (int p = 0; p < e.nqf[k]; p++) {
            const double* baseLeft = &e.baseL[k][p * Nb];
            double Jac = e.JacF[k][p], weight = e.wF[k][p];
            double WL[MAXEQ] = {0, 0, 0, 0, 0};
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
        I += functional_elem(c, e, which, W ? W + c.off[i] : e.W.data());
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
            dIdW[(size_t)c.off[i] + j] = (Iwh - Iw) / h;
            Wh[j] = e.W[j];
        }
    }
}


// Pointwise model functions for the tests: source term and the diffusive flux in x and y of
// the case's model at one state (W, dW[d*nEqs + m]) and wall distance D.
void model_pointwise(const Case& c, const double* W, const double* dW, double D, double* source, double* fvx, double* fvy) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs;
    double w[MAXEQ], dw[2 * MAXEQ];
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
    diffusiveFlux(M, w, dw, ex, q, fvx);
    for (int m = 0; m < nEqs; m++) w[m] = W[m];
    diffusiveFlux(M, w, dw, ey, q, fvy);
}

// numerical flux of one face point and the physical flux of WL; on boundary faces (TT < 0)
// the outer state is the model's boundary value of WL, as in residuumWall
void flux_pointwise(const Case& c, const double* WL, const double* WR, const double* n, int TT, double* fnum, double* fphys) {
    const int nEqs = c.Mc.nEqs;
    double wl[MAXEQ], wr[MAXEQ];
    QPData q;
    q.currentX[0] = q.currentX[1] = q.currentX[2] = 0;
    q.meshVelocity[0] = q.meshVelocity[1] = q.meshVelocity[2] = 0;
    q.wallDistance = 0;
    for (int m = 0; m < nEqs; m++) { wl[m] = WL[m]; wr[m] = WR[m]; }
    if (TT < 0) boundaryValue(c.Mc, wl, n, TT, q, wr);
    numericalConvectiveFlux(c.Mc, wl, wr, n, TT, q, fnum);
    for (int m = 0; m < nEqs; m++) wl[m] = WL[m];
    convectiveFlux(c.Mc, wl, n, fphys);
}
}  // namespace oracle
