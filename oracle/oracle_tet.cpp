// Oracle (test infrastructure only): DGpure on straight tetrahedra, the 3D counterpart of
// oracle_residual.cpp for the models of oracle_model.hpp with dim = 3.
//
// Restates: tetrahedralElement (C/elementType/tetrahedralElement.java:15-68: faces
// {0,2,1}, {0,1,3}, {1,2,3}, {2,0,3}); basis3Dtetra (C/basis/basis3Dtetra.java:23-62: orders 1
// and 2 only, its barycentr returns the 4 vertices, so order >= 3 indexes past the node list
// and the reference throws); transformation3Dtetra (C/transformation/transformation3Dtetra.java:
// affine map from Mat.invert of the vertex matrix, |det| Jacobian, getXs at {0.25, 0.25, 0}
// because of the literal {0.25, 0.25, 0, 25}); faceTransformation3Dtriangle (normal =
// normalize(Js2 x Js1), Jacobian |Js1 x Js2| / 2); Integration (C/Integration.java:40-153,
// neighbour-side basis through the neighbour's inverse map); Element.computeGeometry
// (C/element/Element.java:516-567); DGpure.residuum / residuumWall (C/element/DGpure.java:
// 108-355); FlowProMain.createNeighborMatrix (C/FlowProMain.java:772-873).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle_core.hpp"
#include "oracle_model.hpp"

namespace oracle {

static const int TET_FACE[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};
static const double TET_XI[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
// basis3Dtetra exponents (the first nBasis are used)
static const int XE[19] = {0, 1, 0, 0, 1, 1, 0, 2, 0, 0, 2, 2, 1, 0, 1, 0, 3, 0, 0};
static const int YE[19] = {0, 0, 1, 0, 1, 0, 1, 0, 2, 0, 1, 0, 2, 2, 0, 1, 0, 3, 0};
static const int ZE[19] = {0, 0, 0, 1, 0, 1, 1, 0, 0, 2, 0, 1, 0, 1, 2, 2, 0, 0, 3};

struct TetBasis {
    int nb = 0;
    std::vector<std::vector<double>> coeffs;   // coeffs[i][m]
    double fun(int m, const double* x) const {
        double v = 0;
        for (int i = 0; i < nb; i++) v = v + coeffs[i][m] * std::pow(x[0], XE[i]) * std::pow(x[1], YE[i]) * std::pow(x[2], ZE[i]);
        return v;
    }
    double der(int m, const double* x, int d) const {
        double v = 0;
        for (int i = 0; i < nb; i++) {
            const int e[3] = {XE[i], YE[i], ZE[i]};
            if (e[d] > 0) {
                double t = coeffs[i][m] * e[d];
                for (int q = 0; q < 3; q++) t = t * std::pow(x[q], q == d ? e[q] - 1 : e[q]);
                v = v + t;
            }
        }
        return v;
    }
};

static TetBasis make_tet_basis(int order) {
    TetBasis b;
    b.nb = (3 * (order - 1) * order + 2) / 2;
    if (order == 1) {
        b.coeffs = {{1.0}};
        return b;
    }
    if (order != 2) throw std::runtime_error("basis3Dtetra: the node list has 4 entries, order >= 3 is not available");
    std::vector<std::vector<double>> V(b.nb, std::vector<double>(b.nb));
    for (int i = 0; i < b.nb; i++)
        for (int j = 0; j < b.nb; j++)
            V[i][j] = std::pow(TET_XI[i][0], XE[j]) * std::pow(TET_XI[i][1], YE[j]) * std::pow(TET_XI[i][2], ZE[j]);
    b.coeffs = mat_invert(V);
    return b;
}

struct TetElem {
    int TT[4], TP[4];
    double A[4], B[4], C[4], iA[4], iB[4], iC[4];
    double J[3][3];              // dX/dxi
    std::vector<std::vector<double>> iT;   // Mat.invert(J)
    double jac;                  // |det J|
    int nq = 0;
    std::vector<double> baseV, dbaseV;   // [q*nb+j], [(q*nb+j)*3+d]
    int nqf = 0;
    std::vector<double> xF[4];                       // physical face points [p*3+d]
    std::vector<double> JacF[4], baseL[4], dbaseL[4], baseR[4], dbaseR[4];
    double n[4][3];
    double S[4], Xs[3], Xes[4][3], vol, elemSize;
    std::vector<double> M, Is;
    std::vector<double> W;
    double c_IP = 0;
    double getX(int d, const double* xi) const {
        const double* a = d == 0 ? A : (d == 1 ? B : C);
        return a[0] * xi[0] + a[1] * xi[1] + a[2] * xi[2] + a[3];
    }
};

struct TetCase {
    Props P;
    ModelConst Mc;
    int nPts = 0, nElem = 0, nb = 0, B = 0, dofs = 0;
    std::vector<double> PXY;
    std::vector<int> TP, TT, order, rank;
    TetBasis basis;
    std::vector<double> vx, vw;    // volume rule (xi, eta, zeta), weights
    std::vector<double> fx, fw;    // face rule on the reference triangle
    std::vector<TetElem> E;
};

// FlowProMain.createNeighborMatrix for tetrahedra: sorted face keys; boundary rows
// (type, v0, v1, v2); unmatched faces without a boundary row get -1.
std::vector<int> tet_neighbor_matrix(int nElem, const int* TP, int nBnd, const int* bnd) {
    auto key = [](int a, int b, int c) {
        int v[3] = {a, b, c};
        std::sort(v, v + 3);
        return std::to_string(v[0]) + "$" + std::to_string(v[1]) + "$" + std::to_string(v[2]) + "$";
    };
    std::map<std::string, std::pair<int, int>> faces;
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < 4; k++) {
            const int* f = TET_FACE[k];
            std::string s = key(TP[4 * i + f[0]], TP[4 * i + f[1]], TP[4 * i + f[2]]);
            auto it = faces.find(s);
            if (it == faces.end()) faces.emplace(s, std::make_pair(i, -1000));
            else it->second.second = i;
        }
    std::vector<int> TT((size_t)4 * nElem, -1000);
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < 4; k++) {
            const int* f = TET_FACE[k];
            const auto& fc = faces[key(TP[4 * i + f[0]], TP[4 * i + f[1]], TP[4 * i + f[2]])];
            if (fc.first == i) {
                if (fc.second != -1000) TT[4 * i + k] = fc.second;
            } else {
                TT[4 * i + k] = fc.first;
            }
        }
    std::map<std::string, int> bmap;
    for (int r = 0; r < nBnd; r++) bmap[key(bnd[4 * r + 1], bnd[4 * r + 2], bnd[4 * r + 3])] = bnd[4 * r];
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < 4; k++)
            if (TT[4 * i + k] == -1000) {
                const int* f = TET_FACE[k];
                auto it = bmap.find(key(TP[4 * i + f[0]], TP[4 * i + f[1]], TP[4 * i + f[2]]));
                TT[4 * i + k] = it != bmap.end() ? it->second : -1;
            }
    return TT;
}

static std::vector<double> read_tet_rule(const std::string& path, std::vector<double>& w) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open quadrature file " + path);
    int n;
    f >> n;
    std::vector<double> x(3 * n);
    for (int i = 0; i < 3 * n; i++) f >> x[i];
    w.resize(n);
    for (int i = 0; i < n; i++) f >> w[i];
    if (!f) throw std::runtime_error("bad quadrature file " + path);
    return x;
}

static void inverse_affine(const double* P, double* iA, double* iB, double* iC) {
    // transformation3Dtetra: invVi = Mat.invert([X Y Z 1]); iA = invVi * xi-coords of the vertices
    std::vector<std::vector<double>> V(4, std::vector<double>(4));
    for (int i = 0; i < 4; i++) {
        V[i][0] = P[3 * i];
        V[i][1] = P[3 * i + 1];
        V[i][2] = P[3 * i + 2];
        V[i][3] = 1;
    }
    auto iV = mat_invert(V);
    for (int r = 0; r < 4; r++) {
        iA[r] = iB[r] = iC[r] = 0;
        for (int i = 0; i < 4; i++) {
            iA[r] += iV[r][i] * TET_XI[i][0];
            iB[r] += iV[r][i] * TET_XI[i][1];
            iC[r] += iV[r][i] * TET_XI[i][2];
        }
    }
}

TetCase* build_tet_case(const Props& P, int nPts, const double* PXY, int nElem, const int* TP, int nBnd, const int* bnd,
                        const std::string& rulesDir) {
    auto* c = new TetCase();
    c->P = P;
    c->Mc = model_init(P);
    c->nPts = nPts;
    c->nElem = nElem;
    c->PXY.assign(PXY, PXY + 3 * (size_t)nPts);
    c->TP.assign(TP, TP + 4 * (size_t)nElem);
    c->TT = tet_neighbor_matrix(nElem, TP, nBnd, bnd);
    c->basis = make_tet_basis(P.order);
    c->nb = c->basis.nb;
    c->B = c->nb * c->Mc.nEqs;
    c->dofs = c->B * nElem;
    // QuadratureCentral: qTetra of degree 2k-1 (no swap), qTriangle with the negative-weight swap
    c->vx = read_tet_rule(rulesDir + "/gaussTetra/" + std::to_string(2 * P.order - 1) + ".txt", c->vw);
    Rule fr = load_triangle_rule(rulesDir, P.order);
    c->fx = fr.x;
    c->fw = fr.w;
    // element order (ElementOrdering.reverseCuthillMcKee on the face graph)
    if (P.ordering == 1) c->order = rcm_order(nElem, c->TT.data(), 4);
    else {
        c->order.resize(nElem);
        for (int i = 0; i < nElem; i++) c->order[i] = i;
    }
    c->rank.resize(nElem);
    for (int r = 0; r < nElem; r++) c->rank[c->order[r]] = r;
    const int nb = c->nb;
    const int nq = (int)c->vw.size(), nqf = (int)c->fw.size();
    c->E.resize(nElem);
    // ---- per element maps and volume tables
    for (int i = 0; i < nElem; i++) {
        TetElem& e = c->E[i];
        double Pv[12];
        for (int v = 0; v < 4; v++) {
            e.TP[v] = TP[4 * i + v];
            e.TT[v] = c->TT[4 * i + v];
            for (int d = 0; d < 3; d++) Pv[3 * v + d] = PXY[3 * (size_t)e.TP[v] + d];
        }
        // A = invV * x-coordinates with invV = Mat.invert([xi eta zeta 1] of the reference vertices)
        std::vector<std::vector<double>> V(4, std::vector<double>(4));
        for (int v = 0; v < 4; v++) {
            V[v][0] = TET_XI[v][0];
            V[v][1] = TET_XI[v][1];
            V[v][2] = TET_XI[v][2];
            V[v][3] = 1;
        }
        auto iV = mat_invert(V);
        for (int r = 0; r < 4; r++) {
            e.A[r] = e.B[r] = e.C[r] = 0;
            for (int v = 0; v < 4; v++) {
                e.A[r] += iV[r][v] * Pv[3 * v];
                e.B[r] += iV[r][v] * Pv[3 * v + 1];
                e.C[r] += iV[r][v] * Pv[3 * v + 2];
            }
        }
        inverse_affine(Pv, e.iA, e.iB, e.iC);
        for (int j = 0; j < 3; j++) {
            e.J[0][j] = e.A[j];
            e.J[1][j] = e.B[j];
            e.J[2][j] = e.C[j];
        }
        const double* A = e.A;
        const double* B = e.B;
        const double* C = e.C;
        e.jac = std::fabs(A[0] * B[1] * C[2] + A[1] * B[2] * C[0] + A[2] * B[0] * C[1] - A[2] * B[1] * C[0] - A[1] * B[0] * C[2] -
                          A[0] * B[2] * C[1]);
        e.iT = mat_invert({{e.J[0][0], e.J[0][1], e.J[0][2]}, {e.J[1][0], e.J[1][1], e.J[1][2]}, {e.J[2][0], e.J[2][1], e.J[2][2]}});
        e.nq = nq;
        e.baseV.resize((size_t)nq * nb);
        e.dbaseV.resize((size_t)nq * nb * 3);
        for (int q = 0; q < nq; q++) {
            const double* x = &c->vx[3 * q];
            for (int m = 0; m < nb; m++) {
                e.baseV[q * nb + m] = c->basis.fun(m, x);
                double dxi[3] = {c->basis.der(m, x, 0), c->basis.der(m, x, 1), c->basis.der(m, x, 2)};
                for (int a = 0; a < 3; a++) {
                    double g = 0;
                    for (int b = 0; b < 3; b++) g += e.iT[b][a] * dxi[b];
                    e.dbaseV[(q * nb + m) * 3 + a] = g;
                }
            }
        }
        // Xs at {0.25, 0.25, 0} (the reference's literal {0.25, 0.25, 0, 25})
        const double xs[3] = {0.25, 0.25, 0.0};
        for (int d = 0; d < 3; d++) e.Xs[d] = e.getX(d, xs);
    }
    // ---- faces (needs all maps for the neighbour side)
    for (int i = 0; i < nElem; i++) {
        TetElem& e = c->E[i];
        e.nqf = nqf;
        for (int k = 0; k < 4; k++) {
            const int* f = TET_FACE[k];
            double R0[3], s1[3], s2[3];
            for (int d = 0; d < 3; d++) {
                R0[d] = TET_XI[f[0]][d];
                s1[d] = TET_XI[f[1]][d] - TET_XI[f[0]][d];
                s2[d] = TET_XI[f[2]][d] - TET_XI[f[0]][d];
            }
            double Js1[3] = {0, 0, 0}, Js2[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    Js1[a] += e.J[a][b] * s1[b];
                    Js2[a] += e.J[a][b] * s2[b];
                }
            auto cross = [](const double* u, const double* v, double* w) {
                w[0] = u[1] * v[2] - u[2] * v[1];
                w[1] = u[2] * v[0] - u[0] * v[2];
                w[2] = u[0] * v[1] - u[1] * v[0];
            };
            double c12[3], c21[3];
            cross(Js1, Js2, c12);
            cross(Js2, Js1, c21);
            const double jf = std::sqrt(c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2]) / 2;
            const double nn = std::sqrt(c21[0] * c21[0] + c21[1] * c21[1] + c21[2] * c21[2]);
            for (int d = 0; d < 3; d++) e.n[k][d] = c21[d] / nn;
            const double mid[3] = {R0[0] + s1[0] / 3 + s2[0] / 3, R0[1] + s1[1] / 3 + s2[1] / 3, R0[2] + s1[2] / 3 + s2[2] / 3};
            for (int d = 0; d < 3; d++) e.Xes[k][d] = e.getX(d, mid);
            e.JacF[k].assign(nqf, jf);
            e.xF[k].resize(3 * nqf);
            e.baseL[k].resize((size_t)nqf * nb);
            e.dbaseL[k].resize((size_t)nqf * nb * 3);
            const int t = e.TT[k];
            if (t > -1) {
                e.baseR[k].resize((size_t)nqf * nb);
                e.dbaseR[k].resize((size_t)nqf * nb * 3);
            }
            e.S[k] = 0;
            for (int p = 0; p < nqf; p++) {
                const double u = c->fx[2 * p], v = c->fx[2 * p + 1];
                double xr[3];
                for (int d = 0; d < 3; d++) xr[d] = s1[d] * u + s2[d] * v + R0[d];
                double X[3];
                for (int d = 0; d < 3; d++) X[d] = e.getX(d, xr);
                for (int d = 0; d < 3; d++) e.xF[k][3 * p + d] = X[d];
                for (int m = 0; m < nb; m++) {
                    e.baseL[k][p * nb + m] = c->basis.fun(m, xr);
                    double dxi[3] = {c->basis.der(m, xr, 0), c->basis.der(m, xr, 1), c->basis.der(m, xr, 2)};
                    for (int a = 0; a < 3; a++) {
                        double g = 0;
                        for (int b = 0; b < 3; b++) g += e.iT[b][a] * dxi[b];
                        e.dbaseL[k][(p * nb + m) * 3 + a] = g;
                    }
                }
                if (t > -1) {
                    const TetElem& er = c->E[t];
                    const double xn[3] = {er.iA[0] * X[0] + er.iA[1] * X[1] + er.iA[2] * X[2] + er.iA[3],
                                          er.iB[0] * X[0] + er.iB[1] * X[1] + er.iB[2] * X[2] + er.iB[3],
                                          er.iC[0] * X[0] + er.iC[1] * X[1] + er.iC[2] * X[2] + er.iC[3]};
                    for (int m = 0; m < nb; m++) {
                        e.baseR[k][p * nb + m] = c->basis.fun(m, xn);
                        double dxi[3] = {c->basis.der(m, xn, 0), c->basis.der(m, xn, 1), c->basis.der(m, xn, 2)};
                        for (int a = 0; a < 3; a++) {
                            double g = 0;
                            for (int b = 0; b < 3; b++) g += er.iT[b][a] * dxi[b];
                            e.dbaseR[k][(p * nb + m) * 3 + a] = g;
                        }
                    }
                }
                e.S[k] = e.S[k] + e.JacF[k][p] * c->fw[p];
            }
        }
        e.vol = 0;
        for (int q = 0; q < nq; q++) e.vol = e.vol + e.jac * c->vw[q];
        double sumS = 0;
        for (int k = 0; k < 4; k++) sumS += e.S[k];
        e.elemSize = 4 * e.vol / sumS;
        // DGpure.computeMassMatrixAndBasisWeights
        e.M.assign((size_t)nb * nb, 0.0);
        e.Is.assign(nb, 0.0);
        for (int a = 0; a < nb; a++) {
            for (int b = 0; b < nb; b++)
                for (int q = 0; q < nq; q++) e.M[a * nb + b] = e.M[a * nb + b] + e.jac * c->vw[q] * e.baseV[q * nb + a] * e.baseV[q * nb + b];
            for (int q = 0; q < nq; q++) e.Is[a] = e.Is[a] + e.jac * c->vw[q] * e.baseV[q * nb + a] / e.vol;
        }
        e.c_IP = (P.order > 1 && c->Mc.diffusive) ? P.penalty : 0.0;
        std::vector<double> w0 = const_init_condition(c->Mc);
        e.W.resize(c->B);
        for (int m = 0; m < c->Mc.nEqs; m++)
            for (int j = 0; j < nb; j++) e.W[m * nb + j] = w0[m];
    }
    return c;
}

// DGpure.residuumWall for face k of element e in 3D (static mesh, dampTol = dampConst = 0)
template <class T>
static void tet_face(const TetCase& c, const TetElem& e, int k, const T* Wown, const T* Wnb, T* K, T* KR) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs, dim = 3, nb = c.nb;
    const int TT = e.TT[k];
    if (KR) for (int j = 0; j < nb * nEqs; j++) KR[j] = 0.0;
    double beta0 = c.P.beta0;
    if (c.P.order == 1) beta0 = 0;   // DGpure.java:287-289
    const double* nn = e.n[k];
    double dL = 0;
    for (int d = 0; d < dim; d++) dL += (TT > -1 ? (e.Xs[d] - c.E[TT].Xs[d]) : (e.Xs[d] - e.Xes[k][d])) * nn[d];
    dL = std::fabs(dL);
    for (int p = 0; p < e.nqf; p++) {
        const double* bl = &e.baseL[k][p * nb];
        const double* dbl = &e.dbaseL[k][p * nb * 3];
        const double jw = e.JacF[k][p] * c.fw[p];
        QPData q;
        for (int d = 0; d < 3; d++) {
            q.currentX[d] = e.xF[k][3 * p + d];
            q.meshVelocity[d] = 0;
        }
        q.wallDistance = 0;
        T WL[5], WR[5], dWL[15], dWR[15];
        for (int m = 0; m < nEqs; m++) { WL[m] = 0.0; WR[m] = 0.0; }
        for (int m = 0; m < dim * nEqs; m++) { dWL[m] = 0.0; dWR[m] = 0.0; }
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nb; j++) {
                WL[m] += Wown[m * nb + j] * bl[j];
                for (int d = 0; d < dim; d++) dWL[nEqs * d + m] += Wown[m * nb + j] * dbl[j * 3 + d];
            }
        if (TT > -1) {
            const double* br = &e.baseR[k][p * nb];
            const double* dbr = &e.dbaseR[k][p * nb * 3];
            for (int m = 0; m < nEqs; m++)
                for (int j = 0; j < nb; j++) {
                    WR[m] += Wnb[m * nb + j] * br[j];
                    for (int d = 0; d < dim; d++) dWR[nEqs * d + m] += Wnb[m * nb + j] * dbr[j * 3 + d];
                }
        } else {
            boundaryValue(M, WL, nn, TT, q, WR);
            for (int m = 0; m < dim * nEqs; m++) dWR[m] = dWL[m];
        }
        T Wale[5], fn[5], fvn[5];
        for (int j = 0; j < nEqs; j++) Wale[j] = TT > -1 ? (WL[j] + WR[j]) / 2 : WR[j];
        numericalConvectiveFlux(M, WL, WR, nn, TT, q, fn);
        if (M.diffusive) {
            T Wc[5], dWc[15];
            for (int m = 0; m < nEqs; m++) {
                Wc[m] = TT > -1 ? (WL[m] + WR[m]) / 2 : WR[m];
                for (int d = 0; d < dim; d++)
                    dWc[nEqs * d + m] = (dWL[nEqs * d + m] + dWR[nEqs * d + m]) / 2 + beta0 * (WR[m] - WL[m]) / dL * nn[d];
            }
            numericalDiffusiveFlux(M, Wc, dWc, nn, TT, q, fvn);
        }
        for (int m = 0; m < nEqs; m++) {
            for (int j = 0; j < nb; j++) {
                const double jwb = jw * bl[j];
                K[nb * m + j] -= jwb * (fn[m] - 0.0 * Wale[m]);
                if (M.diffusive) {
                    K[nb * m + j] += jwb * fvn[m];
                    if (TT > -1 || isIPFace(M, TT)) K[nb * m + j] -= e.c_IP * jwb * (WL[m] - WR[m]);
                }
            }
            if (KR) {
                const double* br = &e.baseR[k][p * nb];
                for (int j = 0; j < nb; j++) {
                    const double jwb = jw * br[j];
                    KR[nb * m + j] -= jwb * (fn[m] - 0.0 * Wale[m]);
                    if (M.diffusive) {
                        KR[nb * m + j] += jwb * fvn[m];
                        KR[nb * m + j] -= e.c_IP * jwb * (WL[m] - WR[m]);
                    }
                }
            }
        }
    }
}

// DGpure.residuum in 3D: faces first, then the volume
template <class T>
static void tet_residuum(const TetCase& c, const TetElem& e, const T* Wown, const T* const* Wnb, T* K, T* const* KR) {
    const ModelConst& M = c.Mc;
    const int nEqs = M.nEqs, dim = 3, nb = c.nb;
    for (int k = 0; k < 4; k++) tet_face(c, e, k, Wown, Wnb[k], K, KR[k]);
    if (c.P.order <= 1) return;   // no volume fluxes and no source for the Euler/NS models
    for (int q = 0; q < e.nq; q++) {
        const double* base = &e.baseV[q * nb];
        const double* dBase = &e.dbaseV[q * nb * 3];
        const double jw = e.jac * c.vw[q];
        T WInt[5], dWInt[15];
        for (int m = 0; m < nEqs; m++) WInt[m] = 0.0;
        for (int m = 0; m < dim * nEqs; m++) dWInt[m] = 0.0;
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nb; j++) {
                WInt[m] += Wown[m * nb + j] * base[j];
                for (int d = 0; d < dim; d++) dWInt[nEqs * d + m] += Wown[m * nb + j] * dBase[j * 3 + d];
            }
        T f[3][5], fv[3][5];
        double nor[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) {
            nor[d] = 1;
            convectiveFlux(M, WInt, nor, f[d]);
            nor[d] = 0;
        }
        const QPData qd{};   // static mesh, no wall distance on tetrahedra
        if (M.diffusive)
            for (int d = 0; d < dim; d++) {
                nor[d] = 1;
                diffusiveFlux(M, WInt, dWInt, nor, qd, fv[d]);
                nor[d] = 0;
            }
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < nb; j++) {
                T fsum = 0.0, dWsum = 0.0;
                for (int d = 0; d < dim; d++) {
                    fsum += (f[d][m] - 0.0 * WInt[m]) * dBase[j * 3 + d];
                    dWsum += dWInt[nEqs * d + m] * dBase[j * 3 + d];
                }
                K[nb * m + j] += jw * fsum - (0.0 + 0.0 + 0.0) * jw * dWsum;
                if (M.diffusive) {
                    T fvsum = 0.0;
                    for (int d = 0; d < dim; d++) fvsum += fv[d][m] * dBase[j * 3 + d];
                    K[nb * m + j] -= jw * fvsum;
                }
            }
    }
}

static void tet_residual_all(const TetCase& c, const double* const* Wel, double* R) {
    for (int i = 0; i < c.nElem; i++) {
        const TetElem& e = c.E[i];
        const double* Wnb[4];
        double* KR[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < 4; k++) Wnb[k] = e.TT[k] > -1 ? Wel[e.TT[k]] : nullptr;
        double* K = R + (size_t)i * c.B;
        for (int j = 0; j < c.B; j++) K[j] = 0;
        tet_residuum<double>(c, e, Wel[i], Wnb, K, KR);
    }
}

// Exact (a1 M - dR/dW) of the whole mesh in element-major layout: A[(e*B + j) * dofs + t*B + i]
// = a1 M - dR_e[j] / dW_t[i], column by column with dual numbers.
static void tet_jacobian_dense(const TetCase& c, double a1, double* A) {
    const int B = c.B, nb = c.nb, n = c.dofs;
    std::fill(A, A + (size_t)n * n, 0.0);
    std::vector<std::vector<Dual>> Wd(c.nElem);
    for (int i = 0; i < c.nElem; i++) {
        Wd[i].resize(B);
        for (int j = 0; j < B; j++) Wd[i][j] = Dual(c.E[i].W[j], 0.0);
    }
    std::vector<Dual> K(B);
    for (int t = 0; t < c.nElem; t++)
        for (int i = 0; i < B; i++) {
            Wd[t][i].d = 1.0;
            // residuals that depend on W_t: element t and its face neighbours
            int targets[5] = {t, c.E[t].TT[0], c.E[t].TT[1], c.E[t].TT[2], c.E[t].TT[3]};
            for (int r = 0; r < 5; r++) {
                const int e = targets[r];
                if (e < 0) continue;
                const TetElem& el = c.E[e];
                const Dual* Wnb[4];
                Dual* KR[4] = {nullptr, nullptr, nullptr, nullptr};
                for (int k = 0; k < 4; k++) Wnb[k] = el.TT[k] > -1 ? Wd[el.TT[k]].data() : nullptr;
                for (int j = 0; j < B; j++) K[j] = 0.0;
                tet_residuum<Dual>(c, el, Wd[e].data(), Wnb, K.data(), KR);
                for (int j = 0; j < B; j++) A[((size_t)e * B + j) * n + (size_t)t * B + i] = -K[j].d;
            }
            Wd[t][i].d = 0.0;
        }
    for (int e = 0; e < c.nElem; e++)
        for (int m = 0; m < c.Mc.nEqs; m++)
            for (int a = 0; a < nb; a++)
                for (int b = 0; b < nb; b++)
                    A[((size_t)e * B + m * nb + a) * n + (size_t)e * B + m * nb + b] += a1 * c.E[e].M[a * nb + b];
}

// State.getTimeStep without the CFL factor: min_e elemSize / lambda(mean W)
static double tet_size_over_lambda(const TetCase& c) {
    double best = 1.7976931348623157e308;
    for (const auto& e : c.E) {
        double Wavg[5] = {0, 0, 0, 0, 0};
        for (int j = 0; j < c.Mc.nEqs; j++)
            for (int m = 0; m < c.nb; m++) Wavg[j] = Wavg[j] + e.W[j * c.nb + m] * e.Is[m];
        QPData q;
        for (int d = 0; d < 3; d++) q.meshVelocity[d] = 0;
        double lam = maxEigenvalue(c.Mc, Wavg, q);
        if (lam > 0) best = std::min(best, e.elemSize / lam);
    }
    return best;
}

}  // namespace oracle

using namespace oracle;

extern "C" {
void* ot_create(const Props* P, int nPts, const double* PXY, int nElem, const int* TP, int nBnd, const int* bnd,
                const char* rulesDir) {
    try {
        return build_tet_case(*P, nPts, PXY, nElem, TP, nBnd, bnd, rulesDir);
    } catch (...) {
        return nullptr;
    }
}
void ot_destroy(void* h) { delete (TetCase*)h; }
int ot_nbasis(void* h) { return ((TetCase*)h)->nb; }
int ot_neqs(void* h) { return ((TetCase*)h)->Mc.nEqs; }
void ot_tt(void* h, int* out) { TetCase* c = (TetCase*)h; std::memcpy(out, c->TT.data(), sizeof(int) * c->TT.size()); }
void ot_order(void* h, int* out) { TetCase* c = (TetCase*)h; std::memcpy(out, c->order.data(), sizeof(int) * c->nElem); }
void ot_tet_neighbor_matrix(int nElem, const int* TP, int nBnd, const int* bnd, int* TT) {
    auto t = tet_neighbor_matrix(nElem, TP, nBnd, bnd);
    std::memcpy(TT, t.data(), sizeof(int) * t.size());
}
void ot_rcm_order(int nElem, const int* TT, int nF, int* out) {
    auto o = rcm_order(nElem, TT, nF);
    std::memcpy(out, o.data(), sizeof(int) * nElem);
}
void ot_set_w(void* h, const double* W) {
    TetCase* c = (TetCase*)h;
    for (int i = 0; i < c->nElem; i++) std::memcpy(c->E[i].W.data(), W + (size_t)i * c->B, sizeof(double) * c->B);
}
void ot_get_w(void* h, double* W) {
    TetCase* c = (TetCase*)h;
    for (int i = 0; i < c->nElem; i++) std::memcpy(W + (size_t)i * c->B, c->E[i].W.data(), sizeof(double) * c->B);
}
void ot_residual(void* h, double* R) {
    TetCase* c = (TetCase*)h;
    std::vector<const double*> Wel(c->nElem);
    for (int i = 0; i < c->nElem; i++) Wel[i] = c->E[i].W.data();
    tet_residual_all(*c, Wel.data(), R);
}
void ot_jacobian_dense(void* h, double a1, double* A) { tet_jacobian_dense(*(TetCase*)h, a1, A); }
void ot_mass(void* h, int i, double* M) {
    TetCase* c = (TetCase*)h;
    std::memcpy(M, c->E[i].M.data(), sizeof(double) * c->nb * c->nb);
}
double ot_size_over_lambda(void* h) { return tet_size_over_lambda(*(TetCase*)h); }
void ot_geometry(void* h, int i, double* vol, double* elemSize, double* Xs, double* normals, double* S) {
    TetCase* c = (TetCase*)h;
    const TetElem& e = c->E[i];
    *vol = e.vol;
    *elemSize = e.elemSize;
    for (int d = 0; d < 3; d++) Xs[d] = e.Xs[d];
    for (int k = 0; k < 4; k++) {
        S[k] = e.S[k];
        for (int d = 0; d < 3; d++) normals[3 * k + d] = e.n[k][d];
    }
}
}
