// Oracle (test infrastructure only): setup restatement -- model constants,
// quadrature rule files, Lagrange bases, affine maps, integration tables,
// geometry, mass matrices, neighbour matrix and RCM numbering.
#include <limits>
#include "oracle_core.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace oracle {

// ------------------------------------------------------------------ Mat
// API/Mat.java:502-563 gaussian (scaled partial pivoting), :605-642 invert
static std::vector<int> gaussian(std::vector<std::vector<double>>& A) {
    int n = (int)A.size();
    std::vector<int> index(n);
    std::vector<double> c(n);
    for (int i = 0; i < n; ++i) index[i] = i;
    for (int i = 0; i < n; ++i) {
        double c1 = 0;
        for (int j = 0; j < n; ++j) {
            double c0 = std::fabs(A[i][j]);
            if (c0 > c1) c1 = c0;
        }
        c[i] = c1;
    }
    for (int j = 0; j < n - 1; ++j) {
        int k = j;
        double pi1 = 0;
        for (int i = j; i < n; ++i) {
            double pi0 = std::fabs(A[index[i]][j]);
            pi0 /= c[index[i]];
            if (pi0 > pi1) { pi1 = pi0; k = i; }
        }
        int itmp = index[j]; index[j] = index[k]; index[k] = itmp;
        for (int i = j + 1; i < n; ++i) {
            double pj = A[index[i]][j] / A[index[j]][j];
            A[index[i]][j] = pj;
            for (int l = j + 1; l < n; ++l) A[index[i]][l] -= pj * A[index[j]][l];
        }
    }
    return index;
}

std::vector<std::vector<double>> mat_invert(const std::vector<std::vector<double>>& A) {
    int n = (int)A.size();
    std::vector<std::vector<double>> a = A, x(n, std::vector<double>(n, 0.0)), b(n, std::vector<double>(n, 0.0));
    for (int i = 0; i < n; ++i) b[i][i] = 1;
    std::vector<int> index = gaussian(a);
    for (int i = 0; i < n - 1; ++i)
        for (int j = i + 1; j < n; ++j)
            for (int k = 0; k < n; ++k) b[index[j]][k] -= a[index[j]][i] * b[index[i]][k];
    for (int i = 0; i < n; ++i) {
        x[n - 1][i] = b[index[n - 1]][i] / a[index[n - 1]][n - 1];
        for (int j = n - 2; j >= 0; --j) {
            x[j][i] = b[index[j]][i];
            for (int k = j + 1; k < n; ++k) x[j][i] -= a[index[j]][k] * x[k][i];
            x[j][i] /= a[index[j]][j];
        }
    }
    return x;
}

// flat row-major variant used by BlockILU0 (same arithmetic)
void mat_invert_flat(const double* A, int n, double* out) {
    std::vector<std::vector<double>> a(n, std::vector<double>(n));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) a[i][j] = A[i * n + j];
    auto x = mat_invert(a);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[i * n + j] = x[i][j];
}

// ------------------------------------------------------------------ model init
static double stagP(const ModelConst& M, double mach) {   // AERO/Aerodynamics.java:314-316
    return std::pow(1 + (M.kapa - 1) / 2 * mach * mach, M.kapa / (M.kapa - 1));
}
static double stagRho(const ModelConst& M, double mach) {  // AERO/Aerodynamics.java:318-320
    return std::pow(1 + (M.kapa - 1) / 2 * mach * mach, 1 / (M.kapa - 1));
}

ModelConst model_init(const Props& P) {
    ModelConst M;
    std::memset(&M, 0, sizeof(M));
    M.dim = P.dimension == 3 ? 3 : 2;
    if (P.model == 2) {   // T/user/equation/TestScalarEquation.java:28-37
        M.kind = 2;
        M.nEqs = 1;
        M.a[0] = P.scalar_a0;
        M.a[1] = P.scalar_a1;
        M.k = P.scalar_k;
        double norm = 0;
        for (int d = 0; d < 2; d++) norm += M.a[d] * M.a[d];
        M.lambda = std::sqrt(norm);
        M.diffusive = M.k > 0;
        return M;
    }
    M.kind = P.model == 1 ? 1 : (P.model == 3 ? 3 : (P.model == 4 ? 4 : 0));
    // AdiabaticEulers.init: Aerodynamics.init(props, dim, dim + 1, false) (AERO/AdiabaticEulers.java:13-16)
    // KOmegaSST.init: Aerodynamics.init(props, dim, dim + 4, true) (AERO/KOmegaSST.java:44-47)
    M.nEqs = M.kind == 1 ? M.dim + 3 : (M.kind == 3 ? M.dim + 1 : (M.kind == 4 ? M.dim + 4 : M.dim + 2));
    // SpalartAllmaras.init calls Aerodynamics.init(props, dim, dim + 3, true): always diffusive
    M.diffusive = (M.kind == 1 || M.kind == 4) ? true : (M.kind == 3 ? false : P.viscous != 0);
    M.vanLeer = M.kind == 0 && P.numerical_flux == 1;   // AERO/Aerodynamics.java:300-303
    if (M.kind == 1) {   // AERO/SpalartAllmaras.java:36-61
        M.vtIn = P.sa_vtIn;
        M.sigma = P.sa_sigma;
        M.cb1 = P.sa_cb1;
        M.cb2 = P.sa_cb2;
        M.ka = P.sa_ka;
        M.cw1 = M.cb1 / (M.ka * M.ka) + (1 + M.cb2) / M.sigma;
        M.cw2 = P.sa_cw2;
        M.cw3 = P.sa_cw3;
        M.cv1 = P.sa_cv1;
        M.Prt = P.sa_prt;
    }
    if (M.kind == 4) {   // AERO/KOmegaSST.java:45-71 (subsonic inlet: WIn is not used)
        M.kIn = P.kw_kIn;
        M.omIn = P.kw_omIn;
        M.suth = P.kw_suth;
        M.a1kw = P.kw_a1;
        M.sk1 = P.kw_sk1;
        M.som1 = P.kw_som1;
        M.alpha1 = P.kw_alpha1;
        M.beta1 = P.kw_beta1;
        M.betast = P.kw_betast;
        M.sk2 = P.kw_sk2;
        M.som2 = P.kw_som2;
        M.alpha2 = P.kw_alpha2;
        M.beta2 = P.kw_beta2;
        M.Prt = P.kw_prt;
    }
    // kappa + cv given (AERO/Aerodynamics.java:125-128)
    M.kapa = P.kappa;
    M.cv = P.cv;
    M.cp = M.cv * M.kapa;
    // subsonic stagnation inlet (AERO/Aerodynamics.java:169-182)
    M.pRef = P.pIn0;
    M.rhoRef = P.rhoIn0;
    M.pOut = P.pOut / M.pRef;
    M.lRef = 1;
    M.velocityRef = std::sqrt(M.pRef / M.rhoRef);
    M.tRef = M.lRef / M.velocityRef;
    M.hasAttack = P.has_attack_angle != 0;
    M.attack = P.attack_angle;
    if (M.diffusive) {   // AERO/Aerodynamics.java:262-281
        const double pIn0 = 1;
        double machInf = std::sqrt(2 / (M.kapa - 1) * (std::pow((pIn0 / M.pOut), (M.kapa - 1) / M.kapa) - 1));
        double rhoInf = M.rhoRef / stagRho(M, machInf);
        double pInf = M.pRef / stagP(M, machInf);
        double aInf = std::sqrt(M.kapa * pInf / rhoInf);
        double uInf = machInf * aInf;
        M.Re = P.reynolds * (M.rhoRef * M.velocityRef) / (rhoInf * uInf);
        M.Pr = P.prandtl;
    } else {
        M.Re = -1;
        M.Pr = -1;
    }
    return M;
}

std::vector<double> const_init_condition(const ModelConst& M) {
    if (M.kind == 2) return {0.0};
    // AERO/NavierStokes.java:105-137 (subsonic inlet branch)
    double kapa = M.kapa;
    double machIn = std::sqrt(2 / (kapa - 1) * (std::pow((1 / M.pOut), (kapa - 1) / kapa) - 1));
    double rhoIn = std::pow(1 + ((kapa - 1) / 2) * machIn * machIn, 1 / (1 - kapa));
    double VIn = machIn * std::sqrt((kapa * M.pOut) / rhoIn);
    double alpha = M.hasAttack ? M.attack : 0;
    double vIn[3] = {VIn * std::cos(alpha), VIn * std::sin(alpha), 0.0};
    std::vector<double> W(M.nEqs);
    W[0] = rhoIn;
    for (int d = 0; d < M.dim; ++d) W[d + 1] = rhoIn * vIn[d];
    if (M.kind == 3) return W;   // AERO/AdiabaticEulers.java:97-126: no energy
    W[M.dim + 1] = M.pOut / (kapa - 1) + 0.5 * rhoIn * VIn * VIn;
    if (M.kind == 1) W[M.dim + 2] = rhoIn * M.vtIn;   // AERO/SpalartAllmaras.java:135
    if (M.kind == 4) {   // AERO/KOmegaSST.java:109-151 (subsonic branch)
        W[M.dim + 2] = rhoIn * M.kIn;
        W[M.dim + 3] = rhoIn * M.omIn;
    }
    return W;
}

// ------------------------------------------------------------------ quadrature
static Rule read_rule(const std::string& path, int dimCols) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open quadrature file " + path);
    Rule r;
    f >> r.n;
    if (dimCols == 1) {
        r.x.resize(r.n);
        for (int i = 0; i < r.n; i++) f >> r.x[i];
    } else {
        // barycentric -> Cartesian on the reference triangle (0,0),(1,0),(0,1)
        // (C/quadrature/QuadratureTriangle.java:26-62)
        r.x.resize(2 * r.n);
        const double X[3] = {0, 1, 0}, Y[3] = {0, 0, 1};
        for (int i = 0; i < r.n; i++) {
            double b[3];
            f >> b[0] >> b[1] >> b[2];
            double cx = 0, cy = 0;
            for (int j = 0; j < 3; j++) { cx += X[j] * b[j]; cy += Y[j] * b[j]; }
            r.x[2 * i] = cx;
            r.x[2 * i + 1] = cy;
        }
    }
    r.w.resize(r.n);
    for (int i = 0; i < r.n; i++) f >> r.w[i];
    if (!f) throw std::runtime_error("bad quadrature file " + path);
    return r;
}

Rule load_line_rule(const std::string& dir, int order) {
    return read_rule(dir + "/gauss/" + std::to_string(order) + ".txt", 1);
}

// C/quadrature/QuadratureSquare.java:7-25: the tensor product of the Gauss rule of the order on
// [0, 1]^2, point (i, j) -> ((1 + x_i) / 2, (1 + x_j) / 2), weight w_i w_j / 4, i outer
Rule load_square_rule(const std::string& dir, int order) {
    Rule l = load_line_rule(dir, order), r;
    r.n = l.n * l.n;
    r.x.resize(2 * r.n);
    r.w.resize(r.n);
    int s = 0;
    for (int i = 0; i < l.n; i++)
        for (int j = 0; j < l.n; j++) {
            r.x[2 * s] = (1 + l.x[i]) / 2;
            r.x[2 * s + 1] = (1 + l.x[j]) / 2;
            r.w[s] = l.w[i] * l.w[j] / 4;
            ++s;
        }
    return r;
}

Rule load_triangle_rule(const std::string& dir, int order) {
    int deg = 2 * order - 1;   // C/quadrature/QuadratureCentral.java:46-60
    Rule r = read_rule(dir + "/gaussTri/" + std::to_string(deg) + ".txt", 2);
    bool neg = false;
    for (double w : r.w) if (w < 0) neg = true;
    if (neg) r = read_rule(dir + "/gaussTri/" + std::to_string(deg + 1) + ".txt", 2);
    return r;
}

// ------------------------------------------------------------------ basis
// C/basis/basis2Dtriangle.java:22-118; with square = true C/basis/basis2Dsquare.java:17-104 (the
// same monomial exponents, all order^2 of them, and the Vandermonde matrix at the tensor nodes
// X = (b_i, b_j), b = linspace(0, 1, order), i outer)
struct BasisTri {
    int order, nBasis;
    std::vector<int> xExp, yExp;
    std::vector<std::vector<double>> coeffs;
    explicit BasisTri(int ord, bool square = false) : order(ord) {
        nBasis = square ? order * order : order * (1 + order) / 2;
        xExp.assign(order * order, 0);
        yExp.assign(order * order, 0);
        int s = 0;
        for (int i = 0; i < order; i++)
            for (int j = 0; j <= i; j++) { xExp[s] = i - j; yExp[s] = j; ++s; }
        for (int i = order - 1; i > 0; i--)
            for (int j = 0; j < i; j++) { xExp[s] = order - 1 - j; yExp[s] = order - i + j; ++s; }
        if (order == 1) {
            coeffs = {{1.0}};
        } else if (square) {
            auto b = linspace(0, 1, order);
            std::vector<std::vector<double>> V(nBasis, std::vector<double>(nBasis));
            int r = 0;
            for (int i = 0; i < order; i++)
                for (int j = 0; j < order; j++) {
                    for (int k = 0; k < nBasis; k++) V[r][k] = std::pow(b[i], xExp[k]) * std::pow(b[j], yExp[k]);
                    ++r;
                }
            coeffs = mat_invert(V);
        } else {
            auto b = barycentr(order, nBasis);
            std::vector<std::vector<double>> V(nBasis, std::vector<double>(nBasis));
            for (int i = 0; i < nBasis; i++) {
                double X0 = b[i][1], X1 = b[i][2];
                for (int j = 0; j < nBasis; j++) V[i][j] = std::pow(X0, xExp[j]) * std::pow(X1, yExp[j]);
            }
            coeffs = mat_invert(V);
        }
    }
    static std::vector<double> linspace(double a, double b, int n) {   // API/Mat.java:42-53
        std::vector<double> x(n);
        if (n == 1) x[0] = (a + b) / 2;
        else
            for (int i = 0; i < n; i++) {
                double xi = (double)i / (n - 1);
                x[i] = (1 - xi) * a + xi * b;
            }
        return x;
    }
    static std::vector<std::array<double, 3>> barycentr(int rad, int ne) {
        std::vector<std::array<double, 3>> b(ne);
        auto x = linspace(0, 1, rad);
        int s = 0;
        for (int i = 0; i < rad; i++) {
            auto y = linspace(0, x[i], i + 1);
            for (int j = 0; j < i + 1; j++) {
                b[s][0] = 1 - x[i];
                b[s][1] = y[j];
                b[s][2] = x[i] - y[j];
                s++;
            }
        }
        return b;
    }
    double fun(int m, const double* Xi) const {
        double v = 0;
        for (int i = 0; i < nBasis; i++) v = v + coeffs[i][m] * std::pow(Xi[0], xExp[i]) * std::pow(Xi[1], yExp[i]);
        return v;
    }
    double der(int m, const double* Xi, int dim) const {
        double der = 0;
        if (dim == 0) {
            for (int i = 0; i < nBasis; i++)
                if (xExp[i] > 0)
                    der = der + coeffs[i][m] * xExp[i] * std::pow(Xi[0], xExp[i] - 1) * std::pow(Xi[1], yExp[i]);
        } else {
            for (int i = 0; i < nBasis; i++)
                if (yExp[i] > 0)
                    der = der + coeffs[i][m] * yExp[i] * std::pow(Xi[0], xExp[i]) * std::pow(Xi[1], yExp[i] - 1);
        }
        return der;
    }
};

// ------------------------------------------------------------------ transformation
static const double CX_TRI[3][2] = {{0, 0}, {1, 0}, {0, 1}};            // transformation2Dtriangle.coordsXi
static const double CX_SQ[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};     // transformation2Dsquare.coordsXi
const double* elem_ref_vertex(const Elem& e, int j) { return e.nF == 4 ? CX_SQ[j] : CX_TRI[j]; }

// C/transformation/transformation2Dsquare.java:12-48: rows (xi eta, xi, eta, 1) at the reference
// vertices, A = invV x, B = invV y
static void transform_init_square(Elem& e) {
    std::vector<std::vector<double>> V(4, std::vector<double>(4));
    for (int i = 0; i < 4; i++) {
        V[i][0] = CX_SQ[i][0] * CX_SQ[i][1];
        V[i][1] = CX_SQ[i][0];
        V[i][2] = CX_SQ[i][1];
        V[i][3] = 1;
    }
    auto invV = mat_invert(V);
    for (int i = 0; i < 4; i++) {
        double a = 0, b = 0;
        for (int j = 0; j < 4; j++) { a += invV[i][j] * e.vert[j][0]; b += invV[i][j] * e.vert[j][1]; }
        e.A[i] = a;
        e.B[i] = b;
    }
}
static void getX_square(const Elem& e, const double* Xi, double* X) {   // transformation2Dsquare.java:50-53
    X[0] = e.A[0] * Xi[0] * Xi[1] + e.A[1] * Xi[0] + e.A[2] * Xi[1] + e.A[3];
    X[1] = e.B[0] * Xi[0] * Xi[1] + e.B[1] * Xi[0] + e.B[2] * Xi[1] + e.B[3];
}
// getDX / jacobiMatrix (transformation2Dsquare.java:94-125): J[i][j] = dX_i / dXi_j
static void jacobi_square(const Elem& e, const double* Xi, double J[2][2]) {
    J[0][0] = e.A[0] * Xi[1] + e.A[1];
    J[0][1] = e.A[0] * Xi[0] + e.A[2];
    J[1][0] = e.B[0] * Xi[1] + e.B[1];
    J[1][1] = e.B[0] * Xi[0] + e.B[2];
}
// transformation2Dsquare.getXi (:64-91): Newton from (0.5, 0.5) with a forward-difference
// Jacobian (h = 1e-8), at most 50 steps, stop when |Xi_new - Xi| < 1e-13
static void getXi_square(const Elem& e, const double* X, double* out) {
    double Xi[2] = {0.5, 0.5}, Xin[2] = {0.5, 0.5};
    const double h = 1e-8;
    for (int i = 0; i < 50; i++) {
        std::vector<std::vector<double>> J(2, std::vector<double>(2));
        double X0[2];
        getX_square(e, Xi, X0);
        for (int j = 0; j < 2; j++) {
            double Xp[2] = {Xi[0], Xi[1]}, Xh[2];
            Xp[j] = Xi[j] + h;   // Mat.plusVec(Xi, V) with V = h e_j
            getX_square(e, Xp, Xh);
            for (int d = 0; d < 2; d++) J[j][d] = (Xh[d] - X0[d]) * (1 / h);
        }
        std::swap(J[0][1], J[1][0]);   // Mat.transposeInPlace
        auto iJ = mat_invert(J);
        double r[2] = {X0[0] - X[0], X0[1] - X[1]};
        for (int d = 0; d < 2; d++) Xin[d] = Xi[d] - (iJ[d][0] * r[0] + iJ[d][1] * r[1]);
        double dd = std::sqrt((Xin[0] - Xi[0]) * (Xin[0] - Xi[0]) + (Xin[1] - Xi[1]) * (Xin[1] - Xi[1]));
        if (dd < 1e-13) break;
        if (std::isnan(Xin[0]) || std::isnan(Xin[1])) { Xin[0] = Xin[1] = 0.5; break; }
        Xi[0] = Xin[0];
        Xi[1] = Xin[1];
    }
    out[0] = Xin[0];
    out[1] = Xin[1];
}

// C/transformation/transformation2Dtriangle.java:40-119
static void transform_init(Elem& e) {
    static const double CX[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    std::vector<std::vector<double>> V(3, std::vector<double>(3)), Vi(3, std::vector<double>(3));
    for (int i = 0; i < 3; i++) {
        V[i][0] = CX[i][0]; V[i][1] = CX[i][1]; V[i][2] = 1;
        Vi[i][0] = e.vert[i][0]; Vi[i][1] = e.vert[i][1]; Vi[i][2] = 1;
    }
    auto invV = mat_invert(V);
    auto invVi = mat_invert(Vi);
    auto times = [](const std::vector<std::vector<double>>& A, const double* b, double* out) {
        for (int i = 0; i < 3; i++) {
            double s = 0;
            for (int j = 0; j < 3; j++) s += A[i][j] * b[j];
            out[i] = s;
        }
    };
    double aux[3];
    for (int i = 0; i < 3; i++) aux[i] = e.vert[i][0];
    times(invV, aux, e.A);
    for (int i = 0; i < 3; i++) aux[i] = e.vert[i][1];
    times(invV, aux, e.B);
    for (int i = 0; i < 3; i++) aux[i] = CX[i][0];
    times(invVi, aux, e.iA);
    for (int i = 0; i < 3; i++) aux[i] = CX[i][1];
    times(invVi, aux, e.iB);
}
static void getX(const Elem& e, const double* Xi, double* X) {
    if (e.nF == 4) { getX_square(e, Xi, X); return; }
    X[0] = e.A[0] * Xi[0] + e.A[1] * Xi[1] + e.A[2];
    X[1] = e.B[0] * Xi[0] + e.B[1] * Xi[1] + e.B[2];
}
void elem_getX(const Elem& e, const double* Xi, double* X) { getX(e, Xi, X); }
static void getXi(const Elem& e, const double* X, double* Xi) {
    if (e.nF == 4) { getXi_square(e, X, Xi); return; }
    Xi[0] = e.iA[0] * X[0] + e.iA[1] * X[1] + e.iA[2];
    Xi[1] = e.iB[0] * X[0] + e.iB[1] * X[1] + e.iB[2];
}
// Jacobian of the map at Xi: |det| / 2 on the triangle (transformation2Dtriangle), |det| on the
// square (transformation2Dsquare.jacobian, :128-134)
static double jacobianVol(const Elem& e, const double* Xi) {
    if (e.nF == 4) {
        double J[2][2];
        jacobi_square(e, Xi, J);
        return std::fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    }
    return std::fabs(e.A[0] * e.B[1] - e.A[1] * e.B[0]) / 2;
}
static void jacobi_matrix(const Elem& e, const double* Xi, double J[2][2]) {
    if (e.nF == 4) { jacobi_square(e, Xi, J); return; }
    J[0][0] = e.A[0]; J[0][1] = e.A[1];
    J[1][0] = e.B[0]; J[1][1] = e.B[1];
}
// Transformation.transformBasis: iT = invert(jacobiMatrix(Xi)), d_x = sum_j iT[j][i] d_xi_j
// (C/transformation/Transformation.java:146-163)
static void transform_basis(const Elem& e, const double* Xi, const double* dXi, double* dX, int nb) {
    double Jm[2][2];
    jacobi_matrix(e, Xi, Jm);
    std::vector<std::vector<double>> J = {{Jm[0][0], Jm[0][1]}, {Jm[1][0], Jm[1][1]}};
    auto iT = mat_invert(J);
    for (int m = 0; m < nb; m++)
        for (int i = 0; i < 2; i++) {
            double s = 0;
            for (int j = 0; j < 2; j++) s += iT[j][i] * dXi[m * 2 + j];
            dX[m * 2 + i] = s;
        }
}

// ------------------------------------------------------------------ neighbour matrix
std::vector<int> neighbor_matrix(int nElem, const int* TP, int nBnd, const int* bnd, int nF) {
    // key = sorted vertex pair (C/FlowProMain.java:786-806 uses the string "a$b$")
    std::map<std::pair<int, int>, std::pair<int, int>> faces;
    auto key = [&](int i, int k) {
        int a = TP[nF * i + k], b = TP[nF * i + (k + 1) % nF];
        return a < b ? std::make_pair(a, b) : std::make_pair(b, a);
    };
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nF; k++) {
            auto kk = key(i, k);
            auto it = faces.find(kk);
            if (it == faces.end()) faces[kk] = {i, -1000};
            else it->second.second = i;
        }
    std::vector<int> TT((size_t)nF * nElem, -1000);
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nF; k++) {
            auto& f = faces[key(i, k)];
            if (f.first == i) {
                if (f.second != -1000) TT[nF * i + k] = f.second;
            } else {
                TT[nF * i + k] = f.first;
            }
        }
    std::map<std::pair<int, int>, int> bmap;   // :834-848
    for (int r = 0; r < nBnd; r++) {
        int a = bnd[3 * r + 1], b = bnd[3 * r + 2];
        bmap[a < b ? std::make_pair(a, b) : std::make_pair(b, a)] = bnd[3 * r];
    }
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nF; k++)
            if (TT[nF * i + k] == -1000) {
                auto it = bmap.find(key(i, k));
                TT[nF * i + k] = it != bmap.end() ? it->second : -1;
            }
    return TT;
}

// ------------------------------------------------------------------ RCM
static int pseudo_peripheral(const std::vector<std::vector<int>>& adj, const std::vector<char>& excluded, int start) {
    int node = start;
    int ecc = -1;
    int n = (int)adj.size();
    std::vector<int> level(n);
    for (int iter = 0; iter < 10; iter++) {
        std::fill(level.begin(), level.end(), -1);
        std::deque<int> q;
        q.push_back(node);
        level[node] = 0;
        int last = node;
        while (!q.empty()) {
            int e = q.front(); q.pop_front();
            for (int t : adj[e]) {
                if (!excluded[t] && level[t] == -1) {
                    level[t] = level[e] + 1;
                    q.push_back(t);
                    if (level[t] > level[last] || (level[t] == level[last] && adj[t].size() < adj[last].size())) last = t;
                }
            }
        }
        if (level[last] <= ecc) break;
        ecc = level[last];
        node = last;
    }
    return node;
}

std::vector<int> rcm_order(int n, const int* TT, int nF) {
    std::vector<std::vector<int>> adj(n);
    std::vector<int> degree(n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < nF; k++) {
            int t = TT[nF * i + k];
            if (t > -1 && t != i && std::find(adj[i].begin(), adj[i].end(), t) == adj[i].end()) adj[i].push_back(t);
        }
        degree[i] = (int)adj[i].size();
    }
    for (int i = 0; i < n; i++)   // stable sort by degree (Stream.sorted is stable)
        std::stable_sort(adj[i].begin(), adj[i].end(), [&](int a, int b) { return degree[a] < degree[b]; });
    std::vector<int> order(n);
    std::vector<char> visited(n, 0);
    int count = 0;
    while (count < n) {
        int start = -1;
        for (int i = 0; i < n; i++)
            if (!visited[i] && (start == -1 || degree[i] < degree[start])) start = i;
        start = pseudo_peripheral(adj, visited, start);
        std::deque<int> q;
        q.push_back(start);
        visited[start] = 1;
        while (!q.empty()) {
            int e = q.front(); q.pop_front();
            order[count++] = e;
            for (int t : adj[e])
                if (!visited[t]) { visited[t] = 1; q.push_back(t); }
        }
    }
    for (int i = 0; i < n / 2; i++) std::swap(order[i], order[n - 1 - i]);
    return order;
}

// ------------------------------------------------------------------ case build
// C/Mesh.java:69-224 (constructor and init) for straight triangles, one global order.
std::vector<double> wall_distance(int nPts, const double* PXY, int nElem, const int* TP, const int* TT, int nF) {
    const double INF = std::numeric_limits<double>::infinity();
    std::vector<double> distance(nPts, INF);
    // triangleElement / squareElement.getFaceIndexes(k) = {k, (k + 1) % nF}
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nF; k++)
            if (TT[nF * i + k] == -1) {
                distance[TP[nF * i + k]] = 0;
                distance[TP[nF * i + (k + 1) % nF]] = 0;
            }
    bool changed = true;
    for (int iter = 0; changed && iter <= nPts; iter++) {
        changed = false;
        for (int i = 0; i < nElem; i++)
            for (int j = 0; j < nF; j++) {
                const int pj = TP[nF * i + j];
                for (int k = 0; k < nF; k++) {
                    const int pk = TP[nF * i + k];
                    if (j != k && distance[pk] < distance[pj]) {
                        double L = 0;
                        for (int d = 0; d < 2; d++) L += (PXY[2 * pj + d] - PXY[2 * pk + d]) * (PXY[2 * pj + d] - PXY[2 * pk + d]);
                        const double y = distance[pk] + std::sqrt(L);
                        if (y < distance[pj]) {
                            distance[pj] = y;
                            changed = true;
                        }
                    }
                }
            }
    }
    for (int i = 0; i < nPts; i++)
        if (std::isinf(distance[i])) {
            std::fill(distance.begin(), distance.end(), 1e30);
            break;
        }
    return distance;
}

// FlowProMain.createNeighborMatrix with a vertex count per element (elementType.txt: 3 triangle,
// 4 square; face k between vertices k and (k + 1) % nV); TP and TT rows have `stride` entries,
// the unused entries of TT are -1000
static std::vector<int> neighbor_matrix_v(int nElem, const int* TP, int stride, const std::vector<int>& nV, int nBnd,
                                          const int* bnd) {
    std::map<std::pair<int, int>, std::pair<int, int>> faces;
    auto key = [&](int i, int k) {
        int a = TP[stride * i + k], b = TP[stride * i + (k + 1) % nV[i]];
        return a < b ? std::make_pair(a, b) : std::make_pair(b, a);
    };
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nV[i]; k++) {
            auto kk = key(i, k);
            auto it = faces.find(kk);
            if (it == faces.end()) faces[kk] = {i, -1000};
            else it->second.second = i;
        }
    std::vector<int> TT((size_t)stride * nElem, -1000);
    std::vector<char> open((size_t)stride * nElem, 0);
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nV[i]; k++) {
            auto& f = faces[key(i, k)];
            open[stride * i + k] = 1;
            if (f.first == i) {
                if (f.second != -1000) { TT[stride * i + k] = f.second; open[stride * i + k] = 0; }
            } else {
                TT[stride * i + k] = f.first;
                open[stride * i + k] = 0;
            }
        }
    std::map<std::pair<int, int>, int> bmap;
    for (int r = 0; r < nBnd; r++) {
        int a = bnd[3 * r + 1], b = bnd[3 * r + 2];
        bmap[a < b ? std::make_pair(a, b) : std::make_pair(b, a)] = bnd[3 * r];
    }
    for (int i = 0; i < nElem; i++)
        for (int k = 0; k < nV[i]; k++)
            if (open[stride * i + k]) {
                auto it = bmap.find(key(i, k));
                TT[stride * i + k] = it != bmap.end() ? it->second : -1;
            }
    return TT;
}

// C/Mesh.java:69-224 (constructor and init) for straight triangles and squares.  types / orders
// (nullable) give every element's type (3, 4) and order (order.txt); with either of them TP has 4
// entries per element (the fourth unused on triangles).  Quadrature orders follow the element
// order (Parameters.volumeQuadratureOrder "auto"); a face takes the larger order of its two sides
// (Integration.Face, C/Integration.java:114-117).
Case* build_case(const Props& P, int nPts, const double* PXY, int nElem, const int* TP,
                 int nBnd, const int* bnd, const double* wallDist, const std::string& rulesDir,
                 const int* types, const int* orders) {
    Case* c = new Case();
    c->P = P;
    c->Mc = model_init(P);
    c->state_dt = P.fixed_dt ? P.dt : 0.0;   // State constructor (C/State.java:75-79)
    c->nPts = nPts;
    c->nElem = nElem;
    const bool mixed = types || orders;
    std::vector<int> nV(nElem), ord(nElem);
    for (int i = 0; i < nElem; i++) {
        nV[i] = types ? (types[i] == 4 ? 4 : 3) : (P.elem_type == 4 ? 4 : 3);
        ord[i] = orders ? orders[i] : P.order;
    }
    const int stride = mixed ? 4 : nV[0];
    c->nF = stride;
    c->mixed = mixed;
    c->PXY.assign(PXY, PXY + 2 * nPts);
    c->TP.assign(TP, TP + (size_t)stride * nElem);
    c->TT = neighbor_matrix_v(nElem, TP, stride, nV, nBnd, bnd);
    const int nEqs = c->Mc.nEqs;
    // bases and rules per (square, order)
    std::map<std::pair<bool, int>, BasisTri*> bases;
    std::map<std::pair<bool, int>, Rule> vrules;
    std::map<int, Rule> lrules;
    auto basis_of = [&](bool sq, int k) -> const BasisTri& {
        auto it = bases.find({sq, k});
        if (it == bases.end()) it = bases.emplace(std::make_pair(sq, k), new BasisTri(k, sq)).first;
        return *it->second;
    };
    auto vrule_of = [&](bool sq, int k) -> const Rule& {
        auto it = vrules.find({sq, k});
        if (it == vrules.end())
            it = vrules.emplace(std::make_pair(sq, k), sq ? load_square_rule(rulesDir, k) : load_triangle_rule(rulesDir, k)).first;
        return it->second;
    };
    auto lrule_of = [&](int k) -> const Rule& {
        auto it = lrules.find(k);
        if (it == lrules.end()) it = lrules.emplace(k, load_line_rule(rulesDir, k)).first;
        return it->second;
    };
    c->vol = vrule_of(nV[0] == 4, ord[0]);
    c->face = lrule_of(ord[0]);
    c->E.resize(nElem);
    for (int i = 0; i < nElem; i++) {
        Elem& e = c->E[i];
        e.index = i;
        e.order = ord[i];
        e.nF = nV[i];
        e.nBasis = basis_of(e.nF == 4, e.order).nBasis;
        for (int k = 0; k < e.nF; k++) {
            e.TT[k] = c->TT[stride * i + k];
            e.TP[k] = TP[stride * i + k];
            e.vert[k][0] = PXY[2 * TP[stride * i + k]];
            e.vert[k][1] = PXY[2 * TP[stride * i + k] + 1];
            e.wallDist[k] = wallDist ? wallDist[TP[stride * i + k]] : 0.0;
        }
        if (e.nF == 4) transform_init_square(e);
        else transform_init(e);
        e.dampInner.assign(nEqs, 0.0);
        e.innerIndicatorOld.assign(nEqs, 0.0);
    }
    // DOF numbering (C/Mesh.java:181-187, C/element/Element.java:241-248)
    if (P.ordering == 1) c->order = rcm_order(nElem, c->TT.data(), stride);
    else {
        c->order.resize(nElem);
        for (int i = 0; i < nElem; i++) c->order[i] = i;
    }
    int dofs0 = 0;
    for (int idx : c->order) { c->E[idx].gIndex0 = dofs0; dofs0 += c->E[idx].nBasis * nEqs; }
    c->dofs = dofs0;
    c->off.assign(nElem + 1, 0);
    for (int i = 0; i < nElem; i++) c->off[i + 1] = c->off[i] + c->E[i].nBasis * nEqs;
    // integration tables, geometry, mass (C/Integration.java, C/element/Element.java:516-567)
    for (int i = 0; i < nElem; i++) {
        Elem& e = c->E[i];
        const bool square = e.nF == 4;
        const BasisTri& basis = basis_of(square, e.order);
        const Rule& qv = vrule_of(square, e.order);
        const int Nb = e.nBasis;
        e.nq = qv.n;
        e.xiV = qv.x;
        e.JacV.resize(qv.n);
        for (int p = 0; p < qv.n; p++) e.JacV[p] = jacobianVol(e, &qv.x[2 * p]);
        e.wV = qv.w;
        e.baseV.resize(qv.n * Nb);
        e.dbaseV.resize(qv.n * Nb * 2);
        std::vector<double> dXiRef(Nb * 2);
        for (int p = 0; p < qv.n; p++) {
            for (int m = 0; m < Nb; m++) e.baseV[p * Nb + m] = basis.fun(m, &qv.x[2 * p]);
            for (int d = 0; d < 2; d++)
                for (int m = 0; m < Nb; m++) dXiRef[m * 2 + d] = basis.der(m, &qv.x[2 * p], d);
            transform_basis(e, &qv.x[2 * p], dXiRef.data(), &e.dbaseV[p * Nb * 2], Nb);
        }
        // getInterpolant: barycentric on the triangle, bilinear on the square
        // (transformation2Dsquare.java:137-150)
        e.interpV.resize(qv.n * e.nF);
        for (int p = 0; p < qv.n; p++) {
            double xi = qv.x[2 * p], eta = qv.x[2 * p + 1];
            if (square) {
                e.interpV[4 * p] = (1 - xi) * (1 - eta);
                e.interpV[4 * p + 1] = xi * (1 - eta);
                e.interpV[4 * p + 2] = xi * eta;
                e.interpV[4 * p + 3] = (1 - xi) * eta;
            } else {
                e.interpV[3 * p] = 1 - xi - eta;
                e.interpV[3 * p + 1] = xi;
                e.interpV[3 * p + 2] = eta;
            }
        }
    }
    for (int i = 0; i < nElem; i++) {
        Elem& e = c->E[i];
        const int Nb = e.nBasis;
        const BasisTri& basis = basis_of(e.nF == 4, e.order);
        for (int k = 0; k < e.nF; k++) {
            // faceTransformation2Dline (C/transformation/faceTransformation2Dline.java:23-60) with the
            // face rule of the larger order of the two sides
            int f0 = k, f1 = (k + 1) % e.nF;
            const double* CX0 = elem_ref_vertex(e, f0);
            const double* CX1 = elem_ref_vertex(e, f1);
            double ref0[2] = {CX0[0], CX0[1]}, ref1[2] = {CX1[0], CX1[1]};
            const int t = e.TT[k];
            const Rule& qf = lrule_of(t > -1 ? std::max(e.order, c->E[t].order) : e.order);
            int nf = qf.n;
            e.nqf[k] = nf;
            e.wF[k] = qf.w;
            e.JacF[k].resize(nf);
            e.baseL[k].resize(nf * Nb);
            e.dbaseL[k].resize(nf * Nb * 2);
            e.interpF[k].resize(nf * 2);
            e.xiF[k].resize(nf);
            double s[2] = {ref1[0] - ref0[0], ref1[1] - ref0[1]};
            for (int p = 0; p < nf; p++) {
                double xi = qf.x[p];
                e.xiF[k][p] = xi;
                double XiRef[2];
                XiRef[0] = (1 + xi) / 2 * ref1[0] + (1 - xi) / 2 * ref0[0];
                XiRef[1] = (1 + xi) / 2 * ref1[1] + (1 - xi) / 2 * ref0[1];
                // Js = jacobiMatrix(XiRef) s (Transformation.jacobianMatrixDirection)
                double Jm[2][2];
                jacobi_matrix(e, XiRef, Jm);
                double Js[2] = {0, 0};
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++) Js[a] += Jm[a][b] * s[b];
                double Jsn = std::sqrt(Js[0] * Js[0] + Js[1] * Js[1]);
                e.JacF[k][p] = Jsn / 2;
                double nrm[2] = {Js[1], -Js[0]};
                double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]);
                e.n[k][p][0] = nrm[0] / nn;
                e.n[k][p][1] = nrm[1] / nn;
                e.interpF[k][2 * p] = (1 - xi) / 2;
                e.interpF[k][2 * p + 1] = (1 + xi) / 2;
                for (int m = 0; m < Nb; m++) e.baseL[k][p * Nb + m] = basis.fun(m, XiRef);
                double dXi[128];
                for (int d = 0; d < 2; d++)
                    for (int m = 0; m < Nb; m++) dXi[m * 2 + d] = basis.der(m, XiRef, d);
                transform_basis(e, XiRef, dXi, &e.dbaseL[k][p * Nb * 2], Nb);
            }
            if (t > -1) {
                // the neighbour's own basis at the same physical points (elemRight.basis at
                // elemRight.transform.getXi)
                Elem& er = c->E[t];
                const BasisTri& bR = basis_of(er.nF == 4, er.order);
                const int nR = er.nBasis;
                e.baseR[k].resize(nf * nR);
                e.dbaseR[k].resize(nf * nR * 2);
                for (int p = 0; p < nf; p++) {
                    double xi = qf.x[p];
                    double XiRef[2];
                    XiRef[0] = (1 + xi) / 2 * ref1[0] + (1 - xi) / 2 * ref0[0];
                    XiRef[1] = (1 + xi) / 2 * ref1[1] + (1 - xi) / 2 * ref0[1];
                    double X[2], XiR[2];
                    getX(e, XiRef, X);
                    getXi(er, X, XiR);
                    for (int m = 0; m < nR; m++) e.baseR[k][p * nR + m] = bR.fun(m, XiR);
                    double dXi[128];
                    for (int d = 0; d < 2; d++)
                        for (int m = 0; m < nR; m++) dXi[m * 2 + d] = bR.der(m, XiR, d);
                    transform_basis(er, XiR, dXi, &e.dbaseR[k][p * nR * 2], nR);
                }
            }
        }
        // computeGeometry
        for (int k = 0; k < e.nF; k++) {
            double mid[2] = {0.5 * (0), 0};
            int f0 = k, f1 = (k + 1) % e.nF;
            const double* CX0 = elem_ref_vertex(e, f0);
            const double* CX1 = elem_ref_vertex(e, f1);
            double XiRef[2];
            // faceTransform.getXs = getX(Xi = 0): (1+0)/2*ref1 + (1-0)/2*ref0
            XiRef[0] = (1 + 0.0) / 2 * CX1[0] + (1 - 0.0) / 2 * CX0[0];
            XiRef[1] = (1 + 0.0) / 2 * CX1[1] + (1 - 0.0) / 2 * CX0[1];
            getX(e, XiRef, mid);
            e.Xes[k][0] = mid[0];
            e.Xes[k][1] = mid[1];
            e.S[k] = 0;
            for (int p = 0; p < e.nqf[k]; p++) e.S[k] = e.S[k] + e.JacF[k][p] * e.wF[k][p];
        }
        // transform.getXs: the centroid of the reference triangle, the centre of the square
        const bool square = e.nF == 4;
        double centre[2] = {square ? 0.5 : 1.0 / 3, square ? 0.5 : 1.0 / 3};
        getX(e, centre, e.Xs);
        e.area = 0;
        for (int p = 0; p < e.nq; p++) e.area = e.area + e.JacV[p] * e.wV[p];
        e.elemSize = 0;
        for (int k = 0; k < e.nF; k++) e.elemSize += e.S[k];
        e.elemSize = e.nF * e.area / e.elemSize;
        // mass matrix and basis weights (C/element/DGpure.java:547-573)
        e.M.assign(Nb * Nb, 0.0);
        e.Is.assign(Nb, 0.0);
        for (int a = 0; a < Nb; a++) {
            for (int b = 0; b < Nb; b++)
                for (int p = 0; p < e.nq; p++)
                    e.M[a * Nb + b] = e.M[a * Nb + b] + e.JacV[p] * e.wV[p] * e.baseV[p * Nb + a] * e.baseV[p * Nb + b];
            for (int p = 0; p < e.nq; p++) e.Is[a] = e.Is[a] + e.JacV[p] * e.wV[p] * e.baseV[p * Nb + a] / e.area;
        }
        // DGpure.computeOrderTruncationMatrix (C/element/DGpure.java:588-626): TrunOrd =
        // M^-1 T^T Mhat^-1 T with the Taylor basis {1, x - Xs, y - Xs} at the physical points
        if (e.order > 2) {
            const int nl = 3;
            std::vector<std::vector<double>> Mhat(nl, std::vector<double>(nl, 0.0)), T(nl, std::vector<double>(Nb, 0.0));
            std::vector<double> Xq(2 * e.nq);
            for (int p = 0; p < e.nq; p++) getX(e, &e.xiV[2 * p], &Xq[2 * p]);
            auto taylor = [&](int a, const double* X) { return a > 0 ? X[a - 1] - e.Xs[a - 1] : 1.0; };
            for (int a = 0; a < nl; a++) {
                for (int b = 0; b < nl; b++)
                    for (int p = 0; p < e.nq; p++)
                        Mhat[a][b] += e.JacV[p] * e.wV[p] * taylor(a, &Xq[2 * p]) * taylor(b, &Xq[2 * p]);
                for (int b = 0; b < Nb; b++)
                    for (int p = 0; p < e.nq; p++) T[a][b] += e.JacV[p] * e.wV[p] * taylor(a, &Xq[2 * p]) * e.baseV[p * Nb + b];
            }
            auto times = [](const std::vector<std::vector<double>>& A, const std::vector<std::vector<double>>& Bm) {
                std::vector<std::vector<double>> C(A.size(), std::vector<double>(Bm[0].size(), 0.0));
                for (size_t i = 0; i < A.size(); i++)
                    for (size_t j = 0; j < Bm[0].size(); j++)
                        for (size_t k = 0; k < Bm.size(); k++) C[i][j] += A[i][k] * Bm[k][j];
                return C;
            };
            std::vector<std::vector<double>> Tt(Nb, std::vector<double>(nl)), Mm(Nb, std::vector<double>(Nb));
            for (int a = 0; a < nl; a++)
                for (int b = 0; b < Nb; b++) Tt[b][a] = T[a][b];
            for (int a = 0; a < Nb; a++)
                for (int b = 0; b < Nb; b++) Mm[a][b] = e.M[a * Nb + b];
            auto tr = times(mat_invert(Mhat), T);
            tr = times(Tt, tr);
            tr = times(mat_invert(Mm), tr);
            e.TrunOrd.resize((size_t)Nb * Nb);
            for (int a = 0; a < Nb; a++)
                for (int b = 0; b < Nb; b++) e.TrunOrd[a * Nb + b] = tr[a][b];
        }
        const int B = Nb * nEqs;
        e.ADiag.assign(B * B, 0.0);
        for (int k = 0; k < e.nF; k++)
            if (e.TT[k] > -1) e.ANeigh[k].assign((size_t)c->E[e.TT[k]].nBasis * nEqs * B, 0.0);   // [B_nbr][B]
        e.RHS.assign(B, 0.0);
        e.spatialR.assign(B, 0.0);
    }
    for (int i = 0; i < nElem; i++) {
        Elem& e = c->E[i];
        for (int k = 0; k < e.nF; k++) {
            e.faceIndexReverse[k] = 0;
            if (e.TT[k] > -1)
                for (int s = 0; s < c->E[e.TT[k]].nF; s++)
                    if (c->E[e.TT[k]].TT[s] == i) { e.faceIndexReverse[k] = s; break; }
        }
    }
    // initial condition (C/element/DGpure.java:74-107, lagrange branch)
    auto W0 = const_init_condition(c->Mc);
    for (auto& e : c->E) {
        const int B = e.nBasis * nEqs;
        e.W.assign(B, 0.0);
        for (int j = 0; j < e.nBasis; j++)
            for (int m = 0; m < nEqs; m++) e.W[m * e.nBasis + j] = W0[m];
        e.Wo = e.W;
        e.Wo2 = e.W;
    }
    for (auto& kv : bases) delete kv.second;
    return c;
}

}  // namespace oracle
