// Oracle (test infrastructure only): CRS matrix, preconditioners, GMRES, the
// Newton/BDF driver and the extern "C" entry points used by tests through ctypes.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "oracle_core.hpp"

namespace oracle {

std::vector<std::vector<double>> mat_invert(const std::vector<std::vector<double>>& A);
void mat_invert_flat(const double* A, int n, double* out);

// ------------------------------------------------------------- CRS
// SparseMatrix.scanMatrixStructure + buildCRSformat (C/LinearSolvers/SparseMatrix.java:28-76):
// rows sorted by (row, column); H[row gE+i, col gN+j] = An[j][i] (transposed element blocks).
static void build_crs(Case& c) {
    const int nEqs = c.Mc.nEqs;
    const int n = c.dofs;
    c.Icrs.assign(n + 1, 0);
    std::vector<int> rowLen(n, 0);
    for (const auto& e : c.E) {
        int B = e.nBasis * nEqs;
        int len = B;
        for (int k = 0; k < e.nF; k++)
            if (e.TT[k] > -1) len += c.E[e.TT[k]].nBasis * nEqs;
        for (int i = 0; i < B; i++) rowLen[e.gIndex0 + i] = len;
    }
    for (int r = 0; r < n; r++) c.Icrs[r + 1] = c.Icrs[r] + rowLen[r];
    c.Jcrs.assign(c.Icrs[n], 0);
    c.H.assign(c.Icrs[n], 0.0);
    for (const auto& e : c.E) {
        int B = e.nBasis * nEqs;
        std::vector<std::pair<int, int>> blk = {{e.gIndex0, B}};   // (first DOF, block size)
        for (int k = 0; k < e.nF; k++)
            if (e.TT[k] > -1) blk.push_back({c.E[e.TT[k]].gIndex0, c.E[e.TT[k]].nBasis * nEqs});
        std::sort(blk.begin(), blk.end());
        for (int i = 0; i < B; i++) {
            int pos = c.Icrs[e.gIndex0 + i];
            for (const auto& g : blk)
                for (int j = 0; j < g.second; j++) c.Jcrs[pos++] = g.first + j;
        }
    }
}

// SparseMatrix.updateData (C/LinearSolvers/SparseMatrix.java:113-139)
static void update_data(Case& c) {
    const int nEqs = c.Mc.nEqs;
    for (const auto& e : c.E) {
        int B = e.nBasis * nEqs;
        for (int i = 0; i < B; i++) {
            int row = e.gIndex0 + i;
            int s = c.Icrs[row], t = c.Icrs[row + 1];
            auto put = [&](int col, double v) {
                int* p = std::lower_bound(&c.Jcrs[s], &c.Jcrs[0] + t, col);
                c.H[p - &c.Jcrs[0]] = v;
            };
            for (int j = 0; j < B; j++) put(e.gIndex0 + j, e.ADiag[(size_t)j * B + i]);
            for (int k = 0; k < e.nF; k++)
                if (e.TT[k] > -1) {
                    const Elem& nb = c.E[e.TT[k]];
                    for (int j = 0; j < nb.nBasis * nEqs; j++) put(nb.gIndex0 + j, e.ANeigh[k][(size_t)j * B + i]);
                }
        }
    }
}

static void spmv(const Case& c, double* y, const double* x) {   // SparseMatrix.Mult (:177-185)
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.dofs; i++) {
        double sum = 0;
        for (int j = c.Icrs[i]; j < c.Icrs[i + 1]; j++) sum += c.H[j] * x[c.Jcrs[j]];
        y[i] = sum;
    }
}
static void substr_mult(const Case& c, double* y, const double* b, const double* x) {   // :196-204
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.dofs; i++) {
        double sum = b[i];
        for (int j = c.Icrs[i]; j < c.Icrs[i + 1]; j++) sum -= c.H[j] * x[c.Jcrs[j]];
        y[i] = sum;
    }
}

// ------------------------------------------------------------- BlockILU0
// C/LinearSolvers/preconditioners/BlockILU0.java:47-95 setMatrix
static void ilu_set_matrix(Case& c) {
    int n = c.nElem;
    std::vector<int> byDofs(n);
    for (int i = 0; i < n; i++) byDofs[i] = i;
    std::stable_sort(byDofs.begin(), byDofs.end(), [&](int a, int b) { return c.E[a].gIndex0 < c.E[b].gIndex0; });
    c.ilu_elems.assign(n, 0);
    c.ilu_rank.assign(n, 0);
    for (int r = 0; r < n; r++) { c.ilu_elems[r] = byDofs[r]; c.ilu_rank[byDofs[r]] = r; }
    c.ilu_cols.assign(n, {});
    c.ilu_diag.assign(n, 0);
    c.ilu_blocks.assign(n, {});
    c.ilu_dinv.assign(n, {});
    c.ilu_y.assign(n, {});
    const int nEqs = c.Mc.nEqs;
    for (int i = 0; i < n; i++) {
        const Elem& e = c.E[c.ilu_elems[i]];
        std::vector<int> row = {i};
        for (int k = 0; k < e.nF; k++) {
            int t = e.TT[k];
            if (t > -1 && std::find(row.begin(), row.end(), c.ilu_rank[t]) == row.end()) row.push_back(c.ilu_rank[t]);
        }
        std::sort(row.begin(), row.end());
        c.ilu_cols[i] = row;
        c.ilu_diag[i] = (int)(std::lower_bound(row.begin(), row.end(), i) - row.begin());
        int B = e.nBasis * nEqs;
        c.ilu_blocks[i].resize(row.size());   // [B_i][B_col], row-major
        for (size_t q = 0; q < row.size(); q++)
            c.ilu_blocks[i][q].assign((size_t)B * c.E[c.ilu_elems[row[q]]].nBasis * nEqs, 0.0);
        c.ilu_y[i].assign(B, 0.0);
    }
}

// BlockILU0.factor (:98-129) incl. copyBlocks (:187-206)
static bool ilu_factor(Case& c) {
    const int nEqs = c.Mc.nEqs;
    const int n = c.nElem;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < n; i++) {
        const Elem& e = c.E[c.ilu_elems[i]];
        int B = e.nBasis * nEqs;
        for (auto& blk : c.ilu_blocks[i]) std::fill(blk.begin(), blk.end(), 0.0);
        auto& D = c.ilu_blocks[i][c.ilu_diag[i]];
        for (int r = 0; r < B; r++)
            for (int cc = 0; cc < B; cc++) D[(size_t)r * B + cc] += e.ADiag[(size_t)cc * B + r];
        for (int k = 0; k < e.nF; k++) {
            int t = e.TT[k];
            if (t > -1) {
                int pos = (int)(std::lower_bound(c.ilu_cols[i].begin(), c.ilu_cols[i].end(), c.ilu_rank[t]) - c.ilu_cols[i].begin());
                auto& Bk = c.ilu_blocks[i][pos];
                const int Bt = c.E[t].nBasis * nEqs;
                for (int r = 0; r < B; r++)
                    for (int cc = 0; cc < Bt; cc++) Bk[(size_t)r * Bt + cc] += e.ANeigh[k][(size_t)cc * B + r];
            }
        }
    }
    for (int i = 0; i < n; i++) {
        const auto& row = c.ilu_cols[i];
        int B = c.E[c.ilu_elems[i]].nBasis * nEqs;
        for (int kk = 0; kk < c.ilu_diag[i]; kk++) {
            int k = row[kk];
            const int Bk = c.E[c.ilu_elems[k]].nBasis * nEqs;
            const auto& Dk = c.ilu_dinv[k];
            std::vector<double> L((size_t)B * Bk, 0.0);   // Mat.times(blocks[i][kk], diagInverse[k])
            const auto& A = c.ilu_blocks[i][kk];
            for (int r = 0; r < B; r++)
                for (int cc = 0; cc < Bk; cc++) {
                    double s = 0;
                    for (int m = 0; m < Bk; m++) s = s + A[(size_t)r * Bk + m] * Dk[(size_t)m * Bk + cc];
                    L[(size_t)r * Bk + cc] = s;
                }
            c.ilu_blocks[i][kk] = L;
            for (int jj = kk + 1; jj < (int)row.size(); jj++) {
                auto it = std::lower_bound(c.ilu_cols[k].begin(), c.ilu_cols[k].end(), row[jj]);
                if (it != c.ilu_cols[k].end() && *it == row[jj]) {
                    const int Bj = c.E[c.ilu_elems[row[jj]]].nBasis * nEqs;
                    const auto& U = c.ilu_blocks[k][it - c.ilu_cols[k].begin()];   // [Bk][Bj]
                    auto& C = c.ilu_blocks[i][jj];                                 // [B][Bj]
                    for (int r = 0; r < B; r++)                                    // subtractProduct (:222-233)
                        for (int m = 0; m < Bk; m++) {
                            double a = L[(size_t)r * Bk + m];
                            if (a != 0)
                                for (int cc = 0; cc < Bj; cc++) C[(size_t)r * Bj + cc] -= a * U[(size_t)m * Bj + cc];
                        }
                }
            }
        }
        c.ilu_dinv[i].assign((size_t)B * B, 0.0);
        mat_invert_flat(c.ilu_blocks[i][c.ilu_diag[i]].data(), B, c.ilu_dinv[i].data());
        for (double v : c.ilu_dinv[i])
            if (!std::isfinite(v)) return false;   // ArithmeticException
    }
    return true;
}

// BlockILU0.apply (:132-185)
static void ilu_apply(Case& c, double* x, const double* b) {
    const int nEqs = c.Mc.nEqs;
    const int n = c.nElem;
    for (int i = 0; i < n; i++) {
        const auto& row = c.ilu_cols[i];
        auto& yi = c.ilu_y[i];
        int g = c.E[c.ilu_elems[i]].gIndex0;
        int B = (int)yi.size();
        for (int r = 0; r < B; r++) yi[r] = b[g + r];
        for (int kk = 0; kk < c.ilu_diag[i]; kk++) {
            const auto& L = c.ilu_blocks[i][kk];
            const auto& yk = c.ilu_y[row[kk]];
            const int Bk = (int)yk.size();
            for (int r = 0; r < B; r++) {
                double sum = 0;
                for (int cc = 0; cc < Bk; cc++) sum += L[(size_t)r * Bk + cc] * yk[cc];
                yi[r] -= sum;
            }
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        const auto& row = c.ilu_cols[i];
        std::vector<double> rhs = c.ilu_y[i];
        int B = (int)rhs.size();
        for (int jj = c.ilu_diag[i] + 1; jj < (int)row.size(); jj++) {
            const auto& U = c.ilu_blocks[i][jj];
            int gj = c.E[c.ilu_elems[row[jj]]].gIndex0;
            int Bj = c.E[c.ilu_elems[row[jj]]].nBasis * nEqs;
            for (int r = 0; r < B; r++) {
                double sum = 0;
                for (int cc = 0; cc < Bj; cc++) sum += U[(size_t)r * Bj + cc] * x[gj + cc];
                rhs[r] -= sum;
            }
        }
        const auto& iU = c.ilu_dinv[i];
        int gi = c.E[c.ilu_elems[i]].gIndex0;
        for (int r = 0; r < B; r++) {
            double sum = 0;
            for (int cc = 0; cc < B; cc++) sum += iU[(size_t)r * B + cc] * rhs[cc];
            x[gi + r] = sum;
        }
    }
}

// M^-T b with the BlockILU0 factors M = L U (no reference counterpart: the preconditioner of the
// transposed (adjoint) system): U^T z = b forward with the transposed diagonal inverses, then
// L^T x = z backward (unit diagonal), both as column sweeps over the stored row blocks.
static void ilu_apply_transposed(Case& c, double* x, const double* b) {
    const int n = c.nElem;
    std::vector<std::vector<double>> acc(n), z(n);
    for (int i = 0; i < n; i++) {
        const int g = c.E[c.ilu_elems[i]].gIndex0, B = (int)c.ilu_y[i].size();
        acc[i].assign(b + g, b + g + B);
    }
    for (int i = 0; i < n; i++) {
        const auto& row = c.ilu_cols[i];
        const int B = (int)acc[i].size();
        const auto& iU = c.ilu_dinv[i];
        z[i].assign(B, 0.0);
        for (int r = 0; r < B; r++) {
            double sum = 0;
            for (int cc = 0; cc < B; cc++) sum += iU[(size_t)cc * B + r] * acc[i][cc];
            z[i][r] = sum;
        }
        for (int jj = c.ilu_diag[i] + 1; jj < (int)row.size(); jj++) {
            const auto& U = c.ilu_blocks[i][jj];
            auto& aj = acc[row[jj]];
            const int Bj = (int)aj.size();
            for (int cc = 0; cc < Bj; cc++) {
                double sum = 0;
                for (int r = 0; r < B; r++) sum += U[(size_t)r * Bj + cc] * z[i][r];
                aj[cc] -= sum;
            }
        }
    }
    acc = z;
    for (int i = n - 1; i >= 0; i--) {
        const auto& row = c.ilu_cols[i];
        const int B = (int)acc[i].size(), g = c.E[c.ilu_elems[i]].gIndex0;
        for (int r = 0; r < B; r++) x[g + r] = acc[i][r];
        for (int kk = 0; kk < c.ilu_diag[i]; kk++) {
            const auto& L = c.ilu_blocks[i][kk];
            auto& ak = acc[row[kk]];
            const int Bk = (int)ak.size();
            for (int cc = 0; cc < Bk; cc++) {
                double sum = 0;
                for (int r = 0; r < B; r++) sum += L[(size_t)r * Bk + cc] * x[g + r];
                ak[cc] -= sum;
            }
        }
    }
}

// ------------------------------------------------------------- blockJacobi
// C/LinearSolvers/preconditioners/blockJacobi.java:74-90 + API/Mat.java:1307-1329 lsolveT.
// The elimination works IN PLACE on the element's ADiag (kept in bj_work between applies
// of one linear solve): only the first apply is an exact block solve.
static void lsolveT(double* A, int n, double* b, double* x) {
    for (int p = 0; p < n; p++)
        for (int i = p + 1; i < n; i++) {
            double alpha = A[(size_t)p * n + i] / A[(size_t)p * n + p];
            b[i] -= alpha * b[p];
            for (int j = p; j < n; j++) A[(size_t)j * n + i] -= alpha * A[(size_t)j * n + p];
        }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0.0;
        for (int j = i + 1; j < n; j++) sum += A[(size_t)j * n + i] * x[j];
        x[i] = (b[i] - sum) / A[(size_t)i * n + i];
    }
}
static void bj_apply(Case& c, double* x, const double* b) {
    const int nEqs = c.Mc.nEqs;
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
    for (int i = 0; i < c.nElem; i++) {
        const Elem& e = c.E[i];
        int B = e.nBasis * nEqs;
        std::vector<double> bp(B), y(B);
        for (int j = 0; j < B; j++) bp[j] = b[e.gIndex0 + j];
        lsolveT(c.bj_work[i].data(), B, bp.data(), y.data());
        for (int j = 0; j < B; j++) x[e.gIndex0 + j] = y[j];
    }
}

// ------------------------------------------------------------- ILU0 (point)
// C/LinearSolvers/preconditioners/ILU0.java:32-100 on the CRS matrix in DOF order, no pivoting;
// SparseMatrix.findDiagonalIndices gives each row's diagonal position
static bool pilu_factor(Case& c) {
    const int n = c.dofs;
    const int* IA = c.Icrs.data();
    const int* JA = c.Jcrs.data();
    if (c.diagind.empty()) {
        c.diagind.assign(n, 0);
        for (int r = 0; r < n; r++)
            for (int j = IA[r]; j < IA[r + 1]; j++)
                if (JA[j] == r) { c.diagind[r] = j; break; }
    }
    c.pilu = c.H;
    double* ilu = c.pilu.data();
    const int* diag = c.diagind.data();
    for (int k = 1; k < n; ++k) {
        for (int i = IA[k]; i < diag[k]; ++i) {
            const int index = JA[i];
            const double LUii = ilu[diag[index]];
            if (LUii == 0) return false;   // the reference throws ArithmeticException
            const double LUki = (ilu[i] /= LUii);
            for (int j = diag[index] + 1, l = IA[k] + 1; j < IA[index + 1]; ++j) {
                while (l < IA[k + 1] && JA[l] < JA[j]) l++;
                if (l < IA[k + 1] && JA[l] == JA[j]) ilu[l] -= LUki * ilu[j];
            }
        }
    }
    return true;
}
static void pilu_apply(const Case& c, double* x, const double* b) {
    const int n = c.dofs;
    const int* IA = c.Icrs.data();
    const int* JA = c.Jcrs.data();
    const int* diag = c.diagind.data();
    const double* ilu = c.pilu.data();
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
        double sum = 0;
        for (int j = IA[i]; j < diag[i]; j++) sum = sum + ilu[j] * y[JA[j]];
        y[i] = b[i] - sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0;
        for (int j = diag[i] + 1; j < IA[i + 1]; j++) sum = sum + ilu[j] * x[JA[j]];
        x[i] = (y[i] - sum) / ilu[diag[i]];
    }
}

static void prec_apply(Case& c, double* x, const double* b) {
    if (c.P.preconditioner == 3) pilu_apply(c, x, b);
    else if (c.P.preconditioner == 2) ilu_apply(c, x, b);
    else if (c.P.preconditioner == 1) bj_apply(c, x, b);
    else std::memcpy(x, b, sizeof(double) * c.dofs);
}

// ------------------------------------------------------------- GMRES
// C/LinearSolvers/Gmres.java:42-199
static double norm(const std::vector<double>& a) {
    double s = 0;
    for (double v : a) s += v * v;
    return std::sqrt(s);
}
static void rotmat(double a, double b, double& c, double& s) {
    double temp;
    if (b == 0) { c = 1; s = 0; }
    else if (std::fabs(b) > std::fabs(a)) { temp = a / b; s = 1 / std::sqrt(1 + temp * temp); c = temp * s; }
    else { temp = b / a; c = 1 / std::sqrt(1 + temp * temp); s = temp * c; }
}
static std::vector<double> lsolveH(std::vector<std::vector<double>>& A, std::vector<double>& b, int N) {
    for (int p = 0; p < N; p++)
        for (int i = p + 1; i < N; i++) {
            double alpha = A[i][p] / A[p][p];
            b[i] -= alpha * b[p];
            for (int j = p; j < N; j++) A[i][j] -= alpha * A[p][j];
        }
    std::vector<double> x(N);
    for (int i = N - 1; i >= 0; i--) {
        double sum = 0.0;
        for (int j = i + 1; j < N; j++) sum += A[i][j] * x[j];
        x[i] = (b[i] - sum) / A[i][i];
    }
    return x;
}

static bool gmres(Case& c, double* x, const double* b) {
    const int n = c.dofs, m = c.P.gmres_m, limit = c.P.gmres_cycles;
    const double tol = c.P.gmres_tol;
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(n)), H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), e1(m + 1, 0.0), w(n), r(n), aux(n), s;
    c.last_iters = 0;
    prec_apply(c, r.data(), b);
    double bnrm2 = norm(r);
    if (bnrm2 == 0) bnrm2 = 1;
    substr_mult(c, aux.data(), b, x);
    prec_apply(c, r.data(), aux.data());
    double error = norm(r) / bnrm2;
    c.last_error = error;
    if (error < tol) return true;
    e1[0] = 1;
    auto update = [&](const std::vector<double>& y, int i) {   // updateSolution (:124-130)
#pragma omp parallel for schedule(static, 1) num_threads(c.P.threads > 0 ? c.P.threads : 1)
        for (int j = 0; j < n; j++)
            for (int k = 0; k <= i; k++) x[j] += V[k][j] * y[k];
    };
    for (int iter = 0; iter < limit; iter++) {
        double norm_r = norm(r);
        for (int j = 0; j < n; j++) V[0][j] = r[j] / norm_r;
        s.assign(std::max(n, m + 1), 0.0);   // vectorScalarProduct allocates n entries
        for (int i = 0; i < m + 1; i++) s[i] = e1[i] * norm_r;
        for (int i = 0; i < m; i++) {
            spmv(c, aux.data(), V[i].data());
            prec_apply(c, w.data(), aux.data());
            c.last_iters++;
            for (int k = 0; k <= i; k++) {
                double sp = 0;
                for (int j = 0; j < n; j++) sp += w[j] * V[k][j];
                H[k][i] = sp;
                for (int j = 0; j < n; j++) w[j] = w[j] - V[k][j] * H[k][i];
            }
            H[i + 1][i] = norm(w);
            double hNorm = H[i + 1][i];
            for (int j = 0; j < n; j++) V[i + 1][j] = hNorm > 0 ? w[j] / hNorm : 0;
            for (int k = 0; k <= i - 1; k++) {
                double temp = cs[k] * H[k][i] + sn[k] * H[k + 1][i];
                H[k + 1][i] = -sn[k] * H[k][i] + cs[k] * H[k + 1][i];
                H[k][i] = temp;
            }
            rotmat(H[i][i], H[i + 1][i], cs[i], sn[i]);
            double temp = cs[i] * s[i];
            s[i + 1] = -sn[i] * s[i];
            s[i] = temp;
            H[i][i] = cs[i] * H[i][i] + sn[i] * H[i + 1][i];
            H[i + 1][i] = 0;
            error = std::fabs(s[i + 1]) / bnrm2;
            if (error <= tol) {
                auto y = lsolveH(H, s, i + 1);
                update(y, i);
                break;
            }
        }
        if (error <= tol) break;
        auto y = lsolveH(H, s, m);
        update(y, m - 1);
        substr_mult(c, aux.data(), b, x);
        prec_apply(c, r.data(), aux.data());
        s[m] = norm(r);
        error = s[m] / bnrm2;
        if (error <= tol) break;
    }
    c.last_error = error;
    return error <= tol;
}

// LinearSolver.solve (C/LinearSolvers/LinearSolver.java:91-111) incl. factory-time setup
bool linear_solve(Case& c, double* x) {
    if (c.Icrs.empty()) {
        build_crs(c);
        if (c.P.preconditioner == 2) ilu_set_matrix(c);
    }
    update_data(c);
    std::vector<double> b(c.dofs);
    for (const auto& e : c.E)   // updateB (:141-152)
        for (int i = 0; i < (int)e.RHS.size(); i++) b[e.gIndex0 + i] = e.RHS[i];
    if (c.P.preconditioner == 3) {
        if (!pilu_factor(c)) return false;
    } else if (c.P.preconditioner == 2) {
        if (!ilu_factor(c)) return false;
    } else if (c.P.preconditioner == 1) {
        c.bj_work.assign(c.nElem, {});
        for (int i = 0; i < c.nElem; i++) c.bj_work[i] = c.E[i].ADiag;
    }
    bool flag = gmres(c, x, b.data());
    for (int i = 0; i < c.dofs; i++)
        if (!std::isfinite(x[i])) return false;
    return flag;
}

// ------------------------------------------------------------- norms, driver
// Element.l2NormsSquared (C/element/Element.java:357-373)
static void l2_norms_sq(const Case& c, const Elem& e, const std::vector<double>& R, double* out) {
    int Nb = e.nBasis;
    std::vector<std::vector<double>> M(Nb, std::vector<double>(Nb));
    for (int i = 0; i < Nb; i++)
        for (int j = 0; j < Nb; j++) M[i][j] = e.M[i * Nb + j];
    auto invM = mat_invert(M);
    for (int m = 0; m < c.Mc.nEqs; m++) {
        int s = m * Nb;
        double sum = 0;
        for (int i = 0; i < Nb; i++) {
            double r = 0;
            for (int j = 0; j < Nb; j++) r += invM[i][j] * R[s + j];
            sum += R[s + i] * r;
        }
        out[m] = sum;
    }
}

// LocalImplicitSolver.rhsNorm (C/solver/LocalImplicitSolver.java:87-97)
double rhs_norm(const Case& c) {
    double sum = 0;
    double tmp[MAXEQ];
    for (const auto& e : c.E) {
        l2_norms_sq(c, e, e.RHS, tmp);
        for (int m = 0; m < c.Mc.nEqs; m++) sum += tmp[m];
    }
    return std::sqrt(sum);
}

// MasterSolver.residuumNorms (C/solver/MasterSolver.java:113-137)
std::vector<double> residuum_norms(const Case& c) {
    std::vector<double> norms(c.Mc.nEqs, 0.0);
    double tmp[MAXEQ];
    for (const auto& e : c.E) {
        l2_norms_sq(c, e, e.spatialR, tmp);
        for (int m = 0; m < c.Mc.nEqs; m++) norms[m] += tmp[m];
    }
    for (auto& v : norms) v = std::sqrt(v);
    return norms;
}

// State.setResiduum (C/State.java:145-175), N_NORMALISATION_STEPS = 5
static void set_residuum(Case& c, const std::vector<double>& norms) {
    int nEqs = (int)norms.size();
    if ((int)c.resNorm.size() != nEqs) { c.resNorm.assign(nEqs, 0.0); c.nNormSteps = 0; }
    if (c.nNormSteps < 5) {
        for (int m = 0; m < nEqs; m++) c.resNorm[m] = std::max(c.resNorm[m], norms[m]);
        c.nNormSteps++;
    }
    double maxN = 0;
    for (double r : c.resNorm) maxN = std::max(maxN, r);
    c.eqResiduum.assign(nEqs, 0.0);
    c.residuum = 0;
    for (int m = 0; m < nEqs; m++) {
        double nrm = c.resNorm[m] > 0 ? c.resNorm[m] : maxN;
        c.eqResiduum[m] = nrm > 0 ? norms[m] / nrm : norms[m];
        if (std::isnan(c.eqResiduum[m])) c.residuum = NAN;
        else if (!std::isnan(c.residuum)) c.residuum = std::max(c.residuum, c.eqResiduum[m]);
    }
}

// CFLSetup.getCFL / reduceCFL (C/solver/CFLSetup.java:34-63)
static double get_cfl(Case& c, double actual, double residuum) {
    if (!c.P.vary_cfl) return std::min(actual + c.P.cfl / 20, c.P.cfl);
    bool known = residuum > 0 && residuum < 1.7976931348623157e308;
    if (actual <= 0) { c.cflResOld = known ? residuum : -1; return c.P.cfl; }
    double factor = 1;
    if (known && c.cflResOld > 0) factor = std::max(1 / 2.0, std::min(c.cflResOld / residuum, 2.0));
    c.cflResOld = known ? residuum : -1;
    double lower = std::min(actual, c.P.cfl);
    return std::max(lower, std::min(actual * factor, c.P.max_cfl));
}

// State.getTimeStep (C/State.java:177-192) with Element.calculateMaxEigenvalue
double time_step(Case& c) {
    double temp = 1.7976931348623157e308;
    const int nEqs = c.Mc.nEqs;
    for (auto& e : c.E) {
        double Wavg[MAXEQ] = {0, 0, 0, 0, 0};
        for (int j = 0; j < nEqs; j++)
            for (int m = 0; m < e.nBasis; m++) Wavg[j] = Wavg[j] + e.W[j * e.nBasis + m] * e.Is[m];
        double lam;
        if (c.Mc.kind == 2) lam = c.Mc.lambda;
        else {
            // AERO/Aerodynamics.java:323-337 on the element mean
            double rho = Wavg[0] < 1e-2 ? 1e-2 * std::exp((Wavg[0] - 1e-2) / 1e-2) : Wavg[0];
            Wavg[0] = rho;
            double m2 = 0;
            for (int d = 0; d < 2; d++) m2 += Wavg[d + 1] * Wavg[d + 1];
            rho = Wavg[0] < 1e-2 ? 1e-2 * std::exp((Wavg[0] - 1e-2) / 1e-2) : Wavg[0];
            Wavg[0] = rho;
            // AdiabaticEulers.pressure: rho^kappa (AERO/AdiabaticEulers.java:25-33)
            double p = c.Mc.kind == 3 ? std::pow(Wavg[0], c.Mc.kapa) : (c.Mc.kapa - 1) * (Wavg[3] - m2 / (2 * Wavg[0]));
            const double ptol = (c.Mc.kind == 1 || c.Mc.kind == 3) ? 1e-1 : 1e-2;   // SA and AdiabaticEulers P_TOL
            if (p < ptol) p = ptol * std::exp((p - ptol) / ptol);
            double a = std::sqrt(c.Mc.kapa * p / Wavg[0]);
            double v = 0;
            for (int d = 0; d < 2; d++) {
                double vr = Wavg[d + 1] / Wavg[0] - 0.0;
                v += vr * vr;
            }
            lam = std::sqrt(v) + a;
        }
        if (lam > 0) temp = std::min(temp, e.elemSize / lam / 1.0);
    }
    if (c.P.fixed_dt) {
        c.currentCFL = c.state_dt / temp;
        return c.state_dt;
    }
    c.state_dt = c.currentCFL * temp;
    return c.state_dt;
}

// Element.positivityLimiter (C/element/Element.java:385-476) for NavierStokes; AdiabaticEulers
// admits every positive density (AERO/AdiabaticEulers.java:19-22)
static int g_admissible_kind = 0;
static bool admissible(const double* W) {
    if (!(W[0] > 1e-10)) return false;
    if (g_admissible_kind == 3) return true;
    double m2 = 0;
    for (int d = 0; d < 2; d++) m2 += W[d + 1] * W[d + 1];
    return W[3] - m2 / (2 * W[0]) > 1e-10;
}
static bool positivity_limit(const Case& c, Elem& e) {
    const int Nb = e.nBasis, nEqs = c.Mc.nEqs;
    if (Nb == 1 || c.Mc.kind == 2) return false;
    g_admissible_kind = c.Mc.kind;
    double Wavg[MAXEQ] = {0, 0, 0, 0, 0};
    for (int j = 0; j < nEqs; j++)
        for (int m = 0; m < Nb; m++) Wavg[j] = Wavg[j] + e.W[j * Nb + m] * e.Is[m];
    if (!admissible(Wavg)) return false;
    auto theta_at = [&](const double* base) {
        double Wx[MAXEQ] = {0, 0, 0, 0, 0};
        for (int m = 0; m < nEqs; m++)
            for (int j = 0; j < Nb; j++) Wx[m] += e.W[m * Nb + j] * base[j];
        if (admissible(Wx)) return 1.0;
        double lo = 0, hi = 1, Wt[MAXEQ];
        for (int it = 0; it < 30; it++) {
            double mid = 0.5 * (lo + hi);
            for (int m = 0; m < nEqs; m++) Wt[m] = Wavg[m] + mid * (Wx[m] - Wavg[m]);
            if (admissible(Wt)) lo = mid; else hi = mid;
        }
        return lo;
    };
    double theta = 1;
    for (int p = 0; p < e.nq; p++) theta = std::min(theta, theta_at(&e.baseV[p * Nb]));
    for (int k = 0; k < e.nF; k++)
        for (int p = 0; p < e.nqf[k]; p++) theta = std::min(theta, theta_at(&e.baseL[k][p * Nb]));
    if (theta >= 1) return false;
    std::vector<std::vector<double>> Mg(Nb, std::vector<double>(Nb, 0.0));
    std::vector<double> integ(Nb, 0.0);
    for (int p = 0; p < e.nq; p++) {
        double jw = e.JacV[p] * e.wV[p];
        const double* base = &e.baseV[p * Nb];
        for (int i = 0; i < Nb; i++) {
            integ[i] += jw * base[i];
            for (int j = 0; j < Nb; j++) Mg[i][j] += jw * base[i] * base[j];
        }
    }
    auto iMg = mat_invert(Mg);
    std::vector<double> one(Nb, 0.0);
    for (int i = 0; i < Nb; i++)
        for (int j = 0; j < Nb; j++) one[i] = one[i] + iMg[i][j] * integ[j];
    for (int m = 0; m < nEqs; m++)
        for (int i = 0; i < Nb; i++) {
            double mean = one[i] * Wavg[m];
            e.W[m * Nb + i] = mean + theta * (e.W[m * Nb + i] - mean);
        }
    return true;
}

// LocalImplicitSolver.solve main loop (C/solver/LocalImplicitSolver.java:107-285) for a static mesh.
int run(Case& c, int nsteps) {
    std::vector<double> x(c.dofs);
    double dto = -1;   // LocalImplicitSolver.solve: the previous level is unknown at the start
    bool converges = true;
    int total = c.steps_done + nsteps;
    int completed = 0;
    for (++c.steps_done; c.steps_done <= total && c.residuum > c.P.residuum_stop && c.t < c.P.t_end; ++c.steps_done) {
        if (converges) {
            if (c.P.fixed_dt) c.state_dt = std::min(2 * c.state_dt, c.P.dt);
            else c.currentCFL = get_cfl(c, c.currentCFL, c.residuum);
        }
        if (c.P.fixed_dt ? c.state_dt < c.P.dt / 1024 : (c.currentCFL > 0 && c.currentCFL < c.P.cfl / 20)) break;
        double dt = time_step(c);
        // LocalImplicitSolver.java:151-159: stop at tEnd up to 1e-6 dt, clip the last step
        if (c.P.t_end - c.t <= 1e-6 * dt) break;
        if (c.t + dt > c.P.t_end) dt = c.P.t_end - c.t;
        limiter_all(c, completed == 0);
        double rhsNorm0 = 0;
        converges = true;
        for (int s = 0; s < c.P.newton_iters; s++) {
            if (s > 0) {
                assemble_residuum(c, dt, dto);
                double rel = rhsNorm0 > 0 ? rhs_norm(c) / rhsNorm0 : 0;
                c.newton_history.push_back(rel);
                if (rel < c.P.newton_tol) break;
            }
            assemble(c, dt, dto);
            if (s == 0) {
                set_residuum(c, residuum_norms(c));
                rhsNorm0 = rhs_norm(c);
            }
            std::fill(x.begin(), x.end(), 0.0);
            converges = linear_solve(c, x.data());
            if (!converges) {
                for (auto& e : c.E) e.W = e.Wo;   // copyWo2W
                if (c.P.fixed_dt) c.state_dt /= 2;
                else { c.cflResOld = -1; c.currentCFL = c.currentCFL / 1.5; }
                --c.steps_done;
                c.lin_fail++;
                break;
            }
            for (auto& e : c.E) {   // updateW
                for (int i = 0; i < (int)e.W.size(); i++) e.W[i] += x[e.gIndex0 + i];
            }
            if (c.P.positivity_limiter)
                for (auto& e : c.E) positivity_limit(c, e);
        }
        if (!converges) continue;
        if (!std::isfinite(c.residuum)) break;
        c.t += dt;
        dto = dt;
        c.history.push_back(c.residuum);
        for (auto& e : c.E) { e.Wo2 = e.Wo; e.Wo = e.W; }   // copyW2Wo
        completed++;
    }
    --c.steps_done;
    return completed;
}

}  // namespace oracle

// ===================================================================== C API
using namespace oracle;

extern "C" {

// types / orders: per-element element type (3, 4) and order, nullable; with either, TP has 4 entries
// per element
void* or_create(const Props* P, int nPts, const double* PXY, int nElem, const int* TP, int nBnd, const int* bnd,
                const double* wallDist, const char* rulesDir, const int* types, const int* orders) {
    try {
        return build_case(*P, nPts, PXY, nElem, TP, nBnd, bnd, wallDist, rulesDir, types, orders);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "oracle: %s\n", ex.what());
        return nullptr;
    }
}
void or_destroy(void* h) { delete (Case*)h; }
int or_dofs(void* h) { return ((Case*)h)->dofs; }
int or_nbasis(void* h) { return ((Case*)h)->E[0].nBasis; }
int or_neqs(void* h) { return ((Case*)h)->Mc.nEqs; }
void or_model_const(void* h, double* out) {   // kapa, pOut, Re, Pr, pRef, rhoRef, velocityRef
    const ModelConst& M = ((Case*)h)->Mc;
    out[0] = M.kapa; out[1] = M.pOut; out[2] = M.Re; out[3] = M.Pr;
    out[4] = M.pRef; out[5] = M.rhoRef; out[6] = M.velocityRef;
}
int or_nfaces(void* h) { return ((Case*)h)->nF; }
void or_tt(void* h, int* out) { Case* c = (Case*)h; std::memcpy(out, c->TT.data(), sizeof(int) * c->TT.size()); }
void or_order(void* h, int* out) { Case* c = (Case*)h; std::memcpy(out, c->order.data(), sizeof(int) * c->nElem); }
void or_gindex(void* h, int* out) { Case* c = (Case*)h; for (int i = 0; i < c->nElem; i++) out[i] = c->E[i].gIndex0; }
void or_set_w(void* h, const double* W, const double* Wo, const double* Wo2) {
    Case* c = (Case*)h;
    for (int i = 0; i < c->nElem; i++) {
        int B = (int)c->E[i].W.size();
        const size_t o = c->off[i];
        std::memcpy(c->E[i].W.data(), W + o, sizeof(double) * B);
        std::memcpy(c->E[i].Wo.data(), (Wo ? Wo : W) + o, sizeof(double) * B);
        std::memcpy(c->E[i].Wo2.data(), (Wo2 ? Wo2 : (Wo ? Wo : W)) + o, sizeof(double) * B);
    }
}
void or_get_w(void* h, double* W) {
    Case* c = (Case*)h;
    for (int i = 0; i < c->nElem; i++) {
        int B = (int)c->E[i].W.size();
        std::memcpy(W + c->off[i], c->E[i].W.data(), sizeof(double) * B);
    }
}
void or_limiter(void* h) { limiter_all(*(Case*)h); }
// eps[e] and dampInner[e][m] of every element (mesh order)
void or_artificial_viscosity(void* h, double* eps, double* dampInner) {
    Case* c = (Case*)h;
    for (int i = 0; i < c->nElem; i++) {
        eps[i] = c->E[i].eps;
        for (int m = 0; m < c->Mc.nEqs; m++) dampInner[i * c->Mc.nEqs + m] = c->E[i].dampInner[m];
    }
}
// TrunOrd of element i (zeros when the order is <= 2)
void or_trunc(void* h, int i, double* out) {
    Case* c = (Case*)h;
    const Elem& e = c->E[i];
    const int nb = e.nBasis;
    for (int q = 0; q < nb * nb; q++) out[q] = e.TrunOrd.empty() ? 0.0 : e.TrunOrd[q];
}
void or_residual(void* h, double* R) { residual_all(*(Case*)h, R); }
void or_assemble(void* h, double dt, double dto) { assemble(*(Case*)h, dt, dto); }
void or_assemble_residuum(void* h, double dt, double dto) { assemble_residuum(*(Case*)h, dt, dto); }
// blocks in the reference's transposed layout: diag[e][i][j], neigh[e][k][i][j] (zero for boundary faces)
void or_get_blocks(void* h, double* diag, double* neigh) {
    Case* c = (Case*)h;
    for (int i = 0; i < c->nElem; i++) {
        const Elem& e = c->E[i];
        size_t BB = e.ADiag.size();
        std::memcpy(diag + i * BB, e.ADiag.data(), sizeof(double) * BB);
        for (int k = 0; k < c->nF; k++) {
            double* dst = neigh + ((size_t)i * c->nF + k) * BB;
            if (e.TT[k] > -1) std::memcpy(dst, e.ANeigh[k].data(), sizeof(double) * BB);
            else std::memset(dst, 0, sizeof(double) * BB);
        }
    }
}
// blocks with the elements' own sizes, flat: diag[e] (B_e x B_e) one after another, then for every
// element and face with a neighbour t the [B_t][B_e] block, in element and face order
void or_get_blocks_flat(void* h, double* diag, double* neigh) {
    Case* c = (Case*)h;
    size_t od = 0, on = 0;
    for (int i = 0; i < c->nElem; i++) {
        const Elem& e = c->E[i];
        std::memcpy(diag + od, e.ADiag.data(), sizeof(double) * e.ADiag.size());
        od += e.ADiag.size();
        for (int k = 0; k < e.nF; k++)
            if (e.TT[k] > -1) {
                std::memcpy(neigh + on, e.ANeigh[k].data(), sizeof(double) * e.ANeigh[k].size());
                on += e.ANeigh[k].size();
            }
    }
}
int or_nbasis_of(void* h, int i) { return ((Case*)h)->E[i].nBasis; }
void or_mass(void* h, int i, double* M) {
    const Elem& e = ((Case*)h)->E[i];
    std::memcpy(M, e.M.data(), sizeof(double) * e.M.size());
}
void or_get_rhs(void* h, double* rhs, double* spatial) {
    Case* c = (Case*)h;
    for (int i = 0; i < c->nElem; i++) {
        const Elem& e = c->E[i];
        std::memcpy(rhs + c->off[i], e.RHS.data(), sizeof(double) * e.RHS.size());
        if (spatial) std::memcpy(spatial + c->off[i], e.spatialR.data(), sizeof(double) * e.RHS.size());
    }
}
int or_linear_solve(void* h, double* x, int* iters, double* err) {
    Case* c = (Case*)h;
    bool ok = linear_solve(*c, x);
    *iters = c->last_iters;
    *err = c->last_error;
    return ok ? 1 : 0;
}
// y = A x with the assembled CRS matrix (x, y in DOF numbering)
void or_spmv(void* h, const double* x, double* y) { spmv(*(Case*)h, y, x); }
double or_rhs_norm(void* h) { return rhs_norm(*(Case*)h); }
void or_residuum_norms(void* h, double* out) {
    auto v = residuum_norms(*(Case*)h);
    for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
}
int or_run(void* h, int nsteps, double* history, int maxHist) {
    Case* c = (Case*)h;
    size_t before = c->history.size();
    run(*c, nsteps);
    int n = 0;
    for (size_t i = before; i < c->history.size() && n < maxHist; i++) history[n++] = c->history[i];
    return n;
}
void or_state(void* h, double* out) {   // t, CFL, dt, residuum, lin_fail, steps_done
    Case* c = (Case*)h;
    out[0] = c->t; out[1] = c->currentCFL; out[2] = c->state_dt; out[3] = c->residuum;
    out[4] = c->lin_fail; out[5] = c->steps_done;
}
double or_time_step(void* h) { return time_step(*(Case*)h); }
void or_set_cfl(void* h, double cfl) { ((Case*)h)->currentCFL = cfl; }
void or_jvp(void* h, double a1, const double* v, double* out) { jvp(*(Case*)h, a1, v, out); }
void or_vjp(void* h, double a1, const double* v, double* out) { vjp(*(Case*)h, a1, v, out); }
double or_functional(void* h, int which) { return functional(*(Case*)h, which, nullptr); }
void or_functional_grad_fd(void* h, int which, double* out) { functional_grad_fd(*(Case*)h, which, out); }
// geometry for the closed-form pins: area, elemSize, Xs[2], n[k][p][2] (p < nqf), JacF[k][p], M (Nb*Nb)
void or_elem_geometry(void* h, int i, double* area, double* elemSize, double* Xs, double* normals, double* jacF, double* M) {
    Case* c = (Case*)h;
    const Elem& e = c->E[i];
    *area = e.area;
    *elemSize = e.elemSize;
    Xs[0] = e.Xs[0]; Xs[1] = e.Xs[1];
    for (int k = 0; k < e.nF; k++)
        for (int p = 0; p < e.nqf[k]; p++) {
            normals[(k * e.nqf[k] + p) * 2] = e.n[k][p][0];
            normals[(k * e.nqf[k] + p) * 2 + 1] = e.n[k][p][1];
            jacF[k * e.nqf[k] + p] = e.JacF[k][p];
        }
    std::memcpy(M, e.M.data(), sizeof(double) * e.M.size());
}
// TestSimulation.l2Error (T/core/TestSimulation.java:141-155) against sin(2x + y)
double or_l2_error_scalar(void* h) {
    Case* c = (Case*)h;
    double error = 0;
    for (const auto& e : c->E) {
        for (int p = 0; p < e.nq; p++) {
            double Xi[2] = {e.xiV[2 * p], e.xiV[2 * p + 1]};
            double X[2];
            elem_getX(e, Xi, X);
            double u = 0;
            for (int j = 0; j < e.nBasis; j++) u += e.W[j] * e.baseV[p * e.nBasis + j];
            double diff = u - std::sin(2 * X[0] + X[1]);
            error += e.JacV[p] * e.wV[p] * diff * diff;
        }
    }
    return std::sqrt(error);
}
int or_nqf(void* h) { return ((Case*)h)->face.n; }
int or_nqv(void* h) { return ((Case*)h)->vol.n; }
void or_flux_pointwise(void* h, const double* WL, const double* WR, const double* n, int TT, double* fnum, double* fphys) {
    flux_pointwise(*(Case*)h, WL, WR, n, TT, fnum, fphys);
}
void or_model_pointwise(void* h, const double* W, const double* dW, double D, double* s, double* fx, double* fy) {
    model_pointwise(*(Case*)h, W, dW, D, s, fx, fy);
}
// ---- driver pieces, for the pins of tests/test_oracle_driver_pins.py
void or_bdf_coefficients(void* h, double dt, double dto, double* a) { bdf_coefficients(*(Case*)h, dt, dto, a); }
// State.updateCFL with CFLSetup.getCFL on the case's CFL state; returns the new CFL number
double or_cfl_update(void* h, double residuum) {
    Case* c = (Case*)h;
    c->currentCFL = get_cfl(*c, c->currentCFL, residuum);
    return c->currentCFL;
}
// State.reduceCFL with CFLSetup.reduceCFL
double or_cfl_reduce(void* h) {
    Case* c = (Case*)h;
    c->cflResOld = -1;
    c->currentCFL = c->currentCFL / 1.5;
    return c->currentCFL;
}
// State.setResiduum for one step's norms: returns the residuum, eq[m] = equationResiduum
double or_set_residuum(void* h, const double* norms, double* eq) {
    Case* c = (Case*)h;
    set_residuum(*c, std::vector<double>(norms, norms + c->Mc.nEqs));
    for (int m = 0; m < c->Mc.nEqs; m++) eq[m] = c->eqResiduum[m];
    return c->residuum;
}
// Element.l2NormsSquared of the vector R (Java layout), summed over the elements
void or_l2_norms_sq(void* h, const double* R, double* out) {
    Case* c = (Case*)h;
    const int nEqs = c->Mc.nEqs;
    for (int m = 0; m < nEqs; m++) out[m] = 0;
    double tmp[MAXEQ];
    for (int i = 0; i < c->nElem; i++) {
        const Elem& e = c->E[i];
        std::vector<double> Re(R + c->off[i], R + c->off[i] + (size_t)e.nBasis * nEqs);
        l2_norms_sq(*c, e, Re, tmp);
        for (int m = 0; m < nEqs; m++) out[m] += tmp[m];
    }
}
// Element.positivityLimiter on every element (the state of or_set_w); returns the number limited
int or_positivity(void* h) {
    Case* c = (Case*)h;
    int n = 0;
    for (auto& e : c->E) n += positivity_limit(*c, e) ? 1 : 0;
    return n;
}
// blockJacobi: working copies of the diagonal blocks as at the start of a linear solve
// (LinearSolver.solve -> factor), then applies that eliminate in place (b, x in DOF numbering)
void or_bj_start(void* h) {
    Case* c = (Case*)h;
    c->bj_work.assign(c->nElem, {});
    for (int i = 0; i < c->nElem; i++) c->bj_work[i] = c->E[i].ADiag;
}
void or_bj_apply(void* h, const double* b, double* x) { bj_apply(*(Case*)h, x, b); }
// block-ILU0 M^-1 b (trans = 0) or M^-T b (trans = 1) with the factors of the last linear solve
void or_ilu_apply(void* h, int trans, const double* b, double* x) {
    Case* c = (Case*)h;
    if (trans) ilu_apply_transposed(*c, x, b);
    else ilu_apply(*c, x, b);
}
void or_wall_distance(int nPts, const double* PXY, int nElem, const int* TP, const int* TT, int nF, double* out) {
    std::vector<double> d = wall_distance(nPts, PXY, nElem, TP, TT, nF);
    std::memcpy(out, d.data(), sizeof(double) * nPts);
}
}
