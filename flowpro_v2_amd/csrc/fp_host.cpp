// Host side of libflowpro_hip.so: case setup (neighbour matrix, RCM numbering,
// Lagrange basis tables, element geometry, BSR pattern, ILU level sets), device
// memory, GMRES orchestration and the extern "C" entry points of flowpro_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/flowpro_hip.h"
#include "fp_common.hpp"

namespace fp {
// launchers (fp_element.hip, fp_linalg.hip)
bool element_supported(const ElemLaunch& L);
bool team_supported(const ElemLaunch& L);
hipError_t launch_av_limiter(const ElemLaunch&, hipStream_t, const ModelParams&, const Tables&, const Mesh&, const AvParams&,
                             const double*, const double*, const double*, const double*, double*, double*);
hipError_t launch_residual(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&, const Mesh&,
                           const double*, const double*, const double*, double, double, double, double*, double*, double*);
template <typename BT>
hipError_t launch_jacobian(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&, const Mesh&,
                           const double*, double, BT*, const double*, double*, int);
hipError_t launch_residual_team(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&, const Mesh&,
                                const double*, const double*, const double*, double, double, double, double*, double*, double*);
hipError_t launch_jacobian_team(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&, const Mesh&,
                                const double*, double, double*, float*);
hipError_t launch_jtv_team(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&, const Mesh&,
                           const double*, const double*, double, double*);
hipError_t launch_jvp(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const Tables&, const Mesh&,
                      const double*, const double*, double, double*);
hipError_t launch_size_over_lambda(const ElemLaunch&, hipStream_t, const ModelParams&, const Tables&, const Mesh&,
                                   const double*, const double*, const double*, double*);
hipError_t launch_functional(const ElemLaunch&, hipStream_t, const Tables&, const Mesh&, const double*, double, int, double*, double*);
hipError_t launch_positivity(const ElemLaunch&, hipStream_t, const Tables&, const Mesh&, double*, const double*, int*);
// tetrahedra (fp_tet.hip)
bool tet_supported(const ElemLaunch& L);
hipError_t launch_residual_tet(const ElemLaunch&, hipStream_t, const ModelParams&, const TetTables&, const TetMesh&, const double*,
                               const double*, const double*, double, double, double, double*, double*, double*);
hipError_t launch_jacobian_tet(const ElemLaunch&, hipStream_t, const ModelParams&, const TetTables&, const TetMesh&, const double*,
                               double, double*, const double*, double*, int);
hipError_t launch_jvp_tet(const ElemLaunch&, hipStream_t, const ModelParams&, const TetTables&, const TetMesh&, const double*,
                          const double*, double, double*);
hipError_t launch_jacobian_tet_team(const ElemLaunch&, hipStream_t, const ModelParams&, const TetTables&, const TetMesh&, const double*,
                                    double, double*, float*);
hipError_t launch_size_over_lambda_tet(const ElemLaunch&, hipStream_t, const ModelParams&, const TetMesh&, const double*,
                                       const double*, const double*, double*);
hipError_t launch_positivity_tet(const ElemLaunch&, hipStream_t, const TetTables&, const TetMesh&, double*, const double*, int*);
hipError_t launch_spmv(hipStream_t, int, int, const int*, const int*, const double*, const double*, const double*, double*, int,
                       const int* skip = nullptr);
template <typename TF> hipError_t launch_ilu_copy(hipStream_t, size_t, const double*, TF*);
template <typename TF>
hipError_t launch_ilu_factor_level(hipStream_t, int, const int*, int, const int*, const int*, const int*, TF*, int*);
template <typename TF>
hipError_t launch_ilu_factor_rows(hipStream_t, int, int, const int*, int, const int*, const int*, const int*, const int*, TF*, int*);
template <typename TF>
hipError_t launch_ilu_factor_flow(hipStream_t, int, int, const int*, int, const int*, const int*, const int*, const int*, TF*, int*,
                                  unsigned*, unsigned*, int);
bool ilu_factor_rows_supported(int, int);
int ilu_factor_rows_variants();
int ilu_factor_rows_default_variant(int B);
void ilu_factor_rows_shape(int variant, int* nw, int* ngw);
template <typename TF>
hipError_t launch_ilu_fwd_level(hipStream_t, int, const int*, int, const TF*, const IluIO*, TF*);
template <typename TF>
hipError_t launch_ilu_sweep_flow(hipStream_t, int, int, const int*, int, const TF*, IluIO*, const unsigned long long*,
                                 unsigned long long*, int, const int* skip = nullptr);
template <typename TF>
hipError_t launch_ilu_bwd_level(hipStream_t, int, const int*, int, const TF*, const TF*, TF*, const IluIO*);
template <typename TF>
hipError_t launch_ilu_tfwd_level(hipStream_t, int, const int*, int, const TF*, const IluIO*, TF*);
bool ilu_sweep_spmv_supported(int B);
template <typename TF>
hipError_t launch_ilu_sweep_flow_spmv(hipStream_t, int, const int*, int, const TF*, IluIO*, unsigned long long*, int, const int*,
                                      const int*, const int*, const double*);
template <typename TF>
hipError_t launch_ilu_tbwd_level(hipStream_t, int, const int*, int, const TF*, const TF*, TF*, const IluIO*);
hipError_t launch_set_io(hipStream_t, IluIO*, const double*, double*, const int* skip = nullptr);
hipError_t launch_block_jacobi(hipStream_t, int, int, double*, const double*, double*, const int* skip = nullptr);
hipError_t launch_gather_diag(hipStream_t, int, int, const int*, const double*, double*);
hipError_t launch_min(hipStream_t, size_t, const double*, double*, double*);
int min_reduction_partials();
hipError_t launch_colsum(hipStream_t, int, int, const double*, double*, double*);
bool quad_supported(const ElemLaunch&);
bool mixed_supported(const ElemLaunch&);
hipError_t launch_residual_mixed(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const MixedMesh&, const double*,
                                 const double*, const double*, double, double, double, double*, double*, double*);
hipError_t launch_jacobian_mixed(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const MixedMesh&, const double*,
                                 double, double*, float*, const double*, double*, int);
hipError_t launch_jvp_mixed(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const MixedMesh&, const double*,
                            const double*, double, double*);
hipError_t launch_size_over_lambda_mixed(const ElemLaunch&, hipStream_t, const ModelParams&, const MixedMesh&, const double*,
                                         const double*, double*);
hipError_t launch_positivity_mixed(const ElemLaunch&, hipStream_t, const MixedMesh&, double*, int*);
hipError_t launch_functional_mixed(const ElemLaunch&, hipStream_t, const MixedMesh&, const double*, double, int, double*, double*);
hipError_t launch_functional_quad(const ElemLaunch&, hipStream_t, const QuadTables&, const QuadMesh&, const double*, double, int, double*,
                                  double*);
hipError_t launch_residual_quad(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const QuadTables&, const QuadMesh&,
                                const double*, const double*, const double*, double, double, double, double*, double*, double*);
hipError_t launch_jacobian_quad(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const QuadTables&, const QuadMesh&,
                                const double*, double, double*, float*, const double*, double*, int);
hipError_t launch_jvp_quad(const ElemLaunch&, hipStream_t, const ModelParams&, const DGParams&, const QuadTables&, const QuadMesh&,
                           const double*, const double*, double, double*);
hipError_t launch_size_over_lambda_quad(const ElemLaunch&, hipStream_t, const ModelParams&, const QuadMesh&, const double*,
                                        const double*, double*);
hipError_t launch_positivity_quad(const ElemLaunch&, hipStream_t, const QuadTables&, const QuadMesh&, double*, int*);
hipError_t launch_av_limiter_quad(const ElemLaunch&, hipStream_t, const ModelParams&, const QuadTables&, const QuadMesh&,
                                  const AvParams&, const double*, const double*, const double*, double*, double*);
hipError_t launch_av_limiter_mixed(const ElemLaunch&, hipStream_t, const ModelParams&, const MixedMesh&, const AvParams&,
                                   const double*, const double*, const double*, double*, double*);
hipError_t launch_dot(hipStream_t, size_t, const double*, const double*, double*, double*, int, const int* skip = nullptr);
hipError_t launch_axpy_dot(hipStream_t, size_t, double*, const double*, const double*, const double*, double*, double*, int,
                           const int* skip = nullptr);
hipError_t launch_scale_div(hipStream_t, size_t, double*, const double*, const double*, const int* skip = nullptr);
hipError_t launch_update_solution(hipStream_t, size_t, int, const double*, size_t, const double*, double*,
                                  const int* nvp = nullptr);
hipError_t launch_gmres_cycle_begin(hipStream_t, GmresCtl*, double*, double*, int);
hipError_t launch_gmres_givens(hipStream_t, GmresCtl*, int, int, const double*, double*, double*, double*, double*);
hipError_t launch_gmres_lsolve(hipStream_t, const GmresCtl*, int, const double*, const double*, double*, double*, double*, bool);
hipError_t launch_axpy(hipStream_t, size_t, double, const double*, double*);
hipError_t launch_fill(hipStream_t, size_t, double, double*);
hipError_t launch_copy(hipStream_t, size_t, const double*, double*, const int* skip = nullptr);
hipError_t launch_nonfinite(hipStream_t, size_t, const double*, int*);
}  // namespace fp

using namespace fp;

struct fp_ctx {
    std::string err;
    int device = 0;                 // the GPU the context lives on
    // sizes
    int nElem = 0, order = 0, NB = 0, NEQ = 0, B = 0, NQV = 0, NQF = 0;
    int dim = 2, NF = 3;            // triangles (3 faces), quadrilaterals or tetrahedra (4 faces)
    bool quad = false;              // 2D quadrilaterals (squareElement)
    bool mixed = false;             // per-element types or orders: elements padded to NB functions
    std::vector<int> m_nb;          // mixed: basis functions of every mesh element
    std::vector<long long> m_off;   // mixed: offset of every mesh element in the reference's element layout
    std::vector<long long> m_g;     // mixed: the reference's gIndex of the element at DOF position r
    long long ref_dofs = 0;         // mixed: the reference's number of unknowns
    long long dofs = 0;
    ElemLaunch L{};
    ModelParams mp{};
    DGParams dg{};
    AvParams avp{};
    double *av = nullptr, *avold = nullptr, *trunc = nullptr;   // damping state, sensor memory, truncation
    std::vector<double> h_isref;    // Is_j = int phi_j / |K| (the element mean of W is sum_j W_j Is_j)
    std::vector<double> h_qis;      // quadrilaterals: Is of every element, mesh order [e*NB + j]
    int precond = 2, bits = 64, m = 30, cycles = 5, positivity = 1;
    double tol = 1e-2;
    // host copies
    std::vector<int> order_;        // order_[r] = mesh element at position r
    std::vector<int> rank;          // rank[mesh element] = r
    std::vector<int> blk_own_nb;    // block (row r, column of face k neighbour), -1 on boundary
    std::vector<int> h_nbr;         // mixed: face neighbours by position (nbr on the device)
    int nnzb = 0;
    std::vector<int> lvl_f_ptr, lvl_b_ptr;
    // device
    hipStream_t s = nullptr;
    double *geo = nullptr, *fgeo = nullptr, *tab = nullptr, *esize = nullptr, *isref = nullptr, *wdist = nullptr;
    int *nbr = nullptr, *nbrf = nullptr, *blkn = nullptr, *blkd = nullptr;
    int *rowptr = nullptr, *colidx = nullptr, *diagpos = nullptr, *lvl_f = nullptr, *lvl_b = nullptr;
    int *ftptr = nullptr, *ftask = nullptr;   // per-row task lists of the ILU0 factor
    int *meta_f = nullptr, *meta_b = nullptr; // per-row records of the two sweeps, in level order
    int *meta_tf = nullptr, *meta_tb = nullptr;   // the same for the transposed sweeps
    bool adj_prec = false;                    // the factors in F belong to the adjoint system being solved
    bool ilu_rows = false;                    // factor with the register-resident row kernel
    int ilu_variant = -1;                     // launch shape of the row kernel (-1: not tuned yet)
    std::vector<double> tune_ms;              // the tuner's time of every shape (ms per factorisation)
    bool f_fresh = false;                     // F holds the FP32 copy of the current matrix (written by the Jacobian)
    double* val = nullptr;        // FP64 matrix (nnzb * B * B)
    void* F = nullptr;            // preconditioner factors (FP64 or FP32)
    void *iy = nullptr, *ixs = nullptr;
    // dataflow sweeps: published rows {value bits, epoch} of y and x (FP64: two words a value)
    unsigned long long *ypk = nullptr, *xpk = nullptr;
    bool ilu_flow = true;                     // one dataflow launch per sweep (else one launch per level)
    bool spmv_sweep = false;                  // GMRES: A v folded into the forward sweep of M^-1 A v
    bool lsolve_one_thread = false;           // GMRES: the triangular solve in one thread on global memory
    bool factor_flow = false;                 // one dataflow launch for the row-kernel factorisation
    unsigned *fflag = nullptr, *fctl = nullptr;   // per-row completion epochs, {epoch, ticket}
    int factor_groups = 2048;                 // workgroups of the dataflow factorisation
    int flow_groups = 1024;                   // workgroups of the dataflow sweeps
    double* bjw = nullptr;        // block Jacobi working copies
    double *W = nullptr, *Wo = nullptr, *Wo2 = nullptr, *R = nullptr, *RHS = nullptr, *enorm = nullptr, *tmp = nullptr;
    double *x = nullptr, *V = nullptr, *w = nullptr, *r = nullptr, *aux = nullptr, *partial = nullptr, *scal = nullptr;
    double* y = nullptr;
    double* gm = nullptr;         // GMRES on the device: H, its working copy, cs, sn, s, the copy of s
    GmresCtl* gctl = nullptr;
    int gmres_chunk = 4;          // Arnoldi steps enqueued between two reads of the control block
    bool gmres_graph = false;     // replay each chunk from a HIP graph (FLOWPRO_GMRES_GRAPH=1)
    int* istat = nullptr;
    IluIO* io = nullptr;
    bool use_graph = true;
    hipGraphExec_t g_factor = nullptr, g_apply = nullptr;
    std::vector<hipGraphExec_t> g_arnoldi;   // the GMRES steps [k chunk, (k + 1) chunk) of a cycle, captured once
    Tables tb{};
    Mesh ms{};
    TetTables ttb{};
    TetMesh tms{};
    QuadTables qtb{};
    QuadMesh qms{};
    MixedMesh mms{};
    int *m_nvd = nullptr, *m_nqvd = nullptr, *m_nbd = nullptr, *m_nqfd = nullptr;
    double *m_vol = nullptr, *m_face = nullptr;
    double *qv = nullptr, *qf = nullptr, *qn = nullptr, *xq = nullptr, *qmass = nullptr, *qminv = nullptr, *qis = nullptr;
    bool jac_valid = false;
    double last_a1 = 0;
    double kappa = 1.4;
    int op = 0;          // GMRES operator: 0 assembled matrix, 1 matrix-free J, 2 matrix-free J^T
    double op_a1 = 0;
};

#define HIPCHK(ctx, expr)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                \
            return FP_EHIP;                                                                                 \
        }                                                                                                   \
    } while (0)

// element kernels of the case's cell type
static hipError_t run_residual(fp_ctx* c, const double* a) {
    if (c->mixed)
        return launch_residual_mixed(c->L, c->s, c->mp, c->dg, c->mms, c->W, c->Wo, c->Wo2, a[0], a[1], a[2], c->R, c->RHS, c->enorm);
    if (c->quad)
        return launch_residual_quad(c->L, c->s, c->mp, c->dg, c->qtb, c->qms, c->W, c->Wo, c->Wo2, a[0], a[1], a[2], c->R, c->RHS,
                                    c->enorm);
    if (c->dim == 3)
        return launch_residual_tet(c->L, c->s, c->mp, c->ttb, c->tms, c->W, c->Wo, c->Wo2, a[0], a[1], a[2], c->R, c->RHS, c->enorm);
    return launch_residual_team(c->L, c->s, c->mp, c->dg, c->tb, c->ms, c->W, c->Wo, c->Wo2, a[0], a[1], a[2], c->R, c->RHS,
                                c->enorm);
}
static hipError_t run_jacobian(fp_ctx* c, double a1) {
    // with FP32 block-ILU0 factors the Jacobian kernel also writes the factorisation's starting copy
    float* f32 = (c->precond == 2 && c->bits == 32) ? (float*)c->F : nullptr;
    c->f_fresh = f32 != nullptr;
    if (c->mixed) return launch_jacobian_mixed(c->L, c->s, c->mp, c->dg, c->mms, c->W, a1, c->val, f32, nullptr, nullptr, 0);
    if (c->quad) return launch_jacobian_quad(c->L, c->s, c->mp, c->dg, c->qtb, c->qms, c->W, a1, c->val, f32, nullptr, nullptr, 0);
    if (c->dim == 3) return launch_jacobian_tet_team(c->L, c->s, c->mp, c->ttb, c->tms, c->W, a1, c->val, f32);
    return launch_jacobian_team(c->L, c->s, c->mp, c->dg, c->tb, c->ms, c->W, a1, c->val, f32);
}
static hipError_t run_jvp(fp_ctx* c, const double* v, double a1, double* out) {
    if (c->mixed) return launch_jvp_mixed(c->L, c->s, c->mp, c->dg, c->mms, c->W, v, a1, out);
    if (c->quad) return launch_jvp_quad(c->L, c->s, c->mp, c->dg, c->qtb, c->qms, c->W, v, a1, out);
    if (c->dim == 3) return launch_jvp_tet(c->L, c->s, c->mp, c->ttb, c->tms, c->W, v, a1, out);
    return launch_jvp(c->L, c->s, c->mp, c->dg, c->tb, c->ms, c->W, v, a1, out);
}
static hipError_t run_jtv(fp_ctx* c, const double* v, double a1, double* out) {
    if (c->mixed) return launch_jacobian_mixed(c->L, c->s, c->mp, c->dg, c->mms, c->W, a1, nullptr, nullptr, v, out, 1);
    if (c->quad) return launch_jacobian_quad(c->L, c->s, c->mp, c->dg, c->qtb, c->qms, c->W, a1, nullptr, nullptr, v, out, 1);
    if (c->dim == 3) return launch_jacobian_tet(c->L, c->s, c->mp, c->ttb, c->tms, c->W, a1, nullptr, v, out, 1);
    if (team_supported(c->L)) return launch_jtv_team(c->L, c->s, c->mp, c->dg, c->tb, c->ms, c->W, v, a1, out);
    return launch_jacobian<double>(c->L, c->s, c->mp, c->dg, c->tb, c->ms, c->W, a1, nullptr, v, out, 1);
}
// elemSize / lambda_max of every element into out[nElem]
static hipError_t run_size_over_lambda(fp_ctx* c, double* out) {
    if (c->mixed) return launch_size_over_lambda_mixed(c->L, c->s, c->mp, c->mms, c->W, c->esize, out);
    if (c->quad) return launch_size_over_lambda_quad(c->L, c->s, c->mp, c->qms, c->W, c->esize, out);
    if (c->dim == 3) return launch_size_over_lambda_tet(c->L, c->s, c->mp, c->tms, c->W, c->esize, c->isref, out);
    return launch_size_over_lambda(c->L, c->s, c->mp, c->tb, c->ms, c->W, c->esize, c->isref, out);
}
static hipError_t run_positivity(fp_ctx* c, int* count) {
    if (c->mixed) return launch_positivity_mixed(c->L, c->s, c->mms, c->W, count);
    if (c->quad) return launch_positivity_quad(c->L, c->s, c->qtb, c->qms, c->W, count);
    if (c->dim == 3) return launch_positivity_tet(c->L, c->s, c->ttb, c->tms, c->W, c->isref, count);
    return launch_positivity(c->L, c->s, c->tb, c->ms, c->W, c->isref, count);
}

// ============================================================== integer setup
// 2D elements with nF faces, face k between vertices k and (k + 1) % nF (triangleElement,
// squareElement.getFaceIndexes)
static int neighbor_matrix_2d(int n, const int* TP, int nF, int nb, const int* bnd, int* TT) {
    if (n < 0 || !TP || !TT || (nb > 0 && !bnd)) return FP_EINVAL;
    auto key = [&](int i, int k) {
        long long a = TP[nF * i + k], b = TP[nF * i + (k + 1) % nF];
        if (a > b) std::swap(a, b);
        return (a << 32) | b;
    };
    std::unordered_map<long long, std::pair<int, int>> faces;
    faces.reserve((size_t)n * nF);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < nF; k++) {
            auto it = faces.find(key(i, k));
            if (it == faces.end()) faces.emplace(key(i, k), std::make_pair(i, -1000));
            else it->second.second = i;
        }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < nF; k++) {
            const auto& f = faces[key(i, k)];
            int t = -1000;
            if (f.first == i) {
                if (f.second != -1000) t = f.second;
            } else {
                t = f.first;
            }
            TT[nF * i + k] = t;
        }
    std::unordered_map<long long, int> bmap;
    for (int r = 0; r < nb; r++) {
        long long a = bnd[3 * r + 1], b = bnd[3 * r + 2];
        if (a > b) std::swap(a, b);
        bmap[(a << 32) | b] = bnd[3 * r];
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < nF; k++)
            if (TT[nF * i + k] == -1000) {
                auto it = bmap.find(key(i, k));
                TT[nF * i + k] = it != bmap.end() ? it->second : -1;
            }
    return FP_OK;
}
extern "C" int fp_neighbor_matrix(int n, const int* TP, int nb, const int* bnd, int* TT) {
    return neighbor_matrix_2d(n, TP, 3, nb, bnd, TT);
}
extern "C" int fp_neighbor_matrix_quad(int n, const int* TP, int nb, const int* bnd, int* TT) {
    return neighbor_matrix_2d(n, TP, 4, nb, bnd, TT);
}
// mixed triangles and squares: TP rows of 4 (the fourth unused on triangles), types[e] = 3 or 4;
// TT rows of 4, the fourth entry of a triangle is -1000
extern "C" int fp_neighbor_matrix_mixed(int n, const int* TP, const int* types, int nb, const int* bnd, int* TT) {
    if (n < 0 || !TP || !types || !TT || (nb > 0 && !bnd)) return FP_EINVAL;
    auto nv = [&](int i) { return types[i] == 4 ? 4 : 3; };
    auto key = [&](int i, int k) {
        long long a = TP[4 * i + k], b = TP[4 * i + (k + 1) % nv(i)];
        if (a > b) std::swap(a, b);
        return (a << 32) | b;
    };
    std::unordered_map<long long, std::pair<int, int>> faces;
    faces.reserve((size_t)n * 4);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < nv(i); k++) {
            auto it = faces.find(key(i, k));
            if (it == faces.end()) faces.emplace(key(i, k), std::make_pair(i, -1000));
            else it->second.second = i;
        }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) {
            if (k >= nv(i)) { TT[4 * i + k] = -1000; continue; }
            const auto& f = faces[key(i, k)];
            int t = -1000;
            if (f.first == i) {
                if (f.second != -1000) t = f.second;
            } else {
                t = f.first;
            }
            TT[4 * i + k] = t == -1000 ? -2000 : t;   // -2000: open face, resolved below
        }
    std::unordered_map<long long, int> bmap;
    for (int r = 0; r < nb; r++) {
        long long a = bnd[3 * r + 1], b = bnd[3 * r + 2];
        if (a > b) std::swap(a, b);
        bmap[(a << 32) | b] = bnd[3 * r];
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < nv(i); k++)
            if (TT[4 * i + k] == -2000) {
                auto it = bmap.find(key(i, k));
                TT[4 * i + k] = it != bmap.end() ? it->second : -1;
            }
    return FP_OK;
}

// breadth-first search from `node` over unvisited elements; returns the last element of the
// last level (smallest degree among ties), the pseudo-peripheral search of ElementOrdering.
static int peripheral(const std::vector<int>& adjp, const std::vector<int>& adj, const std::vector<char>& excl, int start,
                      std::vector<int>& level) {
    int node = start, ecc = -1;
    std::vector<int> q;
    for (int iter = 0; iter < 10; iter++) {
        q.clear();
        q.push_back(node);
        level[node] = 0;
        std::vector<int> touched = {node};
        int last = node;
        for (size_t h = 0; h < q.size(); h++) {
            int e = q[h];
            for (int p = adjp[e]; p < adjp[e + 1]; p++) {
                int t = adj[p];
                if (!excl[t] && level[t] == -1) {
                    level[t] = level[e] + 1;
                    touched.push_back(t);
                    q.push_back(t);
                    int dt = adjp[t + 1] - adjp[t], dl = adjp[last + 1] - adjp[last];
                    if (level[t] > level[last] || (level[t] == level[last] && dt < dl)) last = t;
                }
            }
        }
        int lv = level[last];
        for (int t : touched) level[t] = -1;
        if (lv <= ecc) break;
        ecc = lv;
        node = last;
    }
    return node;
}

static int rcm_impl(int n, const int* TT, int nF, int* order) {
    if (n < 0 || !TT || !order) return FP_EINVAL;
    std::vector<int> deg(n, 0), adjp(n + 1, 0), adj;
    adj.reserve((size_t)n * nF);
    for (int i = 0; i < n; i++) {
        int cnt = 0;
        int nb[4];
        for (int k = 0; k < nF; k++) {
            int t = TT[(size_t)nF * i + k];
            bool dup = false;
            for (int c = 0; c < cnt; c++) dup |= nb[c] == t;
            if (t > -1 && t != i && !dup) nb[cnt++] = t;
        }
        for (int c = 0; c < cnt; c++) adj.push_back(nb[c]);
        adjp[i + 1] = (int)adj.size();
        deg[i] = cnt;
    }
    for (int i = 0; i < n; i++)   // neighbours by increasing degree, stable
        std::stable_sort(adj.begin() + adjp[i], adj.begin() + adjp[i + 1], [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<char> visited(n, 0);
    std::vector<int> level(n, -1), q;
    int count = 0;
    // the first unvisited element of minimal degree is found by scanning; keep a cursor per degree
    while (count < n) {
        int start = -1;
        for (int i = 0; i < n; i++)
            if (!visited[i] && (start == -1 || deg[i] < deg[start])) {
                start = i;
                if (deg[i] == 0) break;
            }
        start = peripheral(adjp, adj, visited, start, level);
        q.clear();
        q.push_back(start);
        visited[start] = 1;
        for (size_t h = 0; h < q.size(); h++) {
            int e = q[h];
            order[count++] = e;
            for (int p = adjp[e]; p < adjp[e + 1]; p++)
                if (!visited[adj[p]]) { visited[adj[p]] = 1; q.push_back(adj[p]); }
        }
    }
    std::reverse(order, order + n);
    return FP_OK;
}

extern "C" int fp_rcm_order(int n, const int* TT, int* order) { return rcm_impl(n, TT, 3, order); }
extern "C" int fp_rcm_order_tet(int n, const int* TT, int* order) { return rcm_impl(n, TT, 4, order); }
extern "C" int fp_rcm_order_quad(int n, const int* TT, int* order) { return rcm_impl(n, TT, 4, order); }

// local faces of tetrahedralElement (vertex triples of each face)
static const int TET_FACE[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};

// FlowProMain.createNeighborMatrix for tetrahedra: faces are matched by their sorted vertex
// triple; the element scanned first (element-major, then local face) owns the key, the last
// other element sharing it is its neighbour; unmatched faces take the type of the boundary row
// (type, v0, v1, v2) with the same vertex set, else -1.
extern "C" int fp_neighbor_matrix_tet(int n, const int* TP, int nb, const int* bnd, int* TT) {
    if (n < 0 || !TP || !TT || (nb > 0 && !bnd)) return FP_EINVAL;
    struct F {
        int v[3];
        int idx;   // 4 * element + local face: the scan order
    };
    auto sorted = [](int a, int b, int c, int* v) {
        v[0] = a; v[1] = b; v[2] = c;
        std::sort(v, v + 3);
    };
    std::vector<F> faces((size_t)4 * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) {
            F& f = faces[(size_t)4 * i + k];
            sorted(TP[4 * i + TET_FACE[k][0]], TP[4 * i + TET_FACE[k][1]], TP[4 * i + TET_FACE[k][2]], f.v);
            f.idx = 4 * i + k;
        }
    auto less = [](const F& a, const F& b) {
        if (a.v[0] != b.v[0]) return a.v[0] < b.v[0];
        if (a.v[1] != b.v[1]) return a.v[1] < b.v[1];
        if (a.v[2] != b.v[2]) return a.v[2] < b.v[2];
        return a.idx < b.idx;
    };
    std::sort(faces.begin(), faces.end(), less);
    auto same = [](const F& a, const F& b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2]; };
    for (size_t a = 0; a < faces.size();) {
        size_t b = a + 1;
        while (b < faces.size() && same(faces[a], faces[b])) b++;
        const int first = faces[a].idx / 4, second = b - a > 1 ? faces[b - 1].idx / 4 : -1000;
        for (size_t q = a; q < b; q++) {
            const int i = faces[q].idx / 4;
            TT[faces[q].idx] = i == first ? second : first;
        }
        a = b;
    }
    std::vector<F> bf(nb > 0 ? (size_t)nb : 0);
    for (int r = 0; r < nb; r++) {
        sorted(bnd[4 * r + 1], bnd[4 * r + 2], bnd[4 * r + 3], bf[r].v);
        bf[r].idx = r;
    }
    std::sort(bf.begin(), bf.end(), less);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++)
            if (TT[4 * i + k] == -1000) {
                F key;
                sorted(TP[4 * i + TET_FACE[k][0]], TP[4 * i + TET_FACE[k][1]], TP[4 * i + TET_FACE[k][2]], key.v);
                key.idx = std::numeric_limits<int>::max();
                // the last boundary row with this vertex set wins, as in a map filled in row order
                auto it = std::upper_bound(bf.begin(), bf.end(), key, less);
                TT[4 * i + k] = (it != bf.begin() && same(*(it - 1), key)) ? bnd[4 * (it - 1)->idx] : -1;
            }
    return FP_OK;
}

// FlowProMain.generateDistanceFunction (FlowProMain.java:951-1014) with the wall faces as
// sources: Bellman-Ford relaxation along the element edges, in the reference's sweep order
// (elements, then vertex pairs j != k), until a sweep changes nothing.
extern "C" int fp_wall_distance(int n_points, const double* pxy, int n_elems, const int* tp, const int* tt, double* dist) {
    if (n_points < 0 || n_elems < 0 || !pxy || !tp || !tt || !dist) return FP_EINVAL;
    const double inf = std::numeric_limits<double>::infinity();
    std::fill(dist, dist + n_points, inf);
    for (int i = 0; i < n_elems; i++)
        for (int k = 0; k < 3; k++)
            if (tt[3 * i + k] == -1) {   // Aerodynamics.isIPFace: the wall
                dist[tp[3 * i + k]] = 0;
                dist[tp[3 * i + (k + 1) % 3]] = 0;
            }
    bool changed = true;
    for (int iter = 0; changed && iter <= n_points; iter++) {
        changed = false;
        for (int i = 0; i < n_elems; i++)
            for (int j = 0; j < 3; j++) {
                const int pj = tp[3 * i + j];
                for (int k = 0; k < 3; k++) {
                    const int pk = tp[3 * i + k];
                    if (j == k || !(dist[pk] < dist[pj])) continue;
                    const double dx = pxy[2 * pj] - pxy[2 * pk], dy = pxy[2 * pj + 1] - pxy[2 * pk + 1];
                    const double y = dist[pk] + std::sqrt(dx * dx + dy * dy);
                    if (y < dist[pj]) {
                        dist[pj] = y;
                        changed = true;
                    }
                }
            }
    }
    for (int i = 0; i < n_points; i++)
        if (std::isinf(dist[i])) {   // no wall face: the reference fills 1e30
            std::fill(dist, dist + n_points, 1e30);
            break;
        }
    return FP_OK;
}

// ============================================================== basis (basis2Dtriangle)
namespace {
struct Basis {
    int order, nb;
    std::vector<int> xe, ye;
    std::vector<double> coef;   // coef[i*nb + m]
};

bool invert(std::vector<double>& A, int n, std::vector<double>& inv) {
    // Gauss-Jordan with partial pivoting in long double
    std::vector<long double> a(A.begin(), A.end()), b((size_t)n * n, 0.0L);
    for (int i = 0; i < n; i++) b[(size_t)i * n + i] = 1;
    for (int p = 0; p < n; p++) {
        int best = p;
        for (int r = p + 1; r < n; r++)
            if (std::fabs((double)a[(size_t)r * n + p]) > std::fabs((double)a[(size_t)best * n + p])) best = r;
        if (a[(size_t)best * n + p] == 0) return false;
        for (int c = 0; c < n; c++) {
            std::swap(a[(size_t)p * n + c], a[(size_t)best * n + c]);
            std::swap(b[(size_t)p * n + c], b[(size_t)best * n + c]);
        }
        long double d = a[(size_t)p * n + p];
        for (int c = 0; c < n; c++) { a[(size_t)p * n + c] /= d; b[(size_t)p * n + c] /= d; }
        for (int r = 0; r < n; r++)
            if (r != p) {
                long double f = a[(size_t)r * n + p];
                if (f != 0)
                    for (int c = 0; c < n; c++) {
                        a[(size_t)r * n + c] -= f * a[(size_t)p * n + c];
                        b[(size_t)r * n + c] -= f * b[(size_t)p * n + c];
                    }
            }
    }
    inv.assign((size_t)n * n, 0.0);
    for (size_t i = 0; i < inv.size(); i++) inv[i] = (double)b[i];
    return true;
}

// Lagrange basis on the equispaced barycentric nodes, monomial coefficients from the
// inverse Vandermonde matrix (FlowPro/src/main/java/flowpro/core/basis/basis2Dtriangle.java:22-62)
Basis make_basis(int order) {
    Basis b;
    b.order = order;
    b.nb = order * (order + 1) / 2;
    b.xe.assign(order * order, 0);
    b.ye.assign(order * order, 0);
    int s = 0;
    for (int i = 0; i < order; i++)
        for (int j = 0; j <= i; j++) { b.xe[s] = i - j; b.ye[s] = j; s++; }
    for (int i = order - 1; i > 0; i--)
        for (int j = 0; j < i; j++) { b.xe[s] = order - 1 - j; b.ye[s] = order - i + j; s++; }
    if (order == 1) {
        b.coef = {1.0};
        return b;
    }
    std::vector<double> nodes;   // (xi, eta) of the nodes in the reference's order
    for (int i = 0; i < order; i++) {
        double xi_ = (double)i / (order - 1);
        for (int j = 0; j <= i; j++) {
            double y = i == 0 ? 0.0 : (double)j / i * xi_;
            nodes.push_back(y);          // b[1]
            nodes.push_back(xi_ - y);    // b[2]
        }
    }
    std::vector<double> V((size_t)b.nb * b.nb), inv;
    for (int i = 0; i < b.nb; i++)
        for (int j = 0; j < b.nb; j++) V[(size_t)i * b.nb + j] = std::pow(nodes[2 * i], b.xe[j]) * std::pow(nodes[2 * i + 1], b.ye[j]);
    invert(V, b.nb, inv);
    b.coef = inv;
    return b;
}
// Mat.invert (FlowProAPI/src/main/java/flowpro/api/Mat.java:502-642) in double: Gaussian
// elimination with scaled partial pivoting (row scale = largest |entry|, the pivot row
// maximises |a_ij| / scale, rows exchanged through an index vector), then forward and back
// substitution for each column of the identity.  The square basis takes its coefficients from it
// so that they carry the reference's rounding: the Vandermonde matrix of order 4 has
// cond ~ 1e4, and another elimination order moves the coefficients by ~1e-12.
bool invert_scaled_pivot(const std::vector<double>& A0, int n, std::vector<double>& X) {
    std::vector<double> a(A0), c(n);
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) {
        idx[i] = i;
        double c1 = 0;
        for (int j = 0; j < n; j++) c1 = std::max(c1, std::fabs(a[(size_t)i * n + j]));
        c[i] = c1;
    }
    for (int j = 0; j < n - 1; j++) {
        double best = 0;
        int k = j;
        for (int i = j; i < n; i++) {
            const double r = std::fabs(a[(size_t)idx[i] * n + j]) / c[idx[i]];
            if (r > best) { best = r; k = i; }
        }
        std::swap(idx[j], idx[k]);
        for (int i = j + 1; i < n; i++) {
            const double pj = a[(size_t)idx[i] * n + j] / a[(size_t)idx[j] * n + j];
            a[(size_t)idx[i] * n + j] = pj;
            for (int l = j + 1; l < n; l++) a[(size_t)idx[i] * n + l] -= pj * a[(size_t)idx[j] * n + l];
        }
    }
    std::vector<double> b((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) b[(size_t)i * n + i] = 1;
    for (int i = 0; i < n - 1; i++)
        for (int j = i + 1; j < n; j++)
            for (int k = 0; k < n; k++) b[(size_t)idx[j] * n + k] -= a[(size_t)idx[j] * n + i] * b[(size_t)idx[i] * n + k];
    X.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        const double last = a[(size_t)idx[n - 1] * n + n - 1];
        if (last == 0) return false;
        X[(size_t)(n - 1) * n + i] = b[(size_t)idx[n - 1] * n + i] / last;
        for (int j = n - 2; j >= 0; j--) {
            double x = b[(size_t)idx[j] * n + i];
            for (int k = j + 1; k < n; k++) x -= a[(size_t)idx[j] * n + k] * X[(size_t)k * n + i];
            X[(size_t)j * n + i] = x / a[(size_t)idx[j] * n + j];
        }
    }
    return true;
}

// basis2Dsquare (FlowPro/src/main/java/flowpro/core/basis/basis2Dsquare.java:17-73): the same
// monomial exponents, all order^2 of them; nodes (b_i, b_j), b = linspace(0, 1, order), i outer
Basis make_basis_square(int order) {
    Basis b;
    b.order = order;
    b.nb = order * order;
    b.xe.assign(order * order, 0);
    b.ye.assign(order * order, 0);
    int s = 0;
    for (int i = 0; i < order; i++)
        for (int j = 0; j <= i; j++) { b.xe[s] = i - j; b.ye[s] = j; s++; }
    for (int i = order - 1; i > 0; i--)
        for (int j = 0; j < i; j++) { b.xe[s] = order - 1 - j; b.ye[s] = order - i + j; s++; }
    if (order == 1) {
        b.coef = {1.0};
        return b;
    }
    std::vector<double> V((size_t)b.nb * b.nb), inv;
    int r = 0;
    for (int i = 0; i < order; i++)
        for (int j = 0; j < order; j++, r++) {
            const double x = (double)i / (order - 1), y = (double)j / (order - 1);
            for (int k = 0; k < b.nb; k++) V[(size_t)r * b.nb + k] = std::pow(x, b.xe[k]) * std::pow(y, b.ye[k]);
        }
    invert_scaled_pivot(V, b.nb, inv);
    b.coef = inv;
    return b;
}
double basis_fun(const Basis& b, int m, double x, double y) {
    double v = 0;
    for (int i = 0; i < b.nb; i++) v += b.coef[(size_t)i * b.nb + m] * std::pow(x, b.xe[i]) * std::pow(y, b.ye[i]);
    return v;
}
double basis_der(const Basis& b, int m, double x, double y, int d) {
    double v = 0;
    for (int i = 0; i < b.nb; i++) {
        int ex = b.xe[i], ey = b.ye[i];
        if (d == 0 && ex > 0) v += b.coef[(size_t)i * b.nb + m] * ex * std::pow(x, ex - 1) * std::pow(y, ey);
        if (d == 1 && ey > 0) v += b.coef[(size_t)i * b.nb + m] * ey * std::pow(x, ex) * std::pow(y, ey - 1);
    }
    return v;
}

template <typename T> hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}
template <typename T> hipError_t upload(T** p, const std::vector<T>& h) {
    hipError_t e = dalloc(p, h.size());
    if (e != hipSuccess) return e;
    if (h.empty()) return hipSuccess;
    return hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}
}  // namespace

// ============================================================== create / destroy
// TrunOrd = M^-1 T^T Mhat^-1 T with the Taylor basis {1, xi - 1/3, eta - 1/3} on the reference
// triangle, from the reference tables (volume basis values, weights, points) held in c->tab
static bool truncation_matrix(fp_ctx* c, std::vector<double>& tr) {
    const int NB = c->NB, NQV = c->NQV;
    std::vector<double> h((size_t)NQV * NB + NQV + 2 * (size_t)NQV);
    if (hipMemcpy(h.data(), c->tb.vb, sizeof(double) * NQV * NB, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h.data() + NQV * NB, c->tb.vw, sizeof(double) * NQV, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h.data() + NQV * NB + NQV, c->tb.vx, sizeof(double) * 2 * NQV, hipMemcpyDeviceToHost) != hipSuccess) {
        c->err = "reading the reference tables back failed";
        return false;
    }
    const double* vb = h.data();
    const double* vw = vb + NQV * NB;
    const double* vx = vw + NQV;
    auto taylor = [&](int a, int q) { return a == 0 ? 1.0 : vx[2 * q + a - 1] - 1.0 / 3; };
    std::vector<double> Mh(9, 0.0), T((size_t)3 * NB, 0.0), M((size_t)NB * NB, 0.0), Mhi, Mi;
    for (int q = 0; q < NQV; q++) {
        for (int a = 0; a < 3; a++) {
            for (int b = 0; b < 3; b++) Mh[a * 3 + b] += vw[q] * taylor(a, q) * taylor(b, q);
            for (int b = 0; b < NB; b++) T[a * NB + b] += vw[q] * taylor(a, q) * vb[q * NB + b];
        }
        for (int a = 0; a < NB; a++)
            for (int b = 0; b < NB; b++) M[a * NB + b] += vw[q] * vb[q * NB + a] * vb[q * NB + b];
    }
    if (!invert(Mh, 3, Mhi) || !invert(M, NB, Mi)) {
        c->err = "singular matrix in the truncation operator";
        return false;
    }
    std::vector<double> X((size_t)3 * NB, 0.0), Y((size_t)NB * NB, 0.0);   // X = Mhat^-1 T, Y = T^T X
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < NB; b++)
            for (int k = 0; k < 3; k++) X[a * NB + b] += Mhi[a * 3 + k] * T[k * NB + b];
    for (int a = 0; a < NB; a++)
        for (int b = 0; b < NB; b++)
            for (int k = 0; k < 3; k++) Y[a * NB + b] += T[k * NB + a] * X[k * NB + b];
    for (int a = 0; a < NB; a++)
        for (int b = 0; b < NB; b++) {
            double s = 0;
            for (int k = 0; k < NB; k++) s += Mi[a * NB + k] * Y[k * NB + b];
            tr[(size_t)a * NB + b] = s;
        }
    return true;
}

static void free_all(fp_ctx* c) {
    if (c->g_factor) (void)hipGraphExecDestroy(c->g_factor);
    if (c->g_apply) (void)hipGraphExecDestroy(c->g_apply);
    for (hipGraphExec_t g : c->g_arnoldi)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->io) (void)hipFree(c->io);
    void* ptrs[] = {c->m_nvd, c->m_nqvd, c->m_nbd, c->m_nqfd, c->m_vol, c->m_face, c->qv, c->qf, c->qn, c->xq, c->qmass, c->qminv, c->qis, c->av, c->avold, c->trunc, c->wdist, c->geo, c->fgeo, c->tab, c->esize, c->isref, c->nbr, c->nbrf, c->blkn, c->blkd, c->rowptr, c->colidx,
                    c->diagpos, c->lvl_f, c->lvl_b, c->ftptr, c->ftask, c->meta_f, c->meta_b, c->meta_tf, c->meta_tb, c->val, c->F, c->iy, c->ixs, c->ypk, c->xpk, c->fflag, c->fctl, c->bjw, c->W, c->Wo, c->Wo2, c->R, c->RHS,
                    c->enorm, c->tmp, c->x, c->V, c->w, c->r, c->aux, c->partial, c->scal, c->y, c->gm, c->gctl, c->istat};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->s) (void)hipStreamDestroy(c->s);
}

// triangles: neighbour matrix, numbering, basis2Dtriangle tables, transformation2Dtriangle /
// faceTransformation2Dline geometry (Element.computeGeometry)
static int setup_tri(const fp_case* cs, fp_ctx* c, std::vector<int>& nbr) {
    const int n = c->nElem;
    Basis bs = make_basis(cs->order);
    // --- topology and numbering
    std::vector<int> TT((size_t)3 * n);
    int rc = fp_neighbor_matrix(n, cs->tp, cs->n_bnd, cs->bnd, TT.data());
    if (rc) return rc;
    c->order_.resize(n);
    if (cs->ordering == 1) fp_rcm_order(n, TT.data(), c->order_.data());
    else
        for (int i = 0; i < n; i++) c->order_[i] = i;
    c->rank.resize(n);
    for (int r = 0; r < n; r++) c->rank[c->order_[r]] = r;

    // --- reference tables
    const int NB = c->NB, NQV = c->NQV, NQF = c->NQF;
    std::vector<double> vb(NQV * NB), vd(NQV * NB * 2), vw(cs->vol_w, cs->vol_w + NQV), vx(cs->vol_xi, cs->vol_xi + 2 * NQV);
    for (int q = 0; q < NQV; q++)
        for (int j = 0; j < NB; j++) {
            double x = cs->vol_xi[2 * q], y = cs->vol_xi[2 * q + 1];
            vb[q * NB + j] = basis_fun(bs, j, x, y);
            vd[(q * NB + j) * 2] = basis_der(bs, j, x, y, 0);
            vd[(q * NB + j) * 2 + 1] = basis_der(bs, j, x, y, 1);
        }
    static const double CX[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    std::vector<double> fb(3 * NQF * NB), fd(3 * NQF * NB * 2), fw(cs->face_w, cs->face_w + NQF), fx(3 * NQF * 2),
        fint(2 * NQF);
    for (int p = 0; p < NQF; p++) {   // Element.interpolateVelocityAndFillElementDataObjectOnFace (Element.java:299-312)
        fint[2 * p] = (1 - cs->face_xi[p]) / 2;
        fint[2 * p + 1] = (1 + cs->face_xi[p]) / 2;
    }
    for (int k = 0; k < 3; k++)
        for (int p = 0; p < NQF; p++) {
            double t = cs->face_xi[p];
            int a = k, b = (k + 1) % 3;
            double x = (1 + t) / 2 * CX[b][0] + (1 - t) / 2 * CX[a][0];
            double y = (1 + t) / 2 * CX[b][1] + (1 - t) / 2 * CX[a][1];
            fx[(k * NQF + p) * 2] = x;
            fx[(k * NQF + p) * 2 + 1] = y;
            for (int j = 0; j < NB; j++) {
                fb[(k * NQF + p) * NB + j] = basis_fun(bs, j, x, y);
                fd[((k * NQF + p) * NB + j) * 2] = basis_der(bs, j, x, y, 0);
                fd[((k * NQF + p) * NB + j) * 2 + 1] = basis_der(bs, j, x, y, 1);
            }
        }
    std::vector<double> mref(NB * NB, 0.0), mrefinv, isref(NB, 0.0);
    double wsum = 0;
    for (int q = 0; q < NQV; q++) wsum += vw[q];
    for (int i = 0; i < NB; i++) {
        for (int j = 0; j < NB; j++)
            for (int q = 0; q < NQV; q++) mref[i * NB + j] += vw[q] * vb[q * NB + i] * vb[q * NB + j];
        for (int q = 0; q < NQV; q++) isref[i] += vw[q] * vb[q * NB + i];
        isref[i] /= wsum;   // Is_i = int phi_i / area (DGpure.java:561-563)
    }
    invert(mref, NB, mrefinv);
    std::vector<double> tab;
    auto put = [&](const std::vector<double>& v) {
        size_t off = tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        return off;
    };
    size_t o_vb = put(vb), o_vd = put(vd), o_vw = put(vw), o_vx = put(vx), o_fb = put(fb), o_fd = put(fd), o_fw = put(fw),
           o_fx = put(fx), o_m = put(mref), o_mi = put(mrefinv), o_fi = put(fint);
    HIPCHK(c, upload(&c->tab, tab));
    HIPCHK(c, upload(&c->isref, isref));
    c->h_isref = isref;
    c->tb = Tables{c->tab + o_vb, c->tab + o_vd, c->tab + o_vw, c->tab + o_vx, c->tab + o_fb,
                   c->tab + o_fd, c->tab + o_fw, c->tab + o_fx, c->tab + o_m,  c->tab + o_mi, c->tab + o_fi};

    // --- geometry in the DOF order (transformation2Dtriangle, faceTransformation2Dline,
    //     Element.computeGeometry)
    std::vector<double> geo((size_t)n * GEO), fgeo((size_t)n * 12), esize(n), Xs(2 * (size_t)n);
    std::vector<double> wdist(cs->wall_distance ? 3 * (size_t)n : 0);
    nbr.assign(3 * (size_t)n, 0);
    std::vector<int> nbrf(3 * (size_t)n);
    for (int r = 0; r < n; r++) {
        int e = c->order_[r];
        const double* p0 = cs->pxy + 2 * cs->tp[3 * e];
        const double* p1 = cs->pxy + 2 * cs->tp[3 * e + 1];
        const double* p2 = cs->pxy + 2 * cs->tp[3 * e + 2];
        double A0 = p1[0] - p0[0], A1 = p2[0] - p0[0], B0 = p1[1] - p0[1], B1 = p2[1] - p0[1];
        double det = A0 * B1 - A1 * B0;
        double* g = &geo[(size_t)r * GEO];
        g[0] = B1 / det;    // iT00 = d xi / d x
        g[1] = -A1 / det;   // iT01 = d xi / d y
        g[2] = -B0 / det;   // iT10 = d eta / d x
        g[3] = A0 / det;    // iT11 = d eta / d y
        g[4] = std::fabs(det) / 2;
        g[5] = A0; g[6] = A1; g[7] = p0[0];
        g[8] = B0; g[9] = B1; g[10] = p0[1];
        g[11] = 0;
        if (cs->wall_distance)
            for (int v = 0; v < 3; v++) wdist[3 * (size_t)r + v] = cs->wall_distance[cs->tp[3 * e + v]];
        Xs[2 * r] = A0 / 3 + A1 / 3 + p0[0];
        Xs[2 * r + 1] = B0 / 3 + B1 / 3 + p0[1];
    }
    for (int r = 0; r < n; r++) {
        int e = c->order_[r];
        const double* g = &geo[(size_t)r * GEO];
        double perim = 0;
        for (int k = 0; k < 3; k++) {
            int a = k, b = (k + 1) % 3;
            double s0 = CX[b][0] - CX[a][0], s1 = CX[b][1] - CX[a][1];
            double Js0 = g[5] * s0 + g[6] * s1, Js1 = g[8] * s0 + g[9] * s1;
            double len = std::sqrt(Js0 * Js0 + Js1 * Js1);
            double nx = Js1 / len, ny = -Js0 / len;
            int t = TT[3 * (size_t)e + k];
            double dL;
            if (t > -1) {
                int rt = c->rank[t];
                dL = (Xs[2 * r] - Xs[2 * rt]) * nx + (Xs[2 * r + 1] - Xs[2 * rt + 1]) * ny;
                nbr[3 * r + k] = rt;
                int kr = 0;
                for (int s = 0; s < 3; s++)
                    if (TT[3 * (size_t)t + s] == e) { kr = s; break; }
                nbrf[3 * r + k] = kr;
            } else {
                double mx = g[5] * ((CX[a][0] + CX[b][0]) / 2) + g[6] * ((CX[a][1] + CX[b][1]) / 2) + g[7];
                double my = g[8] * ((CX[a][0] + CX[b][0]) / 2) + g[9] * ((CX[a][1] + CX[b][1]) / 2) + g[10];
                dL = (Xs[2 * r] - mx) * nx + (Xs[2 * r + 1] - my) * ny;
                nbr[3 * r + k] = t;
                nbrf[3 * r + k] = 0;
            }
            double* f = &fgeo[((size_t)r * 3 + k) * 4];
            f[0] = nx; f[1] = ny; f[2] = len / 2; f[3] = std::fabs(dL);
            perim += len;
        }
        esize[r] = 3 * g[4] / perim;   // elemSize = nFaces * area / sum S (Element.java:549-554)
    }
    HIPCHK(c, upload(&c->geo, geo));
    HIPCHK(c, upload(&c->fgeo, fgeo));
    HIPCHK(c, upload(&c->esize, esize));
    HIPCHK(c, upload(&c->nbrf, nbrf));
    if (!wdist.empty()) HIPCHK(c, upload(&c->wdist, wdist));
    return FP_OK;
}

// transformation2Dsquare.getX (FlowPro/src/main/java/flowpro/core/transformation/
// transformation2Dsquare.java:50-53) with the reference's evaluation order and no fused
// multiply-add, ab = (A0..A3, B0..B3)
static void quad_getX(const double* ab, const double* xi, double* X) {
#pragma clang fp contract(off)
    X[0] = ab[0] * xi[0] * xi[1] + ab[1] * xi[0] + ab[2] * xi[1] + ab[3];
    X[1] = ab[4] * xi[0] * xi[1] + ab[5] * xi[0] + ab[6] * xi[1] + ab[7];
}

// transformation2Dsquare.getXi (:64-91): Newton from (0.5, 0.5) with the forward-difference
// Jacobian J[j] = (getX(Xi + h e_j) - getX(Xi)) * (1 / h), h = 1e-8, transposed and inverted by
// Mat.invert, at most 50 steps, stopping when |Xi_new - Xi| < 1e-13.  The neighbour-side basis of a
// face point is evaluated at this iterate, as Integration.Face does; its last step leaves the
// point off by up to ~1e-15, which moves a residual at order 4 by ~3e-12 relative, so the product
// takes the same iterate rather than the exact point.
static void quad_getXi(const double* ab, const double* X, double* out) {
#pragma clang fp contract(off)
    double Xi[2] = {0.5, 0.5}, Xin[2] = {0.5, 0.5};
    const double h = 1e-8;
    for (int i = 0; i < 50; i++) {
        double X0[2];
        quad_getX(ab, Xi, X0);
        std::vector<double> J(4), iJ;
        for (int j = 0; j < 2; j++) {
            double Xp[2] = {Xi[0] + (j == 0 ? h : 0.0), Xi[1] + (j == 1 ? h : 0.0)}, Xh[2];
            quad_getX(ab, Xp, Xh);
            for (int d = 0; d < 2; d++) J[(size_t)d * 2 + j] = (Xh[d] - X0[d]) * (1 / h);   // row j, transposed
        }
        invert_scaled_pivot(J, 2, iJ);
        const double r0 = X0[0] - X[0], r1 = X0[1] - X[1];
        for (int d = 0; d < 2; d++) {
            double t = 0;
            t += iJ[(size_t)d * 2] * r0;
            t += iJ[(size_t)d * 2 + 1] * r1;
            Xin[d] = Xi[d] - t;
        }
        double n2 = 0;
        n2 += (Xin[0] - Xi[0]) * (Xin[0] - Xi[0]);
        n2 += (Xin[1] - Xi[1]) * (Xin[1] - Xi[1]);
        if (std::sqrt(n2) < 1e-13) break;
        if (std::isnan(Xin[0]) || std::isnan(Xin[1])) {
            Xin[0] = Xin[1] = 0.5;
            break;
        }
        Xi[0] = Xin[0];
        Xi[1] = Xin[1];
    }
    out[0] = Xin[0];
    out[1] = Xin[1];
}

// quadrilaterals: neighbour matrix, numbering, basis2Dsquare tables, transformation2Dsquare
// (x = A0 xi eta + A1 xi + A2 eta + A3 over the reference square (0,0), (1,0), (1,1), (0,1)) and
// faceTransformation2Dline geometry (Element.computeGeometry), per point where the bilinear map
// makes it vary: |det J| w, the physical basis gradients d_x = iT^T d_xi, the element mass matrix
static int setup_quad(const fp_case* cs, fp_ctx* c, std::vector<int>& nbr) {
    const int n = c->nElem;
    Basis bs = make_basis_square(cs->order);
    std::vector<int> TT((size_t)4 * n);
    int rc = fp_neighbor_matrix_quad(n, cs->tp, cs->n_bnd, cs->bnd, TT.data());
    if (rc) return rc;
    c->order_.resize(n);
    if (cs->ordering == 1) fp_rcm_order_quad(n, TT.data(), c->order_.data());
    else
        for (int i = 0; i < n; i++) c->order_[i] = i;
    c->rank.resize(n);
    for (int r = 0; r < n; r++) c->rank[c->order_[r]] = r;

    const int NB = c->NB, NQV = c->NQV, NQF = c->NQF, NQT = NQV + 4 * NQF;
    static const double CX[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};   // transformation2Dsquare.coordsXi
    // reference points: volume (the tensor rule on [0,1]^2), then face k point p
    std::vector<double> xr((size_t)NQT * 2);
    for (int q = 0; q < NQV; q++) { xr[2 * q] = cs->vol_xi[2 * q]; xr[2 * q + 1] = cs->vol_xi[2 * q + 1]; }
    for (int k = 0; k < 4; k++)
        for (int p = 0; p < NQF; p++) {
            const double t = cs->face_xi[p];
            const int a = k, b = (k + 1) % 4;
            const size_t i = NQV + (size_t)k * NQF + p;
            xr[2 * i] = (1 + t) / 2 * CX[b][0] + (1 - t) / 2 * CX[a][0];
            xr[2 * i + 1] = (1 + t) / 2 * CX[b][1] + (1 - t) / 2 * CX[a][1];
        }
    std::vector<double> phi((size_t)NQT * NB), dxi((size_t)NQT * NB * 2);
    for (int i = 0; i < NQT; i++)
        for (int j = 0; j < NB; j++) {
            phi[(size_t)i * NB + j] = basis_fun(bs, j, xr[2 * i], xr[2 * i + 1]);
            dxi[((size_t)i * NB + j) * 2] = basis_der(bs, j, xr[2 * i], xr[2 * i + 1], 0);
            dxi[((size_t)i * NB + j) * 2 + 1] = basis_der(bs, j, xr[2 * i], xr[2 * i + 1], 1);
        }
    std::vector<double> vb(phi.begin(), phi.begin() + (size_t)NQV * NB), fb(phi.begin() + (size_t)NQV * NB, phi.end());
    std::vector<double> vw(cs->vol_w, cs->vol_w + NQV), fw(cs->face_w, cs->face_w + NQF);
    std::vector<double> tab;
    auto put = [&](const std::vector<double>& v) {
        size_t off = tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        return off;
    };
    size_t o_vb = put(vb), o_vw = put(vw), o_fb = put(fb), o_fw = put(fw);
    HIPCHK(c, upload(&c->tab, tab));
    c->qtb = QuadTables{c->tab + o_vb, c->tab + o_vw, c->tab + o_fb, c->tab + o_fw};

    // per element, DOF order
    std::vector<double> qv((size_t)n * NQV * (1 + 2 * NB)), qf((size_t)n * 4 * NQF * 2 * NB), xq((size_t)n * NQT * 2);
    std::vector<double> fgeo((size_t)n * 16), esize(n), mass((size_t)n * NB * NB), minv, qis((size_t)n * NB), Xs(2 * (size_t)n);
    std::vector<double> AB((size_t)n * 8);
    c->h_qis.assign((size_t)n * NB, 0.0);
    auto getX = [&](const double* ab, double x, double y, double* X) {
        X[0] = ab[0] * x * y + ab[1] * x + ab[2] * y + ab[3];
        X[1] = ab[4] * x * y + ab[5] * x + ab[6] * y + ab[7];
    };
    auto jacobi = [&](const double* ab, double x, double y, double J[2][2]) {
        J[0][0] = ab[0] * y + ab[1];
        J[0][1] = ab[0] * x + ab[2];
        J[1][0] = ab[4] * y + ab[5];
        J[1][1] = ab[4] * x + ab[6];
    };
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        double vx[4], vy[4];
        for (int v = 0; v < 4; v++) {
            vx[v] = cs->pxy[2 * cs->tp[4 * e + v]];
            vy[v] = cs->pxy[2 * cs->tp[4 * e + v] + 1];
        }
        // the solution of (xi eta, xi, eta, 1) . A = x at the four reference vertices
        double* ab = &AB[(size_t)r * 8];
        ab[0] = vx[0] - vx[1] + vx[2] - vx[3]; ab[1] = vx[1] - vx[0]; ab[2] = vx[3] - vx[0]; ab[3] = vx[0];
        ab[4] = vy[0] - vy[1] + vy[2] - vy[3]; ab[5] = vy[1] - vy[0]; ab[6] = vy[3] - vy[0]; ab[7] = vy[0];
        const double centre[2] = {0.5, 0.5};
        quad_getX(ab, centre, &Xs[2 * (size_t)r]);   // transformation2Dsquare.getXs
        for (int i = 0; i < NQT; i++) {
            double J[2][2];
            jacobi(ab, xr[2 * i], xr[2 * i + 1], J);
            const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            const double iT[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
            double* g = i < NQV ? &qv[((size_t)r * NQV + i) * (1 + 2 * NB) + 1] : &qf[((size_t)r * 4 * NQF + (i - NQV)) * 2 * NB];
            if (i < NQV) g[-1] = std::fabs(det) * vw[i];
            for (int j = 0; j < NB; j++) {
                const double* d = &dxi[((size_t)i * NB + j) * 2];
                g[2 * j] = iT[0][0] * d[0] + iT[1][0] * d[1];
                g[2 * j + 1] = iT[0][1] * d[0] + iT[1][1] * d[1];
            }
            quad_getX(ab, &xr[2 * i], &xq[((size_t)r * NQT + i) * 2]);
        }
        double area = 0;
        for (int q = 0; q < NQV; q++) area += qv[((size_t)r * NQV + q) * (1 + 2 * NB)];
        double* M = &mass[(size_t)r * NB * NB];
        for (int a = 0; a < NB; a++) {
            for (int b = 0; b < NB; b++) {
                double sum = 0;
                for (int q = 0; q < NQV; q++) sum += qv[((size_t)r * NQV + q) * (1 + 2 * NB)] * vb[(size_t)q * NB + a] * vb[(size_t)q * NB + b];
                M[a * NB + b] = sum;
            }
            double is = 0;
            for (int q = 0; q < NQV; q++) is += qv[((size_t)r * NQV + q) * (1 + 2 * NB)] * vb[(size_t)q * NB + a];
            qis[(size_t)r * NB + a] = is / area;
            c->h_qis[(size_t)e * NB + a] = is / area;
        }
        esize[r] = area;   // completed below with the perimeter
    }
    minv.assign(mass.size(), 0.0);
    for (int r = 0; r < n; r++) {
        std::vector<double> M(mass.begin() + (size_t)r * NB * NB, mass.begin() + (size_t)(r + 1) * NB * NB), Mi;
        if (!invert(M, NB, Mi)) {
            c->err = "singular element mass matrix";
            return FP_EINVAL;
        }
        std::copy(Mi.begin(), Mi.end(), minv.begin() + (size_t)r * NB * NB);
    }
    nbr.assign(4 * (size_t)n, 0);
    std::vector<int> nbrf(4 * (size_t)n, 0);
    // the neighbour's basis and physical gradients at the own face points (Integration.Face:
    // coordinatesFaceXiRight = elemRight.transform.getXi(coordinatesFace))
    std::vector<double> qn((size_t)n * 4 * NQF * 3 * NB, 0.0);
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        for (int k = 0; k < 4; k++) {
            const int t = TT[4 * (size_t)e + k];
            if (t < 0) continue;
            const double* abt = &AB[(size_t)c->rank[t] * 8];
            for (int p = 0; p < NQF; p++) {
                const double* X = &xq[((size_t)r * NQT + NQV + (size_t)k * NQF + p) * 2];
                double xi[2], J[2][2];
                quad_getXi(abt, X, xi);
                jacobi(abt, xi[0], xi[1], J);
                const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                const double iT[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
                double* o = &qn[(((size_t)r * 4 + k) * NQF + p) * 3 * NB];
                for (int j = 0; j < NB; j++) {
                    const double d0 = basis_der(bs, j, xi[0], xi[1], 0), d1 = basis_der(bs, j, xi[0], xi[1], 1);
                    o[j] = basis_fun(bs, j, xi[0], xi[1]);
                    o[NB + 2 * j] = iT[0][0] * d0 + iT[1][0] * d1;
                    o[NB + 2 * j + 1] = iT[0][1] * d0 + iT[1][1] * d1;
                }
            }
        }
    }
    double fwsum = 0;
    for (int p = 0; p < NQF; p++) fwsum += fw[p];
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        const double* ab = &AB[(size_t)r * 8];
        double perim = 0;
        for (int k = 0; k < 4; k++) {
            const int a = k, b = (k + 1) % 4;
            const double s0 = CX[b][0] - CX[a][0], s1 = CX[b][1] - CX[a][1];
            // J s is the same at every point of a straight edge of the square
            double J[2][2];
            jacobi(ab, CX[a][0], CX[a][1], J);
            const double Js0 = J[0][0] * s0 + J[0][1] * s1, Js1 = J[1][0] * s0 + J[1][1] * s1;
            const double len = std::sqrt(Js0 * Js0 + Js1 * Js1);
            const double nx = Js1 / len, ny = -Js0 / len;
            const int t = TT[4 * (size_t)e + k];
            double dL;
            if (t > -1) {
                const int rt = c->rank[t];
                dL = (Xs[2 * r] - Xs[2 * rt]) * nx + (Xs[2 * r + 1] - Xs[2 * rt + 1]) * ny;
                nbr[4 * r + k] = rt;
                int kr = 0;
                for (int s2 = 0; s2 < 4; s2++)
                    if (TT[4 * (size_t)t + s2] == e) { kr = s2; break; }
                nbrf[4 * r + k] = kr;
            } else {
                double mid[2], xm[2] = {(CX[a][0] + CX[b][0]) / 2, (CX[a][1] + CX[b][1]) / 2};
                quad_getX(ab, xm, mid);
                dL = (Xs[2 * r] - mid[0]) * nx + (Xs[2 * r + 1] - mid[1]) * ny;
                nbr[4 * r + k] = t;
                nbrf[4 * r + k] = 0;
            }
            double* f = &fgeo[((size_t)r * 4 + k) * 4];
            f[0] = nx; f[1] = ny; f[2] = len / 2; f[3] = std::fabs(dL);
            perim += len / 2 * fwsum;   // S_k = sum_p JacobianFace w_p
        }
        esize[r] = 4 * esize[r] / perim;   // elemSize = nFaces * area / sum S (Element.java:549-554)
    }
    if (c->dg.av && cs->order > 2) {
        // DGpure.computeOrderTruncationMatrix (FlowPro/src/main/java/flowpro/core/element/DGpure.java:
        // 595-626) per element: TrunOrd = M^-1 T^T Mhat^-1 T with the Taylor basis {1, x - Xs, y - Xs}
        // at the physical volume points, weights |det J| w, M the element's mass matrix
        std::vector<double> trunc((size_t)n * NB * NB);
        for (int r = 0; r < n; r++) {
            auto taylor = [&](int a, int q) {
                return a == 0 ? 1.0 : xq[((size_t)r * NQT + q) * 2 + a - 1] - Xs[2 * (size_t)r + a - 1];
            };
            auto jw = [&](int q) { return qv[((size_t)r * NQV + q) * (1 + 2 * NB)]; };
            std::vector<double> Mh(9, 0.0), T((size_t)3 * NB, 0.0), Mhi, Mi;
            for (int a = 0; a < 3; a++) {
                for (int b = 0; b < 3; b++)
                    for (int q = 0; q < NQV; q++) Mh[a * 3 + b] += jw(q) * taylor(a, q) * taylor(b, q);
                for (int b = 0; b < NB; b++)
                    for (int q = 0; q < NQV; q++) T[a * NB + b] += jw(q) * taylor(a, q) * vb[(size_t)q * NB + b];
            }
            std::vector<double> M(mass.begin() + (size_t)r * NB * NB, mass.begin() + (size_t)(r + 1) * NB * NB);
            if (!invert_scaled_pivot(Mh, 3, Mhi) || !invert_scaled_pivot(M, NB, Mi)) {
                c->err = "singular matrix in the truncation operator";
                return FP_EINVAL;
            }
            std::vector<double> X((size_t)3 * NB, 0.0), Y((size_t)NB * NB, 0.0);   // X = Mhat^-1 T, Y = T^T X
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < NB; b++)
                    for (int k = 0; k < 3; k++) X[a * NB + b] += Mhi[a * 3 + k] * T[k * NB + b];
            for (int a = 0; a < NB; a++)
                for (int b = 0; b < NB; b++)
                    for (int k = 0; k < 3; k++) Y[a * NB + b] += T[k * NB + a] * X[k * NB + b];
            double* tr = &trunc[(size_t)r * NB * NB];
            for (int a = 0; a < NB; a++)
                for (int b = 0; b < NB; b++) {
                    double v = 0;
                    for (int k = 0; k < NB; k++) v += Mi[a * NB + k] * Y[k * NB + b];
                    tr[a * NB + b] = v;
                }
        }
        HIPCHK(c, upload(&c->trunc, trunc));
    }
    HIPCHK(c, upload(&c->qv, qv));
    HIPCHK(c, upload(&c->qf, qf));
    HIPCHK(c, upload(&c->qn, qn));
    HIPCHK(c, upload(&c->xq, xq));
    HIPCHK(c, upload(&c->fgeo, fgeo));
    HIPCHK(c, upload(&c->esize, esize));
    HIPCHK(c, upload(&c->qmass, mass));
    HIPCHK(c, upload(&c->qminv, minv));
    HIPCHK(c, upload(&c->qis, qis));
    HIPCHK(c, upload(&c->nbrf, nbrf));
    return FP_OK;
}

// FlowPro's gaussRules files (QuadratureCentral, FlowPro/src/main/java/flowpro/core/quadrature/
// QuadratureCentral.java:34-91): "n", then n rows of coordinates, then n weights
namespace {
struct FileRule {
    int n = 0;
    std::vector<double> x, w;   // x: n * dim
};
bool read_rule_file(const std::string& path, int cols, FileRule& r) {
    std::ifstream f(path);
    if (!f) return false;
    f >> r.n;
    r.x.resize((size_t)r.n * cols);
    r.w.resize(r.n);
    for (auto& v : r.x) f >> v;
    for (auto& v : r.w) f >> v;
    return (bool)f;
}
// the triangle rule of degree 2k - 1 (the next degree when it has a negative weight), barycentric
// (b0, b1, b2) -> (b1, b2) on the reference triangle (QuadratureTriangle)
bool triangle_rule_file(const std::string& dir, int k, FileRule& r) {
    FileRule b;
    int deg = 2 * k - 1;
    if (!read_rule_file(dir + "/gaussTri/" + std::to_string(deg) + ".txt", 3, b)) return false;
    for (double w : b.w)
        if (w < 0) {
            if (!read_rule_file(dir + "/gaussTri/" + std::to_string(deg + 1) + ".txt", 3, b)) return false;
            break;
        }
    r.n = b.n;
    r.w = b.w;
    r.x.resize(2 * (size_t)b.n);
    for (int i = 0; i < b.n; i++) {
        r.x[2 * i] = b.x[3 * i + 1];
        r.x[2 * i + 1] = b.x[3 * i + 2];
    }
    return true;
}
// QuadratureSquare: the tensor Gauss rule of order k on [0, 1]^2, i outer
bool square_rule_file(const std::string& dir, int k, FileRule& r) {
    FileRule l;
    if (!read_rule_file(dir + "/gauss/" + std::to_string(k) + ".txt", 1, l)) return false;
    r.n = l.n * l.n;
    r.x.resize(2 * (size_t)r.n);
    r.w.resize(r.n);
    int s = 0;
    for (int i = 0; i < l.n; i++)
        for (int j = 0; j < l.n; j++, s++) {
            r.x[2 * s] = (1 + l.x[i]) / 2;
            r.x[2 * s + 1] = (1 + l.x[j]) / 2;
            r.w[s] = l.w[i] * l.w[j] / 4;
        }
    return true;
}
}  // namespace

// meshes mixing triangles and squares and element orders: every element is padded to c->NB basis
// functions; the tables of each element are computed with its own basis, rule and map, the face
// rule is the line rule of the larger order of the two sides (Integration.Face), and the
// neighbour-side basis at a face point is the neighbour's at the point its inverse map returns
// (affine on triangles, the Newton iteration of transformation2Dsquare.getXi on squares)
static int setup_mixed(const fp_case* cs, fp_ctx* c, std::vector<int>& nbr) {
    const int n = c->nElem, NB = c->NB, NEQ = c->NEQ;
    if ((long long)n * NB * NEQ > std::numeric_limits<int>::max()) {
        c->err = "mixed meshes: more than 2^31 padded unknowns";
        return FP_UNSUPPORTED;
    }
    if (!cs->gauss_rules_dir) {
        c->err = "per-element types or orders need gauss_rules_dir";
        return FP_EINVAL;
    }
    const std::string dir = cs->gauss_rules_dir;
    std::vector<int> nV(n), ord(n), typ(n, cs->elem_type == 4 ? 4 : 3);
    for (int i = 0; i < n; i++) {
        if (cs->elem_types) typ[i] = cs->elem_types[i] == 4 ? 4 : 3;
        nV[i] = typ[i];
        ord[i] = cs->elem_orders ? cs->elem_orders[i] : cs->order;
    }
    // TP with 4 entries per element
    std::vector<int> tp4((size_t)4 * n, -1);
    const int stride = (cs->elem_types || cs->elem_type == 4) ? 4 : 3;
    for (int i = 0; i < n; i++)
        for (int v = 0; v < nV[i]; v++) tp4[4 * (size_t)i + v] = cs->tp[(size_t)stride * i + v];
    std::vector<int> TT((size_t)4 * n);
    int rc = fp_neighbor_matrix_mixed(n, tp4.data(), typ.data(), cs->n_bnd, cs->bnd, TT.data());
    if (rc) return rc;
    c->order_.resize(n);
    if (cs->ordering == 1) fp_rcm_order_quad(n, TT.data(), c->order_.data());
    else
        for (int i = 0; i < n; i++) c->order_[i] = i;
    c->rank.resize(n);
    for (int r = 0; r < n; r++) c->rank[c->order_[r]] = r;
    // bases and rules per (type, order)
    std::map<std::pair<int, int>, Basis> bases;
    std::map<std::pair<int, int>, FileRule> vrules;
    std::map<int, FileRule> lrules;
    auto basis_of = [&](int t, int k) -> const Basis& {
        auto it = bases.find({t, k});
        if (it == bases.end()) it = bases.emplace(std::make_pair(t, k), t == 4 ? make_basis_square(k) : make_basis(k)).first;
        return it->second;
    };
    auto vrule_of = [&](int t, int k) -> const FileRule* {
        auto it = vrules.find({t, k});
        if (it == vrules.end()) {
            FileRule r;
            if (!(t == 4 ? square_rule_file(dir, k, r) : triangle_rule_file(dir, k, r))) return nullptr;
            it = vrules.emplace(std::make_pair(t, k), r).first;
        }
        return &it->second;
    };
    auto lrule_of = [&](int k) -> const FileRule* {
        auto it = lrules.find(k);
        if (it == lrules.end()) {
            FileRule r;
            if (!read_rule_file(dir + "/gauss/" + std::to_string(k) + ".txt", 1, r)) return nullptr;
            it = lrules.emplace(k, r).first;
        }
        return &it->second;
    };
    c->m_nb.resize(n);
    int nqvm = 1, nqfm = 1;
    for (int i = 0; i < n; i++) {
        c->m_nb[i] = basis_of(typ[i], ord[i]).nb;
        const FileRule* vr = vrule_of(typ[i], ord[i]);
        if (!vr) {
            c->err = "missing quadrature rule file under gauss_rules_dir";
            return FP_EINVAL;
        }
        nqvm = std::max(nqvm, vr->n);
        for (int k = 0; k < nV[i]; k++) {
            const int t = TT[4 * (size_t)i + k];
            const int qk = t > -1 ? std::max(ord[i], ord[t]) : ord[i];
            const FileRule* lr = lrule_of(qk);
            if (!lr) {
                c->err = "missing quadrature rule file under gauss_rules_dir";
                return FP_EINVAL;
            }
            nqfm = std::max(nqfm, lr->n);
        }
    }
    // maps: triangle x = x0 + A0 xi + A1 eta (affine), square transformation2Dsquare (bilinear)
    static const double CXT[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    static const double CXS[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    std::vector<double> AB((size_t)n * 8);
    for (int i = 0; i < n; i++) {
        double vx[4], vy[4];
        for (int v = 0; v < nV[i]; v++) {
            vx[v] = cs->pxy[2 * tp4[4 * (size_t)i + v]];
            vy[v] = cs->pxy[2 * tp4[4 * (size_t)i + v] + 1];
        }
        double* ab = &AB[(size_t)i * 8];
        if (typ[i] == 4) {
            ab[0] = vx[0] - vx[1] + vx[2] - vx[3]; ab[1] = vx[1] - vx[0]; ab[2] = vx[3] - vx[0]; ab[3] = vx[0];
            ab[4] = vy[0] - vy[1] + vy[2] - vy[3]; ab[5] = vy[1] - vy[0]; ab[6] = vy[3] - vy[0]; ab[7] = vy[0];
        } else {   // the bilinear form with a zero twist term
            ab[0] = 0; ab[1] = vx[1] - vx[0]; ab[2] = vx[2] - vx[0]; ab[3] = vx[0];
            ab[4] = 0; ab[5] = vy[1] - vy[0]; ab[6] = vy[2] - vy[0]; ab[7] = vy[0];
        }
    }
    auto getX = [&](int i, const double* xi, double* X) {
        const double* ab = &AB[(size_t)i * 8];
        if (typ[i] == 4) { quad_getX(ab, xi, X); return; }
        X[0] = ab[1] * xi[0] + ab[2] * xi[1] + ab[3];
        X[1] = ab[5] * xi[0] + ab[6] * xi[1] + ab[7];
    };
    auto jac = [&](int i, const double* xi, double J[2][2]) {
        const double* ab = &AB[(size_t)i * 8];
        J[0][0] = ab[0] * xi[1] + ab[1];
        J[0][1] = ab[0] * xi[0] + ab[2];
        J[1][0] = ab[4] * xi[1] + ab[5];
        J[1][1] = ab[4] * xi[0] + ab[6];
    };
    auto getXi = [&](int i, const double* X, double* xi) {
        const double* ab = &AB[(size_t)i * 8];
        if (typ[i] == 4) { quad_getXi(ab, X, xi); return; }
        const double det = ab[1] * ab[6] - ab[2] * ab[5];
        const double dx = X[0] - ab[3], dy = X[1] - ab[7];
        xi[0] = (ab[6] * dx - ab[2] * dy) / det;
        xi[1] = (-ab[5] * dx + ab[1] * dy) / det;
    };
    // the reference basis (value, d/dxi, d/deta) at a reference point; the rule points every
    // element of a (type, order) shares are evaluated once (cache = true)
    std::map<std::tuple<int, int, unsigned long long, unsigned long long>, std::vector<double>> refcache;
    auto ref_values = [&](int t, int k, const double* xi, bool cache, std::vector<double>& tmp) -> const std::vector<double>& {
        const Basis& b = basis_of(t, k);
        unsigned long long bx, by;
        std::memcpy(&bx, &xi[0], 8);
        std::memcpy(&by, &xi[1], 8);
        const auto key = std::make_tuple(t, k, bx, by);
        if (cache) {
            auto it = refcache.find(key);
            if (it != refcache.end()) return it->second;
        }
        tmp.resize(3 * (size_t)b.nb);
        for (int j = 0; j < b.nb; j++) {
            tmp[3 * j] = basis_fun(b, j, xi[0], xi[1]);
            tmp[3 * j + 1] = basis_der(b, j, xi[0], xi[1], 0);
            tmp[3 * j + 2] = basis_der(b, j, xi[0], xi[1], 1);
        }
        if (!cache) return tmp;
        return refcache.emplace(key, tmp).first->second;
    };
    std::vector<double> reftmp;
    // basis values and physical gradients of element i at reference point xi, padded to NB
    auto tabulate = [&](int i, const double* xi, double* phi, double* gr, bool cache) {
        const Basis& b = basis_of(typ[i], ord[i]);
        const std::vector<double>& rv = ref_values(typ[i], ord[i], xi, cache, reftmp);
        double J[2][2];
        jac(i, xi, J);
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double iT[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
        for (int j = 0; j < NB; j++) {
            phi[j] = 0;
            gr[2 * j] = gr[2 * j + 1] = 0;
            if (j >= b.nb) continue;
            const double d0 = rv[3 * j + 1], d1 = rv[3 * j + 2];
            phi[j] = rv[3 * j];
            gr[2 * j] = iT[0][0] * d0 + iT[1][0] * d1;
            gr[2 * j + 1] = iT[0][1] * d0 + iT[1][1] * d1;
        }
        return std::fabs(det) / (typ[i] == 4 ? 1.0 : 2.0);   // transformation jacobian (|det| / 2 on the triangle)
    };
    const int VS = 1 + 3 * NB, FS = 3 + 6 * NB;
    std::vector<double> vol((size_t)n * nqvm * VS, 0.0), face((size_t)n * 4 * nqfm * FS, 0.0), fgeo((size_t)n * 16, 0.0);
    std::vector<double> mass((size_t)n * NB * NB, 0.0), minv((size_t)n * NB * NB, 0.0), qis((size_t)n * NB, 0.0), esize(n), Xs(2 * (size_t)n);
    std::vector<int> nvd(n), nqvd(n), nbd(n), nqfd((size_t)4 * n, 0);
    std::vector<double> trunc(c->dg.av ? (size_t)n * NB * NB : 0, 0.0);
    nbr.assign(4 * (size_t)n, -1000);
    c->h_qis.assign((size_t)n * NB, 0.0);
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        const FileRule& vr = *vrule_of(typ[e], ord[e]);
        nvd[r] = nV[e];
        nqvd[r] = vr.n;
        nbd[r] = c->m_nb[e];
        const double centre[2] = {typ[e] == 4 ? 0.5 : 1.0 / 3, typ[e] == 4 ? 0.5 : 1.0 / 3};
        getX(e, centre, &Xs[2 * (size_t)r]);
        double area = 0;
        for (int q = 0; q < vr.n; q++) {
            double* v = &vol[((size_t)r * nqvm + q) * VS];
            const double jacv = tabulate(e, &vr.x[2 * q], v + 1, v + 1 + NB, true);
            v[0] = jacv * vr.w[q];
            area += v[0];
        }
        const int nb = c->m_nb[e];
        double* M = &mass[(size_t)r * NB * NB];
        for (int a = 0; a < NB; a++) {
            for (int b = 0; b < NB; b++) {
                double sum = 0;
                if (a < nb && b < nb)
                    for (int q = 0; q < vr.n; q++) {
                        const double* v = &vol[((size_t)r * nqvm + q) * VS];
                        sum += v[0] * v[1 + a] * v[1 + b];
                    }
                M[a * NB + b] = a < nb || b < nb ? sum : (a == b ? 1.0 : 0.0);   // identity on the padded unknowns
            }
            double is = 0;
            if (a < nb)
                for (int q = 0; q < vr.n; q++) is += vol[((size_t)r * nqvm + q) * VS] * vol[((size_t)r * nqvm + q) * VS + 1 + a];
            qis[(size_t)r * NB + a] = is / area;
            c->h_qis[(size_t)e * NB + a] = is / area;
        }
        std::vector<double> Mr((size_t)nb * nb), Mi;
        for (int a = 0; a < nb; a++)
            for (int b = 0; b < nb; b++) Mr[(size_t)a * nb + b] = M[a * NB + b];
        if (!invert(Mr, nb, Mi)) {
            c->err = "singular element mass matrix";
            return FP_EINVAL;
        }
        for (int a = 0; a < NB; a++)
            for (int b = 0; b < NB; b++)
                minv[(size_t)r * NB * NB + a * NB + b] = (a < nb && b < nb) ? Mi[(size_t)a * nb + b] : (a == b ? 1.0 : 0.0);
        esize[r] = area;
        if (c->dg.av && ord[e] > 2) {
            // DGpure.computeOrderTruncationMatrix (FlowPro/src/main/java/flowpro/core/element/DGpure.java:
            // 595-626): TrunOrd = M^-1 T^T Mhat^-1 T with the Taylor basis {1, x - Xs, y - Xs} at the
            // element's physical volume points, padded with zeros to NB x NB
            std::vector<double> Xq(2 * (size_t)vr.n);
            for (int q = 0; q < vr.n; q++) getX(e, &vr.x[2 * q], &Xq[2 * q]);
            auto taylor = [&](int a, int q) { return a == 0 ? 1.0 : Xq[2 * q + a - 1] - Xs[2 * (size_t)r + a - 1]; };
            auto vq = [&](int q) { return &vol[((size_t)r * nqvm + q) * VS]; };
            std::vector<double> Mh(9, 0.0), T((size_t)3 * nb, 0.0), Mhi, Mri;
            for (int a = 0; a < 3; a++) {
                for (int b = 0; b < 3; b++)
                    for (int q = 0; q < vr.n; q++) Mh[a * 3 + b] += vq(q)[0] * taylor(a, q) * taylor(b, q);
                for (int b = 0; b < nb; b++)
                    for (int q = 0; q < vr.n; q++) T[(size_t)a * nb + b] += vq(q)[0] * taylor(a, q) * vq(q)[1 + b];
            }
            if (!invert_scaled_pivot(Mh, 3, Mhi) || !invert_scaled_pivot(Mr, nb, Mri)) {
                c->err = "singular matrix in the truncation operator";
                return FP_EINVAL;
            }
            std::vector<double> X((size_t)3 * nb, 0.0), Y((size_t)nb * nb, 0.0);   // X = Mhat^-1 T, Y = T^T X
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < nb; b++)
                    for (int kk = 0; kk < 3; kk++) X[(size_t)a * nb + b] += Mhi[a * 3 + kk] * T[(size_t)kk * nb + b];
            for (int a = 0; a < nb; a++)
                for (int b = 0; b < nb; b++)
                    for (int kk = 0; kk < 3; kk++) Y[(size_t)a * nb + b] += T[(size_t)kk * nb + a] * X[(size_t)kk * nb + b];
            double* tr = &trunc[(size_t)r * NB * NB];
            for (int a = 0; a < nb; a++)
                for (int b = 0; b < nb; b++) {
                    double v = 0;
                    for (int kk = 0; kk < nb; kk++) v += Mri[(size_t)a * nb + kk] * Y[(size_t)kk * nb + b];
                    tr[a * NB + b] = v;
                }
        }
    }
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        double perim = 0;
        for (int k = 0; k < nV[e]; k++) {
            const int t = TT[4 * (size_t)e + k];
            const FileRule& lr = *lrule_of(t > -1 ? std::max(ord[e], ord[t]) : ord[e]);
            nqfd[4 * r + k] = lr.n;
            const double* CA = typ[e] == 4 ? CXS[k] : CXT[k];
            const double* CB = typ[e] == 4 ? CXS[(k + 1) % 4] : CXT[(k + 1) % 3];
            const double s0 = CB[0] - CA[0], s1 = CB[1] - CA[1];
            double nx = 0, ny = 0, S = 0;
            for (int p = 0; p < lr.n; p++) {
                const double tt = lr.x[p];
                double xi[2] = {(1 + tt) / 2 * CB[0] + (1 - tt) / 2 * CA[0], (1 + tt) / 2 * CB[1] + (1 - tt) / 2 * CA[1]};
                double J[2][2];
                jac(e, xi, J);
                const double Js0 = J[0][0] * s0 + J[0][1] * s1, Js1 = J[1][0] * s0 + J[1][1] * s1;
                const double len = std::sqrt(Js0 * Js0 + Js1 * Js1);
                nx = Js1 / len;
                ny = -Js0 / len;
                double* f = &face[(((size_t)r * 4 + k) * nqfm + p) * FS];
                f[0] = len / 2 * lr.w[p];
                S += len / 2 * lr.w[p];
                getX(e, xi, f + 1);
                tabulate(e, xi, f + 3, f + 3 + NB, true);
                if (t > -1) {
                    double xin[2];
                    getXi(t, f + 1, xin);
                    tabulate(t, xin, f + 3 + 3 * NB, f + 3 + 4 * NB, false);
                }
            }
            double dL;
            if (t > -1) {
                const int rt = c->rank[t];
                dL = (Xs[2 * r] - Xs[2 * rt]) * nx + (Xs[2 * r + 1] - Xs[2 * rt + 1]) * ny;
                nbr[4 * r + k] = rt;
            } else {
                double mid[2], xm[2] = {(CA[0] + CB[0]) / 2, (CA[1] + CB[1]) / 2};
                getX(e, xm, mid);
                dL = (Xs[2 * r] - mid[0]) * nx + (Xs[2 * r + 1] - mid[1]) * ny;
                nbr[4 * r + k] = t;
            }
            double* g = &fgeo[((size_t)r * 4 + k) * 4];
            g[0] = nx; g[1] = ny; g[2] = 0; g[3] = std::fabs(dL);
            perim += S;
        }
        esize[r] = nV[e] * esize[r] / perim;   // elemSize = nFaces * area / sum S
    }
    // the reference's element layout and DOF numbering
    c->m_off.assign(n + 1, 0);
    for (int i = 0; i < n; i++) c->m_off[i + 1] = c->m_off[i] + (long long)c->m_nb[i] * NEQ;
    c->m_g.assign(n, 0);
    long long g = 0;
    for (int r = 0; r < n; r++) {
        c->m_g[r] = g;
        g += (long long)c->m_nb[c->order_[r]] * NEQ;
    }
    c->ref_dofs = g;
    HIPCHK(c, upload(&c->m_vol, vol));
    if (c->dg.av) HIPCHK(c, upload(&c->trunc, trunc));
    HIPCHK(c, upload(&c->m_face, face));
    HIPCHK(c, upload(&c->fgeo, fgeo));
    HIPCHK(c, upload(&c->esize, esize));
    HIPCHK(c, upload(&c->qmass, mass));
    HIPCHK(c, upload(&c->qminv, minv));
    HIPCHK(c, upload(&c->qis, qis));
    HIPCHK(c, upload(&c->m_nvd, nvd));
    HIPCHK(c, upload(&c->m_nqvd, nqvd));
    HIPCHK(c, upload(&c->m_nbd, nbd));
    HIPCHK(c, upload(&c->m_nqfd, nqfd));
    c->mms.nElem = n;
    c->mms.nqvm = nqvm;
    c->mms.nqfm = nqfm;
    return FP_OK;
}

// tetrahedra: neighbour matrix, numbering, basis3Dtetra tables (orders 1 and 2: the constant,
// then the Lagrange basis on the vertices), transformation3Dtetra / faceTransformation3Dtriangle
// geometry (Element.computeGeometry).  The face normal is normalize(Js2 x Js1) as in
// faceTransformation3Dtriangle: outward on elements whose vertex matrix has a negative
// determinant, the orientation the reference's meshes carry.
static int setup_tet(const fp_case* cs, fp_ctx* c, std::vector<int>& nbr) {
    const int n = c->nElem;
    std::vector<int> TT((size_t)4 * n);
    int rc = fp_neighbor_matrix_tet(n, cs->tp, cs->n_bnd, cs->bnd, TT.data());
    if (rc) return rc;
    c->order_.resize(n);
    if (cs->ordering == 1) rcm_impl(n, TT.data(), 4, c->order_.data());
    else
        for (int i = 0; i < n; i++) c->order_[i] = i;
    c->rank.resize(n);
    for (int r = 0; r < n; r++) c->rank[c->order_[r]] = r;

    // --- reference tables
    const int NB = c->NB, NQV = c->NQV, NQF = c->NQF;
    static const double XI[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    static const double DPHI[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    auto phi = [&](const double* x, double* out) {
        if (NB == 1) out[0] = 1.0;
        else {
            out[0] = 1.0 - x[0] - x[1] - x[2];
            out[1] = x[0];
            out[2] = x[1];
            out[3] = x[2];
        }
    };
    std::vector<double> vb((size_t)NQV * NB), vd((size_t)NQV * NB * 3, 0.0), vw(cs->vol_w, cs->vol_w + NQV);
    for (int q = 0; q < NQV; q++) {
        phi(cs->vol_xi + 3 * q, &vb[(size_t)q * NB]);
        if (NB == 4)
            for (int j = 0; j < NB; j++)
                for (int d = 0; d < 3; d++) vd[((size_t)q * NB + j) * 3 + d] = DPHI[j][d];
    }
    std::vector<double> fb((size_t)4 * NQF * NB), fw(cs->face_w, cs->face_w + NQF), fxr((size_t)4 * NQF * 3);
    for (int k = 0; k < 4; k++) {
        const double* R0 = XI[TET_FACE[k][0]];
        for (int p = 0; p < NQF; p++) {
            const double u = cs->face_xi[2 * p], v = cs->face_xi[2 * p + 1];
            double* xr = &fxr[((size_t)k * NQF + p) * 3];
            for (int d = 0; d < 3; d++)
                xr[d] = (XI[TET_FACE[k][1]][d] - R0[d]) * u + (XI[TET_FACE[k][2]][d] - R0[d]) * v + R0[d];
            phi(xr, &fb[((size_t)k * NQF + p) * NB]);
        }
    }
    std::vector<double> mref((size_t)NB * NB, 0.0), mrefinv, isref(NB, 0.0);
    double wsum = 0;
    for (int q = 0; q < NQV; q++) wsum += vw[q];
    for (int i = 0; i < NB; i++) {
        for (int j = 0; j < NB; j++)
            for (int q = 0; q < NQV; q++) mref[i * NB + j] += vw[q] * vb[q * NB + i] * vb[q * NB + j];
        for (int q = 0; q < NQV; q++) isref[i] += vw[q] * vb[q * NB + i];
        isref[i] /= wsum;
    }
    if (!invert(mref, NB, mrefinv)) {
        c->err = "singular reference mass matrix (volume rule)";
        return FP_EINVAL;
    }
    std::vector<double> tab;
    auto put = [&](const std::vector<double>& v) {
        size_t off = tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        return off;
    };
    size_t o_vb = put(vb), o_vd = put(vd), o_vw = put(vw), o_fb = put(fb), o_fw = put(fw), o_fx = put(fxr), o_m = put(mref),
           o_mi = put(mrefinv);
    HIPCHK(c, upload(&c->tab, tab));
    HIPCHK(c, upload(&c->isref, isref));
    c->h_isref = isref;
    c->ttb = TetTables{c->tab + o_vb, c->tab + o_vd, c->tab + o_vw, c->tab + o_fb, c->tab + o_fw, c->tab + o_fx, c->tab + o_m,
                       c->tab + o_mi};

    // --- geometry in the DOF order
    std::vector<double> geo((size_t)n * TGEO), fgeo((size_t)n * 4 * TFGEO), esize(n), Xs(3 * (size_t)n);
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        const double* P[4];
        for (int v = 0; v < 4; v++) P[v] = cs->pxy + 3 * (size_t)cs->tp[4 * e + v];
        double J[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) J[a][b] = P[b + 1][a] - P[0][a];
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                     c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        if (!(std::fabs(det) > 0)) {
            c->err = "degenerate tetrahedron " + std::to_string(e);
            return FP_EINVAL;
        }
        double* g = &geo[(size_t)r * TGEO];
        // iT = J^-1 by cofactors: iT[b][a] = cof(J)[a][b] / det
        g[0] = c00 / det;
        g[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        g[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        g[3] = c01 / det;
        g[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        g[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        g[6] = c02 / det;
        g[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        g[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        g[9] = std::fabs(det);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) g[10 + 3 * a + b] = J[a][b];
        for (int a = 0; a < 3; a++) g[19 + a] = P[0][a];
        // Xs = X(0.25, 0.25, 0): transformation3Dtetra.getXs evaluates the literal {0.25, 0.25, 0, 25}
        for (int a = 0; a < 3; a++) Xs[3 * (size_t)r + a] = J[a][0] * 0.25 + J[a][1] * 0.25 + P[0][a];
    }
    double fwsum = 0, vwsum = 0;
    for (int p = 0; p < NQF; p++) fwsum += fw[p];
    for (int q = 0; q < NQV; q++) vwsum += vw[q];
    nbr.assign(4 * (size_t)n, 0);
    for (int r = 0; r < n; r++) {
        const int e = c->order_[r];
        const double* g = &geo[(size_t)r * TGEO];
        double sumS = 0;
        for (int k = 0; k < 4; k++) {
            const double* R0 = XI[TET_FACE[k][0]];
            double s1[3], s2[3], Js1[3], Js2[3];
            for (int d = 0; d < 3; d++) {
                s1[d] = XI[TET_FACE[k][1]][d] - R0[d];
                s2[d] = XI[TET_FACE[k][2]][d] - R0[d];
            }
            for (int a = 0; a < 3; a++) {
                Js1[a] = g[10 + 3 * a] * s1[0] + g[11 + 3 * a] * s1[1] + g[12 + 3 * a] * s1[2];
                Js2[a] = g[10 + 3 * a] * s2[0] + g[11 + 3 * a] * s2[1] + g[12 + 3 * a] * s2[2];
            }
            const double cr[3] = {Js2[1] * Js1[2] - Js2[2] * Js1[1], Js2[2] * Js1[0] - Js2[0] * Js1[2],
                                  Js2[0] * Js1[1] - Js2[1] * Js1[0]};
            const double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            const double nn[3] = {cr[0] / len, cr[1] / len, cr[2] / len};
            const int t = TT[4 * (size_t)e + k];
            const double* xs = &Xs[3 * (size_t)r];
            double dL = 0;
            if (t > -1) {
                const int rt = c->rank[t];
                for (int d = 0; d < 3; d++) dL += (xs[d] - Xs[3 * (size_t)rt + d]) * nn[d];
                nbr[4 * r + k] = rt;
            } else {
                double X[3];
                for (int a = 0; a < 3; a++) {
                    double m = 0;
                    for (int b = 0; b < 3; b++) m += g[10 + 3 * a + b] * (R0[b] + s1[b] / 3 + s2[b] / 3);
                    X[a] = m + g[19 + a];
                }
                for (int d = 0; d < 3; d++) dL += (xs[d] - X[d]) * nn[d];
                nbr[4 * r + k] = t;
            }
            double* f = &fgeo[((size_t)r * 4 + k) * TFGEO];
            f[0] = nn[0]; f[1] = nn[1]; f[2] = nn[2];
            f[3] = len / 2;   // |Js1 x Js2| / 2
            f[4] = std::fabs(dL);
            sumS += len / 2 * fwsum;
        }
        esize[r] = 4 * (g[9] * vwsum) / sumS;   // elemSize = nFaces * volume / sum S (Element.java:549-554)
    }
    HIPCHK(c, upload(&c->geo, geo));
    HIPCHK(c, upload(&c->fgeo, fgeo));
    HIPCHK(c, upload(&c->esize, esize));
    return FP_OK;
}

// ============================================================== allocation-free time step
// fp_time_step_noalloc works in `scratch`, a file-scope device array that every GPU the library
// runs on holds until the process ends, instead of the context's buffers: per-element values
// [0, nElem), the block minima [nElem, nElem + P), the result at nElem + P (P = the reduction's
// block count).  Host code never passes the array's name to a kernel: on the host it is only a
// stand-in for the symbol; the device address of each GPU's copy is taken once per device with
// hipGetSymbolAddress (at the first fp_create on that device) and kept in g_scratch.
constexpr size_t FP_SCRATCH_BYTES = size_t(32) << 20;   // 32 MiB: up to ~4.19 M elements
constexpr int FP_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[FP_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch[FP_MAX_DEVICES] = {};

static int scratch_setup(int device, std::string* err) {
    if (device < 0 || device >= FP_MAX_DEVICES) {
        *err = "device index beyond the scratch table";
        return FP_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return FP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));   // the current device's copy
    if (e != hipSuccess || !p) {
        *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return FP_EHIP;
    }
    g_scratch[device] = p;
    return FP_OK;
}

static int create_impl(const fp_case* cs, fp_ctx* c) {
    const int n = cs->n_elems;
    c->nElem = n;
    c->order = cs->order;
    c->dim = cs->dimension == 3 ? 3 : 2;
    c->mixed = c->dim == 2 && (cs->elem_types || cs->elem_orders);
    c->quad = c->dim == 2 && cs->elem_type == 4 && !c->mixed;
    c->NF = c->quad || c->mixed ? 4 : c->dim + 1;
    if (cs->elem_type != 0 && cs->elem_type != 3 && cs->elem_type != 4) {
        c->err = "elem_type: 3 (or 0) triangles, 4 quadrilaterals";
        return FP_EINVAL;
    }
    if (c->mixed) {
        if (cs->model == MODEL_SA || cs->model == MODEL_SCALAR || cs->model == MODEL_KOMEGA) {
            c->err = "mixed element types or orders: device kernels for the NavierStokes and AdiabaticEulers models";
            return FP_UNSUPPORTED;
        }
        int maxo = 0;
        for (int i = 0; i < n; i++) {
            const int o = cs->elem_orders ? cs->elem_orders[i] : cs->order;
            const int t = cs->elem_types ? cs->elem_types[i] : (cs->elem_type == 4 ? 4 : 3);
            if (o < 1 || o > 5 || (t != 0 && t != 3 && t != 4)) {
                c->err = "mixed meshes: element orders 1-5, element types 3 or 4";
                return FP_UNSUPPORTED;
            }
            maxo = std::max(maxo, o);
        }
        if (maxo != cs->order) {   // Parameters.order: the largest element order (beta0, c_IP)
            c->err = "order must be the largest element order";
            return FP_EINVAL;
        }
    }
    if (c->quad && (cs->model == MODEL_SA || cs->model == MODEL_SCALAR || cs->model == MODEL_KOMEGA)) {
        c->err = "quadrilaterals: device kernels for the NavierStokes and AdiabaticEulers models";
        return FP_UNSUPPORTED;
    }
    if (c->dim == 3) {
        if (cs->model != MODEL_NAVIER_STOKES || cs->viscous) {
            c->err = "tetrahedra: only the inviscid NavierStokes model has device kernels";
            return FP_UNSUPPORTED;
        }
        if (cs->has_attack) {   // the device inlet of the 3D model blows along -n
            c->err = "tetrahedra: no attack angle";
            return FP_UNSUPPORTED;
        }
        if (cs->order > 2) {   // basis3Dtetra's node list has the 4 vertices only
            c->err = "tetrahedra: basis3Dtetra builds orders 1 and 2 only";
            return FP_UNSUPPORTED;
        }
        c->NB = cs->order == 1 ? 1 : 4;
        c->NEQ = 5;
    } else if (c->quad) {
        c->NB = cs->order * cs->order;
        c->NEQ = cs->model == MODEL_ADIABATIC ? 3 : 4;
    } else if (c->mixed) {
        // padded basis size: the largest basis of the mesh rounded up to 4, 9 or 16
        int nbmax = 0;
        for (int i = 0; i < n; i++) {
            const int o = cs->elem_orders ? cs->elem_orders[i] : cs->order;
            const bool sq = cs->elem_types ? cs->elem_types[i] == 4 : cs->elem_type == 4;
            nbmax = std::max(nbmax, sq ? o * o : o * (o + 1) / 2);
        }
        c->NB = nbmax <= 4 ? 4 : nbmax <= 9 ? 9 : nbmax <= 16 ? 16 : 25;
        c->NEQ = cs->model == MODEL_ADIABATIC ? 3 : 4;
    } else {
        c->NB = make_basis(cs->order).nb;
        c->NEQ = cs->model == MODEL_SCALAR ? 1
                 : cs->model == MODEL_SA ? 5
                 : cs->model == MODEL_ADIABATIC ? 3
                 : cs->model == MODEL_KOMEGA ? 6 : 4;
    }
    if ((cs->model == MODEL_SA || cs->model == MODEL_KOMEGA) && !cs->wall_distance) {
        c->err = "the Spalart-Allmaras and k-omega SST models need the vertex wall distance";
        return FP_EINVAL;
    }
    if (cs->model == MODEL_KOMEGA && (!(cs->cv > 0) || !(cs->p_ref > 0) || !(cs->rho_ref > 0))) {
        c->err = "KOmegaSST needs cv, p_ref and rho_ref (the Sutherland reference temperature)";
        return FP_EINVAL;
    }
    c->B = c->NB * c->NEQ;
    c->NQV = cs->n_vol;
    c->NQF = cs->n_face;
    c->dofs = (long long)n * c->B;
    // SpalartAllmaras.init and KOmegaSST.init pass isDiffusive = true to Aerodynamics.init
    // (SpalartAllmaras.java:38, KOmegaSST.java:47); AdiabaticEulers.init passes false (AdiabaticEulers.java:13-16)
    const bool diffusive = cs->model == MODEL_SCALAR ? cs->scalar_k > 0
                                                     : (cs->model == MODEL_SA || cs->model == MODEL_KOMEGA ||
                                                        (cs->model != MODEL_ADIABATIC && cs->viscous != 0));
    if (cs->model < 0 || cs->model > MODEL_KOMEGA) {
        c->err = "unknown model";
        return FP_EINVAL;
    }
    const bool av = cs->damp_tol > 0 || cs->damp_const != 0 || cs->damp_inner_tol > 0;
    if (av && c->dim == 3) {
        c->err = "artificial viscosity: device kernels in 2D";
        return FP_UNSUPPORTED;
    }
    if (cs->n_damp_inner_coef < 0 || cs->n_damp_inner_coef > 5 ||
        (cs->n_damp_inner_coef > 1 && cs->n_damp_inner_coef < c->NEQ)) {
        c->err = "dampInnerCoef needs one value or one per equation";
        return FP_EINVAL;
    }
    c->L = ElemLaunch{cs->model, c->NB, c->NEQ, c->NQV, c->NQF, diffusive ? 1 : 0, av ? 1 : 0};
    if (!(c->dim == 3    ? tet_supported(c->L)
          : c->mixed     ? mixed_supported(c->L)
          : c->quad      ? quad_supported(c->L)
                         : element_supported(c->L) && team_supported(c->L))) {
        c->err = "no element kernel for this model / order / quadrature";
        return FP_UNSUPPORTED;
    }
    if (cs->preconditioner < 0 || cs->preconditioner > 3 || (cs->precond_bits != 64 && cs->precond_bits != 32)) {
        c->err = "bad preconditioner settings";
        return FP_EINVAL;
    }
    // ilu0 (ILU0.java: point ILU(0) on the CRS matrix) is the same incomplete factorisation as
    // BlockILU0 on a pattern of dense element blocks, so it runs on the block factorisation
    c->precond = cs->preconditioner == 3 ? 2 : cs->preconditioner;
    c->bits = cs->precond_bits;
    c->m = cs->gmres_restart > 0 ? cs->gmres_restart : 30;
    c->cycles = cs->gmres_cycles > 0 ? cs->gmres_cycles : 5;
    c->tol = cs->gmres_tol;
    c->positivity = cs->positivity_limiter;
    ModelParams& P = c->mp;
    P.kind = cs->model;
    P.nEqs = c->NEQ;
    P.diffusive = diffusive;
    P.has_attack = cs->has_attack;
    P.kapa = cs->kappa;
    c->kappa = cs->kappa;
    P.pOut = cs->p_out;
    P.Re = cs->reynolds;
    P.Pr = cs->prandtl;
    P.attack = cs->attack;
    P.a0 = cs->scalar_a0;
    P.a1 = cs->scalar_a1;
    P.k = cs->scalar_k;
    P.lambda = std::sqrt(cs->scalar_a0 * cs->scalar_a0 + cs->scalar_a1 * cs->scalar_a1);
    P.vtIn = cs->sa_vt_in;
    P.sigma = cs->sa_sigma;
    P.cb1 = cs->sa_cb1;
    P.cb2 = cs->sa_cb2;
    P.ka = cs->sa_ka;
    P.cw1 = P.ka != 0 && P.sigma != 0 ? P.cb1 / (P.ka * P.ka) + (1 + P.cb2) / P.sigma : 0.0;   // SpalartAllmaras.java:53
    P.cw2 = cs->sa_cw2;
    P.cw3 = cs->sa_cw3;
    P.cv1 = cs->sa_cv1;
    P.Prt = cs->model == MODEL_KOMEGA ? cs->kw_prt : cs->sa_prt;
    if (cs->model == MODEL_KOMEGA) {   // KOmegaSST.init (KOmegaSST.java:45-71) and sutherland (:566-570)
        P.kIn = cs->kw_k_in;
        P.omIn = cs->kw_om_in;
        const double TRef = 1 / (cs->cv * (cs->kappa - 1)) * cs->p_ref / cs->rho_ref;
        P.suthRatio = cs->kw_suth / TRef;
        P.kwA1 = cs->kw_a1;
        P.sk1 = cs->kw_sk1;
        P.som1 = cs->kw_som1;
        P.alpha1 = cs->kw_alpha1;
        P.beta1 = cs->kw_beta1;
        P.betast = cs->kw_betast;
        P.sk2 = cs->kw_sk2;
        P.som2 = cs->kw_som2;
        P.alpha2 = cs->kw_alpha2;
        P.beta2 = cs->kw_beta2;
    }
    P.flux = cs->model == MODEL_NAVIER_STOKES && cs->numerical_flux == 1 ? 1 : 0;   // Aerodynamics.java:300-303
    if (cs->numerical_flux < 0 || cs->numerical_flux > 1) {
        c->err = "numerical_flux: 0 (default) or 1 (vanLeer)";
        return FP_EINVAL;
    }
    // DGpure.java:287-289 zeroes beta0 inside the isDiffusive branch; the damping flux of an
    // inviscid model keeps it
    c->dg.beta0 = (cs->order == 1 && diffusive) ? 0.0 : cs->beta0;
    c->dg.dampConst = cs->damp_const;
    c->dg.av = av ? 1 : 0;
    c->avp.dampTol = cs->damp_tol;
    c->avp.dampInnerTol = cs->damp_inner_tol;
    c->avp.nCoef = cs->n_damp_inner_coef;
    for (int i = 0; i < cs->n_damp_inner_coef; i++) c->avp.coef[i] = cs->damp_inner_coef[i];
    c->avp.order = cs->order;
    c->dg.cIP = (cs->order > 1 && diffusive) ? cs->penalty : 0.0;   // DGpure.limiter :396-398

    HIPCHK(c, hipGetDevice(&c->device));
    {
        const int rc = scratch_setup(c->device, &c->err);
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    std::vector<int> nbr;
    int rc = c->dim == 3 ? setup_tet(cs, c, nbr) : c->mixed ? setup_mixed(cs, c, nbr) : c->quad ? setup_quad(cs, c, nbr) : setup_tri(cs, c, nbr);
    if (rc) return rc;
    const int NF = c->NF;
    if (c->mixed) c->h_nbr = nbr;

    // --- BSR pattern: row r = own block + face neighbours, sorted by column
    std::vector<int> rowptr(n + 1, 0), colidx, diagpos(n);
    for (int r = 0; r < n; r++) {
        std::vector<int> cols = {r};
        for (int k = 0; k < NF; k++)
            if (nbr[NF * r + k] > -1 && std::find(cols.begin(), cols.end(), nbr[NF * r + k]) == cols.end())
                cols.push_back(nbr[NF * r + k]);
        std::sort(cols.begin(), cols.end());
        for (size_t i = 0; i < cols.size(); i++) {
            if (cols[i] == r) diagpos[r] = (int)colidx.size();
            colidx.push_back(cols[i]);
        }
        rowptr[r + 1] = (int)colidx.size();
    }
    c->nnzb = (int)colidx.size();
    auto find_blk = [&](int row, int col) {
        for (int s = rowptr[row]; s < rowptr[row + 1]; s++)
            if (colidx[s] == col) return s;
        return -1;
    };
    std::vector<int> blkn(NF * (size_t)n, -1), blkd(n);
    c->blk_own_nb.assign(NF * (size_t)n, -1);
    for (int r = 0; r < n; r++) {
        blkd[r] = diagpos[r];
        for (int k = 0; k < NF; k++)
            if (nbr[NF * r + k] > -1) {
                blkn[NF * r + k] = find_blk(nbr[NF * r + k], r);
                c->blk_own_nb[NF * r + k] = find_blk(r, nbr[NF * r + k]);
            }
    }
    // --- ILU level sets (forward: lower blocks, backward: upper blocks)
    std::vector<int> lf(n, 0), lb(n, 0);
    int nlf = 0, nlb = 0;
    for (int r = 0; r < n; r++) {
        int lv = 0;
        for (int s = rowptr[r]; s < diagpos[r]; s++) lv = std::max(lv, lf[colidx[s]] + 1);
        lf[r] = lv;
        nlf = std::max(nlf, lv + 1);
    }
    for (int r = n - 1; r >= 0; r--) {
        int lv = 0;
        for (int s = diagpos[r] + 1; s < rowptr[r + 1]; s++) lv = std::max(lv, lb[colidx[s]] + 1);
        lb[r] = lv;
        nlb = std::max(nlb, lv + 1);
    }
    auto bucket = [&](const std::vector<int>& lv, int nl, std::vector<int>& ptr, std::vector<int>& rows) {
        ptr.assign(nl + 1, 0);
        for (int r = 0; r < n; r++) ptr[lv[r] + 1]++;
        for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
        rows.assign(n, 0);
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int r = 0; r < n; r++) rows[fill[lv[r]]++] = r;
    };
    std::vector<int> rowsf, rowsb;
    bucket(lf, nlf, c->lvl_f_ptr, rowsf);
    bucket(lb, nlb, c->lvl_b_ptr, rowsb);
    // sweep records (layout documented at k_ilu_fwd_level): {i, n, diag, 0, blocks[6], rows[6]}
    std::vector<int> meta_f((size_t)16 * n, 0), meta_b((size_t)16 * n, 0);
    for (int pos = 0; pos < n; pos++) {
        for (int side = 0; side < 2; side++) {
            const int r = side == 0 ? rowsf[pos] : rowsb[pos];
            int* m = (side == 0 ? meta_f.data() : meta_b.data()) + (size_t)16 * pos;
            const int s0 = side == 0 ? rowptr[r] : diagpos[r] + 1, s1 = side == 0 ? diagpos[r] : rowptr[r + 1];
            if (s1 - s0 > 6) {
                c->err = "more than 6 neighbours on one side of a block row";
                return FP_UNSUPPORTED;
            }
            m[0] = r;
            m[1] = s1 - s0;
            m[2] = diagpos[r];
            for (int q = s0; q < s1; q++) {
                m[4 + q - s0] = q;
                m[10 + q - s0] = colidx[q];
            }
        }
    }
    // records of the transposed sweeps (adjoint preconditioning), same layout: in forward level
    // order the blocks (k, i) of the lower neighbours' rows, in backward level order the blocks
    // (j, i) of the upper neighbours' rows
    std::vector<int> meta_tf((size_t)16 * n, 0), meta_tb((size_t)16 * n, 0);
    for (int pos = 0; pos < n; pos++) {
        for (int side = 0; side < 2; side++) {
            const int r = side == 0 ? rowsf[pos] : rowsb[pos];
            int* m = (side == 0 ? meta_tf.data() : meta_tb.data()) + (size_t)16 * pos;
            const int s0 = side == 0 ? rowptr[r] : diagpos[r] + 1, s1 = side == 0 ? diagpos[r] : rowptr[r + 1];
            m[0] = r;
            m[1] = s1 - s0;
            m[2] = diagpos[r];
            for (int q = s0; q < s1; q++) {
                const int tb = find_blk(colidx[q], r);
                if (tb < 0) {
                    c->err = "block pattern is not symmetric";
                    return FP_EINVAL;
                }
                m[4 + q - s0] = tb;
                m[10 + q - s0] = colidx[q];
            }
        }
    }
    // --- ILU0 factor tasks per row, in the order of BlockILU0.factor (BlockILU0.java:101-127):
    // for each lower k: {slot of A_ik | k << 9, block of D_k^-1}, then for each later block j of
    // row i that row k holds: {slot of A_ij | 256, block U_kj}
    std::vector<int> ftptr(n + 1, 0), ftask;
    int max_row = 0;
    for (int r = 0; r < n; r++) {
        const int rs = rowptr[r], re = rowptr[r + 1];
        max_row = std::max(max_row, re - rs);
        for (int kk = rs; kk < diagpos[r]; kk++) {
            const int k = colidx[kk];
            ftask.push_back((kk - rs) | (k < (1 << 22) ? k << 9 : 0));
            ftask.push_back(diagpos[k]);
            for (int jj = kk + 1; jj < re; jj++) {
                const int kj = find_blk(k, colidx[jj]);
                if (kj < 0) continue;
                ftask.push_back((jj - rs) | 256);
                ftask.push_back(kj);
            }
        }
        ftptr[r + 1] = (int)ftask.size() / 2;
    }
    {
        const char* g = std::getenv("FLOWPRO_ILU_ROWS");   // "0" selects the per-level block kernels
        c->ilu_rows = ilu_factor_rows_supported(c->B, max_row) && !(g && g[0] == '0');
        const char* f = std::getenv("FLOWPRO_ILU_FLOW");   // "0": one launch per level (factor and sweeps)
        c->factor_flow = c->ilu_rows && n < (1 << 22) && !(f && f[0] == '0');
    }

    HIPCHK(c, upload(&c->nbr, nbr));
    HIPCHK(c, upload(&c->blkn, blkn));
    HIPCHK(c, upload(&c->blkd, blkd));
    HIPCHK(c, upload(&c->rowptr, rowptr));
    HIPCHK(c, upload(&c->colidx, colidx));
    HIPCHK(c, upload(&c->diagpos, diagpos));
    HIPCHK(c, upload(&c->lvl_f, rowsf));
    HIPCHK(c, upload(&c->lvl_b, rowsb));
    HIPCHK(c, upload(&c->ftptr, ftptr));
    HIPCHK(c, upload(&c->meta_f, meta_f));
    HIPCHK(c, upload(&c->meta_b, meta_b));
    HIPCHK(c, upload(&c->meta_tf, meta_tf));
    HIPCHK(c, upload(&c->meta_tb, meta_tb));
    HIPCHK(c, upload(&c->ftask, ftask));
    if (c->dg.av) {
        HIPCHK(c, dalloc(&c->av, (size_t)n * (1 + c->NEQ)));
        HIPCHK(c, dalloc(&c->avold, (size_t)n * c->NEQ));
        HIPCHK(c, hipMemsetAsync(c->av, 0, sizeof(double) * (size_t)n * (1 + c->NEQ), c->s));
        HIPCHK(c, hipMemsetAsync(c->avold, 0, sizeof(double) * (size_t)n * c->NEQ, c->s));
        // DGpure.computeOrderTruncationMatrix on the reference triangle: for an affine map the
        // projection onto linear functions in the nodal basis does not depend on the element
        // (quadrilaterals hold one per element, built by setup_quad)
        if (!c->quad && !c->mixed) {
            std::vector<double> tr((size_t)c->NB * c->NB, 0.0);
            if (c->order > 2 && !truncation_matrix(c, tr)) return FP_EINVAL;
            HIPCHK(c, upload(&c->trunc, tr));
        }
    }
    c->ms = Mesh{n, c->geo, c->fgeo, c->nbr, c->nbrf, c->blkn, c->blkd, c->wdist, c->av};
    c->tms = TetMesh{n, c->geo, c->fgeo, c->nbr, c->blkn, c->blkd};
    c->qms = QuadMesh{n, c->qv, c->qf, c->qn, c->xq, c->fgeo, c->qmass, c->qminv, c->qis, c->nbr, c->nbrf, c->blkn, c->blkd, c->av};
    if (c->mixed)
        c->mms = MixedMesh{n, c->mms.nqvm, c->mms.nqfm, c->m_nvd, c->m_nqvd, c->m_nbd, c->m_nqfd, c->m_vol, c->m_face,
                           c->fgeo, c->qmass, c->qminv, c->qis, c->nbr, c->blkn, c->blkd, c->av};

    // --- state, matrix, Krylov workspace
    const size_t N = (size_t)c->dofs, BB = (size_t)c->B * c->B;
    HIPCHK(c, dalloc(&c->W, N));
    HIPCHK(c, dalloc(&c->Wo, N));
    HIPCHK(c, dalloc(&c->Wo2, N));
    HIPCHK(c, dalloc(&c->R, N));
    HIPCHK(c, dalloc(&c->RHS, N));
    HIPCHK(c, dalloc(&c->tmp, std::max(N, (size_t)n)));
    HIPCHK(c, dalloc(&c->enorm, (size_t)n * 2 * c->NEQ));
    HIPCHK(c, dalloc(&c->val, (size_t)c->nnzb * BB));
    if (c->precond == 2) {
        size_t es = c->bits == 32 ? sizeof(float) : sizeof(double);
        HIPCHK(c, hipMalloc(&c->F, (size_t)c->nnzb * BB * es));
        HIPCHK(c, hipMalloc(&c->iy, N * es));
        HIPCHK(c, hipMalloc(&c->ixs, N * es));
        const size_t words = N * (es / 4);
        HIPCHK(c, hipMalloc(&c->ypk, words * sizeof(unsigned long long)));
        HIPCHK(c, hipMalloc(&c->xpk, words * sizeof(unsigned long long)));
        HIPCHK(c, hipMemsetAsync(c->ypk, 0, words * sizeof(unsigned long long), c->s));
        HIPCHK(c, hipMemsetAsync(c->xpk, 0, words * sizeof(unsigned long long), c->s));
        HIPCHK(c, hipMalloc(&c->fflag, (size_t)n * sizeof(unsigned)));
        HIPCHK(c, hipMalloc(&c->fctl, 4 * sizeof(unsigned)));
        HIPCHK(c, hipMemsetAsync(c->fflag, 0, (size_t)n * sizeof(unsigned), c->s));
        HIPCHK(c, hipMemsetAsync(c->fctl, 0, 4 * sizeof(unsigned), c->s));
    } else if (c->precond == 1) {
        HIPCHK(c, dalloc(&c->bjw, (size_t)n * BB));
    }
    HIPCHK(c, dalloc(&c->x, N));
    HIPCHK(c, dalloc(&c->V, N * (c->m + 1)));
    HIPCHK(c, dalloc(&c->w, N));
    HIPCHK(c, dalloc(&c->r, N));
    HIPCHK(c, dalloc(&c->aux, N));
    HIPCHK(c, dalloc(&c->partial, 4096));
    HIPCHK(c, dalloc(&c->scal, 64 + 2 * (size_t)c->m + 8));
    HIPCHK(c, dalloc(&c->y, (size_t)c->m + 1));
    HIPCHK(c, dalloc(&c->gm, 2 * (size_t)(c->m + 1) * c->m + 4 * (size_t)c->m + 4));
    HIPCHK(c, hipMalloc((void**)&c->gctl, sizeof(GmresCtl)));
    {
        const char* k = std::getenv("FLOWPRO_GMRES_CHUNK");   // Arnoldi steps per read of the control block
        if (k && std::atoi(k) > 0) c->gmres_chunk = std::atoi(k);
        const char* gg = std::getenv("FLOWPRO_GMRES_GRAPH");
        c->gmres_graph = gg && gg[0] == '1';
    }
    HIPCHK(c, dalloc(&c->istat, 4));
    HIPCHK(c, hipMalloc((void**)&c->io, sizeof(IluIO)));
    HIPCHK(c, hipMemsetAsync(c->io, 0, sizeof(IluIO), c->s));
    {
        const char* g = std::getenv("FLOWPRO_GRAPH");
        c->use_graph = !(g && g[0] == '0');
        const char* f = std::getenv("FLOWPRO_ILU_FLOW");   // "0": one sweep launch per level
        // the dataflow sweeps keep a block row in one wavefront: blocks above 64 rows sweep per level
        c->ilu_flow = !(f && f[0] == '0') && c->B <= 64;
        // "1": the Arnoldi step's product formed inside the forward sweep (bit-identical; off by default:
        // at NS p=3, 200k triangles the rows' matrix streams did not stay ahead of the dependency front,
        // 47.6 ms per step against 45.6 ms with the product launch)
        const char* sv = std::getenv("FLOWPRO_SPMV_SWEEP");
        c->spmv_sweep = sv && sv[0] == '1' && c->ilu_flow && ilu_sweep_spmv_supported(c->B);
        const char* lt = std::getenv("FLOWPRO_GMRES_LSOLVE_ONE_THREAD");   // "1": the one-thread form (tests)
        c->lsolve_one_thread = lt && lt[0] == '1';
        const char* fg = std::getenv("FLOWPRO_FLOW_GROUPS");
        if (fg && std::atoi(fg) > 0) c->flow_groups = std::atoi(fg);
        const char* ff = std::getenv("FLOWPRO_FACTOR_GROUPS");
        if (ff && std::atoi(ff) > 0) c->factor_groups = std::atoi(ff);
    }
    HIPCHK(c, hipMemsetAsync(c->W, 0, N * sizeof(double), c->s));
    HIPCHK(c, hipMemsetAsync(c->Wo, 0, N * sizeof(double), c->s));
    HIPCHK(c, hipMemsetAsync(c->Wo2, 0, N * sizeof(double), c->s));
    HIPCHK(c, hipMemsetAsync(c->x, 0, N * sizeof(double), c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

extern "C" int fp_create(const fp_case* cs, fp_ctx** out) {
    if (!cs || !out || cs->n_elems <= 0 || !cs->tp || !cs->pxy || cs->order < 1)
        return FP_EINVAL;
    const bool mixed = cs->dimension != 3 && (cs->elem_types || cs->elem_orders);   // rules read from gauss_rules_dir
    if (!mixed && (cs->n_vol <= 0 || cs->n_face <= 0 || !cs->vol_xi || !cs->vol_w || !cs->face_xi || !cs->face_w))
        return FP_EINVAL;
    *out = nullptr;
    fp_ctx* c = new fp_ctx();
    int rc = create_impl(cs, c);
    if (rc != FP_OK) {
        std::fprintf(stderr, "fp_create: %s\n", c->err.c_str());
        free_all(c);
        delete c;
        return rc;
    }
    *out = c;
    return FP_OK;
}

extern "C" int fp_destroy(fp_ctx* c) {
    if (!c) return FP_EINVAL;
    if (c->s) (void)hipStreamSynchronize(c->s);
    free_all(c);
    delete c;
    return FP_OK;
}

extern "C" const char* fp_last_error(const fp_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int fp_info(const fp_ctx* c, long long* info) {
    if (!c || !info) return FP_EINVAL;
    info[0] = c->NB;
    info[1] = c->NEQ;
    info[2] = c->mixed ? c->ref_dofs : c->dofs;
    info[3] = (long long)c->lvl_f_ptr.size() - 1;
    info[4] = (long long)c->lvl_b_ptr.size() - 1;
    info[5] = c->nnzb;
    return FP_OK;
}

extern "C" int fp_element_order(const fp_ctx* c, int* order) {
    if (!c || !order) return FP_EINVAL;
    std::memcpy(order, c->order_.data(), sizeof(int) * c->nElem);
    return FP_OK;
}

extern "C" void* fp_stream(fp_ctx* c) { return c ? (void*)c->s : nullptr; }

extern "C" int fp_autotune_report(const fp_ctx* c, int* variant, int* shapes, double* ms, int n) {
    if (!c || n < 0) return FP_EINVAL;
    const int nv = ilu_factor_rows_variants();
    if (variant) *variant = c->ilu_rows ? c->ilu_variant : -1;
    for (int v = 0; v < std::min(n, nv); v++) {
        if (shapes) ilu_factor_rows_shape(v, &shapes[2 * v], &shapes[2 * v + 1]);
        if (ms) ms[v] = v < (int)c->tune_ms.size() ? c->tune_ms[v] : 0.0;
    }
    return nv;
}

// ============================================================== state
// element layout, mesh order <-> device (DOF order)
// mixed meshes: the reference's layout holds element e at m_off[e] as [m][nb_e]; the device pads
// every element to [m][NB] with zeros
static int to_device(fp_ctx* c, const double* src, double* dst) {
    const size_t B = c->B;
    std::vector<double> h((size_t)c->dofs, 0.0);
    if (c->mixed) {
        for (int r = 0; r < c->nElem; r++) {
            const int e = c->order_[r], nb = c->m_nb[e];
            for (int m = 0; m < c->NEQ; m++)
                std::memcpy(&h[r * B + (size_t)m * c->NB], src + c->m_off[e] + (size_t)m * nb, nb * sizeof(double));
        }
    } else {
        for (int r = 0; r < c->nElem; r++) std::memcpy(&h[r * B], src + (size_t)c->order_[r] * B, B * sizeof(double));
    }
    HIPCHK(c, hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}
static int from_device(fp_ctx* c, const double* src, double* dst) {
    const size_t B = c->B;
    std::vector<double> h((size_t)c->dofs);
    HIPCHK(c, hipMemcpyAsync(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->mixed) {
        for (int r = 0; r < c->nElem; r++) {
            const int e = c->order_[r], nb = c->m_nb[e];
            for (int m = 0; m < c->NEQ; m++)
                std::memcpy(dst + c->m_off[e] + (size_t)m * nb, &h[r * B + (size_t)m * c->NB], nb * sizeof(double));
        }
        return FP_OK;
    }
    for (int r = 0; r < c->nElem; r++) std::memcpy(dst + (size_t)c->order_[r] * B, &h[r * B], B * sizeof(double));
    return FP_OK;
}

extern "C" int fp_set_state(fp_ctx* c, const double* W, const double* Wo, const double* Wo2) {
    if (!c || !W) return FP_EINVAL;
    int rc = to_device(c, W, c->W);
    if (rc) return rc;
    rc = to_device(c, Wo ? Wo : W, c->Wo);
    if (rc) return rc;
    return to_device(c, Wo2 ? Wo2 : (Wo ? Wo : W), c->Wo2);
}

extern "C" int fp_get_state(fp_ctx* c, double* W) {
    if (!c || !W) return FP_EINVAL;
    return from_device(c, c->W, W);
}

extern "C" int fp_next_level(fp_ctx* c) {
    if (!c) return FP_EINVAL;
    const size_t N = (size_t)c->dofs * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(c->Wo2, c->Wo, N, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->Wo, c->W, N, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

extern "C" int fp_restore_level(fp_ctx* c) {
    if (!c) return FP_EINVAL;
    HIPCHK(c, hipMemcpyAsync(c->W, c->Wo, (size_t)c->dofs * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

// ============================================================== Newton step
extern "C" int fp_time_step(fp_ctx* c, double* out) {
    if (!c || !out) return FP_EINVAL;
    HIPCHK(c, run_size_over_lambda(c, c->tmp));
    HIPCHK(c, launch_min(c->s, (size_t)c->nElem, c->tmp, c->partial, c->scal));
    HIPCHK(c, hipMemcpyAsync(out, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

extern "C" int fp_time_step_noalloc(fp_ctx* c, double* out) {
    if (!c || !out) return FP_EINVAL;
    const size_t P = (size_t)min_reduction_partials();
    const size_t need = sizeof(double) * ((size_t)c->nElem + P + 1);
    if (need > FP_SCRATCH_BYTES) {
        c->err = "fp_time_step_noalloc: " + std::to_string(c->nElem) + " elements need " + std::to_string(need) +
                 " bytes of scratch, more than its " + std::to_string(FP_SCRATCH_BYTES);
        return FP_EINVAL;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = c->device >= 0 && c->device < FP_MAX_DEVICES ? g_scratch[c->device] : nullptr;
    }
    if (!base) {
        c->err = "fp_time_step_noalloc: no scratch address for the context's device";
        return FP_EHIP;
    }
    double* per_elem = static_cast<double*>(base);
    double* partial = per_elem + c->nElem;
    double* result = partial + P;
    HIPCHK(c, run_size_over_lambda(c, per_elem));
    HIPCHK(c, launch_min(c->s, (size_t)c->nElem, per_elem, partial, result));
    HIPCHK(c, hipMemcpyAsync(out, result, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

// BDF2.coefficients (FlowPro/src/main/java/flowpro/core/element/BDF2.java:53-61)
static void bdf_coefficients(int time_order, double dt, double dto, double* a) {
    if (time_order == 2 && dto > 0) {
        a[0] = (2 * dt + dto) / (dt * (dt + dto));
        a[1] = -(dt + dto) / (dt * dto);
        a[2] = dt / (dto * (dt + dto));
    } else {
        a[0] = 1 / dt;
        a[1] = -1 / dt;
        a[2] = 0;
    }
}

static int assemble_impl(fp_ctx* c, double dt, double dto, int time_order, int with_jac, double* eq_norms, double* rhs_norm) {
    double a[3];
    bdf_coefficients(time_order, dt, dto, a);
    HIPCHK(c, run_residual(c, a));
    if (with_jac) {
        HIPCHK(c, run_jacobian(c, a[0]));
        c->jac_valid = true;
        c->last_a1 = a[0];
    }
    if (eq_norms || rhs_norm) {
        const int ne = c->NEQ;
        HIPCHK(c, launch_colsum(c->s, c->nElem, 2 * ne, c->enorm, c->partial, c->scal));   // 256 x 2 NEQ <= 4096 partials
        double h[16];   // 2 nEqs <= 12 (k-omega SST: 6 equations)
        HIPCHK(c, hipMemcpyAsync(h, c->scal, sizeof(double) * 2 * ne, hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (eq_norms)
            for (int m = 0; m < ne; m++) eq_norms[m] = std::sqrt(h[m]);
        if (rhs_norm) {
            double s = 0;
            for (int m = 0; m < ne; m++) s += h[ne + m];
            *rhs_norm = std::sqrt(s);
        }
    }
    return FP_OK;
}

extern "C" int fp_assemble(fp_ctx* c, double dt, double dto, int time_order, int with_jacobian, double* eq_norms,
                           double* rhs_norm) {
    if (!c || !(dt > 0)) return FP_EINVAL;
    int rc = assemble_impl(c, dt, dto, time_order, with_jacobian, eq_norms, rhs_norm);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

// --- preconditioner
// The level sequences of the block-ILU0 factorisation and of the two sweeps are captured
// once per context into HIP graphs and replayed (the per-level launches dominate otherwise).
template <typename TF> static int ilu_copy_launch(fp_ctx* c) {
    HIPCHK(c, launch_ilu_copy<TF>(c->s, (size_t)c->nnzb * c->B * c->B, c->val, (TF*)c->F));
    return FP_OK;
}
template <typename TF> static int ilu_level_launches(fp_ctx* c) {
    TF* F = (TF*)c->F;
    if (c->factor_flow) {
        HIPCHK(c, launch_ilu_factor_flow<TF>(c->s, c->B, c->ilu_variant >= 0 ? c->ilu_variant : ilu_factor_rows_default_variant(c->B),
                                             c->lvl_f, c->nElem, c->rowptr, c->diagpos, c->ftptr, c->ftask, F, c->istat,
                                             c->fflag, c->fctl, c->factor_groups));
        return FP_OK;
    }
    for (size_t l = 0; l + 1 < c->lvl_f_ptr.size(); l++) {
        int b = c->lvl_f_ptr[l], e = c->lvl_f_ptr[l + 1];
        if (c->ilu_rows)
            HIPCHK(c, launch_ilu_factor_rows<TF>(c->s, c->B, c->ilu_variant >= 0 ? c->ilu_variant : ilu_factor_rows_default_variant(c->B),
                                                 c->lvl_f + b, e - b, c->rowptr, c->diagpos, c->ftptr, c->ftask, F, c->istat));
        else
            HIPCHK(c, launch_ilu_factor_level<TF>(c->s, c->B, c->lvl_f + b, e - b, c->rowptr, c->colidx, c->diagpos, F,
                                                  c->istat));
    }
    return FP_OK;
}
template <typename TF> static int ilu_apply_launches(fp_ctx* c) {
    const TF* F = (const TF*)c->F;
    if (c->ilu_flow) {
        HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 0, c->meta_f, c->nElem, F, c->io, nullptr, c->ypk, c->flow_groups));
        HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 1, c->meta_b, c->nElem, F, c->io, c->ypk, c->xpk, c->flow_groups));
        return FP_OK;
    }
    for (size_t l = 0; l + 1 < c->lvl_f_ptr.size(); l++) {
        int b = c->lvl_f_ptr[l], e = c->lvl_f_ptr[l + 1];
        HIPCHK(c, launch_ilu_fwd_level<TF>(c->s, c->B, c->meta_f + 16 * (size_t)b, e - b, F, c->io, (TF*)c->iy));
    }
    for (size_t l = 0; l + 1 < c->lvl_b_ptr.size(); l++) {
        int b = c->lvl_b_ptr[l], e = c->lvl_b_ptr[l + 1];
        HIPCHK(c, launch_ilu_bwd_level<TF>(c->s, c->B, c->meta_b + 16 * (size_t)b, e - b, F, (const TF*)c->iy,
                                           (TF*)c->ixs, c->io));
    }
    return FP_OK;
}
template <typename TF> static int run_captured(fp_ctx* c, hipGraphExec_t* exec, int (*body)(fp_ctx*)) {
    if (!c->use_graph) return body(c);
    if (!*exec) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        int rc = body(c);
        hipError_t e = hipStreamEndCapture(c->s, &g);
        if (rc) return rc;
        HIPCHK(c, e);
        e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(c, e);
    }
    HIPCHK(c, hipGraphLaunch(*exec, c->s));
    return FP_OK;
}
// Autotuner of the row kernel's launch shape (task waves, inversion waves).  On the first
// factorisation of a context every shape factors the case's own assembled matrix twice, each
// time from a fresh copy of the values, and the second run is timed with HIP events on the
// library's stream; the fastest shape is kept for the context and captured into the factor
// graph.  The shapes give bit-identical factors, so the choice changes time only.
// FLOWPRO_AUTOTUNE=0 keeps the default shape.
template <typename TF> static int ilu_autotune(fp_ctx* c) {
    const int nv = ilu_factor_rows_variants();
    c->tune_ms.assign(nv, 0.0);
    c->ilu_variant = ilu_factor_rows_default_variant(c->B);
    const char* fv = std::getenv("FLOWPRO_ILU_VARIANT");   // a fixed shape (tests of bit-identity)
    if (fv && std::atoi(fv) >= 0 && std::atoi(fv) < nv) {
        c->ilu_variant = std::atoi(fv);
        return FP_OK;
    }
    const char* g = std::getenv("FLOWPRO_AUTOTUNE");
    if (!c->ilu_rows || (g && g[0] == '0')) return FP_OK;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    int best = c->ilu_variant, rc = FP_OK;
    double best_ms = std::numeric_limits<double>::infinity();
    for (int v = 0; v < nv && rc == FP_OK; v++) {
        c->ilu_variant = v;
        for (int rep = 0; rep < 2 && rc == FP_OK; rep++) {
            if ((rc = ilu_copy_launch<TF>(c))) break;
            if (hipEventRecord(e0, c->s) != hipSuccess) { rc = FP_EHIP; break; }
            if ((rc = ilu_level_launches<TF>(c))) break;
            float ms = 0;
            if (hipEventRecord(e1, c->s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = FP_EHIP;
                break;
            }
            c->tune_ms[v] = ms;
        }
        if (rc == FP_OK && c->tune_ms[v] < best_ms) {
            best_ms = c->tune_ms[v];
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->ilu_variant = best;
    c->f_fresh = false;   // F holds a factor now: the factorisation starts from a fresh copy
    if (rc != FP_OK && c->err.empty()) c->err = "ILU autotuning failed";
    return rc;
}

template <typename TF> static int ilu_factor_t(fp_ctx* c) {
    if (c->ilu_variant < 0) {
        int rc = ilu_autotune<TF>(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemsetAsync(c->istat, 0, sizeof(int), c->s));
    // the factorisation overwrites F: it starts from a copy of the assembled values, which the
    // Jacobian kernel has already written in FP32 after an assembly
    if (!c->f_fresh) {
        int rc = ilu_copy_launch<TF>(c);
        if (rc) return rc;
    }
    c->f_fresh = false;
    // the dataflow factorisation is two launches: replayed from a graph only in the per-level mode
    if (c->factor_flow) return ilu_level_launches<TF>(c);
    return run_captured<TF>(c, &c->g_factor, ilu_level_launches<TF>);
}
template <typename TF> static int ilu_apply_t(fp_ctx* c, double* out, const double* in) {
    HIPCHK(c, launch_set_io(c->s, c->io, in, out));
    if (c->ilu_flow) return ilu_apply_launches<TF>(c);   // two launches, no graph needed
    return run_captured<TF>(c, &c->g_apply, ilu_apply_launches<TF>);
}
// the dataflow application with every launch returning at once when *skip is set (GMRES steps
// enqueued past the end of a cycle)
template <typename TF> static int ilu_apply_skip_t(fp_ctx* c, double* out, const double* in, const int* skip) {
    const TF* F = (const TF*)c->F;
    HIPCHK(c, launch_set_io(c->s, c->io, in, out, skip));
    HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 0, c->meta_f, c->nElem, F, c->io, nullptr, c->ypk, c->flow_groups, skip));
    HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 1, c->meta_b, c->nElem, F, c->io, c->ypk, c->xpk, c->flow_groups, skip));
    return FP_OK;
}

// M^-1 (A v) in two launches: the forward sweep forms each row's (A v)_i itself (bit-identical to the
// product launch followed by the forward sweep), with every launch returning at once when *skip is set
template <typename TF> static int ilu_apply_spmv_skip_t(fp_ctx* c, double* out, const double* v, const int* skip) {
    const TF* F = (const TF*)c->F;
    HIPCHK(c, launch_set_io(c->s, c->io, v, out, skip));
    HIPCHK(c, launch_ilu_sweep_flow_spmv<TF>(c->s, c->B, c->meta_f, c->nElem, F, c->io, c->ypk, c->flow_groups, skip, c->rowptr,
                                             c->colidx, c->val));
    HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 1, c->meta_b, c->nElem, F, c->io, c->ypk, c->xpk, c->flow_groups, skip));
    return FP_OK;
}

// a wait of a dataflow sweep that exceeded its spin limit (a broken dependency order) voids the
// applications of the solve
static int check_sweep_waits(fp_ctx* c) {
    if (c->precond != 2 || !c->ilu_flow) return FP_OK;
    int e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, &c->io->err, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (e) {
        HIPCHK(c, hipMemsetAsync(&c->io->err, 0, sizeof(int), c->s));
        c->err = "block-ILU0 sweep: a dependency wait exceeded its spin limit";
        return FP_EHIP;
    }
    return FP_OK;
}

static int prec_factor(fp_ctx* c) {
    if (c->precond == 2) return c->bits == 32 ? ilu_factor_t<float>(c) : ilu_factor_t<double>(c);
    if (c->precond == 1) {
        HIPCHK(c, launch_gather_diag(c->s, c->nElem, c->B * c->B, c->blkd, c->val, c->bjw));
    }
    return FP_OK;
}
static int prec_apply(fp_ctx* c, double* out, const double* in) {
    if (c->precond == 2) return c->bits == 32 ? ilu_apply_t<float>(c, out, in) : ilu_apply_t<double>(c, out, in);
    if (c->precond == 1) {
        HIPCHK(c, launch_block_jacobi(c->s, c->nElem, c->B, c->bjw, in, out));
        return FP_OK;
    }
    HIPCHK(c, launch_copy(c->s, (size_t)c->dofs, in, out));
    return FP_OK;
}

static int dnorm(fp_ctx* c, const double* v, double* h) {
    HIPCHK(c, launch_dot(c->s, (size_t)c->dofs, v, v, c->partial, c->scal, 1));
    HIPCHK(c, hipMemcpyAsync(h, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

// Givens rotation (Gmres.rotmat, Gmres.java:132-148)
static void rotmat(double a, double b, double& cs, double& sn) {
    if (b == 0) { cs = 1; sn = 0; }
    else if (std::fabs(b) > std::fabs(a)) { double t = a / b; sn = 1 / std::sqrt(1 + t * t); cs = t * sn; }
    else { double t = b / a; cs = 1 / std::sqrt(1 + t * t); sn = t * cs; }
}
// elimination on the rotated (upper triangular) H (Gmres.lsolve, Gmres.java:177-199)
static std::vector<double> lsolve(std::vector<std::vector<double>> A, std::vector<double> b, int N) {
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

// y = A v with the current operator
static int op_apply(fp_ctx* c, const double* v, double* y) {
    if (c->op == 0) {
        HIPCHK(c, launch_spmv(c->s, c->nElem, c->B, c->rowptr, c->colidx, c->val, v, nullptr, y, 0));
    } else if (c->op == 1) {
        HIPCHK(c, run_jvp(c, v, c->op_a1, y));
    } else {
        HIPCHK(c, run_jtv(c, v, c->op_a1, y));
    }
    return FP_OK;
}
// y = b - A x
static int op_residual(fp_ctx* c, const double* x, const double* b, double* y) {
    if (c->op == 0) {
        HIPCHK(c, launch_spmv(c->s, c->nElem, c->B, c->rowptr, c->colidx, c->val, x, b, y, 1));
        return FP_OK;
    }
    int rc = op_apply(c, x, c->tmp);
    if (rc) return rc;
    HIPCHK(c, launch_copy(c->s, (size_t)c->dofs, b, y));
    HIPCHK(c, launch_axpy(c->s, (size_t)c->dofs, -1.0, c->tmp, y));
    return FP_OK;
}
// M^-T v with the block-ILU0 factors (U^T sweep with D^-T, then the unit L^T sweep): the
// preconditioner of the adjoint system, whose matrix is the transpose of the factored one
template <typename TF> static int ilu_apply_trans_t(fp_ctx* c, double* out, const double* in) {
    const TF* F = (const TF*)c->F;
    HIPCHK(c, launch_set_io(c->s, c->io, in, out));
    if (c->ilu_flow) {
        HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 2, c->meta_tf, c->nElem, F, c->io, nullptr, c->ypk, c->flow_groups));
        HIPCHK(c, launch_ilu_sweep_flow<TF>(c->s, c->B, 3, c->meta_tb, c->nElem, F, c->io, c->ypk, c->xpk, c->flow_groups));
        return FP_OK;
    }
    // one launch per level (blocks above 64 rows, or FLOWPRO_ILU_FLOW=0)
    for (size_t l = 0; l + 1 < c->lvl_f_ptr.size(); l++) {
        const int b = c->lvl_f_ptr[l], e = c->lvl_f_ptr[l + 1];
        HIPCHK(c, launch_ilu_tfwd_level<TF>(c->s, c->B, c->meta_tf + 16 * (size_t)b, e - b, F, c->io, (TF*)c->iy));
    }
    for (size_t l = 0; l + 1 < c->lvl_b_ptr.size(); l++) {
        const int b = c->lvl_b_ptr[l], e = c->lvl_b_ptr[l + 1];
        HIPCHK(c, launch_ilu_tbwd_level<TF>(c->s, c->B, c->meta_tb + 16 * (size_t)b, e - b, F, (const TF*)c->iy, (TF*)c->ixs,
                                            c->io));
    }
    return FP_OK;
}

// blockJacobi overwrites its diagonal blocks on the first apply of a solve (blockJacobi.java:74-90)
static bool prec_changes_on_apply(const fp_ctx* c) { return c->op != 2 && c->precond == 1; }
static int op_prec(fp_ctx* c, double* out, const double* in) {
    if (c->op == 2) {   // adjoint system: transposed block-ILU0 when factored for it, else none
        if (c->adj_prec) return c->bits == 32 ? ilu_apply_trans_t<float>(c, out, in) : ilu_apply_trans_t<double>(c, out, in);
        HIPCHK(c, launch_copy(c->s, (size_t)c->dofs, in, out));
        return FP_OK;
    }
    return prec_apply(c, out, in);
}

// Arnoldi step i of the current cycle (Gmres.java:66-116): w = M^-1 A V_i, modified Gram-Schmidt
// against V_0..V_i with the dots on the device, V_(i+1) = w / h_(i+1,i), then the Givens update of
// the control block.  Predicated (operator 0 with a dataflow or non-ILU preconditioner): every
// launch returns at once when the cycle is over.
static int arnoldi_step(fp_ctx* c, int i, bool predicated) {
    const size_t N = (size_t)c->dofs;
    const int m = c->m;
    double* hdev = c->scal + 8;   // the Gram-Schmidt dots of the step
    double* H = c->gm;
    double* cs = H + 2 * (size_t)(m + 1) * m;
    double* sn = cs + m;
    double* sv = sn + m;
    GmresCtl* ctl = c->gctl;
    const int* sk = predicated ? &ctl->stop : nullptr;
    double* Vi = c->V + (size_t)i * N;
    int rc;
    if (predicated && c->precond == 2 && c->spmv_sweep) {
        rc = c->bits == 32 ? ilu_apply_spmv_skip_t<float>(c, c->w, Vi, sk) : ilu_apply_spmv_skip_t<double>(c, c->w, Vi, sk);
        if (rc) return rc;
    } else if (predicated) {
        HIPCHK(c, launch_spmv(c->s, c->nElem, c->B, c->rowptr, c->colidx, c->val, Vi, nullptr, c->aux, 0, sk));
        if (c->precond == 2) {
            rc = c->bits == 32 ? ilu_apply_skip_t<float>(c, c->w, c->aux, sk) : ilu_apply_skip_t<double>(c, c->w, c->aux, sk);
            if (rc) return rc;
        } else if (c->precond == 1) {
            HIPCHK(c, launch_block_jacobi(c->s, c->nElem, c->B, c->bjw, c->aux, c->w, sk));
        } else {
            HIPCHK(c, launch_copy(c->s, N, c->aux, c->w, sk));
        }
    } else {
        if ((rc = op_apply(c, Vi, c->aux))) return rc;
        if ((rc = op_prec(c, c->w, c->aux))) return rc;
    }
    // MGS: h_k = (w, V_k); w -= h_k V_k  (the next dot is fused into the update)
    HIPCHK(c, launch_dot(c->s, N, c->w, c->V, c->partial, hdev, 0, sk));
    for (int k = 0; k <= i; k++) {
        const double* u = k < i ? c->V + (size_t)(k + 1) * N : nullptr;
        HIPCHK(c, launch_axpy_dot(c->s, N, c->w, c->V + (size_t)k * N, hdev + k, u, c->partial, hdev + k + 1, k < i ? 0 : 1, sk));
    }
    HIPCHK(c, launch_scale_div(c->s, N, c->V + (size_t)(i + 1) * N, c->w, hdev + i + 1, sk));
    HIPCHK(c, launch_gmres_givens(c->s, ctl, i, m, hdev, H, cs, sn, sv));
    return FP_OK;
}
// the steps [i0, i1) of a cycle; with FLOWPRO_GMRES_GRAPH=1 predicated steps replay a HIP graph captured
// at their first use (the launches of a chunk depend only on buffers fixed for the context's lifetime).
// Off by default: on 32 concurrent NS instances of 5k-50k triangles the replays measured no faster than
// the direct launches (156 against 162 M DOF-updates/s)
static int arnoldi_chunk(fp_ctx* c, int i0, int i1, bool predicated) {
    int rc;
    if (!predicated || !c->gmres_graph) {
        for (int i = i0; i < i1; i++)
            if ((rc = arnoldi_step(c, i, predicated))) return rc;
        return FP_OK;
    }
    const size_t slot = (size_t)(i0 / c->gmres_chunk);
    if (c->g_arnoldi.size() <= slot) c->g_arnoldi.resize(slot + 1, nullptr);
    hipGraphExec_t* exec = &c->g_arnoldi[slot];
    if (!*exec) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        rc = FP_OK;
        for (int i = i0; i < i1 && rc == FP_OK; i++) rc = arnoldi_step(c, i, true);
        hipError_t e = hipStreamEndCapture(c->s, &g);
        if (rc) return rc;
        HIPCHK(c, e);
        e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(c, e);
    }
    HIPCHK(c, hipGraphLaunch(*exec, c->s));
    return FP_OK;
}

// Gmres.solve (FlowPro/src/main/java/flowpro/core/LinearSolvers/Gmres.java:42-122): left
// preconditioned, restart m, modified Gram-Schmidt, residual relative to |M^-1 b|.  The scalar
// bookkeeping (Hessenberg column, Givens rotations, convergence test, triangular solve) runs on the
// device (k_gmres_*), and the kernels of an Arnoldi step return at once once the cycle is over, so the
// host enqueues gmres_chunk steps between two reads of the control block instead of waiting for every
// Hessenberg column.  Operators or preconditioners whose kernels do not take the stop flag (the
// matrix-free products, the per-level ILU graphs) read it after every step.
static int gmres(fp_ctx* c, const double* b, double* x, int* iters, double* err_out, bool* ok, bool x_zero) {
    const size_t N = (size_t)c->dofs;
    const int m = c->m;
    int rc;
    *iters = 0;
    double bnrm2, nr, error;
    if ((rc = op_prec(c, c->r, b))) return rc;
    if ((rc = dnorm(c, c->r, &bnrm2))) return rc;
    nr = bnrm2;
    if (bnrm2 == 0) bnrm2 = 1;
    // r = M^-1 (b - A x): every solve starts from x = 0, where b - A x is b itself, so for a
    // preconditioner that is a fixed linear map the preconditioned residual is the vector just
    // computed.  blockJacobi is not: its first apply eliminates the diagonal blocks in place
    // (Mat.lsolveT), so the second apply of the reference (Gmres.java:49-57) sees the upper
    // factors and is repeated here.
    if (!x_zero) {
        if ((rc = op_residual(c, x, b, c->aux))) return rc;
        if ((rc = op_prec(c, c->r, c->aux))) return rc;
        if ((rc = dnorm(c, c->r, &nr))) return rc;
    } else if (prec_changes_on_apply(c)) {
        if ((rc = op_prec(c, c->r, b))) return rc;
        if ((rc = dnorm(c, c->r, &nr))) return rc;
    }
    error = nr / bnrm2;
    if (error < c->tol) { *ok = true; *err_out = error; return FP_OK; }
    double* H = c->gm;
    double* Hw = H + (size_t)(m + 1) * m;
    double* sv = Hw + (size_t)(m + 1) * m + 2 * (size_t)m;
    double* bw = sv + m + 1;
    GmresCtl* ctl = c->gctl;
    const bool predicated = c->op == 0 && (c->precond != 2 || c->ilu_flow);
    const int chunk = predicated ? c->gmres_chunk : 1;
    GmresCtl h{0, 0, 0, 0, bnrm2, error, nr, c->tol};
    HIPCHK(c, hipMemcpyAsync(ctl, &h, sizeof(GmresCtl), hipMemcpyHostToDevice, c->s));
    for (int it = 0; it < c->cycles; it++) {
        HIPCHK(c, launch_gmres_cycle_begin(c->s, ctl, H, sv, m));
        HIPCHK(c, launch_scale_div(c->s, N, c->V, c->r, &ctl->nr));
        for (int i0 = 0; i0 < m; i0 += chunk) {
            if (i0 > 0) {   // read the control block: stop enqueueing once the cycle is over
                HIPCHK(c, hipMemcpyAsync(&h, ctl, sizeof(GmresCtl), hipMemcpyDeviceToHost, c->s));
                HIPCHK(c, hipStreamSynchronize(c->s));
                if (h.stop) break;
            }
            if ((rc = arnoldi_chunk(c, i0, std::min(m, i0 + chunk), predicated))) return rc;
        }
        HIPCHK(c, launch_gmres_lsolve(c->s, ctl, m, H, sv, Hw, bw, c->y, c->lsolve_one_thread));
        HIPCHK(c, launch_update_solution(c->s, N, 0, c->V, N, c->y, x, &ctl->nused));
        HIPCHK(c, hipMemcpyAsync(&h, ctl, sizeof(GmresCtl), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        error = h.error;
        if (h.conv) break;
        if ((rc = op_residual(c, x, b, c->aux))) return rc;
        if ((rc = op_prec(c, c->r, c->aux))) return rc;
        if ((rc = dnorm(c, c->r, &nr))) return rc;
        error = nr / bnrm2;
        if (error <= c->tol) break;
        h.nr = nr;
        h.error = error;
        HIPCHK(c, hipMemcpyAsync(ctl, &h, sizeof(GmresCtl), hipMemcpyHostToDevice, c->s));
    }
    HIPCHK(c, hipMemcpyAsync(&h, ctl, sizeof(GmresCtl), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    *iters = h.iters;
    *ok = error <= c->tol;
    *err_out = error;
    return FP_OK;
}

static int linear_solve_impl(fp_ctx* c, int op, int* iterations, double* rel_residual) {
    if (!c->jac_valid) {
        c->err = "linear solve before fp_assemble(with_jacobian=1)";
        return FP_EINVAL;
    }
    int rc = prec_factor(c);
    if (rc) return rc;
    int st = 0;
    if (c->precond == 2) {
        HIPCHK(c, hipMemcpyAsync(&st, c->istat, sizeof(int), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (st & 2) {
            c->err = "block-ILU0 factorisation: a dependency wait exceeded its spin limit";
            return FP_EHIP;
        }
        if (st) {
            c->err = "singular diagonal block during block ILU";
            return FP_SINGULAR_BLOCK;
        }
    }
    HIPCHK(c, launch_fill(c->s, (size_t)c->dofs, 0.0, c->x));
    int it = 0;
    double err = 0;
    bool ok = false;
    c->op = op;
    c->op_a1 = c->last_a1;
    rc = gmres(c, c->RHS, c->x, &it, &err, &ok, true);
    c->op = 0;
    if (rc) return rc;
    if ((rc = check_sweep_waits(c))) return rc;
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = err;
    HIPCHK(c, hipMemsetAsync(c->istat, 0, sizeof(int), c->s));
    HIPCHK(c, launch_nonfinite(c->s, (size_t)c->dofs, c->x, c->istat));
    HIPCHK(c, hipMemcpyAsync(&st, c->istat, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (st) return FP_NONFINITE;
    return ok ? FP_OK : FP_NOT_CONVERGED;
}

extern "C" int fp_linear_solve(fp_ctx* c, int* iterations, double* rel_residual) {
    if (!c) return FP_EINVAL;
    return linear_solve_impl(c, 0, iterations, rel_residual);
}

extern "C" int fp_linear_solve_jfnk(fp_ctx* c, int* iterations, double* rel_residual) {
    if (!c) return FP_EINVAL;
    return linear_solve_impl(c, 1, iterations, rel_residual);
}

extern "C" int fp_adjoint_solve(fp_ctx* c, double a1, const double* rhs, double* lambda, int* iterations, double* rel_residual) {
    if (!c || !rhs || !lambda || c->mp.kind != MODEL_NAVIER_STOKES) return FP_EINVAL;
    if (c->dim != 2) return FP_UNSUPPORTED;
    int rc = to_device(c, rhs, c->RHS);
    if (rc) return rc;
    // preconditioner: block-ILU0 of the assembled a1 M - dR/dW at the current state, applied
    // transposed (dataflow sweeps, or one launch per level); the assembled matrix and factors of the Newton step are
    // overwritten
    c->adj_prec = false;
    c->jac_valid = false;
    if (c->precond == 2) {
        HIPCHK(c, run_jacobian(c, a1));
        if ((rc = prec_factor(c))) return rc;
        int st = 0;
        HIPCHK(c, hipMemcpyAsync(&st, c->istat, sizeof(int), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (st) {
            c->err = "singular diagonal block in the adjoint preconditioner";
            return (st & 2) ? FP_EHIP : FP_SINGULAR_BLOCK;
        }
        c->adj_prec = true;
    }
    HIPCHK(c, launch_fill(c->s, (size_t)c->dofs, 0.0, c->x));
    int it = 0;
    double err = 0;
    bool ok = false;
    c->op = 2;
    c->op_a1 = a1;
    rc = gmres(c, c->RHS, c->x, &it, &err, &ok, true);
    c->op = 0;
    c->adj_prec = false;
    if (rc) return rc;
    if ((rc = check_sweep_waits(c))) return rc;
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = err;
    rc = from_device(c, c->x, lambda);
    if (rc) return rc;
    return ok ? FP_OK : FP_NOT_CONVERGED;
}

extern "C" int fp_functional(fp_ctx* c, int which, double* value, double* dIdW) {
    if (!c || !value || (which != 0 && which != 1) || c->mp.kind != MODEL_NAVIER_STOKES) return FP_EINVAL;
    if (c->dim != 2) return FP_UNSUPPORTED;   // the force coefficients integrate over the wall edges of 2D meshes
    if (c->quad)
        HIPCHK(c, launch_functional_quad(c->L, c->s, c->qtb, c->qms, c->W, c->kappa, which, c->tmp, dIdW ? c->w : nullptr));
    else if (c->mixed)
        HIPCHK(c, launch_functional_mixed(c->L, c->s, c->mms, c->W, c->kappa, which, c->tmp, dIdW ? c->w : nullptr));
    else
        HIPCHK(c, launch_functional(c->L, c->s, c->tb, c->ms, c->W, c->kappa, which, c->tmp, dIdW ? c->w : nullptr));
    HIPCHK(c, launch_dot(c->s, (size_t)c->nElem, c->tmp, nullptr, c->partial, c->scal, 2));
    HIPCHK(c, hipMemcpyAsync(value, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (dIdW) return from_device(c, c->w, dIdW);
    return FP_OK;
}

extern "C" int fp_update(fp_ctx* c, int* n_limited) {
    if (!c) return FP_EINVAL;
    HIPCHK(c, launch_axpy(c->s, (size_t)c->dofs, 1.0, c->x, c->W));
    int cnt = 0;
    if (c->positivity) {
        HIPCHK(c, hipMemsetAsync(c->istat + 1, 0, sizeof(int), c->s));
        HIPCHK(c, run_positivity(c, c->istat + 1));
        HIPCHK(c, hipMemcpyAsync(&cnt, c->istat + 1, sizeof(int), hipMemcpyDeviceToHost, c->s));
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (n_limited) *n_limited = cnt;
    return FP_OK;
}

// ============================================================== exports
extern "C" int fp_export_residual(fp_ctx* c, double* R, double* RHS) {
    if (!c) return FP_EINVAL;
    int rc = 0;
    if (R && (rc = from_device(c, c->R, R))) return rc;
    if (RHS && (rc = from_device(c, c->RHS, RHS))) return rc;
    return FP_OK;
}

extern "C" int fp_export_jacobian(fp_ctx* c, double* diag, double* neigh) {
    if (!c || !c->jac_valid) return FP_EINVAL;
    const size_t BB = (size_t)c->B * c->B;
    std::vector<double> h((size_t)c->nnzb * BB);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->val, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    std::vector<int> diagpos(c->nElem);
    HIPCHK(c, hipMemcpy(diagpos.data(), c->diagpos, sizeof(int) * c->nElem, hipMemcpyDeviceToHost));
    if (c->mixed) {
        // flat: diag element after element (B_e x B_e), neigh per element and face with a
        // neighbour t (B_t x B_e), both in mesh order; block index [column][row] over the
        // unpadded [m][nb] numbering of the two elements
        const int NB = c->NB, NEQ = c->NEQ, B = c->B;
        auto unpad = [&](const double* blk, int nbc, int nbr_, double* out) {
            for (int mc = 0; mc < NEQ; mc++)
                for (int jc = 0; jc < nbc; jc++)
                    for (int mr = 0; mr < NEQ; mr++)
                        for (int jr = 0; jr < nbr_; jr++)
                            out[(size_t)(mc * nbc + jc) * nbr_ * NEQ + mr * nbr_ + jr] = blk[(size_t)(mc * NB + jc) * B + mr * NB + jr];
        };
        const size_t BB2 = (size_t)B * B;
        size_t od = 0, on = 0;
        for (int e = 0; e < c->nElem; e++) {
            const int r = c->rank[e], nbe = c->m_nb[e];
            if (diag) unpad(&h[(size_t)diagpos[r] * BB2], nbe, nbe, diag + od);
            od += (size_t)nbe * NEQ * nbe * NEQ;
            for (int k = 0; k < 4; k++) {
                const int rt = c->h_nbr[4 * (size_t)r + k];
                if (rt < 0) continue;
                const int nbt = c->m_nb[c->order_[rt]];
                if (neigh) unpad(&h[(size_t)c->blk_own_nb[4 * (size_t)r + k] * BB2], nbt, nbe, neigh + on);
                on += (size_t)nbt * NEQ * nbe * NEQ;
            }
        }
        return FP_OK;
    }
    for (int r = 0; r < c->nElem; r++) {
        const size_t e = c->order_[r];
        if (diag) std::memcpy(diag + e * BB, &h[(size_t)diagpos[r] * BB], BB * sizeof(double));
        if (neigh)
            for (int k = 0; k < c->NF; k++) {
                double* dst = neigh + (e * c->NF + k) * BB;
                int blk = c->blk_own_nb[(size_t)c->NF * r + k];
                if (blk >= 0) std::memcpy(dst, &h[(size_t)blk * BB], BB * sizeof(double));
                else std::memset(dst, 0, BB * sizeof(double));
            }
    }
    return FP_OK;
}

extern "C" int fp_limiter(fp_ctx* c, int first_step) {
    (void)first_step;   // the first-iteration branch of DGpure.limiter has an empty body
    if (!c) return FP_EINVAL;
    if (!c->dg.av || c->dim != 2) return FP_OK;
    if (c->quad)
        HIPCHK(c, launch_av_limiter_quad(c->L, c->s, c->mp, c->qtb, c->qms, c->avp, c->W, c->esize, c->trunc, c->av, c->avold));
    else if (c->mixed)
        HIPCHK(c, launch_av_limiter_mixed(c->L, c->s, c->mp, c->mms, c->avp, c->W, c->esize, c->trunc, c->av, c->avold));
    else
        HIPCHK(c, launch_av_limiter(c->L, c->s, c->mp, c->tb, c->ms, c->avp, c->W, c->esize, c->isref, c->trunc, c->av, c->avold));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

extern "C" int fp_export_artificial_viscosity(fp_ctx* c, double* eps, double* damp_inner) {
    if (!c || !eps || !damp_inner) return FP_EINVAL;
    const int ne = c->NEQ;
    std::vector<double> h((size_t)c->nElem * (1 + ne), 0.0);
    if (c->av) {
        HIPCHK(c, hipMemcpyAsync(h.data(), c->av, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    for (int r = 0; r < c->nElem; r++) {
        const size_t e = c->order_[r];
        eps[e] = h[(size_t)r * (1 + ne)];
        for (int m = 0; m < ne; m++) damp_inner[e * ne + m] = h[(size_t)r * (1 + ne) + 1 + m];
    }
    return FP_OK;
}

// Mesh.getAvgW (Mesh.java:263-271, Element.calculateAvgW): avg[e][m] = sum_j W[e][m Nb + j] Is_j
extern "C" int fp_export_averages(fp_ctx* c, double* avg) {
    if (!c || !avg) return FP_EINVAL;
    std::vector<double> W((size_t)(c->mixed ? c->ref_dofs : c->dofs));
    int rc = from_device(c, c->W, W.data());
    if (rc) return rc;
    for (int e = 0; e < c->nElem; e++) {
        const int nb = c->mixed ? c->m_nb[e] : c->NB;
        const size_t o = c->mixed ? (size_t)c->m_off[e] : (size_t)e * c->B;
        for (int m = 0; m < c->NEQ; m++) {
            double s = 0;
            for (int j = 0; j < nb; j++)
                s = s + W[o + (size_t)m * nb + j] * (c->quad || c->mixed ? c->h_qis[(size_t)e * c->NB + j] : c->h_isref[j]);
            avg[(size_t)e * c->NEQ + m] = s;
        }
    }
    return FP_OK;
}

extern "C" int fp_export_solution(fp_ctx* c, double* x) {
    if (!c || !x) return FP_EINVAL;
    if (c->mixed) {   // the reference's DOF numbering: element at position r from gIndex m_g[r], [m][nb]
        std::vector<double> h((size_t)c->dofs);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->x, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        for (int r = 0; r < c->nElem; r++) {
            const int nb = c->m_nb[c->order_[r]];
            for (int m = 0; m < c->NEQ; m++)
                std::memcpy(x + c->m_g[r] + (size_t)m * nb, &h[(size_t)r * c->B + (size_t)m * c->NB], nb * sizeof(double));
        }
        return FP_OK;
    }
    HIPCHK(c, hipMemcpyAsync(x, c->x, (size_t)c->dofs * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return FP_OK;
}

extern "C" int fp_jv(fp_ctx* c, double a1, const double* v, double* out) {
    if (!c || !v || !out) return FP_EINVAL;
    int rc = to_device(c, v, c->aux);
    if (rc) return rc;
    HIPCHK(c, run_jvp(c, c->aux, a1, c->w));
    return from_device(c, c->w, out);
}

extern "C" int fp_jtv(fp_ctx* c, double a1, const double* v, double* out) {
    if (!c || !v || !out) return FP_EINVAL;
    int rc = to_device(c, v, c->aux);
    if (rc) return rc;
    HIPCHK(c, run_jtv(c, c->aux, a1, c->w));
    return from_device(c, c->w, out);
}

// ============================================================== measurement
extern "C" int fp_time_kernel(fp_ctx* c, int which, int reps, double* avg_ms) {
    if (!c || !avg_ms || reps < 1) return FP_EINVAL;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    double a[3] = {c->last_a1 > 0 ? c->last_a1 : 1.0, -(c->last_a1 > 0 ? c->last_a1 : 1.0), 0.0};
    int rc = FP_OK;
    auto once = [&]() -> int {
        switch (which) {
            case 0:
                HIPCHK(c, run_residual(c, a));
                return FP_OK;
            case 1:
                HIPCHK(c, run_jacobian(c, a[0]));
                c->jac_valid = true;
                return FP_OK;
            case 2:
                HIPCHK(c, launch_spmv(c->s, c->nElem, c->B, c->rowptr, c->colidx, c->val, c->RHS, nullptr, c->aux, 0));
                return FP_OK;
            case 3:
                return prec_factor(c);
            case 4:
                return prec_apply(c, c->aux, c->RHS);
            case 5:
                HIPCHK(c, run_jtv(c, c->RHS, a[0], c->w));
                return FP_OK;
        }
        return FP_EINVAL;
    };
    if (which == 6) {
        // the level launches of the block-ILU0 factorisation alone (the copy of the
        // assembled values into the factor storage precedes each repetition, untimed)
        if (c->precond != 2) return FP_EINVAL;
        double tot = 0;
        for (int i = 0; i <= reps; i++) {
            rc = c->bits == 32 ? ilu_copy_launch<float>(c) : ilu_copy_launch<double>(c);
            if (rc) return rc;
            HIPCHK(c, hipEventRecord(e0, c->s));
            rc = c->bits == 32 ? ilu_level_launches<float>(c) : ilu_level_launches<double>(c);
            if (rc) return rc;
            HIPCHK(c, hipEventRecord(e1, c->s));
            HIPCHK(c, hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            if (i > 0) tot += ms;   // repetition 0 warms up
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_ms = tot / reps;
        return FP_OK;
    }
    if ((rc = once())) return rc;   // warm-up
    HIPCHK(c, hipEventRecord(e0, c->s));
    for (int i = 0; i < reps && rc == FP_OK; i++) rc = once();
    HIPCHK(c, hipEventRecord(e1, c->s));
    HIPCHK(c, hipEventSynchronize(e1));
    if (rc == FP_OK && which == 4) rc = check_sweep_waits(c);
    if (rc) return rc;
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return rc;
}
