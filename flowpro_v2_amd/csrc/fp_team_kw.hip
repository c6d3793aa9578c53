// The k-omega SST element-team kernels (residual, Jacobian blocks, transposed product), in a
// translation unit of their own: build.py compiles this file at -O1.  Built at -O2 or -O3 these
// instantiations returned wrong Jacobian rows (values that changed with the layout of the LDS arrays
// and from run to run), while at -O1 they agree with the CPU restatement of KOmegaSST.java to 1e-10
// and every other model's instantiation agrees at -O3.
#include "fp_team.hpp"

namespace fp {

#define FP_KW_MATCH(MK, NB, NEQ, NQV, NQF, V, A) \
    (L.model == MK && L.nb == NB && L.neq == NEQ && L.nqv == NQV && L.nqf == NQF && (bool)L.visc == V && (bool)L.av == A)

hipError_t launch_residual_team_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                   const Mesh& ms, const double* W, const double* Wo, const double* Wo2, double a1, double a2,
                                   double a3, double* R, double* RHS, double* enorm) {
#define X(MK, NB, NEQ, NQV, NQF, V, A)                                                                                \
    if (FP_KW_MATCH(MK, NB, NEQ, NQV, NQF, V, A)) {                                                                      \
        constexpr int EPB = residual_team_epb<MK, NB, NEQ, NQV, NQF>();                                              \
        hipLaunchKernelGGL((k_residual_team<MK, NB, NEQ, NQV, NQF, V, A>), dim3((ms.nElem + EPB - 1) / EPB), dim3(TEAM_THREADS), 0, s, P, \
                           G, tb, ms, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm);                                          \
        return hipGetLastError();                                                                                      \
    }
    FP_TEAM_KW_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

template <bool JTV>
static hipError_t launch_jacobian_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                     const Mesh& ms, const double* W, double a1, double* blocks, float* fblocks, const double* v,
                                     double* out) {
#define X(MK, NB, NEQ, NQV, NQF, V, A)                                                                                \
    if (FP_KW_MATCH(MK, NB, NEQ, NQV, NQF, V, A)) {                                                                    \
        constexpr int EPB = jacobian_team_epb<MK, NB, NEQ, NQV, NQF, V || A>();                                     \
        hipLaunchKernelGGL((k_jacobian_team<MK, NB, NEQ, NQV, NQF, V, A, JTV>), dim3((ms.nElem + EPB - 1) / EPB), dim3(TEAM_THREADS), \
                           0, s, P, G, tb, ms, W, a1, blocks, fblocks, v, out);                                        \
        return hipGetLastError();                                                                                      \
    }
    FP_TEAM_KW_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_jacobian_team_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                   const Mesh& ms, const double* W, double a1, double* blocks, float* fblocks) {
    return launch_jacobian_kw<false>(L, s, P, G, tb, ms, W, a1, blocks, fblocks, nullptr, nullptr);
}

hipError_t launch_jtv_team_kw(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                              const Mesh& ms, const double* W, const double* v, double a1, double* out) {
    return launch_jacobian_kw<true>(L, s, P, G, tb, ms, W, a1, nullptr, nullptr, v, out);
}

}  // namespace fp
