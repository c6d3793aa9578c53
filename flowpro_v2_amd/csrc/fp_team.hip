// Dispatch of the element-team kernels (fp_team.hpp) over the compiled (model, order, flags) cases.
#include "fp_team.hpp"

namespace fp {

// ------------------------------------------------------------------ dispatch
// (model, Nb, nEqs, volume points, face points, viscous, artificial viscosity)
#define FP_TEAM_CASES(X)                                 \
    X(MODEL_NAVIER_STOKES, 3, 4, 6, 2, false, false)     \
    X(MODEL_NAVIER_STOKES, 6, 4, 7, 3, false, false)     \
    X(MODEL_NAVIER_STOKES, 10, 4, 16, 4, false, false)   \
    X(MODEL_NAVIER_STOKES, 3, 4, 6, 2, true, false)      \
    X(MODEL_NAVIER_STOKES, 6, 4, 7, 3, true, false)      \
    X(MODEL_NAVIER_STOKES, 10, 4, 16, 4, true, false)    \
    X(MODEL_NAVIER_STOKES, 3, 4, 6, 2, false, true)      \
    X(MODEL_NAVIER_STOKES, 6, 4, 7, 3, false, true)      \
    X(MODEL_NAVIER_STOKES, 10, 4, 16, 4, false, true)    \
    X(MODEL_NAVIER_STOKES, 3, 4, 6, 2, true, true)       \
    X(MODEL_NAVIER_STOKES, 6, 4, 7, 3, true, true)       \
    X(MODEL_NAVIER_STOKES, 10, 4, 16, 4, true, true)     \
    X(MODEL_SA, 3, 5, 6, 2, true, false)                 \
    X(MODEL_SA, 6, 5, 7, 3, true, false)                 \
    X(MODEL_SA, 3, 5, 6, 2, true, true)                  \
    X(MODEL_SA, 6, 5, 7, 3, true, true)                  \
    X(MODEL_ADIABATIC, 3, 3, 6, 2, false, false)         \
    X(MODEL_ADIABATIC, 6, 3, 7, 3, false, false)         \
    X(MODEL_ADIABATIC, 10, 3, 16, 4, false, false)       \
    X(MODEL_ADIABATIC, 3, 3, 6, 2, false, true)          \
    X(MODEL_ADIABATIC, 6, 3, 7, 3, false, true)          \
    X(MODEL_ADIABATIC, 10, 3, 16, 4, false, true)        \
    X(MODEL_SCALAR, 1, 1, 1, 1, false, false)            \
    X(MODEL_SCALAR, 3, 1, 6, 2, false, false)            \
    X(MODEL_SCALAR, 6, 1, 7, 3, false, false)            \
    X(MODEL_SCALAR, 10, 1, 16, 4, false, false)          \
    X(MODEL_SCALAR, 3, 1, 6, 2, true, false)             \
    X(MODEL_SCALAR, 6, 1, 7, 3, true, false)             \
    X(MODEL_SCALAR, 10, 1, 16, 4, true, false)           \
    X(MODEL_SCALAR, 3, 1, 6, 2, false, true)             \
    X(MODEL_SCALAR, 6, 1, 7, 3, false, true)

#define FP_TMATCH(MK, NB, NEQ, NQV, NQF, V, A)                                                                    \
    (L.model == MK && L.nb == NB && L.neq == NEQ && L.nqv == NQV && L.nqf == NQF && (bool)L.visc == V && \
     (bool)L.av == A)

bool team_supported(const ElemLaunch& L) {
#define X(MK, NB, NEQ, NQV, NQF, V, A) \
    if (FP_TMATCH(MK, NB, NEQ, NQV, NQF, V, A)) return true;
    FP_TEAM_CASES(X)
    FP_TEAM_KW_CASES(X)
#undef X
    return false;
}

hipError_t launch_residual_team(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                const Mesh& ms, const double* W, const double* Wo, const double* Wo2, double a1, double a2,
                                double a3, double* R, double* RHS, double* enorm) {
#define X(MK, NB, NEQ, NQV, NQF, V, A)                                                                                \
    if (FP_TMATCH(MK, NB, NEQ, NQV, NQF, V, A)) {                                                                      \
        constexpr int EPB = residual_team_epb<MK, NB, NEQ, NQV, NQF>();                                              \
        hipLaunchKernelGGL((k_residual_team<MK, NB, NEQ, NQV, NQF, V, A>), dim3((ms.nElem + EPB - 1) / EPB), dim3(TEAM_THREADS), 0, s, P, \
                           G, tb, ms, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm);                                          \
        return hipGetLastError();                                                                                      \
    }
    FP_TEAM_CASES(X)
#undef X
    if (L.model == MODEL_KOMEGA) return launch_residual_team_kw(L, s, P, G, tb, ms, W, Wo, Wo2, a1, a2, a3, R, RHS, enorm);
    return hipErrorInvalidValue;
}

hipError_t launch_jacobian_team(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                                const Mesh& ms, const double* W, double a1, double* blocks, float* fblocks) {
#define X(MK, NB, NEQ, NQV, NQF, V, A)                                                                                \
    if (FP_TMATCH(MK, NB, NEQ, NQV, NQF, V, A)) {                                                                      \
        constexpr int EPB = jacobian_team_epb<MK, NB, NEQ, NQV, NQF, V || A>();                                     \
        hipLaunchKernelGGL((k_jacobian_team<MK, NB, NEQ, NQV, NQF, V, A, false>), dim3((ms.nElem + EPB - 1) / EPB), dim3(TEAM_THREADS), \
                           0, s, P, G, tb, ms, W, a1, blocks, fblocks, nullptr, nullptr);                              \
        return hipGetLastError();                                                                                      \
    }
    FP_TEAM_CASES(X)
#undef X
    if (L.model == MODEL_KOMEGA) return launch_jacobian_team_kw(L, s, P, G, tb, ms, W, a1, blocks, fblocks);
    return hipErrorInvalidValue;
}

// out = (a1 M - dR/dW)^T v (device layout) by the Jacobian team without storing a block
hipError_t launch_jtv_team(const ElemLaunch& L, hipStream_t s, const ModelParams& P, const DGParams& G, const Tables& tb,
                           const Mesh& ms, const double* W, const double* v, double a1, double* out) {
#define X(MK, NB, NEQ, NQV, NQF, V, A)                                                                                \
    if (FP_TMATCH(MK, NB, NEQ, NQV, NQF, V, A)) {                                                                      \
        constexpr int EPB = jacobian_team_epb<MK, NB, NEQ, NQV, NQF, V || A>();                                     \
        hipLaunchKernelGGL((k_jacobian_team<MK, NB, NEQ, NQV, NQF, V, A, true>), dim3((ms.nElem + EPB - 1) / EPB), dim3(TEAM_THREADS), \
                           0, s, P, G, tb, ms, W, a1, nullptr, nullptr, v, out);                                       \
        return hipGetLastError();                                                                                      \
    }
    FP_TEAM_CASES(X)
#undef X
    if (L.model == MODEL_KOMEGA) return launch_jtv_team_kw(L, s, P, G, tb, ms, W, v, a1, out);
    return hipErrorInvalidValue;
}

}  // namespace fp
