/*
 * flowpro_hip.h -- C ABI of the MI355X implicit-DG Newton step for FlowPro.
 *
 * The library replaces, for one case, the body of the Newton loop of
 * LocalImplicitSolver.solve (FlowPro/src/main/java/flowpro/core/solver/
 * LocalImplicitSolver.java:107-285): residual and element-Jacobian assembly
 * (JacobiAssembler.assemble -> BDF1/BDF2.assembleJacobiMatrix -> DGpure.residuum),
 * the preconditioned GMRES (LinearSolver.solve -> Gmres.solve with BlockILU0 or
 * blockJacobi) and the update W += x with the positivity limiter.  The caller
 * keeps the time-step loop, CFL control and residual bookkeeping (State).
 *
 * Plain C types only.  Arrays passed in are copied; the library never keeps a
 * caller pointer.  Element-wise arrays use the reference's layout: element-major
 * in mesh order, W[e][m*Nb + j] (Element.java:119-121).  Vectors of the linear
 * system use the reference's DOF numbering (gIndex, Element.setGlobalIndex).
 *
 * Return codes: 0 success; > 0 recoverable (the reference's "converges = false"
 * branch: restore Wo, reduce the CFL, repeat the step); < 0 fatal.
 */
#ifndef FLOWPRO_HIP_H
#define FLOWPRO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FP_OK = 0,
    FP_NOT_CONVERGED = 1,   /* Gmres.solve returned false (Gmres.java:121) */
    FP_SINGULAR_BLOCK = 2,  /* BlockILU0.factor ArithmeticException (BlockILU0.java:119-127) */
    FP_NONFINITE = 3,       /* non-finite x (LinearSolver.java:105-110) */
    FP_UNSUPPORTED = 4,     /* no kernel for this model / order / feature: use the Java path */
    FP_EINVAL = -1,
    FP_EHIP = -2,
    FP_ENOMEM = -3
};

typedef struct fp_ctx fp_ctx;

/* Case description: the inputs of FlowProMain.solverFactory and Mesh
 * (FlowProMain.java:486-730, Mesh.java:69-224) for straight triangles of one order, for
 * quadrilaterals (elem_type = 4, elementType.txt type 4; "quad:" shapes), or for straight
 * tetrahedra (dimension = 3; then the array shapes in brackets after "3D:" apply). */
typedef struct fp_case {
    int n_points;
    const double* pxy;            /* [n_points][2]; 3D: [n_points][3] */
    int n_elems;
    const int* tp;                /* [n_elems][3], counter-clockwise; quad: [n_elems][4]; 3D: [n_elems][4] */
    int n_bnd;
    const int* bnd;               /* [n_bnd][3]: boundary type, v0, v1 (boundaryType.txt); 3D: [n_bnd][4] */
    const double* wall_distance;  /* [n_points] or NULL */
    int order;                    /* element order k = p + 1 */
    int n_vol;                    /* volume rule (QuadratureCentral: degree 2k-1, positive weights) */
    const double* vol_xi;         /* [n_vol][2] reference-triangle coordinates; quad: QuadratureSquare of
                                     order k on [0,1]^2 (k^2 points); 3D: [n_vol][3] gaussTetra */
    const double* vol_w;          /* [n_vol] */
    int n_face;                   /* Gauss-Legendre rule of order k on [-1, 1]; 3D: the triangle rule */
    const double* face_xi;        /* [n_face]; 3D: [n_face][2] reference-triangle coordinates */
    const double* face_w;         /* [n_face] */
    int model;                    /* 0 NavierStokes (Euler when viscous == 0), 1 SpalartAllmaras,
                                     2 TestScalarEquation, 3 AdiabaticEulers (subsonic inlet),
                                     4 KOmegaSST (subsonic inlet; fields at the end) */
    int viscous;
    double kappa;                 /* Aerodynamics.init derived constants */
    double p_out;                 /* outlet pressure / pRef */
    double reynolds;              /* Re after the reference's rescaling (Aerodynamics.java:262-281) */
    double prandtl;
    int has_attack;
    double attack;
    double scalar_a0, scalar_a1, scalar_k;  /* TestScalarEquation a[0], a[1], k */
    double beta0;                 /* DDG constant (DGpure, default 2) */
    double penalty;               /* interior penalty (DGpure, default 0) */
    int ordering;                 /* 0 none, 1 rcm (ElementOrdering) */
    int preconditioner;           /* 0 none, 1 blockjacobi, 2 blockilu0, 3 ilu0 (the same factorisation
                                     as blockilu0 on the pattern of dense element blocks) */
    int precond_bits;             /* 64 or 32: precision of the preconditioner factors and sweeps */
    int gmres_restart;            /* 30 */
    int gmres_cycles;             /* 5 */
    double gmres_tol;             /* iterativeSolverTol */
    int positivity_limiter;       /* Parameters.positivityLimiter */
    /* SpalartAllmaras (SpalartAllmaras.java:36-61): inlet turbulence vtIn and the model constants
     * (cw1 is derived); wall_distance above is required for this model */
    double sa_vt_in, sa_sigma, sa_cb1, sa_cb2, sa_ka, sa_cw2, sa_cw3, sa_cv1, sa_prt;
    /* Parameters.dimension: 2 (or 0) triangles, 3 tetrahedra.  Tetrahedra carry the inviscid
     * model (viscous = 0) at orders 1 and 2, the orders basis3Dtetra can build */
    int dimension;
    /* artificial viscosity (DGpure.initMethod, DGpure.java:40-71): dampTol (shock-sensor tolerance
     * of eps), dampConst, dampInnerTol and dampInnerCoef (n_damp_inner_coef = 0: the key is absent;
     * 1: one value for every equation; nEqs: one per equation).  Triangles only */
    double damp_tol, damp_const, damp_inner_tol;
    int n_damp_inner_coef;
    double damp_inner_coef[5];
    /* Aerodynamics.numericalFluxType of NavierStokes: 0 "default" (Lax-Friedrichs), 1 "vanLeer" */
    int numerical_flux;
    /* 2D element type of every element (elementType.txt): 3 (or 0) triangleElement, 4 squareElement
     * (basis2Dsquare, transformation2Dsquare); quadrilaterals carry the NavierStokes and
     * AdiabaticEulers models at orders 1-4, without artificial viscosity */
    int elem_type;
    /* meshes mixing triangles and squares (elementType.txt) or element orders (order.txt): when
     * elem_types or elem_orders is not NULL, tp holds 4 vertices per element (the fourth unused on a
     * triangle), every element takes its own basis, volume rule and map, faces take the line rule of
     * the larger order of their two sides, and the rules are read from the reference's gaussRules
     * folder gauss_rules_dir (gauss/<k>.txt, gaussTri/<d>.txt).  Solution vectors then use the
     * reference's layout: element after element, each [m][nb_e].  Models NavierStokes and
     * AdiabaticEulers, orders 1-5 on triangles and squares, no artificial viscosity; vol_* and
     * face_* are unused */
    const int* elem_types;        /* [n_elems]: 3 or 4, or NULL (all elem_type) */
    const int* elem_orders;       /* [n_elems]: 1..5, or NULL (all order) */
    const char* gauss_rules_dir;
    /* Aerodynamics.init reference values used by KOmegaSST.sutherland (KOmegaSST.java:566-570):
     * cv and the dimensional pRef, rhoRef (TRef = pRef / (rhoRef cv (kappa - 1))) */
    double cv, p_ref, rho_ref;
    /* KOmegaSST (KOmegaSST.java:45-71; straight triangles of orders 2 and 3, wall_distance required,
     * no artificial viscosity): inlet k and log(omega) (kIn, omIn), SuthConst, and the model constants
     * a1, sk1, som1, alpha1, beta1, betast, sk2, som2, alpha2, beta2, Prt */
    double kw_k_in, kw_om_in, kw_suth, kw_a1, kw_sk1, kw_som1, kw_alpha1, kw_beta1, kw_betast, kw_sk2, kw_som2,
        kw_alpha2, kw_beta2, kw_prt;
} fp_case;

/* --- integer setup (bit-exact with the reference) ------------------------------ */
/* FlowProMain.createNeighborMatrix (FlowProMain.java:772-873): tt[n_elems][3] */
int fp_neighbor_matrix(int n_elems, const int* tp, int n_bnd, const int* bnd, int* tt);
/* ElementOrdering.reverseCuthillMcKee (ElementOrdering.java:52-142): order[n_elems] */
int fp_rcm_order(int n_elems, const int* tt, int* order);
/* the same two for tetrahedra (faces {0,2,1}, {0,1,3}, {1,2,3}, {2,0,3} of tetrahedralElement):
 * bnd[n_bnd][4] = type, v0, v1, v2; tt[n_elems][4] */
int fp_neighbor_matrix_tet(int n_elems, const int* tp, int n_bnd, const int* bnd, int* tt);
int fp_rcm_order_tet(int n_elems, const int* tt, int* order);
/* the same two for quadrilaterals (squareElement.getFaceIndexes(k) = {k, (k + 1) % 4}):
 * tp[n_elems][4], bnd[n_bnd][3], tt[n_elems][4] */
int fp_neighbor_matrix_quad(int n_elems, const int* tp, int n_bnd, const int* bnd, int* tt);
/* the same on a mesh of triangles and squares (elementType.txt): types[e] 3 or 4, tp[n_elems][4]
 * (the fourth vertex unused on a triangle), tt[n_elems][4] with -1000 after a triangle's 3 faces */
int fp_neighbor_matrix_mixed(int n_elems, const int* tp, const int* types, int n_bnd, const int* bnd, int* tt);
int fp_rcm_order_quad(int n_elems, const int* tt, int* order);
/* FlowProMain.generateWallDistanceFunction / generateDistanceFunction (FlowProMain.java:931-1014):
 * distance of every vertex from the wall faces (tt == -1) along mesh edges; dist[n_points] */
int fp_wall_distance(int n_points, const double* pxy, int n_elems, const int* tp, const int* tt, double* dist);

/* --- context --------------------------------------------------------------------- */
/* Mesh constructor + Mesh.init + LinearSolver.factory (Mesh.java:69-224, LinearSolver.java:35-89) */
int fp_create(const fp_case* c, fp_ctx** out);
int fp_destroy(fp_ctx* ctx);
const char* fp_last_error(const fp_ctx* ctx);
/* info: [0] Nb (mixed meshes: the padded size), [1] nEqs, [2] dofs (mixed: the reference's count), [3] forward ILU levels, [4] backward ILU levels, [5] block nonzeros */
int fp_info(const fp_ctx* ctx, long long* info);
/* the element order of the DOF numbering: order[r] = mesh element at position r */
int fp_element_order(const fp_ctx* ctx, int* order);

/* --- state (Element.W, Wo, Wo2) ---------------------------------------------------- */
int fp_set_state(fp_ctx* ctx, const double* W, const double* Wo, const double* Wo2);
int fp_get_state(fp_ctx* ctx, double* W);
/* copyW2Wo (Element.java:479-482) and copyWo2W (Element.java:485-487) */
int fp_next_level(fp_ctx* ctx);
int fp_restore_level(fp_ctx* ctx);

/* --- Newton step --------------------------------------------------------------------- */
/* State.getTimeStep without the CFL factor (State.java:177-192): min_e elemSize / lambda_max */
int fp_time_step(fp_ctx* ctx, double* min_size_over_lambda);
/* The same minimum computed without any allocation: no counterpart in the reference.  For callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch (the per-element values and the reduction's partial minima) lives in
 * a 32 MiB device array of the library that each GPU holds until the process ends, so it serves
 * cases of up to about 4.19 M elements (FP_EINVAL with a message beyond that).  Every call on one
 * GPU uses that GPU's array: calls that overlap on one GPU (contexts on several threads or
 * streams) race and must be serialised by the caller. */
int fp_time_step_noalloc(fp_ctx* ctx, double* min_size_over_lambda);
/* BDF1/BDF2.assembleJacobiMatrix for all elements (JacobiAssembler.assemble) or only the
 * right-hand side (assembleResiduum) when with_jacobian == 0.  dto <= 0: first step (BDF1).
 * eq_norms[nEqs] = MasterSolver.residuumNorms, rhs_norm = LocalImplicitSolver.rhsNorm. */
int fp_assemble(fp_ctx* ctx, double dt, double dto, int time_order, int with_jacobian, double* eq_norms,
                double* rhs_norm);
/* DGpure.limiter for all elements (LocalImplicitSolver.java:170-172, DGpure.java:358-400): the
 * artificial viscosity eps and dampInner of the current state, once per time step before the
 * Newton iterations; a no-op without damping.  first_step mirrors isFirstIter. */
int fp_limiter(fp_ctx* ctx, int first_step);
/* LinearSolver.solve: factor the preconditioner, GMRES from x = 0 */
int fp_linear_solve(fp_ctx* ctx, int* iterations, double* rel_residual);
/* the same Newton system solved Jacobian-free: GMRES on the matrix-free product
 * (a1 M - dR/dW) v (fp_jv), preconditioned by the factors of the assembled matrix */
int fp_linear_solve_jfnk(fp_ctx* ctx, int* iterations, double* rel_residual);
/* Element.updateW for all elements, then Element.positivityLimiter (Element.java:334-338, 385-476) */
int fp_update(fp_ctx* ctx, int* n_limited);

/* --- parity and adjoint ---------------------------------------------------------------- */
int fp_export_residual(fp_ctx* ctx, double* R, double* RHS);
/* blocks in the reference's transposed layout: diag[e][i][j] = ADiag, neigh[e][k][i][j] = ANeighs[k].A
 * (k over the element's faces: 3 on triangles, 4 on tetrahedra)
 * (zero for boundary faces).  Mixed types or orders: flat, mesh order, diag = the B_e x B_e blocks
 * one after the other, neigh = for every element and face with a neighbour t the B_t x B_e block
 * (B_e = nb_e nEqs) */
int fp_export_jacobian(fp_ctx* ctx, double* diag, double* neigh);
/* eps[n_elems] and damp_inner[n_elems][nEqs] of the last fp_limiter call, mesh order */
int fp_export_artificial_viscosity(fp_ctx* ctx, double* eps, double* damp_inner);
/* element means of the state, avg[n_elems][nEqs] in mesh order (Mesh.getAvgW: the W.txt content) */
int fp_export_averages(fp_ctx* ctx, double* avg);
/* the last Newton correction x in DOF numbering (mixed meshes: the reference's gIndex numbering,
 * elements in the numbering's order, each [m][nb_e]) */
int fp_export_solution(fp_ctx* ctx, double* x);
/* matrix-free (a1 M - dR/dW) v and its transpose, element layout in and out */
int fp_jv(fp_ctx* ctx, double a1, const double* v, double* out);
int fp_jtv(fp_ctx* ctx, double a1, const double* v, double* out);

/* functional of interest at the current state and its exact gradient dI/dW (element layout,
 * NULL to skip): which 0 DragCoefficient, 1 LiftCoefficient (FunctionalsOfInterest
 * DragCoefficient.java:31-48, LiftCoefficient.java:32-34; replaces the FD gradient of
 * OptimisationToolExport.exportFunctionalDerivative, OptimisationToolExport.java:180-217) */
int fp_functional(fp_ctx* ctx, int which, double* value, double* dIdW);
/* adjoint system (a1 M - dR/dW)^T lambda = rhs by GMRES on the matrix-free fp_jtv (element
 * layout in and out); a1 = 0 gives the steady adjoint lambda = (-dR/dW)^-T dI/dW that the
 * reference computes in MATLAB from J.txt (toolbox/optimisation/getFunctionalGradient.m:5-12).
 * With blockilu0 the system is preconditioned by the block-ILU0 of the assembled a1 M - dR/dW at
 * the current state applied transposed (U^T, then L^T sweeps); this reassembles and refactors,
 * so the Newton step's matrix must be assembled again (fp_assemble) before the next fp_linear_solve */
int fp_adjoint_solve(fp_ctx* ctx, double a1, const double* rhs, double* lambda, int* iterations, double* rel_residual);

/* --- measurement ------------------------------------------------------------------------ */
/* average device time (ms) of one launch of a kernel family, timed with HIP events on the
 * library's stream: 0 residual, 1 Jacobian, 2 SpMV, 3 ILU factor, 4 ILU apply, 5 J^T v,
 * 6 the level launches of the ILU factor without the value copy (all levels together;
 * divide by the forward level count of fp_info for the average level launch) */
int fp_time_kernel(fp_ctx* ctx, int which, int reps, double* avg_ms);
/* hipStream_t of the context (for callers that record their own events) */
void* fp_stream(fp_ctx* ctx);
/* the launch-shape autotuner of the block-ILU0 factor kernel (no reference counterpart: the
 * reference has one CPU loop): on the first factorisation of a context every shape
 * (shapes[v] = {task waves, inversion waves}; inversion waves < 0: all task waves split the block's
 * columns in the inversion) factors the case's own matrix, timed with HIP
 * events, and the fastest is kept; the shapes give bit-identical factors.  variant = the chosen
 * shape (-1 before the first factorisation or when the row kernel does not serve this block
 * size), ms[v] = its time per factorisation (0 when FLOWPRO_AUTOTUNE=0).  Fills min(n, count)
 * entries and returns the number of shapes */
int fp_autotune_report(const fp_ctx* ctx, int* variant, int* shapes, double* ms, int n);

#ifdef __cplusplus
}
#endif
#endif
