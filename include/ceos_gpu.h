/*
 * ceos_gpu.h -- C ABI of the MI355X element hot path of CEOS (libceos_gpu.so).
 *
 * Drop-in for the reference's element layer (paths relative to SourceCode/ of the reference):
 * a Fortran caller binds these entry points with ISO_C_BINDING and keeps its Newton loop; its
 * Dirichlet handling and PARDISO solve either stay on the host or move to the device too
 * (ceos_bc_set, ceos_residual_bc, ceos_linear_solve).  All host arrays are caller-owned, contiguous and use
 * the reference's layouts: nodal vectors DOF = 3*(node-1)+k (k = 1,2,3), element connectivity
 * node-major with 1-based node ids, CSR 1-based with sorted columns and RowMap(n+1) = nnz+1.
 *
 * Return value of every int function: 0 on success, else one of the CEOS_ERR_* codes; when a
 * ceos_status is passed it is filled like the reference's ClassStatus
 * (Tools/ModStatus.f90:3-16): error flag, error id, description.
 * A handle is not reentrant: one caller thread at a time, one HIP stream per handle.
 */
#ifndef CEOS_GPU_H
#define CEOS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define CEOS_OK 0
#define CEOS_ERR_DETJ (-1)      /* detJ <= 1e-13 in some element (ModElement.f90:554-558, 1254-1257) */
#define CEOS_ERR_LOCAL_NR (-2)  /* material local Newton did not converge (ModViscoelasticMatrix.f90:565) */
#define CEOS_ERR_SLOT (-3)      /* CSR slot missing (SubAssembleGlobalMatrixUpperTriangular.f90:58 stops) */
#define CEOS_ERR_ARG (-4)       /* invalid argument / unsupported input */
#define CEOS_ERR_HIP (-5)       /* HIP runtime error (no device, out of memory, launch failure) */

/* constitutive model ids, ConstitutiveLibrary/ConstitutiveModelsLibrary/ModConstitutiveModelLibrary.f90:39-56 */
#define CEOS_MODEL_GENERALIZED_HOOKES_LAW 1 /* props: E, nu */
#define CEOS_MODEL_J2_PLASTICITY 2          /* props: E, nu, yield stress, hardening modulus */
#define CEOS_MODEL_NEO_HOOKEAN 3            /* props: C10, bulk modulus */
#define CEOS_MODEL_NEO_HOOKEAN_Q1P0 4       /* props: C10, bulk modulus; CEOS_ELEM_MEAN_DILATATION only */
#define CEOS_MODEL_ST_VENANT_KIRCHHOFF 5    /* props: E, nu */
#define CEOS_MODEL_COMPRESSIBLE_NEO_HOOKEAN 7 /* props: mu, lambda */
#define CEOS_MODEL_VISCOELASTIC_MATRIX 12   /* props: K_inf, Mu_inf, Lambda_inf, K_e, Mu_e, Ni_v */
/* No counterpart in the reference's library: an FE^2 macro Gauss point whose second Piola stress
 * and material tangent come from the caller (ceos_set_external_gp); no props. */
#define CEOS_MODEL_EXTERNAL 99

/* Gauss-point state rows exchanged by ceos_get_state / ceos_set_state (per Gauss point):
 *   J2:   [0..5] plastic strain (engineering shear), [6] accumulated plastic strain,
 *         [7] PlasticState flag of the last update
 *   visco: [0..8] Fv (column-major), [9..17] dv (column-major), [18] gamma, [19] time */
#define CEOS_STATE_STRIDE 20

/* element technology, FEM/Modules/ModAnalysis.f90:47-50 (reduced integration is not provided).
 * Mean dilatation (Q1P0): every Gauss point of an element sees the element's volume ratio
 * Jbar = V/V0 (ElementVolume, ModElement.f90:1313-1427) and the tangent gains
 * d2Psi/dJbar2 / V0 Bdiv^T Bdiv (ModElement.f90:429-467); it goes with model 4 only, as the
 * reference's library pairs them (ModConstitutiveModelLibrary.f90:504). */
#define CEOS_ELEM_FULL_INTEGRATION 1
#define CEOS_ELEM_MEAN_DILATATION 2

/* tangent storage (ceos_set_precision) */
#define CEOS_K_FP64 0
#define CEOS_K_FP32_STORED 1

typedef struct ceos_handle ceos_handle;

typedef struct ceos_status {
    int error;             /* 1 if an error occurred (ClassStatus%Error) */
    int error_id;          /* ClassStatus%ErrorID */
    int n_failed;          /* number of elements that failed (detJ) */
    char description[256]; /* ClassStatus%ErrorDescription */
} ceos_status;

/* Mesh setup. Replaces ElementConstructor/NodeConstructor + AllocateKgSparseUpperTriangular
 * (FEM/ModFEMAnalysis.f90:209-273): X is 3*nn node-major reference coordinates, conn is
 * npe*ne (npe = 8, hex8 node order of ModElementHexa8.f90), 1-based; mat_id per element
 * (NULL = all 1). */
int ceos_mesh_create(int nn, const double *X, int ne, int npe, const int *conn, const int *mat_id,
                     ceos_handle **h);
void ceos_mesh_destroy(ceos_handle *h);

/* ReadMaterialParameters_* of the model; may be called before or after the first evaluation. */
int ceos_material_set(ceos_handle *h, int mat_id, int model, const double *props, int nprops);

/* AnalysisSettings%NLAnalysis and %ElementTech (FEM/Modules/ModAnalysis.f90:90-118). */
int ceos_set_analysis(ceos_handle *h, int nl_analysis, int elem_tech);

/* CEOS_K_FP64 (default) or CEOS_K_FP32_STORED: element tangents accumulated in FP64 and the
 * global CSR values stored in FP32 (ceos_tangent_f32 / ceos_evaluate_device). */
int ceos_set_precision(ceos_handle *h, int mode);

/* CSR sparsity, FEM/ModFEMAnalysis.f90:209-273 with ModSparseMatrixRoutines.f90:573-609.
 * upper = 1: upper triangle incl. diagonal (Kg of the reference); upper = 0: full pattern.
 * Two-call protocol: pass rowmap = col = NULL to get n and nnz. */
int ceos_csr_pattern(ceos_handle *h, int upper, int *n, long long *nnz, int *rowmap, int *col);

/* SolveConstitutiveModel(ElementList, AnalysisSettings, Time, U, Status)
 * (FEM/Subroutines/SubSolveConstitutiveModel.f90:13-146).  U: 3*nn.  By default the same
 * device pass also integrates Fint and the tangent (see ceos_set_fused). */
int ceos_solve_constitutive(ceos_handle *h, double time, const double *U, ceos_status *st);

/* InternalForce(ElementList, AnalysisSettings, Fint, Status) (FEM/Subroutines/SubInternalForce.f90:13-97).
 * Fint: 3*nn, overwritten (the reference zeroes it first). */
int ceos_internal_force(ceos_handle *h, double *Fint, ceos_status *st);

/* TangentStiffnessMatrix(AnalysisSettings, ElementList, nDOF, Kg) FEM branch
 * (FEM/Subroutines/SubTangentStiffnessMatrix.f90:63,294-310): Kval (nnz of the upper pattern)
 * is overwritten with the assembled upper triangle. */
int ceos_tangent(ceos_handle *h, double *Kval);
int ceos_tangent_f32(ceos_handle *h, float *Kval);

/* UpdateMeshCoordinates(GlobalNodesList, AnalysisSettings, U)
 * (FEM/Subroutines/SubUpdateMeshCoordinates.f90:15-19): x = X + U. */
int ceos_update_mesh(ceos_handle *h, const double *U);

/* SwitchConvergedState of every Gauss point (FEM/ModFEMAnalysis.f90:1349-1353). */
int ceos_switch_converged_state(ceos_handle *h);

/* Ids (1-based) of the elements whose last evaluation hit detJ <= 1e-13 (the reference
 * appends them to ErrorElement.txt, SubSolveConstitutiveModel.f90:69-80). Returns the count. */
int ceos_failed_elements(ceos_handle *h, int *ids, int cap);

/* Gauss-point Cauchy stress of the last evaluation, [ne][8][6] (ClassConstitutiveModel%Stress). */
int ceos_get_gp_stress(ceos_handle *h, double *sig);

/* Committed (which = 0) or last-update (which = 1) Gauss-point state, [ne][8][CEOS_STATE_STRIDE]
 * (zero for models without state). */
int ceos_get_state(ceos_handle *h, int which, double *state);
int ceos_set_state(ceos_handle *h, int which, const double *state);

/* HomogenizeStress / HomogenizeDeformationGradient (MultiScale/ModMultiscaleAnalysis.f90:140-441):
 * volume averages over the reference configuration of P = J sigma F^-T and F, as column-major
 * 9-vectors (Tools/MathTools/ModVoigtNotation.f90:70-86), from the last evaluation. */
int ceos_homogenize(ceos_handle *h, double P[9], double F[9]);

/* Checkpoint of the Gauss-point state and the current coordinates (binary file). */
int ceos_state_save(ceos_handle *h, const char *path);
int ceos_state_load(ceos_handle *h, const char *path);

/* 1 (default): ceos_solve_constitutive also assembles Fint and the tangent in the same pass, so
 * ceos_internal_force / ceos_tangent only copy; 0: each call runs its own pass. */
int ceos_set_fused(ceos_handle *h, int fused);

/* Device-resident evaluation for callers that keep U, Fint and K in HBM (no host copies):
 * dU (3*nn doubles), dFint (3*nn doubles) and dK (nnz doubles, or floats in CEOS_K_FP32_STORED
 * mode; NULL = residual only) are device pointers; stream is a hipStream_t (NULL = the
 * handle's stream).  The detJ check is left in a device counter read by ceos_device_status.
 *
 * Once the handle is prepared (ceos_mesh_prepare with records = 1, or one earlier call of this
 * function) the call is asynchronous: kernel launches on `stream` only, no host
 * synchronisation, no allocation, no memset, so a caller may capture it in a HIP graph.  An
 * unprepared handle is prepared by the call itself, which synchronises and allocates; on a
 * stream under capture the call returns CEOS_ERR_ARG instead.  ceos_set_schedule,
 * ceos_set_pipeline, ceos_set_assembly, a change of material model and a change of element
 * technology make the handle unprepared and invalidate graphs captured before them: prepare and
 * capture again.
 *
 * Ordering: the entries that run on the handle's own stream (ceos_device_status,
 * ceos_switch_converged_state, ceos_get_state, ceos_set_state, ceos_get_gp_stress,
 * ceos_failed_elements, ceos_homogenize, the host-array entries) wait, through an event, for the
 * last eager call of a device entry on a caller's stream.  Work replayed from a captured graph is
 * not seen by the library: synchronise the stream the graph was launched on first. */
int ceos_evaluate_device(ceos_handle *h, double time, const double *dU, double *dFint, void *dK,
                         void *stream);
/* Builds the per-element tables of the evaluation (launch order, slot maps) and, with
 * records = 1, the element-record buffer of the gather assembly (2,592 B per element), so that
 * the next ceos_evaluate_device neither allocates nor synchronises.  No reference counterpart:
 * the reference builds its element data once at input (FEM/Input/ModReadInputFile.f90). */
int ceos_mesh_prepare(ceos_handle *h, int records);

/* ceos_evaluate_device for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and results as ceos_evaluate_device, but the element records of the gather assembly
 * go to a static device buffer of CEOS_STATIC_SCRATCH_MIB MiB that the library holds on every
 * device it runs on (one per device, kept until the process ends), so the call allocates nothing,
 * frees nothing and synchronises nothing.  The handle must be prepared first
 * (ceos_mesh_prepare, records = 0 suffices); otherwise CEOS_ERR_ARG.  A mesh whose records do not
 * fit is refused with CEOS_ERR_ARG: F+K needs 72 B per node block of every 16-element patch (at
 * most 2,592 B per element; about 1,535 B on a structured hex mesh, about 87,000 elements), residual
 * only (dK NULL) 192 B per element (at most 699,050).  Every call on one device shares that
 * buffer: calls that overlap in time on one device (two streams, two threads) race. */
#define CEOS_STATIC_SCRATCH_MIB 128
int ceos_evaluate_device_static(ceos_handle *h, double time, const double *dU, double *dFint, void *dK,
                                void *stream);
int ceos_update_mesh_device(ceos_handle *h, const double *dU, void *stream);
int ceos_device_status(ceos_handle *h, ceos_status *st);

/* Phase timing of the next evaluations (instrumentation, no reference counterpart): after
 * ceos_timing_start, each evaluation (up to max_evaluations) records three HIP events on the
 * stream it runs on -- before the element kernels, between them and the assembly kernel, after
 * the assembly.  ceos_timing_stop waits for the last event and returns the summed element and
 * assembly milliseconds, the number of evaluations timed and of element-kernel launches. */
int ceos_timing_start(ceos_handle *h, int max_evaluations);
int ceos_timing_stop(ceos_handle *h, double *element_ms, double *assembly_ms, int *n_evaluations,
                     long long *element_launches);

/* Execution schedule of the element launches: the id-ordered element list of every colour is
 * cut into `slabs` chunks and chunk s of all colours runs before chunk s+1 (0 or 1 = plain
 * colour order, the default).  Changing it rebuilds the per-element tables at the next
 * evaluation; results agree to rounding for every schedule. */
int ceos_set_schedule(ceos_handle *h, int slabs);

/* Assembly of Fint and the CSR tangent (both atomic-free, deterministic):
 * CEOS_ASSEMBLY_GATHER (default): the element kernel writes per-element records and a second
 *   kernel sums them into each CSR row segment through a precomputed contribution map (element
 *   order, as the reference's loop); every K entry is written once, row-contiguously.
 * CEOS_ASSEMBLY_COLORED: one launch per element colour scatters into K and Fint in place. */
#define CEOS_ASSEMBLY_GATHER 0
#define CEOS_ASSEMBLY_COLORED 1
int ceos_set_assembly(ceos_handle *h, int mode);

/* Tangent modulus used in the stiffness: the model's own (default) or the central-difference
 * tangent of GetTangentModulusBase (ConstitutiveLibrary/ModConstitutiveModel.f90:104-164), the
 * reference's fallback for models without an analytic tangent. */
#define CEOS_TANGENT_ANALYTIC 0
#define CEOS_TANGENT_NUMERICAL 1
int ceos_set_tangent_mode(ceos_handle *h, int mode);

/* Pipelined evaluation with the gather assembly (no reference counterpart: a schedule): the
 * element launches are cut into `chunks` pieces of the slot order, and after each piece the
 * gather of the CSR rows whose contributing elements are all done runs on a second stream of the
 * handle, overlapping the next piece (FP64 compute against HBM traffic).  The caller's stream
 * waits for that stream at the end of the evaluation.  Results are bitwise the same for every
 * chunk count.  chunks 0 or 1 (default): no pipeline.  On the 80^3 cube the overlap does not pay:
 * the element kernel streams 1.8 GB per evaluation itself, and the kernels contend for HBM. */
int ceos_set_pipeline(ceos_handle *h, int chunks);

/* Mesh facts after setup: number of element colours and the nnz of the upper pattern. */
int ceos_mesh_info(ceos_handle *h, int *n_colors, long long *nnz_upper);

/* Facts of the gather assembly of a prepared handle (ceos_mesh_prepare), for traffic accounting.
 * No reference counterpart.  info: [0] patches (16-slot element groups), [1] boundary records
 * (9 doubles each, one contribution word each, summed by the gather), [2] of which diagonal,
 * [3] complete records (written into K by the element kernel), [4] of which diagonal, [5] node
 * blocks of the upper pattern, [6] node blocks the gather writes, [7] of which diagonal.
 * CEOS_ERR_ARG on an unprepared handle. */
int ceos_assembly_info(ceos_handle *h, long long info[8]);

/* FE^2 macro meshes (elements of CEOS_MODEL_EXTERNAL).  The reference has no concurrent FE^2
 * loop; these stand where a macro analysis would call the RVE analysis of
 * MultiScale/ModMultiscaleAnalysis.f90 at every Gauss point.
 *
 * ceos_set_external_gp: device array [ne][8][27] in element order (0-based element, Gauss point
 * in the hex8 order), per point the second Piola stress S (Voigt 11,22,33,12,23,13) then the
 * upper triangle of dS/dE (21, engineering shear).  The element kernel forms sigma = F S F^T / J
 * and pushes the tangent forward (Push_Forward_Voigt, Tools/MathTools/ModMathRoutines.f90:254-311).
 * The pointer is kept until the next call; an evaluation with an external material and no array
 * returns CEOS_ERR_ARG. */
int ceos_set_external_gp(ceos_handle *h, const double *dext);

/* Deformation gradient at every Gauss point, F = I + sum_a u_a (x) grad_X N_a
 * (DeformationGradient, FEM/Elements/ElementMethods/ModElement.f90:1210-1300), the values the
 * element kernel uses: F[(e 8 + g) 9 + 3J + i] = F_iJ.  Device arrays on a caller stream
 * (NULL = the handle's stream), or host arrays (synchronous). */
int ceos_gp_deformation_gradient_device(ceos_handle *h, const double *dU, double *dF, void *stream);
int ceos_gp_deformation_gradient(ceos_handle *h, const double *U, double *F);

/* ------------------------------------------------------------------------------------------
 * The macro Newton step around the element path: Dirichlet boundary conditions, residual norm
 * and the linear solve, for callers that move their Newton loop's data to the device.
 *
 * ceos_bc_set: constrained DOFs of the load case (1-based, dof = 3 (node - 1) + k), prescribed
 * displacements (ClassBoundaryConditions' Presc_Disp_DOF) and fixed supports (FixedSupport%dof).
 * Replaces AllocatePrescDispSparseMapping / AllocateFixedSupportSparseMapping
 * (FEM/BoundaryConditions/ModBoundaryConditions.f90:344-536), which search all nnz entries for
 * every constrained DOF: here a per-DOF mask, O(ndof) to set, is tested as rows are streamed.
 * The call also allocates the linear solver's buffers (ceos_linear_solve_device_async may then
 * be captured). */
int ceos_bc_set(ceos_handle *h, int n_presc, const int *presc_dof, int n_fixed, const int *fixed_dof);

/* EvaluateR + EvaluateKt of Newton iteration `it` (FEM/Modules/ModFEMSystemOfEquations.f90:40-93)
 * with ApplyBoundaryConditionsNEW (ModBoundaryConditions.f90:246-340) and the residual norm of
 * NewtonRaphsonFull_Solve (Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:109-117),
 * after ceos_solve_constitutive at U:
 *   R = Fint - Fext; at it = 0, R += K (Ubar - U) over the prescribed DOFs;
 *   R(prescribed) = -(Ubar - U) at it = 0 and 0 afterwards; R(fixed) = 0;
 *   K (kept in the handle; ceos_tangent returns it afterwards, as Kg after EvaluateKt): rows and
 *   columns of constrained DOFs cleared, diagonal 1.  Like EvaluateKt, which assembles a fresh
 *   tangent every call, a second call at the same state re-evaluates K before the product.
 *   The device call works on the caller's dK as it is: it must come fresh from
 *   ceos_evaluate_device.
 * norm_type 1 = L2, 2 = max |R| (the reference's default).  Fext (NULL = 0), Ubar and R (NULL
 * = not returned) are 3 nn host arrays.
 * _device: the same on device arrays of a ceos_evaluate_device call (dK NULL = residual only);
 * dnorm receives {L2, max |R|, the selected one}.  Asynchronous on `stream`. */
int ceos_residual_bc(ceos_handle *h, int it, const double *Fext, const double *Ubar, double *R, int norm_type,
                     double *normR);
int ceos_residual_bc_device(ceos_handle *h, int it, const double *dU, const double *dFint, void *dK,
                            const double *dFext, const double *dUbar, double *dR, int norm_type, double *dnorm,
                            void *stream);

/* Linear solve of the Newton step, K x = b, replacing the PARDISO call
 * (Tools/MathTools/LinearSolver/ModPardisoSolver.f90:343-380: analysis + factorization + solve
 * every iteration, mtype -2, symmetric indefinite).  Krylov methods from x = 0 on the device with
 * a block-Jacobi (3x3 node blocks) preconditioner (a block that is not positive definite is
 * replaced by its absolute diagonal):
 *   CEOS_SOLVER_PCG     conjugate gradients, stopped at ||r|| <= rtol ||b||; an indefinite K
 *                       stops it with flag 2 (p.Kp <= 0);
 *   CEOS_SOLVER_MINRES  MINRES, for symmetric indefinite K, stopped when the residual in the
 *                       M^-1 norm is <= rtol times that of b;
 *   CEOS_SOLVER_AUTO    (default) PCG, and MINRES from x = 0 when PCG breaks down.
 * K is the handle's tangent after ceos_residual_bc (host call) or the caller's device K (device
 * call), copied once per solve into a neighbour-slot (SELL) layout of the full symmetric matrix.
 * Vectors, products and dot products are FP64; the matrix values follow
 * ceos_set_solver_precision (-1: the handle's precision, or CEOS_K_FP64 / CEOS_K_FP32_STORED).
 * Outputs: iterations (PCG + MINRES), relative residual of the last method, flag (1 converged by
 * PCG, 4 converged by MINRES, 2 breakdown, 3 iteration limit).  Both calls synchronise the
 * stream every 32 iterations to read the state. */
#define CEOS_SOLVER_AUTO 0
#define CEOS_SOLVER_PCG 1
#define CEOS_SOLVER_MINRES 2
int ceos_set_linear_solver(ceos_handle *h, int method);
int ceos_set_solver_precision(ceos_handle *h, int mode);

/* Preconditioner of the PCG (no reference counterpart: the reference factorises with PARDISO,
 * ModPardisoSolver.f90:343-380).
 *   CEOS_PREC_BLOCK_JACOBI (default): the inverse 3x3 diagonal block of every node.
 *   CEOS_PREC_TWO_LEVEL: block Jacobi plus a coarse correction, M^-1 = D^-1 + P (P^T K P)^-1 P^T,
 *   P the 6 rigid-body modes of aggregates of nodes (boxes of the reference coordinates with an
 *   edge of aggregate_size mean element edges, <= 0: 8; grown until the coarse space has at most
 *   12,288 DOFs), zero on the constrained DOFs and orthonormalized per aggregate (stored FP32).
 *   The aggregates are built at the first solve after a ceos_bc_set that changes the constrained
 *   DOFs or a change of aggregate_size.  A solve assembles P^T K P (dense, FP64) and inverts it
 *   (Cholesky, rocSOLVER; the inverse is applied in FP32); that coarse inverse then also serves
 *   the next coarse_reuse solves (0: every solve builds its own; a preconditioner built from an
 *   earlier tangent of the same load step is still symmetric positive definite, PCG only needs
 *   that).  If the factorisation fails (P^T K P not positive definite) the solves use block
 *   Jacobi.  PCG only: MINRES, and ceos_linear_solve_device_async, always use block Jacobi.  All
 *   reductions run in a fixed order (bitwise reproducible).
 * ceos_preconditioner_info: aggregates and coarse DOFs of the last two-level setup (0 before).
 * ceos_coarse_space: the aggregate of every node (nn ints), the rows of P (nn x 3 x 6 doubles,
 * row-major per node) and the last assembled P^T K P (n_coarse^2 doubles, row-major); any of them
 * may be NULL.  For checks; synchronises the device. */
#define CEOS_PREC_BLOCK_JACOBI 0
#define CEOS_PREC_TWO_LEVEL 1
int ceos_set_preconditioner(ceos_handle *h, int kind, double aggregate_size, int coarse_reuse);
int ceos_preconditioner_info(ceos_handle *h, int *n_aggregates, int *n_coarse);
int ceos_coarse_space(ceos_handle *h, int *agg, double *basis, double *Ac);
int ceos_linear_solve(ceos_handle *h, const double *b, double *x, double rtol, int maxit, int *iters,
                      double *relres, int *flag);
int ceos_linear_solve_device(ceos_handle *h, const void *dK, const double *db, double *dx, double rtol, int maxit,
                             int *iters, double *relres, int *flag, void *stream);

/* ceos_linear_solve_device without host synchronisation, for a Newton step captured in a HIP graph
 * (the reference's PARDISO call, ModPardisoSolver.f90:343-380, has no such form).  Block-Jacobi
 * PCG only: the call enqueues the operator copy, the start and exactly maxit iterations on
 * `stream`; once the relative residual test holds (or p.Kp <= 0), the remaining iterations return
 * at once and change nothing.  dresult (3 doubles, device) receives {iterations, relative
 * residual, flag} as ceos_linear_solve_device reports them (1 converged, 2 breakdown, 3 iteration
 * limit), readable once the stream has passed the call.  The solver buffers are allocated by the
 * first call outside a capture (or by ceos_linear_solve_device); on a capturing stream without
 * them, or after ceos_set_solver_precision changed their type, the call returns CEOS_ERR_ARG. */
int ceos_linear_solve_device_async(ceos_handle *h, const void *dK, const double *db, double *dx, double rtol,
                                   int maxit, double *dresult, void *stream);

/* X = X + scale DX and PostUpdate (FEMUpdateMesh: x = X0 + U when the analysis is nonlinear,
 * FEM/Modules/ModFEMSystemOfEquations.f90:99-106) on device arrays. */
int ceos_newton_update_device(ceos_handle *h, double *dU, const double *dDX, double scale, void *stream);

/* ------------------------------------------------------------------------------------------
 * Periodic RVEs on general hex8 meshes: the products with the periodicity matrix T of the
 * reference's periodic system of equations (FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:77-222,
 * EvaluateR / EvaluateKt / ExpandFluctuation: K_red = T^T K T, R_red = T^T R, DeltaU = T DeltaU_red),
 * between a handle `full` of the RVE mesh and a handle `red` of its periodic mesh: the same
 * elements with every node renamed to its reduced node (node_map[A], 0-based, one entry per node
 * of `full`: a slave's master, every corner the first corner; BuildT, :226-895, finds them).
 * `red` supplies the reduced pattern, its boundary conditions (the fixed corner) and the linear
 * solve; its elements are never evaluated.  Both handles store K in FP64.  Sums run in a fixed
 * order: bitwise reproducible.
 * ------------------------------------------------------------------------------------------ */
typedef struct ceos_periodic ceos_periodic;

/* Builds the contribution lists (host, once).  CEOS_ERR_SLOT: a pair of reduced nodes that the
 * map produces is missing from the reduced pattern (node_map and red's connectivity disagree). */
int ceos_periodic_create(ceos_handle *full, ceos_handle *red, const int *node_map, ceos_periodic **p);
void ceos_periodic_destroy(ceos_periodic *p);

/* Reduced nodes, stored contributions to K_red, and the RVE's reference volume V0. */
int ceos_periodic_info(const ceos_periodic *p, int *nn_red, long long *contributions, double *volume);

/* dKred = T^T dK T (red's upper CSR) and dRred = T^T dF (3 per reduced node); either pair may be
 * NULL.  mask_constrained: entries of dRred at DOFs that red's boundary conditions constrain are
 * zeroed (a right-hand side of the constrained system).  Device arrays, asynchronous on stream
 * (NULL = red's stream). */
int ceos_periodic_reduce_device(ceos_periodic *p, const double *dK, const double *dF, double *dKred, double *dRred,
                                int mask_constrained, void *stream);

/* dU += scale T dw (dw: 3 per reduced node, dU: 3 per node of full). */
int ceos_periodic_expand_device(ceos_periodic *p, const double *dw, double scale, double *dU, void *stream);

/* dP = (1/V0) sum_A dF_A (x) X_A (9 values, column-major): for the internal force of a
 * total-Lagrangian hex8 mesh this is the volume average of the first Piola stress that
 * HomogenizeStress forms (MultiScale/ModMultiscaleAnalysis.f90:140-329), since the shape
 * functions reproduce X; for K dU it is the matching increment. */
int ceos_periodic_homogenize_device(ceos_periodic *p, const double *dF, double *dP, void *stream);

/* dy = K dx for the handle's symmetric tangent stored as its upper CSR (the handle's precision).
 * Needs the neighbour slot tables of ceos_bc_set (an empty set of conditions will do). */
int ceos_spmv_device(ceos_handle *h, const void *dK, const double *dx, double *dy, void *stream);

/* ------------------------------------------------------------------------------------------
 * Batched periodic RVEs (FE^2 micro problems).
 *
 * Replaces, for every RVE of a batch at once, the reference's periodic multiscale analysis of
 * one macro step: the periodic system of equations with K_red = T^T K T and R_red = T^T R
 * (FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:77-222), the affine displacement U = (F - I) Y
 * with the cube corners fixed (MultiScale/ModMultiscaleBoundaryConditions.f90:200-289,
 * 632-725), its Newton loop (Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:74-163,
 * max-abs residual, relative test), the golden-ratio cut-back (FEM/ModFEMAnalysis.f90:1270-1340)
 * and the homogenized first Piola stress (MultiScale/ModMultiscaleAnalysis.f90:140-329).
 * PARDISO is replaced by block-Jacobi preconditioned CG on the GPU.
 *
 * RVE i is the unit cube cut into side[i]^3 voxel hex8 elements (side even, 4..16); phase holds
 * one byte per voxel (0 matrix, 1 inclusion), RVEs concatenated, voxel (x,y,z) at
 * x + s (y + s z).  Both phases use `model` (3 Neo-Hookean, 7 compressible Neo-Hookean or
 * 5 St.Venant-Kirchhoff) with props[nprops * phase + k].
 * ------------------------------------------------------------------------------------------ */
typedef struct ceos_rve_batch ceos_rve_batch;

int ceos_rve_batch_create(int n_rve, const int *side, const unsigned char *phase, int model,
                          const double *props, int nprops, ceos_rve_batch **b);
void ceos_rve_batch_destroy(ceos_rve_batch *b);

/* Newton: converged when max|R_red| <= max(newton_rtol * max|R_red at the predictor|,
 * newton_atol), at most newton_maxit iterations per substep, at most max_cutbacks cut-backs;
 * CG: ||r||_2 <= pcg_rtol ||b||_2, at most pcg_maxit iterations.
 * Defaults: 1e-10, 1e-13, 12, 1e-10, 5000, 20. */
int ceos_rve_set_solver(ceos_rve_batch *b, double newton_rtol, double newton_atol, int newton_maxit,
                        double pcg_rtol, int pcg_maxit, int max_cutbacks);

/* 1 (default): inexact Newton, the CG tolerance of a Newton step is
 * min(1e-2, max(pcg_rtol, 0.1 max|R| / max|R_0|, 0.5 target / max|R|)), target =
 * max(newton_rtol max|R_0|, newton_atol) (the last step is not solved past the Newton test);
 * 0: every CG solve to pcg_rtol.  The Newton convergence test is the same either way; the
 * tangent solves always use their own tolerance. */
int ceos_rve_set_inexact(ceos_rve_batch *b, int inexact);

/* Storage of the reduced tangent in the CG solves (Newton steps and condensation):
 * CEOS_K_FP64 (default) or CEOS_K_FP32_STORED, FP32 storage read into FP64 products and sums.
 * Residuals, the Newton test, dot products and CG convergence tests stay FP64; the FP64 Newton
 * residual corrects the perturbed steps (inexact Newton), so P converges to the same tolerance.
 * The condensed tangent carries the storage rounding (about 1e-6 relative). */
int ceos_rve_set_precision(ceos_rve_batch *b, int mode);

/* CG tolerance of the nine condensation solves of the homogenized tangent (default: pcg_rtol).
 * The nine solves run three at a time, each reading the reduced tangent once for its three
 * right-hand sides (search directions stored FP32; residuals and tolerances FP64). */
int ceos_rve_set_tangent_rtol(ceos_rve_batch *b, double rtol);

/* Preconditioner of the RVE CG solves (Newton steps and condensation).  1 (default): two-level,
 * block Jacobi plus a coarse correction on the three translations of every aggregate of 4 x 4 x 4
 * torus nodes (node 0 left out; 192 coarse unknowns at side 16), z = D^-1 r + P A_c^-1 P^T r,
 * A_c = P^T K P assembled from the reduced tangent and inverted at the first Newton step of
 * every substep and once per condensation (FP64, stored FP32).  0: block Jacobi alone.  The solves meet the same tolerances either way;
 * only the iteration counts (info[2]) differ.  Replaces the PARDISO solve of the periodic system
 * (FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:77-175) like the CG it preconditions; no
 * counterpart of its own in the reference. */
int ceos_rve_set_coarse(ceos_rve_batch *b, int on);

/* 1 (default): the nine condensation solves of an RVE start from that RVE's solutions of its last
 * condensation (kept per node in FP32; a nearby state, so the initial residual is a fraction of
 * the right-hand side), and, being dw/dF at the converged state, those solutions also predict
 * the next solve's first iterate, w = w_conv + sum_m v_m (F - F_conv)_m (first attempt only);
 * 0: condensation from zero, Newton from w_conv.  The stopping tests are unchanged, so results
 * agree to the solver tolerances and only the iteration counts (info[1], info[2]) differ.  ceos_rve_reset forgets the solutions;
 * ceos_rve_state_save/load carry them.  No counterpart in the reference (PARDISO is direct). */
int ceos_rve_set_warm_start(ceos_rve_batch *b, int on);

/* Tile of the Newton steps' CG vector sweeps: every thread loads the operands of g nodes (2, 4 or
 * 8) before computing them.  ceos_rve_autotune times, for each g, one linear solve of that CG per
 * RVE of a sample of the batch (up to max_rves, largest first, at their converged states, to
 * 1e-4: the time per iteration is what is compared; the states, P, A and info are not changed),
 * keeps the fastest and returns it (chosen) with the three times in seconds (may be NULL).
 * Results agree to the solver tolerances across g. */
int ceos_rve_autotune(ceos_rve_batch *b, int max_rves, int *chosen, double *seconds);
int ceos_rve_set_cg_group(ceos_rve_batch *b, int g);

/* Every RVE back to its undeformed state (F = I, zero fluctuation). */
int ceos_rve_reset(ceos_rve_batch *b);

/* Equilibrium of every RVE at its macro F (Fbar: 9 doubles per RVE, column-major like the
 * reference's Voigt 9-vectors, ModVoigtNotation.f90:70-86), reached from the RVE's last
 * converged state with the cut-back substepping.  Outputs per RVE: P (9, column-major),
 * A (81, A[9*(3J+i) + 3L+k] = dP_iJ/dF_kL by static condensation; NULL = not computed) and
 * info (4 ints: status 0 / -1, Newton iterations, CG iterations, cut-backs).  Returns
 * CEOS_ERR_LOCAL_NR if some RVE did not converge (its state is left at its last converged
 * substep).  Host arrays; the call synchronises. */
int ceos_rve_solve(ceos_rve_batch *b, const double *Fbar, double *P, double *A, int *info);

/* Same with device arrays on a caller stream (NULL = the batch's stream), asynchronous. */
int ceos_rve_solve_device(ceos_rve_batch *b, const double *dFbar, double *dP, double *dA, int *dinfo,
                          void *stream);

/* Instrumentation (no reference counterpart): per RVE, the device wall-clock ticks (100 MHz)
 * spent in the Newton assemblies and in the Newton steps' linear solves (preconditioner set-up
 * included) of the solves since the last reset. */
int ceos_rve_profile(ceos_rve_batch *b, long long *ticks);

/* Converged fluctuation of RVE `rve` (3 side^3 doubles, torus node x + s (y + s z)). */
int ceos_rve_get_fluctuation(ceos_rve_batch *b, int rve, double *w);

/* Converged state of the whole batch (F_conv, fluctuation, last homogenized P and dP/dF, and the
 * condensation solutions that warm-start the next condensation) to and
 * from a file, for restarting a multiscale analysis (FEM/ModFEMAnalysis.f90:2827-2956 rebuilds
 * the same state by replaying the end-of-step macro F through the RVE analysis).  The file carries
 * the RVE count and sides; loading into a batch of other RVEs returns CEOS_ERR_ARG.  Both calls
 * synchronise the device. */
int ceos_rve_state_save(ceos_rve_batch *b, const char *path);
int ceos_rve_state_load(ceos_rve_batch *b, const char *path);

/* The same solve for the RVEs listed in sel only (distinct ids), each from its last converged
 * state; Fsel / Psel / Asel (may be NULL) / info_sel hold one entry per listed RVE, in the order
 * of sel.  Host arrays; the call synchronises.  The other RVEs keep their state. */
int ceos_rve_solve_subset(ceos_rve_batch *b, int n_sel, const int *sel, const double *Fsel, double *Psel,
                          double *Asel, int *info_sel);

/* RVE results into external material points: for k < n, S = F^-1 P and
 * dS_IJ/dE_KL = sum_ik Finv_Ii Finv_Kk (dP_iJ/dF_kL - delta_ik S_JL)
 * (First_Elasticity_Modulus_To_Second_Voigt, Tools/MathTools/ModMathRoutines.f90:315), averaged
 * over the index symmetries, written to ext row dsel[k] (dsel NULL: row k).  dF, dP: 9 per k
 * (column-major), dA: 81 per k as ceos_rve_solve returns it, or NULL: only S is written and the
 * row keeps its tangent.  Device arrays. */
int ceos_rve_to_external_device(int n, const int *dsel, const double *dF, const double *dP, const double *dA,
                                double *dext, void *stream);

/* ------------------------------------------------------------------------------------------
 * MLP surrogate of the homogenized RVE response (no counterpart in the reference, which has no
 * concurrent FE^2 loop: it stands in for the batched RVE solve at macro points inside its
 * training range).  n_in -> width -> width -> width -> n_out with SiLU; parameters in one FP32
 * array [W1 b1 W2 b2 W3 b3 W4 b4], W row-major [out][in].  Device pointers throughout.
 * ------------------------------------------------------------------------------------------ */
typedef struct ceos_mlp ceos_mlp;

/* Weights and biases uniform in +-1/sqrt(fan_in) from splitmix64(seed); batches up to max_batch. */
int ceos_mlp_create(int n_in, int width, int n_out, int max_batch, unsigned long long seed, ceos_mlp **m);
void ceos_mlp_destroy(ceos_mlp *m);
long long ceos_mlp_num_params(const ceos_mlp *m);
int ceos_mlp_shape(const ceos_mlp *m, int *n_in, int *width, int *n_out, int *max_batch);
/* 0 (default): training GEMMs on BF16 MFMA with FP32 accumulation; 1: FP32 GEMMs throughout (a
 * reference mode for checking the BF16 path's training behaviour). */
int ceos_mlp_set_training_precision(ceos_mlp *m, int fp32);
int ceos_mlp_get_params(ceos_mlp *m, float *host_params);
/* Also resets the Adam moments. */
int ceos_mlp_set_params(ceos_mlp *m, const float *host_params);

/* Mean squared error over a batch (dX [batch][n_in], dT [batch][n_out]) and its gradient with
 * respect to the parameters, BF16 MFMA GEMMs with FP32 accumulation (host_grads / loss may be
 * NULL). */
int ceos_mlp_gradients(ceos_mlp *m, const float *dX, const float *dT, int batch, float *host_grads, float *loss);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the same gradient; loss (may be NULL) is the
 * batch MSE before the step. */
int ceos_mlp_train_step(ceos_mlp *m, const float *dX, const float *dT, int batch, float lr, float *loss);

/* The same two with a weighted mean squared error, sum W (Y - T)^2 / (batch n_out), dW [batch][n_out]
 * (NULL = unweighted).  The surrogate is trained with weights that make the loss the squared
 * relative error of the stress it predicts. */
int ceos_mlp_gradients_weighted(ceos_mlp *m, const float *dX, const float *dT, const float *dW, int batch,
                                float *host_grads, float *loss);
int ceos_mlp_train_step_weighted(ceos_mlp *m, const float *dX, const float *dT, const float *dW, int batch, float lr,
                                 float *loss);

/* FP32 inference: dY [n][n_out]; dJ (may be NULL) the input Jacobian [n][n_out][n_in] by reverse
 * mode.  stream NULL = the model's stream, synchronised before returning. */
int ceos_mlp_forward(ceos_mlp *m, const float *dX, int n, float *dY, float *dJ, void *stream);

/* ------------------------------------------------------------------------------------------
 * Surrogate of the homogenized RVE response at FE^2 macro points (no counterpart in the
 * reference).  Built on a ceos_mlp with n_in = 27, n_out = 6:
 *   x = [(U - I)_Voigt / u_scale, (C_lin - d_mean) / d_std],  U = (F^T F)^(1/2)
 *   S = C_lin : E + (|E|^2 / e_ref^2) (t_mean + t_std MLP(x)),  E = (F^T F - I) / 2
 * C_lin (21 per point, upper Voigt) is the small-strain homogenized stiffness of the point's RVE
 * (its condensed tangent at F = I).  dS/dE follows from the network's input Jacobian.  A point is
 * inside the training hull when every principal stretch lies in [lam_lo, lam_hi] and every
 * normalized descriptor component in [d_lo, d_hi]; flags[k] = 1 marks points outside.
 * ------------------------------------------------------------------------------------------ */
typedef struct ceos_surrogate ceos_surrogate;
typedef struct {
    double u_scale, lam_lo, lam_hi, e_ref;
    double d_mean[21], d_std[21], d_lo[21], d_hi[21];
    double t_mean[6], t_std[6];
} ceos_surrogate_cfg;

/* The surrogate uses mlp (not owned; it must outlive the surrogate) for up to max_points. */
int ceos_surrogate_create(ceos_mlp *mlp, const ceos_surrogate_cfg *cfg, int max_points, ceos_surrogate **s);
void ceos_surrogate_destroy(ceos_surrogate *s);

/* Network inputs (dX [n][27] FP32) and hull flags (may be NULL) of n points. */
int ceos_surrogate_features_device(ceos_surrogate *s, int n, const double *dF, const double *dClin, float *dX,
                                   int *dflags, void *stream);

/* External material points (dext [n][27], as ceos_set_external_gp takes them), hull flags (may
 * be NULL) and first Piola stresses (dP, column-major 9 per point, may be NULL) of n points.
 * Device arrays; stream NULL = synchronous. */
int ceos_surrogate_eval_device(ceos_surrogate *s, int n, const double *dF, const double *dClin, double *dext,
                               int *dflags, double *dP, void *stream);


#ifdef __cplusplus
}
#endif
#endif
