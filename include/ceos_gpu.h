/*
 * ceos_gpu.h -- C ABI of the MI355X element hot path of CEOS (libceos_gpu.so).
 *
 * Drop-in for the reference's element layer (paths relative to SourceCode/ of the reference):
 * a Fortran caller binds these entry points with ISO_C_BINDING and keeps its Newton loop,
 * Dirichlet handling and PARDISO solve.  All host arrays are caller-owned, contiguous and use
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
#define CEOS_MODEL_ST_VENANT_KIRCHHOFF 5    /* props: E, nu */
#define CEOS_MODEL_COMPRESSIBLE_NEO_HOOKEAN 7 /* props: mu, lambda */
#define CEOS_MODEL_VISCOELASTIC_MATRIX 12   /* props: K_inf, Mu_inf, Lambda_inf, K_e, Mu_e, Ni_v */

/* Gauss-point state rows exchanged by ceos_get_state / ceos_set_state (per Gauss point):
 *   J2:   [0..5] plastic strain (engineering shear), [6] accumulated plastic strain,
 *         [7] PlasticState flag of the last update
 *   visco: [0..8] Fv (column-major), [9..17] dv (column-major), [18] gamma, [19] time */
#define CEOS_STATE_STRIDE 20

/* element technology, FEM/Modules/ModAnalysis.f90 (only full integration is provided) */
#define CEOS_ELEM_FULL_INTEGRATION 1

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
 * handle's stream).  The call is asynchronous: no host synchronisation, and the detJ check is
 * left in a device counter read by ceos_device_status. */
int ceos_evaluate_device(ceos_handle *h, double time, const double *dU, double *dFint, void *dK,
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

/* Mesh facts after setup: number of element colours and the nnz of the upper pattern. */
int ceos_mesh_info(ceos_handle *h, int *n_colors, long long *nnz_upper);

#ifdef __cplusplus
}
#endif
#endif
