/*
 * ceos_oracle.h -- CPU restatement of the CEOS element hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path (libceos_gpu.so) never
 * links or calls it.
 *
 * It restates, in plain C and FP64, the reference algorithms named below (all citations
 * are relative to the reference tree, SourceCode/...):
 *   hex8 shape functions / Gauss rule   FEM/Elements/ElementLibrary/ModElementHexa8.f90:150-329
 *   B, G, detJ, J^-1 (current config)   FEM/Elements/ElementMethods/ModElement.f90:787-889
 *   deformation gradient (reference)    FEM/Elements/ElementMethods/ModElement.f90:1210-1300
 *   element volume                      FEM/Elements/ElementMethods/ModElement.f90:1313-1427
 *   element stiffness (NL / linear)     FEM/Elements/ElementMethods/ModElement.f90:200-348
 *   element internal force              FEM/Elements/ElementMethods/ModElement.f90:484-627
 *   S "matrix of stresses"              ConstitutiveLibrary/ModConstitutiveModel.f90:228-314
 *   element loops                       FEM/Subroutines/Sub{SolveConstitutiveModel,InternalForce,
 *                                       TangentStiffnessMatrix,UpdateMeshCoordinates}.f90
 *   CSR upper assembly (linear search)  FEM/Subroutines/SubAssembleGlobalMatrixUpperTriangular.f90:13-69
 *   CSR pattern                         FEM/ModFEMAnalysis.f90:209-273,
 *                                       Tools/MathTools/SparseMatrixRoutines/ModSparseMatrixRoutines.f90:254-288,573-609
 *   materials (3D):
 *     1  generalized Hooke's law        ConstitutiveModelsLibrary/ModGeneralizedHookesLaw .f90:475-560
 *     2  J2 plasticity (radial return)  ConstitutiveModelsLibrary/ModJ2Plasticity.f90:385-633
 *     3  Neo-Hookean (C10, K)           ConstitutiveModelsLibrary/ModNeoHookean.f90:409-615
 *     5  St.Venant-Kirchhoff            ConstitutiveModelsLibrary/ModStVenantKirchhoff.f90:643-773
 *     7  compressible Neo-Hookean       ConstitutiveModelsLibrary/ModCompressibleNeoHookean.f90:422-546
 *     12 viscoelastic matrix            ConstitutiveModelsLibrary/ModViscoelasticMatrix.f90:276-1028
 *   numerical tangent                   ConstitutiveLibrary/ModConstitutiveModel.f90:104-164
 *   Voigt push-forward                  Tools/MathTools/ModMathRoutines.f90:254-311
 *   homogenization                      MultiScale/ModMultiscaleAnalysis.f90:140-441
 *
 * Conventions (as in the reference): DOF = 3*(node-1)+k (1-based in the CSR arrays),
 * element-local DOFs node-major, Voigt order 11,22,33,12,23,13 with engineering shear
 * in strain, D and S read from their upper triangle in the nonlinear branch.
 */
#ifndef CEOS_ORACLE_H
#define CEOS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_OK 0
#define OR_ERR_DETJ (-1)
#define OR_ERR_LOCAL_NR (-2)
#define OR_ERR_SLOT (-3)
#define OR_ERR_ARG (-4)

/* model enum, ConstitutiveModelsLibrary/ModConstitutiveModelLibrary.f90:39-56 */
#define OR_MODEL_GHL 1
#define OR_MODEL_J2 2
#define OR_MODEL_NEOHOOKEAN 3
#define OR_MODEL_NH_Q1P0 4   /* mean-dilatation Neo-Hookean, needs nl analysis */
#define OR_MODEL_SVK 5
#define OR_MODEL_COMP_NEOHOOKEAN 7
#define OR_MODEL_VISCO_MATRIX 12

typedef struct or_mesh or_mesh;

/* ---- hex8 element primitives (all 0-based, row-major) ---- */
void or_hex8_gauss(double nat[8][3], double w[8]);
void or_hex8_shape(const double nat[3], double N[8]);
void or_hex8_dshape(const double nat[3], double dN[8][3]);
/* B (6x24), G (9x24), detJ from nodal coordinates; returns OR_ERR_DETJ if detJ<=1e-13 */
int or_matrix_bg(const double coord[8][3], const double nat[3], double B[6][24], double G[9][24],
                 double *detJ);
int or_deformation_gradient(const double X[8][3], const double nat[3], const double Ue[24],
                            double F[3][3]);

/* ---- material point: one update + tangent.
 * props: model dependent (see or_material_set); F row-major 3x3.
 * state_old / state_new: model dependent sizes (or_model_nstate). D: 6x6 row-major, only the
 * upper triangle is meaningful for models whose reference tangent fills only that part. */
int or_model_nstate(int model);
int or_material_point(int model, const double *props, const double F[9], double time,
                      const double *state_old, double *sigma6, double *D36, double *state_new);

/* ---- mesh level, mirrors the reference element loops ---- */
or_mesh *or_mesh_create(int nn, const double *X, int ne, const int *conn1, const int *mat_id);
void or_mesh_destroy(or_mesh *m);
/* props: GHL/SVK: E, nu; J2: E, nu, sigma_y, H; NH: C10, K; CNH: mu, lambda;
 * visco: see ceos_oracle.c (or_visco_*) */
int or_material_set(or_mesh *m, int mat_id, int model, const double *props, int nprops);
void or_set_analysis(or_mesh *m, int nl_analysis);
/* 0: the model's own tangent; 1: central-difference tangent of GetTangentModulusBase */
void or_set_tangent_mode(or_mesh *m, int numerical);
void or_set_threads(or_mesh *m, int nthreads);
int or_ndof(const or_mesh *m);
long or_csr_nnz(or_mesh *m);
int or_csr_pattern(or_mesh *m, int *rowmap, int *col);
int or_solve_constitutive(or_mesh *m, double time, const double *U);
int or_internal_force(or_mesh *m, double *Fint);
int or_tangent(or_mesh *m, double *Kval);
void or_update_mesh(or_mesh *m, const double *U);
void or_switch_converged_state(or_mesh *m);
int or_failed_elements(const or_mesh *m, int *ids, int cap);
void or_get_gp_stress(const or_mesh *m, double *out);      /* [ne][8][6] */
void or_get_gp_F(const or_mesh *m, double *out);           /* [ne][8][9] */
#define OR_STATE_STRIDE 20
void or_get_state(const or_mesh *m, int which, double *out); /* which 0=old 1=new; [ne][8][20] */
void or_set_state(or_mesh *m, int which, const double *in);
int or_element_stiffness(or_mesh *m, int e, double *Ke);   /* 24x24 row-major, e 0-based */
int or_element_internal_force(or_mesh *m, int e, double *Fe);
void or_element_volume(or_mesh *m, int e, double *V, double *VX);
int or_homogenize(or_mesh *m, double P[9], double F[9]);   /* column-major 9-vectors */

#ifdef __cplusplus
}
#endif
#endif
