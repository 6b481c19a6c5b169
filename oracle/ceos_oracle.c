/*
 * ceos_oracle.c -- CPU restatement of the CEOS element hot path (TEST INFRASTRUCTURE ONLY).
 *
 * The checker for libceos_gpu.so: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are the only users.  Every routine follows the reference file:line
 * given in its comment, written with the same operation structure (dense B/G, D read from
 * its upper triangle, Ke = sum_gp w detJ (B^T D B + G^T S G), CSR linear-search assembly)
 * so that it is an independent check of the structure-exploiting GPU kernels.
 *
 * Parity pins (tests/test_oracle_*.py): closed forms and invariants for the shape functions,
 * B/G, F and volume; finite-difference checks of every material tangent and of Ke = dFe/du;
 * brute-force and closed-form CSR patterns; the reference's own golden data from
 * Checker/Cases/{NonLinear_FEA_3D_Hexa8_StVenant, Linear_FEA_3D_Hexa8_GeneralizedHookesLaw}
 * (Newton residual histories, Marc and CEOS displacements, Gauss-point stresses).
 */
#include "ceos_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NGP 8
#define NPE 8
#define NDE 24
#define MAXMAT 64
#define MAXSTATE 32

/* ------------------------------------------------------------------------------------------
 * small 3x3 helpers, Tools/MathTools/ModMathRoutines.f90:16-62 (det = rule of Sarrus in the
 * same term order; inverse = adjugate / det)
 * ---------------------------------------------------------------------------------------- */
static double det3(const double M[3][3]) {
    return M[0][0] * M[1][1] * M[2][2] + M[0][1] * M[1][2] * M[2][0] + M[0][2] * M[1][0] * M[2][1] -
           M[0][2] * M[1][1] * M[2][0] - M[0][1] * M[1][0] * M[2][2] - M[0][0] * M[1][2] * M[2][1];
}

static void inv3(const double A[3][3], double B[3][3]) {
    double d = det3(A);
    B[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    B[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    B[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    B[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    B[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    B[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    B[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    B[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    B[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B[i][j] /= d;
}

static void matmul3(const double A[3][3], const double B[3][3], double C[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += A[i][k] * B[k][j];
            C[i][j] = s;
        }
}

static void transpose3(const double A[3][3], double B[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B[i][j] = A[j][i];
}

/* ------------------------------------------------------------------------------------------
 * hex8, FEM/Elements/ElementLibrary/ModElementHexa8.f90:150-329
 * node order (---,+--,++-,-+-,--+,+-+,+++,-++), Gauss points = node signs / sqrt(3), w = 1
 * ---------------------------------------------------------------------------------------- */
static const double HEX8_ID[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void or_hex8_gauss(double nat[8][3], double w[8]) {
    double x = 1.0 / sqrt(3.0);
    for (int g = 0; g < 8; g++) {
        for (int k = 0; k < 3; k++) nat[g][k] = HEX8_ID[g][k] * x;
        w[g] = 1.0;
    }
}

void or_hex8_shape(const double nat[3], double N[8]) {
    for (int i = 0; i < 8; i++)
        N[i] = (1.0 + HEX8_ID[i][0] * nat[0]) * (1.0 + HEX8_ID[i][1] * nat[1]) *
               (1.0 + HEX8_ID[i][2] * nat[2]) / 8.0;
}

void or_hex8_dshape(const double nat[3], double dN[8][3]) {
    double xi = nat[0], eta = nat[1], zeta = nat[2];
    for (int i = 0; i < 8; i++) {
        dN[i][0] = HEX8_ID[i][0] * (1.0 + HEX8_ID[i][1] * eta) * (1.0 + HEX8_ID[i][2] * zeta) / 8.0;
        dN[i][1] = HEX8_ID[i][1] * (1.0 + HEX8_ID[i][0] * xi) * (1.0 + HEX8_ID[i][2] * zeta) / 8.0;
        dN[i][2] = HEX8_ID[i][2] * (1.0 + HEX8_ID[i][0] * xi) * (1.0 + HEX8_ID[i][1] * eta) / 8.0;
    }
}

/* Jacob(i,j) = sum_n dN(n,i) coord(n,j); gradN(n,:) = Jacob^-1 dN(n,:)
 * FEM/Elements/ElementMethods/ModElement.f90:848-866 */
static int grad_shape(const double coord[8][3], const double nat[3], double grad[8][3], double *detJ) {
    double dN[8][3], Jac[3][3], Ji[3][3];
    or_hex8_dshape(nat, dN);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int n = 0; n < 8; n++) s += dN[n][i] * coord[n][j];
            Jac[i][j] = s;
        }
    *detJ = det3(Jac);
    if (*detJ <= 1.0e-13) return OR_ERR_DETJ;
    inv3(Jac, Ji);
    for (int n = 0; n < 8; n++)
        for (int i = 0; i < 3; i++)
            grad[n][i] = Ji[i][0] * dN[n][0] + Ji[i][1] * dN[n][1] + Ji[i][2] * dN[n][2];
    return OR_OK;
}

/* B rows: e11,e22,e33,g12,g23,g13; G rows: du_i/dx_j, i-major.
 * FEM/Elements/ElementMethods/ModElement.f90:867-888 */
int or_matrix_bg(const double coord[8][3], const double nat[3], double B[6][24], double G[9][24],
                 double *detJ) {
    double g[8][3];
    int st = grad_shape(coord, nat, g, detJ);
    if (st) return st;
    memset(B, 0, sizeof(double) * 6 * 24);
    memset(G, 0, sizeof(double) * 9 * 24);
    for (int n = 0; n < 8; n++) {
        int c = 3 * n;
        B[0][c + 0] = g[n][0];
        B[1][c + 1] = g[n][1];
        B[2][c + 2] = g[n][2];
        B[3][c + 0] = g[n][1];
        B[3][c + 1] = g[n][0];
        B[4][c + 2] = g[n][1];
        B[4][c + 1] = g[n][2];
        B[5][c + 2] = g[n][0];
        B[5][c + 0] = g[n][2];
        for (int j = 0; j < 3; j++) {
            G[0 + j][c + 0] = g[n][j];
            G[3 + j][c + 1] = g[n][j];
            G[6 + j][c + 2] = g[n][j];
        }
    }
    return OR_OK;
}

/* F(i,j) = delta_ij + sum_n dN(n,j)/dX u_n,i, FEM/Elements/ElementMethods/ModElement.f90:1210-1300 */
int or_deformation_gradient(const double X[8][3], const double nat[3], const double Ue[24],
                            double F[3][3]) {
    double g[8][3], detJ;
    int st = grad_shape(X, nat, g, &detJ);
    if (st) return st;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int n = 0; n < 8; n++) s += g[n][j] * Ue[3 * n + i];
            F[i][j] = s;
        }
    for (int i = 0; i < 3; i++) F[i][i] += 1.0;
    return OR_OK;
}

/* ------------------------------------------------------------------------------------------
 * Voigt helpers, Tools/MathTools/ModMathRoutines.f90:194-205, 532-660
 * ---------------------------------------------------------------------------------------- */
static void ball_v(const double a[6], const double b[6], double T[6][6]) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) T[i][j] = a[i] * b[j];
}

static void square_v(const double A[6], const double B[6], double T[6][6]) {
    T[0][0] = A[0] * B[0];
    T[0][1] = A[3] * B[3];
    T[0][2] = A[5] * B[5];
    T[0][3] = A[0] * B[3] / 2 + A[3] * B[0] / 2;
    T[0][4] = A[3] * B[5] / 2 + A[5] * B[3] / 2;
    T[0][5] = A[0] * B[5] / 2 + A[5] * B[0] / 2;
    T[1][0] = A[3] * B[3];
    T[1][1] = A[1] * B[1];
    T[1][2] = A[4] * B[4];
    T[1][3] = A[3] * B[1] / 2 + A[1] * B[3] / 2;
    T[1][4] = A[1] * B[4] / 2 + A[4] * B[1] / 2;
    T[1][5] = A[3] * B[4] / 2 + A[4] * B[3] / 2;
    T[2][0] = A[5] * B[5];
    T[2][1] = A[4] * B[4];
    T[2][2] = A[2] * B[2];
    T[2][3] = A[5] * B[4] / 2 + A[4] * B[5] / 2;
    T[2][4] = A[4] * B[2] / 2 + A[2] * B[4] / 2;
    T[2][5] = A[5] * B[2] / 2 + A[2] * B[5] / 2;
    T[3][0] = A[0] * B[3] / 2 + A[3] * B[0] / 2;
    T[3][1] = A[3] * B[1] / 2 + A[1] * B[3] / 2;
    T[3][2] = A[5] * B[4] / 2 + A[4] * B[5] / 2;
    T[3][3] = A[0] * B[1] / 4 + A[3] * B[3] / 2 + A[1] * B[0] / 4;
    T[3][4] = A[3] * B[4] / 4 + A[5] * B[1] / 4 + A[1] * B[5] / 4 + A[4] * B[3] / 4;
    T[3][5] = A[0] * B[4] / 4 + A[5] * B[3] / 4 + A[3] * B[5] / 4 + A[4] * B[0] / 4;
    T[4][0] = A[3] * B[5] / 2 + A[5] * B[3] / 2;
    T[4][1] = A[1] * B[4] / 2 + A[4] * B[1] / 2;
    T[4][2] = A[4] * B[2] / 2 + A[2] * B[4] / 2;
    T[4][3] = A[3] * B[4] / 4 + A[5] * B[1] / 4 + A[1] * B[5] / 4 + A[4] * B[3] / 4;
    T[4][4] = A[1] * B[2] / 4 + A[4] * B[4] / 2 + A[2] * B[1] / 4;
    T[4][5] = A[3] * B[2] / 4 + A[5] * B[4] / 4 + A[4] * B[5] / 4 + A[2] * B[3] / 4;
    T[5][0] = A[0] * B[5] / 2 + A[5] * B[0] / 2;
    T[5][1] = A[3] * B[4] / 2 + A[4] * B[3] / 2;
    T[5][2] = A[5] * B[2] / 2 + A[2] * B[5] / 2;
    T[5][3] = A[0] * B[4] / 4 + A[5] * B[3] / 4 + A[3] * B[5] / 4 + A[4] * B[0] / 4;
    T[5][4] = A[3] * B[2] / 4 + A[5] * B[4] / 4 + A[4] * B[5] / 4 + A[2] * B[3] / 4;
    T[5][5] = A[0] * B[2] / 4 + A[5] * B[5] / 2 + A[2] * B[0] / 4;
}

static double inner_v(const double a[6], const double b[6]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + 2.0 * a[3] * b[3] + 2.0 * a[4] * b[4] +
           2.0 * a[5] * b[5];
}

static void to_voigt(const double A[3][3], double v[6]) {
    v[0] = A[0][0];
    v[1] = A[1][1];
    v[2] = A[2][2];
    v[3] = A[0][1];
    v[4] = A[1][2];
    v[5] = A[0][2];
}

/* T = (1/J) A W_sym A^T with W read from its upper triangle (dsymm 'R','U'),
 * Tools/MathTools/ModMathRoutines.f90:254-311 */
static void push_forward_v(const double W[6][6], const double F[3][3], double T[6][6]) {
    double J = det3(F), Ft[3][3], A[6][6], Ws[6][6], AW[6][6];
    transpose3(F, Ft);
#define FT(i, j) Ft[(i)-1][(j)-1]
    A[0][0] = FT(1, 1) * FT(1, 1);
    A[0][1] = FT(2, 1) * FT(2, 1);
    A[0][2] = FT(3, 1) * FT(3, 1);
    A[0][3] = 2.0 * FT(1, 1) * FT(2, 1);
    A[0][4] = 2.0 * FT(2, 1) * FT(3, 1);
    A[0][5] = 2.0 * FT(1, 1) * FT(3, 1);
    A[1][0] = FT(1, 2) * FT(1, 2);
    A[1][1] = FT(2, 2) * FT(2, 2);
    A[1][2] = FT(3, 2) * FT(3, 2);
    A[1][3] = 2.0 * FT(1, 2) * FT(2, 2);
    A[1][4] = 2.0 * FT(2, 2) * FT(3, 2);
    A[1][5] = 2.0 * FT(1, 2) * FT(3, 2);
    A[2][0] = FT(1, 3) * FT(1, 3);
    A[2][1] = FT(2, 3) * FT(2, 3);
    A[2][2] = FT(3, 3) * FT(3, 3);
    A[2][3] = 2.0 * FT(1, 3) * FT(2, 3);
    A[2][4] = 2.0 * FT(2, 3) * FT(3, 3);
    A[2][5] = 2.0 * FT(1, 3) * FT(3, 3);
    A[3][0] = FT(1, 1) * FT(1, 2);
    A[3][1] = FT(2, 1) * FT(2, 2);
    A[3][2] = FT(3, 1) * FT(3, 2);
    A[3][3] = FT(1, 1) * FT(2, 2) + FT(2, 1) * FT(1, 2);
    A[3][4] = FT(2, 1) * FT(3, 2) + FT(3, 1) * FT(2, 2);
    A[3][5] = FT(3, 1) * FT(1, 2) + FT(1, 1) * FT(3, 2);
    A[4][0] = FT(1, 2) * FT(1, 3);
    A[4][1] = FT(2, 2) * FT(2, 3);
    A[4][2] = FT(3, 2) * FT(3, 3);
    A[4][3] = FT(1, 2) * FT(2, 3) + FT(2, 2) * FT(1, 3);
    A[4][4] = FT(2, 2) * FT(3, 3) + FT(3, 2) * FT(2, 3);
    A[4][5] = FT(3, 2) * FT(1, 3) + FT(1, 2) * FT(3, 3);
    A[5][0] = FT(1, 1) * FT(1, 3);
    A[5][1] = FT(2, 1) * FT(2, 3);
    A[5][2] = FT(3, 1) * FT(3, 3);
    A[5][3] = FT(1, 3) * FT(2, 1) + FT(2, 3) * FT(1, 1);
    A[5][4] = FT(2, 3) * FT(3, 1) + FT(3, 3) * FT(2, 1);
    A[5][5] = FT(3, 3) * FT(1, 1) + FT(1, 3) * FT(3, 1);
#undef FT
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) Ws[i][j] = (i <= j) ? W[i][j] : W[j][i];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += A[i][k] * Ws[k][j];
            AW[i][j] = s;
        }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += AW[i][k] * A[j][k];
            T[i][j] = s / J;
        }
}

/* ------------------------------------------------------------------------------------------
 * materials
 * ---------------------------------------------------------------------------------------- */
static void elastic_D(double E, double nu, double D[6][6]) {
    /* ModGeneralizedHookesLaw .f90:526-560 and ModJ2Plasticity.f90:587-597 */
    double cte = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    memset(D, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) D[i][j] = cte * ((i == j) ? (1.0 - nu) : nu);
    for (int i = 3; i < 6; i++) D[i][i] = cte * (0.5 - nu);
}

static void small_strain(const double F[3][3], double e[6]) {
    e[0] = F[0][0] - 1.0;
    e[1] = F[1][1] - 1.0;
    e[2] = F[2][2] - 1.0;
    e[3] = F[0][1] + F[1][0];
    e[4] = F[1][2] + F[2][1];
    e[5] = F[0][2] + F[2][0];
}

/* model 1: ModGeneralizedHookesLaw .f90:475-520 */
static void mat_ghl(const double *p, const double F[3][3], double s[6], double D[6][6]) {
    double e[6];
    small_strain(F, e);
    elastic_D(p[0], p[1], D);
    for (int i = 0; i < 6; i++) {
        double acc = 0.0;
        for (int j = 0; j < 6; j++) acc += D[i][j] * e[j];
        s[i] = acc;
    }
}

/* model 7: ModCompressibleNeoHookean.f90:422-546 */
static void mat_cnh(const double *p, const double F[3][3], double s[6], double D[6][6]) {
    double mu = p[0], lam = p[1], Ft[3][3], b[3][3];
    transpose3(F, Ft);
    matmul3(F, Ft, b);
    double J = det3(F), lnJ = log(J);
    double S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            S[i][j] = (mu / J) * (b[i][j] - (i == j ? 1.0 : 0.0)) + ((lam / J) * lnJ) * (i == j ? 1.0 : 0.0);
    to_voigt(S, s);
    double a = lam / J, c = (2.0 / J) * (mu - lam * lnJ);
    memset(D, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) D[i][j] = a;
    for (int i = 0; i < 3; i++) D[i][i] += c;
    for (int i = 3; i < 6; i++) D[i][i] = c * 0.5;
}

/* model 3: ModNeoHookean.f90:409-615 */
static void mat_nh(const double *p, const double F[3][3], double s[6], double D[6][6]) {
    double C10 = p[0], K = p[1], Ft[3][3], b[3][3];
    transpose3(F, Ft);
    matmul3(F, Ft, b);
    double trb = b[0][0] + b[1][1] + b[2][2];
    double J = det3(F);
    double pr = 3.0 * K * pow(J, -2.0 / 3.0) * (pow(J, 1.0 / 3.0) - 1.0);
    double S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            S[i][j] = 2.0 * C10 * pow(J, -5.0 / 3.0) * (b[i][j] - (trb / 3.0) * (i == j ? 1.0 : 0.0)) +
                      pr * (i == j ? 1.0 : 0.0);
    to_voigt(S, s);
    /* tangent */
    double C[3][3], Ci[3][3], CV[6], CiV[6];
    matmul3(Ft, F, C);
    inv3(C, Ci);
    to_voigt(C, CV);
    to_voigt(Ci, CiV);
    double Sf[6] = {2.0 * C10, 2.0 * C10, 2.0 * C10, 0, 0, 0};
    double ipSC = inner_v(Sf, CV);
    double devSf[6], Siso[6];
    for (int i = 0; i < 6; i++) devSf[i] = Sf[i] - (1.0 / 3.0) * ipSC * CiV[i];
    double Jm23 = pow(J, -2.0 / 3.0);
    for (int i = 0; i < 6; i++) Siso[i] = Jm23 * devSf[i];
    double sq[6][6], bl[6][6], Pm[6][6], b1[6][6], b2[6][6], Cbar[6][6];
    square_v(CiV, CiV, sq);
    ball_v(CiV, CiV, bl);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) Pm[i][j] = sq[i][j] - (1.0 / 3.0) * bl[i][j];
    double d2 = -K * pow(J, -5.0 / 3.0) * (pow(J, 1.0 / 3.0) - 2.0);
    ball_v(Siso, CiV, b1);
    ball_v(CiV, Siso, b2);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double ciso = (2.0 / 3.0) * Jm23 * ipSC * Pm[i][j] - (2.0 / 3.0) * (b1[i][j] + b2[i][j]);
            double cvol = J * (pr + J * d2) * bl[i][j] - 2.0 * J * pr * sq[i][j];
            Cbar[i][j] = ciso + cvol;
        }
    push_forward_v(Cbar, F, D);
}

/* model 4, the Q1P0 (mean dilatation) Neo-Hookean, 3D: the isochoric part of model 3 at the
 * Gauss point's J and the pressure of the element's Jbar = V/V0
 * (ModNeoHookeanQ1P0.f90:429-484 stress, :540-613 tangent):
 *   pbar = 3 K (Jbar^1/3 - 1) / Jbar^2/3, sigma = 2 C10 J^-5/3 (b - tr b/3 I) + pbar I,
 *   Cbar = Ciso + J pbar (Cinv (x) Cinv - 2 Cinv (.) Cinv), D = push forward of Cbar.
 * The pressure's derivative enters the element stiffness through Bdiv (elem_ke). */
static void mat_nhq1p0(const double *p, const double F[3][3], double Jbar, double s[6], double D[6][6]) {
    double C10 = p[0], K = p[1], Ft[3][3], b[3][3];
    transpose3(F, Ft);
    matmul3(F, Ft, b);
    double trb = b[0][0] + b[1][1] + b[2][2];
    double J = det3(F);
    double pbar = (3.0 * K * (pow(Jbar, 1.0 / 3.0) - 1.0)) / pow(Jbar, 2.0 / 3.0);
    double S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            S[i][j] = 2.0 * C10 * pow(J, -5.0 / 3.0) * (b[i][j] - (trb / 3.0) * (i == j ? 1.0 : 0.0)) +
                      pbar * (i == j ? 1.0 : 0.0);
    to_voigt(S, s);
    double C[3][3], Ci[3][3], CV[6], CiV[6];
    matmul3(Ft, F, C);
    inv3(C, Ci);
    to_voigt(C, CV);
    to_voigt(Ci, CiV);
    double Sf[6] = {2.0 * C10, 2.0 * C10, 2.0 * C10, 0, 0, 0};
    double ipSC = inner_v(Sf, CV);
    double Siso[6], Jm23 = pow(J, -2.0 / 3.0);
    for (int i = 0; i < 6; i++) Siso[i] = Jm23 * (Sf[i] - (1.0 / 3.0) * ipSC * CiV[i]);
    double sq[6][6], bl[6][6], b1[6][6], b2[6][6], Cbar[6][6];
    square_v(CiV, CiV, sq);
    ball_v(CiV, CiV, bl);
    ball_v(Siso, CiV, b1);
    ball_v(CiV, Siso, b2);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double ciso = (2.0 / 3.0) * Jm23 * ipSC * (sq[i][j] - (1.0 / 3.0) * bl[i][j]) -
                          (2.0 / 3.0) * (b1[i][j] + b2[i][j]);
            Cbar[i][j] = ciso + J * pbar * (bl[i][j] - 2.0 * sq[i][j]);
        }
    push_forward_v(Cbar, F, D);
}

/* SecondDerivativesOfPSI_Jbar_NeoHookeanQ1P0 (ModNeoHookeanQ1P0.f90:625-640) */
static double nhq1p0_d2psi(const double *p, double Jbar) {
    return -(p[1] * (pow(Jbar, 1.0 / 3.0) - 2.0)) / pow(Jbar, 5.0 / 3.0);
}

/* model 5: ModStVenantKirchhoff.f90:643-773 (tangent: upper triangle only) */
static void mat_svk(const double *p, const double F[3][3], double s[6], double D[6][6]) {
    double E = p[0], nu = p[1];
    double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = E / (2.0 * (1.0 + nu));
    double Ft[3][3], C[3][3], Ee[3][3], S[3][3], tmp[3][3], sig[3][3];
    transpose3(F, Ft);
    matmul3(Ft, F, C);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ee[i][j] = 0.5 * (C[i][j] - (i == j ? 1.0 : 0.0));
    double trE = Ee[0][0] + Ee[1][1] + Ee[2][2];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i][j] = lam * trE * (i == j ? 1.0 : 0.0) + 2.0 * mu * Ee[i][j];
    double J = det3(F);
    matmul3(S, Ft, tmp);
    matmul3(F, tmp, sig);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) sig[i][j] = (1.0 / J) * sig[i][j];
    to_voigt(sig, s);
    double b[3][3];
    matmul3(F, Ft, b);
    static const int vi[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
    memset(D, 0, sizeof(double) * 36);
    for (int I = 0; I < 6; I++)
        for (int K = I; K < 6; K++) {
            int i = vi[I][0], j = vi[I][1], k = vi[K][0], l = vi[K][1];
            D[I][K] = (lam / J) * b[i][j] * b[k][l] + (mu / J) * (b[i][k] * b[j][l] + b[i][l] * b[j][k]);
        }
}

/* model 2: ModJ2Plasticity.f90:385-633.
 * state layout (nstate = 17): [0..5] plastic strain, [6] acc. plastic strain,
 * [7..12] S_dev_t, [13] q_t, [14] DeltaGamma, [15] PlasticState flag, [16] spare.
 * state_old: only [0..6] are read (the committed values). */
#define J2_NSTATE 17
static void mat_j2(const double *p, const double F[3][3], const double *old, double s[6],
                   double *nw) {
    double E = p[0], nu = p[1], sy = p[2], H = p[3];
    double strain[6];
    small_strain(F, strain);
    double K = E / (3.0 * (1.0 - 2.0 * nu));
    double G = E / (2.0 * (1.0 + nu));
    double Eet[6], Eedev[6], Sdt[6];
    for (int i = 0; i < 6; i++) Eet[i] = strain[i] - old[i];
    double evol = Eet[0] + Eet[1] + Eet[2];
    for (int i = 0; i < 6; i++) Eedev[i] = Eet[i] - evol * (i < 3 ? 1.0 : 0.0) / 3.0;
    double ep_acc_t = old[6];
    double pt = K * evol;
    for (int i = 0; i < 6; i++) Sdt[i] = G * Eedev[i] * (i < 3 ? 2.0 : 1.0);
    double qt = sqrt((3.0 / 2.0) * (Sdt[0] * Sdt[0] + Sdt[1] * Sdt[1] + Sdt[2] * Sdt[2] +
                                    2.0 * Sdt[3] * Sdt[3] + 2.0 * Sdt[4] * Sdt[4] + 2.0 * Sdt[5] * Sdt[5]));
    double phi = qt - (sy + H * ep_acc_t);
    for (int i = 0; i < 6; i++) nw[7 + i] = Sdt[i];
    nw[13] = qt;
    nw[16] = 0.0;
    if (phi <= 0.0) {
        nw[15] = 0.0;
        nw[6] = ep_acc_t;
        for (int i = 0; i < 6; i++) nw[i] = old[i];
        for (int i = 0; i < 6; i++) s[i] = Sdt[i] + pt * (i < 3 ? 1.0 : 0.0);
        nw[14] = 0.0;
    } else {
        nw[15] = 1.0;
        double dg = phi / (3.0 * G + H);
        nw[14] = dg;
        double Sd[6], Ee[6];
        for (int i = 0; i < 6; i++) Sd[i] = (1.0 - dg * 3.0 * G / qt) * Sdt[i];
        for (int i = 0; i < 6; i++) s[i] = Sd[i] + pt * (i < 3 ? 1.0 : 0.0);
        for (int i = 0; i < 6; i++) Ee[i] = Sd[i] / (2.0 * G) + evol * (i < 3 ? 1.0 : 0.0) / 3.0;
        nw[6] = old[6] + dg;
        for (int i = 0; i < 6; i++) nw[i] = strain[i] - Ee[i];
    }
}

static void mat_j2_tangent(const double *p, const double *st, double D[6][6]) {
    double E = p[0], nu = p[1], H = p[3];
    if (st[15] == 0.0) {
        elastic_D(E, nu, D);
        return;
    }
    double K = E / (3.0 * (1.0 - 2.0 * nu));
    double G = E / (2.0 * (1.0 + nu));
    const double *Sdt = st + 7;
    double qt = st[13], dg = st[14];
    double nrm = sqrt(Sdt[0] * Sdt[0] + Sdt[1] * Sdt[1] + Sdt[2] * Sdt[2] + 2.0 * Sdt[3] * Sdt[3] +
                      2.0 * Sdt[4] * Sdt[4] + 2.0 * Sdt[5] * Sdt[5]);
    double N[6];
    for (int i = 0; i < 6; i++) N[i] = Sdt[i] / nrm;
    double a = 2.0 * G * (1.0 - dg * 3.0 * G / qt);
    double b = 6.0 * G * G * (dg / qt - 1.0 / (3.0 * G + H));
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double Is = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
            double I1 = (i < 3 && j < 3) ? 1.0 : 0.0;
            double Id = Is - I1 / 3.0;
            D[i][j] = a * Id + b * (N[i] * N[j]) + K * I1;
        }
}


/* ------------------------------------------------------------------------------------------
 * model 12: viscoelastic matrix, ConstitutiveModelsLibrary/ModViscoelasticMatrix.f90:276-1028.
 * Second-order tensors as column-major 9-vectors p = i + 3j (Tools/MathTools/ModVoigtNotation.f90:
 * Tensor2ToVoigt :70-86); fourth-order tensors as 9x9 matrices T[p][q] = T_ijkl.
 * props: K_inf, Mu_inf, Lambda_inf, K_e, Mu_e, Ni_v (K_inf and K_e are read but not used by the
 * reference either).
 * state (nstate 20): [0..8] Fv, [9..17] dv, [18] gamma, [19] time of the state; the initial
 * state is Fv = I, dv = 0, gamma = 0, time 0 (the reference never initialises Time_old,
 * ModViscoelasticMatrix.f90:47,329).
 * ---------------------------------------------------------------------------------------- */
#define VISCO_NSTATE 20

static void v9_eye(double I[9]) {
    memset(I, 0, sizeof(double) * 9);
    I[0] = I[4] = I[8] = 1.0;
}

static void v9_trans(const double *A, double *T) { /* TransposeT2Voigt :394-409 */
    T[0] = A[0]; T[1] = A[3]; T[2] = A[6];
    T[3] = A[1]; T[4] = A[4]; T[5] = A[7];
    T[6] = A[2]; T[7] = A[5]; T[8] = A[8];
}

static double v9_det(const double *A) { /* DeterminantT2Voigt :381-390 */
    return A[0] * A[4] * A[8] - A[0] * A[7] * A[5] + A[1] * A[5] * A[6] - A[1] * A[3] * A[8] +
           A[2] * A[3] * A[7] - A[2] * A[4] * A[6];
}

static void v9_inv(const double *A, double *B) { /* InverseT2Voigt :413-430 */
    double d = v9_det(A);
    B[0] = A[4] * A[8] - A[7] * A[5];
    B[1] = A[2] * A[7] - A[1] * A[8];
    B[2] = -A[2] * A[4] + A[1] * A[5];
    B[3] = A[5] * A[6] - A[3] * A[8];
    B[4] = -A[2] * A[6] + A[0] * A[8];
    B[5] = A[2] * A[3] - A[0] * A[5];
    B[6] = A[3] * A[7] - A[4] * A[6];
    B[7] = A[1] * A[6] - A[0] * A[7];
    B[8] = -A[1] * A[3] + A[0] * A[4];
    for (int k = 0; k < 9; k++) B[k] /= d;
}

static void v9_mul(const double *A, const double *B, double *C) { /* SingleContractionT2T2Voigt :434-449 */
    double R[9];
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += A[i + 3 * k] * B[k + 3 * j];
            R[i + 3 * j] = s;
        }
    memcpy(C, R, sizeof(R));
}

static void t4_ball(const double *A, const double *B, double T[9][9]) { /* BallVoigt: A_ij B_kl */
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) T[p][q] = A[p] * B[q];
}

static void t4_square(const double *A, const double *B, double T[9][9]) { /* SquareVoigt: A_ik B_jl */
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++)
            for (int l = 0; l < 3; l++)
                for (int k = 0; k < 3; k++) T[i + 3 * j][k + 3 * l] = A[i + 3 * k] * B[j + 3 * l];
}

static int sw9(int p) { return (p % 3) * 3 + p / 3; } /* (i,j) -> (j,i) */

static void t4_rsym(const double A[9][9], double R[9][9]) { /* RightSymmetrizationT4Voigt */
    double T[9][9];
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) T[p][q] = (q == sw9(q)) ? A[p][q] : 0.5 * A[p][q] + 0.5 * A[p][sw9(q)];
    memcpy(R, T, sizeof(T));
}

static void t4_rlsym(const double A[9][9], double R[9][9]) { /* RigthLeftSymmetrizationT4Voigt */
    double T[9][9];
    t4_rsym(A, T);
    double L[9][9];
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) L[p][q] = (p == sw9(p)) ? T[p][q] : 0.5 * T[p][q] + 0.5 * T[sw9(p)][q];
    memcpy(R, L, sizeof(L));
}

static void t4_mul(const double A[9][9], const double B[9][9], double C[9][9]) { /* DoubleContractionT4T4Voigt */
    double T[9][9];
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) {
            double s = 0.0;
            for (int r = 0; r < 9; r++) s += A[p][r] * B[r][q];
            T[p][q] = s;
        }
    memcpy(C, T, sizeof(T));
}

static void t4_isym(double T[9][9]) { /* IdentityT4SymVoigt: (d_ik d_jl + d_il d_jk)/2 */
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) T[p][q] = 0.5 * ((p == q) + (q == sw9(p)));
}

static const int SYM6[6] = {0, 4, 8, 3, 7, 6}; /* Tensor4VoigtToTensor4VoigtSym :324-367 */

static void t4_to_sym(const double A[9][9], double T[6][6]) {
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) T[a][b] = A[SYM6[a]][SYM6[b]];
}

static void v9_to_sym(const double *A, double *T) { /* Tensor2VoigtToTensor2VoigtSym :147-160 */
    T[0] = A[0]; T[1] = A[4]; T[2] = A[8]; T[3] = A[3]; T[4] = A[7]; T[5] = A[2];
}

static void sym_to_v9(const double *A, double *T) { /* Tensor2VoigtSymToTensor2Voigt :164-181 */
    T[0] = A[0]; T[1] = A[3]; T[2] = A[5];
    T[3] = A[3]; T[4] = A[1]; T[5] = A[4];
    T[6] = A[5]; T[7] = A[4]; T[8] = A[2];
}

/* Gaussian elimination with partial pivoting (the dgetrf/dgetrs solve of
 * SolveLinearSystemLU / SolveMultipleLinearSystemLU, ModMathRoutines.f90:1036-1120). */
static int lu_solve(int n, double *A, int nrhs, double *B) {
    int piv[16];
    for (int k = 0; k < n; k++) {
        int pmax = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i * n + k]) > fabs(A[pmax * n + k])) pmax = i;
        piv[k] = pmax;
        if (A[pmax * n + k] == 0.0) return -1;
        if (pmax != k)
            for (int j = 0; j < n; j++) {
                double t = A[k * n + j]; A[k * n + j] = A[pmax * n + j]; A[pmax * n + j] = t;
            }
        for (int i = k + 1; i < n; i++) {
            A[i * n + k] /= A[k * n + k];
            for (int j = k + 1; j < n; j++) A[i * n + j] -= A[i * n + k] * A[k * n + j];
        }
    }
    for (int r = 0; r < nrhs; r++) {
        double x[16];
        for (int i = 0; i < n; i++) x[i] = B[i * nrhs + r];
        for (int k = 0; k < n; k++)
            if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < i; j++) x[i] -= A[i * n + j] * x[j];
        for (int i = n - 1; i >= 0; i--) {
            for (int j = i + 1; j < n; j++) x[i] -= A[i * n + j] * x[j];
            x[i] /= A[i * n + i];
        }
        for (int i = 0; i < n; i++) B[i * nrhs + r] = x[i];
    }
    return 0;
}

typedef struct {
    double Fvm[9], FvFo[9], Jvm, Cem[9], CemInv[9], Sem[9], Mem[9], Dem[9][9];
} visco_kin;

/* intermediate quantities of one local iteration, ModViscoelasticMatrix.f90:452-471 */
static void visco_kin_eval(const double *F, const double *Fvm_old, const double *dvm, double dt,
                           double Mu_e, visco_kin *k) {
    double I[9], A[9], Ai[9], Foi[9], Fvi[9], Fem[9], FemT[9], sq[9][9];
    v9_eye(I);
    for (int q = 0; q < 9; q++) A[q] = I[q] - dt * dvm[q];
    v9_inv(A, Ai);
    v9_mul(Ai, Fvm_old, k->Fvm);
    k->Jvm = v9_det(k->Fvm);
    v9_inv(Fvm_old, Foi);
    v9_mul(k->Fvm, Foi, k->FvFo);
    v9_inv(k->Fvm, Fvi);
    v9_mul(F, Fvi, Fem);
    v9_trans(Fem, FemT);
    v9_mul(FemT, Fem, k->Cem);
    v9_inv(k->Cem, k->CemInv);
    for (int q = 0; q < 9; q++) k->Sem[q] = Mu_e * (I[q] - k->CemInv[q]);
    t4_square(k->CemInv, k->CemInv, sq);
    t4_rsym(sq, k->Dem);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) k->Dem[p][q] *= 2.0 * Mu_e;
    v9_mul(k->Cem, k->Sem, k->Mem);
}

/* Residual_MATRIX, ModViscoelasticMatrix.f90:573-623 */
static void visco_residual(const visco_kin *k, double gama, const double *dvm, double dt, double Ni_v,
                           double Rm[7]) {
    double FvFoT[9], Dp[9], DpT[9], DL[9];
    v9_trans(k->FvFo, FvFoT);
    v9_mul(FvFoT, k->Mem, Dp);
    for (int q = 0; q < 9; q++) Dp[q] *= -dt;
    v9_trans(Dp, DpT);
    for (int q = 0; q < 9; q++) Dp[q] = 0.5 * (Dp[q] + DpT[q]);
    for (int q = 0; q < 9; q++) Dp[q] += dt * (Ni_v * dvm[q]);
    for (int q = 0; q < 9; q++) DL[q] = Dp[q] + gama * (dt * k->Jvm * k->FvFo[q]);
    v9_to_sym(DL, Rm);
    Rm[6] = k->Jvm - 1.0;
}

/* Jacobian_MATRIX, ModViscoelasticMatrix.f90:627-734; also returns DSem_Ddvm */
static void visco_jacobian(const visco_kin *k, const double *Fvm_old, double gama, double dt,
                           double Ni_v, double DSem[9][9], double Km[7][7]) {
    double I[9], CF[9], B[9][9], DCem[9][9], T[9][9], T2[9][9], DMem[9][9];
    v9_eye(I);
    v9_mul(k->Cem, k->FvFo, CF);
    t4_square(CF, I, B);
    t4_rlsym(B, DCem);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) DCem[p][q] *= -2.0 * dt;
    t4_mul(k->Dem, B, T);
    t4_rsym(T, DSem);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) DSem[p][q] *= -dt;
    t4_square(I, k->Sem, T);
    t4_mul(T, DCem, T);
    t4_square(k->Cem, I, T2);
    t4_mul(T2, DSem, T2);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) DMem[p][q] = T[p][q] + T2[p][q];
    double Foi[9], Aux1[9], Aux2[9], tmp[9], FvFoT[9], D2H[9][9];
    v9_inv(Fvm_old, Foi);
    v9_trans(Foi, Aux1);
    v9_mul(Aux1, k->FvFo, Aux1);
    v9_mul(k->Fvm, k->Mem, tmp);
    v9_trans(tmp, Aux2);
    t4_square(Aux1, Aux2, T);
    v9_trans(k->FvFo, FvFoT);
    t4_square(FvFoT, I, T2);
    t4_mul(T2, DMem, T2);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) D2H[p][q] = pow(dt, 2.0) * T[p][q] + dt * T2[p][q];
    t4_rlsym(D2H, D2H);
    double Is[9][9], D2J[9][9], D2L[9][9];
    t4_isym(Is);
    t4_ball(k->FvFo, k->FvFo, T);
    t4_square(k->FvFo, k->FvFo, T2);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) D2J[p][q] = T[p][q] + T2[p][q];
    t4_rlsym(D2J, D2J);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) {
            double d2pinc = -D2H[p][q] + dt * (Ni_v * Is[p][q]);
            D2L[p][q] = d2pinc + gama * (pow(dt, 2.0) * k->Jvm * D2J[p][q]);
        }
    double D2Lg[9], auxs[6], Ks[6][6];
    for (int q = 0; q < 9; q++) D2Lg[q] = dt * k->Jvm * k->FvFo[q];
    memset(Km, 0, sizeof(double) * 49);
    t4_to_sym(D2L, Ks);
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) Km[a][b] = Ks[a][b];
    v9_to_sym(D2Lg, auxs);
    for (int j = 0; j < 6; j++) Km[j][6] = Km[6][j] = auxs[j];
    /* Km(1:6,1:6) = Km(1:6,1:6) * IdentityT4SymInvVoigtSym; Km(7,1:6) likewise */
    for (int a = 0; a < 7; a++)
        for (int b = 3; b < 6; b++) Km[a][b] *= 2.0;
}

/* UpdateStressAndStateVariables_ViscoelasticMatrix_3D + Local_Newton_Raphson_MATRIX
 * (ModViscoelasticMatrix.f90:276-569).  F9: column-major deformation gradient. */
static int visco_update(const double *p, const double *F9, double time, const double *old,
                        double s[6], double *nw) {
    double Mu_inf = p[1], Lambda_inf = p[2], Mu_e = p[4], Ni_v = p[5];
    const double *Fvm_old = old, *gama_old = old + 18;
    double dt = time - old[19];
    double dvm[9] = {0}, X[7], I[9];
    double gama = *gama_old;
    v9_eye(I);
    v9_to_sym(dvm, X);
    X[6] = gama;
    visco_kin k;
    int ok = 0;
    for (int it = 1; it <= 10; it++) {
        visco_kin_eval(F9, Fvm_old, dvm, dt, Mu_e, &k);
        double Rm[7];
        visco_residual(&k, gama, dvm, dt, Ni_v, Rm);
        double nrm = 0.0;
        for (int q = 0; q < 7; q++) nrm += Rm[q] * Rm[q];
        nrm = sqrt(nrm);
        if (nrm < 1.0e-5) { ok = 1; break; }
        double DSem[9][9], Km[7][7], dX[7];
        visco_jacobian(&k, Fvm_old, gama, dt, Ni_v, DSem, Km);
        for (int q = 0; q < 7; q++) dX[q] = -Rm[q];
        if (lu_solve(7, &Km[0][0], 1, dX)) break;
        for (int q = 0; q < 7; q++) X[q] += dX[q];
        sym_to_v9(X, dvm);
        gama = X[6];
    }
    memcpy(nw, k.Fvm, sizeof(double) * 9);
    memcpy(nw + 9, dvm, sizeof(double) * 9);
    nw[18] = gama;
    nw[19] = time;
    /* stress, ModViscoelasticMatrix.f90:364-400 */
    double Ft[9], C[9], Ci[9], Sinf[9], Fvi[9], FviT[9], aux[9], Sm[9];
    double J = v9_det(F9);
    v9_trans(F9, Ft);
    v9_mul(Ft, F9, C);
    v9_inv(C, Ci);
    for (int q = 0; q < 9; q++) Sinf[q] = Mu_inf * (I[q] - Ci[q]) + Lambda_inf * log(J) * Ci[q];
    v9_inv(k.Fvm, Fvi);
    v9_mul(Fvi, k.Sem, aux);
    v9_trans(Fvi, FviT);
    v9_mul(aux, FviT, aux);
    for (int q = 0; q < 9; q++) Sm[q] = Sinf[q] + aux[q];
    double FS[9], sg[9];
    v9_mul(F9, Sm, FS);
    v9_mul(FS, Ft, sg);
    for (int q = 0; q < 9; q++) sg[q] *= 1.0 / J;
    s[0] = sg[0]; s[1] = sg[4]; s[2] = sg[8]; s[3] = sg[3]; s[4] = sg[7]; s[5] = sg[6];
    return ok ? OR_OK : OR_ERR_LOCAL_NR;
}

/* GetTangentModulus_ViscoelasticMatrix_3D, ModViscoelasticMatrix.f90:745-998 */
static void visco_tangent(const double *p, const double *F9, const double *old, const double *nw,
                          const double F[3][3], double D[6][6]) {
    double Mu_inf = p[1], Lambda_inf = p[2], Mu_e = p[4], Ni_v = p[5];
    const double *Fvm_old = old, *dvm = nw + 9;
    double gama = nw[18];
    double dt = nw[19] - old[19];
    double I[9], Ft[9], C[9], Ci[9];
    v9_eye(I);
    double J = v9_det(F9);
    v9_trans(F9, Ft);
    v9_mul(Ft, F9, C);
    v9_inv(C, Ci);
    visco_kin k;
    visco_kin_eval(F9, Fvm_old, dvm, dt, Mu_e, &k);
    double Fvi[9], FviT[9], FiSq[9][9], FiTSq[9][9], Dinf[9][9], T[9][9], T2[9][9];
    v9_inv(k.Fvm, Fvi);
    v9_trans(Fvi, FviT);
    t4_square(Fvi, Fvi, FiSq);
    t4_square(FviT, FviT, FiTSq);
    t4_ball(Ci, Ci, Dinf);
    t4_square(Ci, Ci, T);
    t4_rsym(T, T);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++)
            Dinf[a][b] = Lambda_inf * Dinf[a][b] + 2.0 * (Mu_inf - Lambda_inf * log(J)) * T[a][b];
    double DSem[9][9], Km[7][7];
    visco_jacobian(&k, Fvm_old, gama, dt, Ni_v, DSem, Km);
    double W[9][9], DSmDC[9][9], Z[9][9], DSmDdv[9][9], DCemDC[9][9], DSemDC[9][9], DMemDC[9][9];
    t4_mul(FiSq, k.Dem, W);
    t4_mul(W, FiTSq, W);
    t4_rlsym(W, T);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) DSmDC[a][b] = 0.5 * (Dinf[a][b] + T[a][b]);
    double Foi[9], FviS[9];
    v9_inv(Fvm_old, Foi);
    v9_mul(Fvi, k.Sem, FviS);
    t4_square(Foi, FviS, Z);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) Z[a][b] *= -2.0 * dt;
    t4_mul(FiSq, DSem, T);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) Z[a][b] += T[a][b];
    t4_rlsym(Z, DSmDdv);
    t4_rsym(FiTSq, DCemDC);
    t4_mul(k.Dem, DCemDC, DSemDC);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) DSemDC[a][b] *= 0.5;
    t4_square(I, k.Sem, T);
    t4_mul(T, DCemDC, DMemDC);
    t4_square(k.Cem, I, T);
    t4_mul(T, DSemDC, T2);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) DMemDC[a][b] += T2[a][b];
    double FvFoT[9], G[9][9];
    v9_trans(k.FvFo, FvFoT);
    t4_square(FvFoT, I, T);
    t4_mul(T, DMemDC, G);
    t4_rlsym(G, G);
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) G[a][b] *= -dt;
    double Gs[6][6], rhs[7 * 6];
    t4_to_sym(G, Gs);
    memset(rhs, 0, sizeof(rhs));
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) rhs[a * 6 + b] = -Gs[a][b];
    lu_solve(7, &Km[0][0], 6, rhs);
    double dd[6][6], Ds[6][6], Dm[6][6], DC[6][6];
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) dd[a][b] = rhs[a * 6 + b] * (a >= 3 ? 2.0 : 1.0);
    t4_to_sym(DSmDdv, Ds);
    t4_to_sym(DSmDC, DC);
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            double acc = 0.0;
            for (int c = 0; c < 6; c++) acc += Ds[a][c] * dd[c][b];
            Dm[a][b] = 2.0 * (DC[a][b] + acc);
        }
    push_forward_v(Dm, F, D);
}


/* GetTangentModulusBase, ConstitutiveLibrary/ModConstitutiveModel.f90:104-164: central
 * differences (h = 1e-8) of the first Piola stress over the 9 components of F, converted to the
 * material modulus (First_Elasticity_Modulus_To_Second_Voigt, Tools/MathTools/ModMathRoutines.f90:315)
 * and pushed forward.  A[(i,J)][(k,L)] = dP_kL / dF_iJ (column-major 9-indices);
 * C_IJKL = sum_ik Finv_Ii Finv_Kk (A[(i,J)][(k,L)] - delta_ik S_JL). */
static void first_piola(const double sig[6], const double F[3][3], double P[3][3]) {
    double sg[3][3] = {{sig[0], sig[3], sig[5]}, {sig[3], sig[1], sig[4]}, {sig[5], sig[4], sig[2]}};
    double Ft[3][3], FiT[3][3];
    double J = det3(F);
    transpose3(F, Ft);
    inv3(Ft, FiT);
    matmul3(sg, FiT, P);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) P[i][j] *= J;
}

static int mat_update(int model, const double *p, const double F[3][3], double time, double jbar,
                      const double *old, double s[6], double *nw);

static void mat_tangent_numerical(int model, const double *p, const double F0[3][3], double time, double jbar,
                                  const double *old, const double sig0[6], double D[6][6]) {
    const double h = 1.0e-8;
    double A[9][9], F[3][3], s[6], nw[MAXSTATE], Pf[3][3], Pb[3][3];
    memcpy(F, F0, sizeof(F));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            F[i][j] = F0[i][j] + h;
            mat_update(model, p, F, time, jbar, old, s, nw);
            first_piola(s, F, Pf);
            F[i][j] = F0[i][j] - h;
            mat_update(model, p, F, time, jbar, old, s, nw);
            first_piola(s, F, Pb);
            for (int l = 0; l < 3; l++)
                for (int k = 0; k < 3; k++) A[i + 3 * j][k + 3 * l] = (Pf[k][l] - Pb[k][l]) / (2.0 * h);
            F[i][j] = F0[i][j];
        }
    /* S = F^-1 (J sigma) F^-T from the stress of the unperturbed update */
    double P0[3][3], V[3][3], S[3][3];
    first_piola(sig0, F0, P0);
    inv3(F0, V);
    matmul3(V, P0, S);
    static const int vi[6] = {0, 1, 2, 0, 1, 0}, vj[6] = {0, 1, 2, 1, 2, 2};
    double C[6][6];
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            int I = vi[a], J = vj[a], K = vi[b], L = vj[b];
            double acc = 0.0;
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++)
                    acc += V[I][i] * V[K][k] * (A[i + 3 * J][k + 3 * L] - (i == k ? S[J][L] : 0.0));
            C[a][b] = acc;
        }
    push_forward_v(C, F0, D);
}

int or_model_nstate(int model) {
    switch (model) {
    case OR_MODEL_J2: return J2_NSTATE;
    case OR_MODEL_VISCO_MATRIX: return VISCO_NSTATE;
    default: return 0;
    }
}

static int mat_nprops(int model) {
    switch (model) {
    case OR_MODEL_GHL: return 2;
    case OR_MODEL_J2: return 4;
    case OR_MODEL_NEOHOOKEAN: return 2;
    case OR_MODEL_NH_Q1P0: return 2;
    case OR_MODEL_SVK: return 2;
    case OR_MODEL_COMP_NEOHOOKEAN: return 2;
    case OR_MODEL_VISCO_MATRIX: return 6;
    default: return -1;
    }
}

/* stress update (UpdateStressAndStateVariables); jbar: the element's V/V0 (model 4) */
static int mat_update(int model, const double *p, const double F[3][3], double time, double jbar,
                      const double *old, double s[6], double *nw) {
    double Dtmp[6][6];
    switch (model) {
    case OR_MODEL_GHL: mat_ghl(p, F, s, Dtmp); return OR_OK;
    case OR_MODEL_NH_Q1P0: mat_nhq1p0(p, F, jbar, s, Dtmp); return OR_OK;
    case OR_MODEL_COMP_NEOHOOKEAN: mat_cnh(p, F, s, Dtmp); return OR_OK;
    case OR_MODEL_NEOHOOKEAN: mat_nh(p, F, s, Dtmp); return OR_OK;
    case OR_MODEL_SVK: mat_svk(p, F, s, Dtmp); return OR_OK;
    case OR_MODEL_J2: mat_j2(p, F, old, s, nw); return OR_OK;
    case OR_MODEL_VISCO_MATRIX: {
        double F9[9];
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) F9[i + 3 * j] = F[i][j];
        return visco_update(p, F9, time, old, s, nw);
    }
    default: return OR_ERR_ARG;
    }
}

/* tangent (GetTangentModulus), uses F and the state left by the last update */
static void mat_tangent(int model, const double *p, const double F[3][3], double jbar, const double *old,
                        const double *nw, double D[6][6]) {
    double s[6];
    switch (model) {
    case OR_MODEL_NH_Q1P0: mat_nhq1p0(p, F, jbar, s, D); return;
    case OR_MODEL_VISCO_MATRIX: {
        double F9[9];
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) F9[i + 3 * j] = F[i][j];
        visco_tangent(p, F9, old, nw, F, D);
        return;
    }
    case OR_MODEL_GHL: mat_ghl(p, F, s, D); return;
    case OR_MODEL_COMP_NEOHOOKEAN: mat_cnh(p, F, s, D); return;
    case OR_MODEL_NEOHOOKEAN: mat_nh(p, F, s, D); return;
    case OR_MODEL_SVK: mat_svk(p, F, s, D); return;
    case OR_MODEL_J2: mat_j2_tangent(p, nw, D); return;
    default: memset(D, 0, sizeof(double) * 36); return;
    }
}

int or_material_point(int model, const double *props, const double Fv[9], double time,
                      const double *state_old, double *sigma6, double *D36, double *state_new) {
    double F[3][3];
    double zero[MAXSTATE] = {0};
    double nw[MAXSTATE] = {0};
    memcpy(F, Fv, sizeof(F));
    if (model == OR_MODEL_VISCO_MATRIX) zero[0] = zero[4] = zero[8] = 1.0; /* Fv = I */
    if (!state_old) state_old = zero;
    const double jbar = det3(F);  /* a single point: the element volume ratio is J */
    int st = mat_update(model, props, F, time, jbar, state_old, sigma6, nw);
    if (st) return st;
    if (D36) mat_tangent(model, props, F, jbar, state_old, nw, (double(*)[6])D36);
    if (state_new) memcpy(state_new, nw, sizeof(double) * or_model_nstate(model));
    return OR_OK;
}

/* ------------------------------------------------------------------------------------------
 * mesh
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    int model;
    double p[16];
    int set;
} or_matdef;

struct or_mesh {
    int nn, ne, nl;
    int numerical_tangent;
    int nthreads;
    double *X, *x;          /* [nn][3] reference / current */
    int *conn;              /* [ne][8] 0-based */
    int *mat;               /* [ne] material id */
    or_matdef mats[MAXMAT]; /* indexed by material id */
    double *F;              /* [ne][8][9] */
    double *sig;            /* [ne][8][6] */
    double *st_old, *st_new;/* [ne][8][MAXSTATE] */
    double *vol, *volx;     /* [ne] */
    double time;
    int *failed;            /* [ne] 0/1 */
    /* CSR upper (1-based) */
    long nnz;
    int *rowmap, *col;
    /* element colouring for the threaded mode */
    int ncolor;
    int *color_ptr, *color_elems;
};

or_mesh *or_mesh_create(int nn, const double *X, int ne, const int *conn1, const int *mat_id) {
    if (nn <= 0 || ne <= 0 || !X || !conn1) return NULL;
    or_mesh *m = (or_mesh *)calloc(1, sizeof(or_mesh));
    m->nn = nn;
    m->ne = ne;
    m->nl = 1;
    m->nthreads = 1;
    m->X = (double *)malloc(sizeof(double) * 3 * nn);
    m->x = (double *)malloc(sizeof(double) * 3 * nn);
    memcpy(m->X, X, sizeof(double) * 3 * nn);
    memcpy(m->x, X, sizeof(double) * 3 * nn);
    m->conn = (int *)malloc(sizeof(int) * 8 * ne);
    m->mat = (int *)malloc(sizeof(int) * ne);
    for (int e = 0; e < ne; e++) {
        for (int a = 0; a < 8; a++) {
            int n = conn1[8 * e + a] - 1;
            if (n < 0 || n >= nn) {
                or_mesh_destroy(m);
                return NULL;
            }
            m->conn[8 * e + a] = n;
        }
        m->mat[e] = mat_id ? mat_id[e] : 1;
        if (m->mat[e] < 0 || m->mat[e] >= MAXMAT) {
            or_mesh_destroy(m);
            return NULL;
        }
    }
    m->F = (double *)calloc((size_t)ne * 8 * 9, sizeof(double));
    for (long i = 0; i < (long)ne * 8; i++) m->F[9 * i] = m->F[9 * i + 4] = m->F[9 * i + 8] = 1.0;
    m->sig = (double *)calloc((size_t)ne * 8 * 6, sizeof(double));
    m->st_old = (double *)calloc((size_t)ne * 8 * MAXSTATE, sizeof(double));
    m->st_new = (double *)calloc((size_t)ne * 8 * MAXSTATE, sizeof(double));
    m->vol = (double *)calloc(ne, sizeof(double));
    m->volx = (double *)calloc(ne, sizeof(double));
    m->failed = (int *)calloc(ne, sizeof(int));
    m->nnz = -1;
    return m;
}

void or_mesh_destroy(or_mesh *m) {
    if (!m) return;
    free(m->X); free(m->x); free(m->conn); free(m->mat); free(m->F); free(m->sig);
    free(m->st_old); free(m->st_new); free(m->vol); free(m->volx); free(m->failed);
    free(m->rowmap); free(m->col); free(m->color_ptr); free(m->color_elems);
    free(m);
}

int or_material_set(or_mesh *m, int mat_id, int model, const double *props, int nprops) {
    if (!m || mat_id < 0 || mat_id >= MAXMAT) return OR_ERR_ARG;
    int np = mat_nprops(model);
    if (np < 0 || nprops < np || nprops > 16) return OR_ERR_ARG;
    m->mats[mat_id].model = model;
    memset(m->mats[mat_id].p, 0, sizeof(m->mats[mat_id].p));
    memcpy(m->mats[mat_id].p, props, sizeof(double) * nprops);
    m->mats[mat_id].set = 1;
    if (model == OR_MODEL_VISCO_MATRIX) /* initial viscous state Fv = I (ModViscoelasticMatrix.f90:120-133) */
        for (int e = 0; e < m->ne; e++)
            if (m->mat[e] == mat_id)
                for (int g = 0; g < 8; g++) {
                    long gi = (long)e * 8 + g;
                    memset(&m->st_old[MAXSTATE * gi], 0, sizeof(double) * MAXSTATE);
                    memset(&m->st_new[MAXSTATE * gi], 0, sizeof(double) * MAXSTATE);
                    m->st_old[MAXSTATE * gi + 0] = m->st_old[MAXSTATE * gi + 4] = m->st_old[MAXSTATE * gi + 8] = 1.0;
                    m->st_new[MAXSTATE * gi + 0] = m->st_new[MAXSTATE * gi + 4] = m->st_new[MAXSTATE * gi + 8] = 1.0;
                }
    return OR_OK;
}

void or_set_analysis(or_mesh *m, int nl_analysis) { m->nl = nl_analysis ? 1 : 0; }
void or_set_tangent_mode(or_mesh *m, int numerical) { m->numerical_tangent = numerical ? 1 : 0; }
int or_ndof(const or_mesh *m) { return 3 * m->nn; }

static void elem_coords(const or_mesh *m, int e, const double *src, double c[8][3]) {
    for (int a = 0; a < 8; a++)
        for (int k = 0; k < 3; k++) c[a][k] = src[3 * m->conn[8 * e + a] + k];
}

/* ElementVolume, FEM/Elements/ElementMethods/ModElement.f90:1313-1427 */
static int elem_volume(const or_mesh *m, int e, double *V, double *VX) {
    double nat[8][3], w[8], x[8][3], X[8][3];
    or_hex8_gauss(nat, w);
    elem_coords(m, e, m->x, x);
    elem_coords(m, e, m->X, X);
    *V = 0.0;
    *VX = 0.0;
    for (int g = 0; g < 8; g++) {
        double dN[8][3], Jc[3][3], Jx[3][3];
        or_hex8_dshape(nat[g], dN);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0.0, sx = 0.0;
                for (int n = 0; n < 8; n++) {
                    s += dN[n][i] * x[n][j];
                    sx += dN[n][i] * X[n][j];
                }
                Jc[i][j] = s;
                Jx[i][j] = sx;
            }
        double dj = det3(Jc), djx = det3(Jx);
        if (dj <= 1.0e-13) return OR_ERR_DETJ;
        *VX += w[g] * djx;
        *V += w[g] * dj;
    }
    return OR_OK;
}

void or_element_volume(or_mesh *m, int e, double *V, double *VX) { elem_volume(m, e, V, VX); }

/* one element of SolveConstitutiveModel, FEM/Subroutines/SubSolveConstitutiveModel.f90:56-136 */
static int solve_elem(or_mesh *m, int e, double time, const double *U) {
    double nat[8][3], w[8], X[8][3], Ue[24], V, VX;
    int st = OR_OK;
    or_hex8_gauss(nat, w);
    int sv = elem_volume(m, e, &V, &VX);
    m->failed[e] = (sv != OR_OK);
    m->vol[e] = V;
    m->volx[e] = VX;
    elem_coords(m, e, m->X, X);
    for (int a = 0; a < 8; a++)
        for (int k = 0; k < 3; k++) Ue[3 * a + k] = U[3 * m->conn[8 * e + a] + k];
    const or_matdef *md = &m->mats[m->mat[e]];
    for (int g = 0; g < 8; g++) {
        double F[3][3];
        long gi = (long)e * 8 + g;
        int sf = or_deformation_gradient(X, nat[g], Ue, F);
        if (sf) st = sf;
        memcpy(&m->F[9 * gi], F, sizeof(F));
        int su = mat_update(md->model, md->p, F, time, V / VX, &m->st_old[MAXSTATE * gi], &m->sig[6 * gi],
                            &m->st_new[MAXSTATE * gi]);
        if (su) st = su;
    }
    if (sv) st = sv;
    return st;
}

static void build_colors(or_mesh *m);

int or_solve_constitutive(or_mesh *m, double time, const double *U) {
    for (int i = 0; i < MAXMAT; i++) (void)i;
    for (int e = 0; e < m->ne; e++)
        if (!m->mats[m->mat[e]].set) return OR_ERR_ARG;
    m->time = time;
    int st = OR_OK;
    if (m->nthreads <= 1) {
        for (int e = 0; e < m->ne; e++) {
            int s = solve_elem(m, e, time, U);
            if (s && (st == OR_OK || s == OR_ERR_DETJ)) st = s;
        }
    } else {
        int err = 0, errdet = 0;
#pragma omp parallel for schedule(static) num_threads(m->nthreads) reduction(| : err, errdet)
        for (int e = 0; e < m->ne; e++) {
            int s = solve_elem(m, e, time, U);
            if (s) err |= 1;
            if (s == OR_ERR_DETJ) errdet |= 1;
        }
        if (errdet) st = OR_ERR_DETJ;
        else if (err) st = OR_ERR_LOCAL_NR;
    }
    return st;
}

int or_failed_elements(const or_mesh *m, int *ids, int cap) {
    int n = 0;
    for (int e = 0; e < m->ne; e++)
        if (m->failed[e]) {
            if (ids && n < cap) ids[n] = e + 1;
            n++;
        }
    return n;
}

/* ElementInternalForce, FEM/Elements/ElementMethods/ModElement.f90:484-562 */
static int elem_fe(or_mesh *m, int e, double Fe[24]) {
    double nat[8][3], w[8], x[8][3], B[6][24], G[9][24], detJ;
    or_hex8_gauss(nat, w);
    elem_coords(m, e, m->x, x);
    memset(Fe, 0, sizeof(double) * 24);
    for (int g = 0; g < 8; g++) {
        if (or_matrix_bg(x, nat[g], B, G, &detJ)) return OR_ERR_DETJ;
        const double *s = &m->sig[6 * ((long)e * 8 + g)];
        double a = w[g] * detJ;
        for (int j = 0; j < 24; j++) {
            double acc = 0.0;
            for (int i = 0; i < 6; i++) acc += B[i][j] * s[i];
            Fe[j] += a * acc;
        }
    }
    return OR_OK;
}

int or_element_internal_force(or_mesh *m, int e, double *Fe) { return elem_fe(m, e, Fe); }

/* InternalForce, FEM/Subroutines/SubInternalForce.f90:49-97 */
int or_internal_force(or_mesh *m, double *Fint) {
    int ndof = 3 * m->nn, st = OR_OK;
    memset(Fint, 0, sizeof(double) * ndof);
    if (m->nthreads <= 1) {
        for (int e = 0; e < m->ne; e++) {
            double Fe[24];
            if (elem_fe(m, e, Fe)) {
                m->failed[e] = 1;
                st = OR_ERR_DETJ;
            }
            for (int a = 0; a < 8; a++)
                for (int k = 0; k < 3; k++) Fint[3 * m->conn[8 * e + a] + k] += Fe[3 * a + k];
        }
    } else {
        build_colors(m);
        int err = 0;
        for (int c = 0; c < m->ncolor; c++) {
#pragma omp parallel for schedule(static) num_threads(m->nthreads) reduction(| : err)
            for (int i = m->color_ptr[c]; i < m->color_ptr[c + 1]; i++) {
                int e = m->color_elems[i];
                double Fe[24];
                if (elem_fe(m, e, Fe)) {
                    m->failed[e] = 1;
                    err |= 1;
                }
                for (int a = 0; a < 8; a++)
                    for (int k = 0; k < 3; k++) Fint[3 * m->conn[8 * e + a] + k] += Fe[3 * a + k];
            }
        }
        if (err) st = OR_ERR_DETJ;
    }
    return st;
}

/* ElementStiffnessMatrix, FEM/Elements/ElementMethods/ModElement.f90:200-348.
 * nl = 0: Ke += (B^T D B) w detJ with the full D (:293-308);
 * nl = 1: DB = D_sym B, SG = S_sym G (dsymm 'L','U'), Ke += w detJ (B^T DB + G^T SG) (:310-348) */
static int elem_ke(or_mesh *m, int e, double Ke[24][24]) {
    double nat[8][3], w[8], x[8][3], B[6][24], G[9][24], detJ;
    or_hex8_gauss(nat, w);
    elem_coords(m, e, m->x, x);
    memset(Ke, 0, sizeof(double) * 576);
    const or_matdef *md = &m->mats[m->mat[e]];
    const double jbar = m->vol[e] / m->volx[e];
    const int q1p0 = md->model == OR_MODEL_NH_Q1P0 && m->nl;
    double Bdiv[24] = {0};
    for (int g = 0; g < 8; g++) {
        long gi = (long)e * 8 + g;
        double D[6][6], F[3][3];
        memcpy(F, &m->F[9 * gi], sizeof(F));
        if (m->numerical_tangent)
            mat_tangent_numerical(md->model, md->p, F, m->time, jbar, &m->st_old[MAXSTATE * gi], &m->sig[6 * gi], D);
        else
            mat_tangent(md->model, md->p, F, jbar, &m->st_old[MAXSTATE * gi], &m->st_new[MAXSTATE * gi], D);
        if (or_matrix_bg(x, nat[g], B, G, &detJ)) return OR_ERR_DETJ;
        double a = w[g] * detJ;
        if (!m->nl) {
            double BtD[24][6];
            for (int i = 0; i < 24; i++)
                for (int j = 0; j < 6; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 6; k++) s += B[k][i] * D[k][j];
                    BtD[i][j] = s;
                }
            for (int i = 0; i < 24; i++)
                for (int j = 0; j < 24; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 6; k++) s += BtD[i][k] * B[k][j];
                    Ke[i][j] += s * a;
                }
            continue;
        }
        const double *sg = &m->sig[6 * gi];
        double Ds[6][6], S[9][9], DB[6][24], SG[9][24];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) Ds[i][j] = (i <= j) ? D[i][j] : D[j][i];
        /* GetMatrixOfStressesBase, ConstitutiveLibrary/ModConstitutiveModel.f90:292-310 */
        memset(S, 0, sizeof(S));
        for (int b = 0; b < 3; b++) {
            int o = 3 * b;
            S[o + 0][o + 0] = sg[0];
            S[o + 1][o + 1] = sg[1];
            S[o + 2][o + 2] = sg[2];
            S[o + 0][o + 1] = S[o + 1][o + 0] = sg[3];
            S[o + 1][o + 2] = S[o + 2][o + 1] = sg[4];
            S[o + 0][o + 2] = S[o + 2][o + 0] = sg[5];
        }
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += Ds[i][k] * B[k][j];
                DB[i][j] = s;
            }
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0.0;
                for (int k = 0; k < 9; k++) s += S[i][k] * G[k][j];
                SG[i][j] = s;
            }
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += B[k][i] * DB[k][j];
                Ke[i][j] += a * s;
            }
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0.0;
                for (int k = 0; k < 9; k++) s += G[k][i] * SG[k][j];
                Ke[i][j] += a * s;
            }
        /* mean dilatation: Bdiv = integral of div N over the current element, the rows
         * (1,1), (2,2), (3,3) of G (ModElement.f90:429-461) */
        if (q1p0)
            for (int j = 0; j < 24; j++) Bdiv[j] += (G[0][j] + G[4][j] + G[8][j]) * a;
    }
    /* Ke += d2Psi/dJbar2 / V0 Bdiv^T Bdiv (ModElement.f90:464-467) */
    if (q1p0) {
        const double c = nhq1p0_d2psi(md->p, jbar) / m->volx[e];
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) Ke[i][j] += c * Bdiv[i] * Bdiv[j];
    }
    return OR_OK;
}

int or_element_stiffness(or_mesh *m, int e, double *Ke) { return elem_ke(m, e, (double(*)[24])Ke); }

/* ------------------------------------------------------------------------------------------
 * CSR pattern: each element's DOF clique is inserted into per-row sorted column lists
 * (SparseMatrixSetArray + AddItems, ModSparseMatrixRoutines.f90:254-288 /
 * ModSparseVectorRoutines.f90:222-312), then the upper triangle (col >= row) is kept with
 * 1-based indices and RowMap(n+1) = nz+1 (ConvertToCoordinateFormatUpperTriangular :573-609).
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    int n, cap;
    int *c;
} row_list;

static void row_insert(row_list *r, int c) {
    int lo = 0, hi = r->n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (r->c[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    if (lo < r->n && r->c[lo] == c) return;
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 32;
        r->c = (int *)realloc(r->c, sizeof(int) * r->cap);
    }
    memmove(r->c + lo + 1, r->c + lo, sizeof(int) * (r->n - lo));
    r->c[lo] = c;
    r->n++;
}

static int build_pattern(or_mesh *m) {
    if (m->nnz >= 0) return OR_OK;
    int ndof = 3 * m->nn;
    row_list *rows = (row_list *)calloc(ndof, sizeof(row_list));
    for (int e = 0; e < m->ne; e++) {
        int gm[24];
        for (int a = 0; a < 8; a++)
            for (int k = 0; k < 3; k++) gm[3 * a + k] = 3 * m->conn[8 * e + a] + k + 1;
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) row_insert(&rows[gm[i] - 1], gm[j]);
    }
    long nzfull = 0;
    for (int r = 0; r < ndof; r++) nzfull += rows[r].n;
    long nz = (nzfull - ndof) / 2 + ndof;
    m->rowmap = (int *)malloc(sizeof(int) * (ndof + 1));
    m->col = (int *)malloc(sizeof(int) * (nz > 0 ? nz : 1));
    long n = 0;
    int bad = 0;
    for (int r = 0; r < ndof; r++) {
        m->rowmap[r] = (int)(n + 1);
        if (rows[r].n == 0) bad = 1; /* orphan DOF: reference formula for nz would be wrong */
        for (int i = 0; i < rows[r].n; i++)
            if (rows[r].c[i] >= r + 1) {
                if (n < nz) m->col[n] = rows[r].c[i];
                n++;
            }
        free(rows[r].c);
    }
    free(rows);
    m->rowmap[ndof] = (int)(nz + 1);
    m->nnz = nz;
    if (bad || n != nz) return OR_ERR_ARG;
    return OR_OK;
}

long or_csr_nnz(or_mesh *m) {
    if (build_pattern(m)) return -1;
    return m->nnz;
}

int or_csr_pattern(or_mesh *m, int *rowmap, int *col) {
    int st = build_pattern(m);
    if (st) return st;
    memcpy(rowmap, m->rowmap, sizeof(int) * (3 * m->nn + 1));
    memcpy(col, m->col, sizeof(int) * m->nnz);
    return OR_OK;
}

/* AssembleGlobalMatrixUpperTriangular, FEM/Subroutines/SubAssembleGlobalMatrixUpperTriangular.f90:13-69 */
static int assemble_upper(const or_mesh *m, const int gm[24], double Ke[24][24], double *Kval) {
    for (int i = 0; i < 24; i++) {
        int iG = gm[i];
        for (int j = 0; j < 24; j++) {
            int jG = gm[j];
            if (jG >= iG) {
                int found = 0;
                for (int k = m->rowmap[iG - 1]; k <= m->rowmap[iG] - 1; k++)
                    if (m->col[k - 1] == jG) {
                        found = 1;
                        Kval[k - 1] += Ke[i][j];
                    }
                if (!found) return OR_ERR_SLOT;
            }
        }
    }
    return OR_OK;
}

/* greedy element colouring (elements of one colour share no node) for the threaded mode */
static void build_colors(or_mesh *m) {
    if (m->color_ptr) return;
    int ne = m->ne, nn = m->nn;
    int *col = (int *)malloc(sizeof(int) * ne);
    uint64_t *used = (uint64_t *)calloc(nn, sizeof(uint64_t));
    int nc = 0;
    for (int e = 0; e < ne; e++) {
        uint64_t u = 0;
        for (int a = 0; a < 8; a++) u |= used[m->conn[8 * e + a]];
        int c = 0;
        while (c < 64 && (u >> c) & 1) c++;
        col[e] = c;
        if (c + 1 > nc) nc = c + 1;
        for (int a = 0; a < 8; a++) used[m->conn[8 * e + a]] |= (1ull << c);
    }
    m->ncolor = nc;
    m->color_ptr = (int *)calloc(nc + 1, sizeof(int));
    m->color_elems = (int *)malloc(sizeof(int) * ne);
    for (int e = 0; e < ne; e++) m->color_ptr[col[e] + 1]++;
    for (int c = 0; c < nc; c++) m->color_ptr[c + 1] += m->color_ptr[c];
    int *fill = (int *)malloc(sizeof(int) * nc);
    for (int c = 0; c < nc; c++) fill[c] = m->color_ptr[c];
    for (int e = 0; e < ne; e++) m->color_elems[fill[col[e]]++] = e;
    free(fill);
    free(used);
    free(col);
}

void or_set_threads(or_mesh *m, int nthreads) { m->nthreads = nthreads < 1 ? 1 : nthreads; }

/* TangentStiffnessMatrix FEM branch, FEM/Subroutines/SubTangentStiffnessMatrix.f90:63, 294-310 */
int or_tangent(or_mesh *m, double *Kval) {
    int st = build_pattern(m);
    if (st) return st;
    memset(Kval, 0, sizeof(double) * m->nnz);
    if (m->nthreads <= 1) {
        for (int e = 0; e < m->ne; e++) {
            double Ke[24][24];
            int gm[24];
            elem_ke(m, e, Ke);
            for (int a = 0; a < 8; a++)
                for (int k = 0; k < 3; k++) gm[3 * a + k] = 3 * m->conn[8 * e + a] + k + 1;
            int s = assemble_upper(m, gm, Ke, Kval);
            if (s) return s;
        }
    } else {
        build_colors(m);
        int err = 0;
        for (int c = 0; c < m->ncolor; c++) {
#pragma omp parallel for schedule(static) num_threads(m->nthreads) reduction(| : err)
            for (int i = m->color_ptr[c]; i < m->color_ptr[c + 1]; i++) {
                int e = m->color_elems[i];
                double Ke[24][24];
                int gm[24];
                elem_ke(m, e, Ke);
                for (int a = 0; a < 8; a++)
                    for (int k = 0; k < 3; k++) gm[3 * a + k] = 3 * m->conn[8 * e + a] + k + 1;
                if (assemble_upper(m, gm, Ke, Kval)) err |= 1;
            }
        }
        if (err) return OR_ERR_SLOT;
    }
    return OR_OK;
}

/* UpdateMeshCoordinates, FEM/Subroutines/SubUpdateMeshCoordinates.f90:15-19 */
void or_update_mesh(or_mesh *m, const double *U) {
    for (long i = 0; i < 3L * m->nn; i++) m->x[i] = m->X[i] + U[i];
}

/* SwitchConvergedState loop, FEM/ModFEMAnalysis.f90:1349-1353 (J2: ModJ2Plasticity.f90:625-633) */
void or_switch_converged_state(or_mesh *m) {
    for (int e = 0; e < m->ne; e++) {
        int model = m->mats[m->mat[e]].model;
        if (model == OR_MODEL_J2) {
            for (int g = 0; g < 8; g++) {
                long gi = (long)e * 8 + g;
                memcpy(&m->st_old[MAXSTATE * gi], &m->st_new[MAXSTATE * gi], sizeof(double) * 7);
            }
        } else if (model == OR_MODEL_VISCO_MATRIX) { /* ModViscoelasticMatrix.f90:1005-1028 */
            for (int g = 0; g < 8; g++) {
                long gi = (long)e * 8 + g;
                memcpy(&m->st_old[MAXSTATE * gi], &m->st_new[MAXSTATE * gi], sizeof(double) * VISCO_NSTATE);
            }
        }
    }
}

void or_get_gp_stress(const or_mesh *m, double *out) {
    memcpy(out, m->sig, sizeof(double) * 6 * 8 * (size_t)m->ne);
}

void or_get_gp_F(const or_mesh *m, double *out) {
    memcpy(out, m->F, sizeof(double) * 9 * 8 * (size_t)m->ne);
}

void or_get_state(const or_mesh *m, int which, double *out) {
    const double *src = which ? m->st_new : m->st_old;
    for (long gi = 0; gi < (long)m->ne * 8; gi++)
        memcpy(&out[OR_STATE_STRIDE * gi], &src[MAXSTATE * gi], sizeof(double) * OR_STATE_STRIDE);
}

void or_set_state(or_mesh *m, int which, const double *in) {
    double *dst = which ? m->st_new : m->st_old;
    for (long gi = 0; gi < (long)m->ne * 8; gi++)
        memcpy(&dst[MAXSTATE * gi], &in[OR_STATE_STRIDE * gi], sizeof(double) * OR_STATE_STRIDE);
}

/* HomogenizeStress / HomogenizeDeformationGradient, MultiScale/ModMultiscaleAnalysis.f90:140-441.
 * P = J sigma F^-T (Tools/ModContinuumMechanics.f90:125-129); 9-vectors column-major
 * (Tools/MathTools/ModVoigtNotation.f90:70-86). Uses the element VolumeX stored by the last
 * or_solve_constitutive. */
int or_homogenize(or_mesh *m, double P[9], double Fh[9]) {
    double nat[8][3], w[8];
    or_hex8_gauss(nat, w);
    double tv = 0.0;
    for (int e = 0; e < m->ne; e++) tv += m->volx[e];
    memset(P, 0, sizeof(double) * 9);
    memset(Fh, 0, sizeof(double) * 9);
    for (int e = 0; e < m->ne; e++) {
        double X[8][3];
        elem_coords(m, e, m->X, X);
        for (int g = 0; g < 8; g++) {
            double dN[8][3], Jx[3][3];
            or_hex8_dshape(nat[g], dN);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0.0;
                    for (int n = 0; n < 8; n++) s += dN[n][i] * X[n][j];
                    Jx[i][j] = s;
                }
            double djx = det3(Jx);
            long gi = (long)e * 8 + g;
            double F[3][3], Fi[3][3], FiT[3][3], sg[3][3], Pt[3][3];
            memcpy(F, &m->F[9 * gi], sizeof(F));
            const double *s6 = &m->sig[6 * gi];
            sg[0][0] = s6[0]; sg[1][1] = s6[1]; sg[2][2] = s6[2];
            sg[0][1] = sg[1][0] = s6[3];
            sg[1][2] = sg[2][1] = s6[4];
            sg[0][2] = sg[2][0] = s6[5];
            double J = det3(F);
            double Ft[3][3];
            transpose3(F, Ft);
            inv3(Ft, FiT);
            (void)Fi;
            matmul3(sg, FiT, Pt);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) Pt[i][j] *= J;
            double a = w[g] * djx;
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < 3; i++) {
                    P[3 * j + i] += (Pt[i][j] * a) / tv;
                    Fh[3 * j + i] += (F[i][j] * a) / tv;
                }
        }
    }
    return OR_OK;
}
