// Device-side hex8 kinematics and constitutive updates for gfx950 (FP64).
//
// Conventions follow the reference (file:line relative to SourceCode/):
//   hex8 node order and Gauss rule      FEM/Elements/ElementLibrary/ModElementHexa8.f90:150-329
//   J = dN/dxi^T x, gradN = J^-1 dN      FEM/Elements/ElementMethods/ModElement.f90:848-866
//   F = I + sum_a u_a (x) grad_X N_a     FEM/Elements/ElementMethods/ModElement.f90:1210-1262
//   Voigt order 11,22,33,12,23,13; tangents are used through their upper triangle
//   (dsymm 'L','U' in ElementStiffnessMatrix, ModElement.f90:336-339).
// Every material writes sigma (6) and the upper triangle of its spatial tangent D (21 values,
// row-major packed: (0,0..5),(1,1..5),...,(5,5)).
#pragma once
#include <hip/hip_runtime.h>

namespace ceos {

constexpr double kDetTol = 1.0e-13;  // ModElement.f90:870 (detJ <= 1e-13 is an error)

__device__ __forceinline__ int up_idx(int i, int j) {  // i <= j
    return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

__device__ __forceinline__ double det3(const double M[3][3]) {
    return M[0][0] * M[1][1] * M[2][2] + M[0][1] * M[1][2] * M[2][0] + M[0][2] * M[1][0] * M[2][1] -
           M[0][2] * M[1][1] * M[2][0] - M[0][1] * M[1][0] * M[2][2] - M[0][0] * M[1][2] * M[2][1];
}

// adjugate / det, Tools/MathTools/ModMathRoutines.f90:33-62
__device__ __forceinline__ void inv3(const double A[3][3], double d, double B[3][3]) {
    const double r = 1.0 / d;
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * r;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * r;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * r;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * r;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * r;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * r;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * r;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * r;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * r;
}

// hex8 node sign of node a, component k
__device__ __forceinline__ double hex_sign(int a, int k) {
    // (---,+--,++-,-+-,--+,+-+,+++,-++)
    const int sx = ((a + 1) >> 1) & 1;  // 0,1,1,0,0,1,1,0
    const int sy = (a >> 1) & 1;        // 0,0,1,1,0,0,1,1
    const int sz = (a >> 2) & 1;
    const int s = (k == 0) ? sx : (k == 1 ? sy : sz);
    return s ? 1.0 : -1.0;
}

// dN_a/dxi_k at Gauss point g (natural coordinates = node signs / sqrt(3))
__device__ __forceinline__ void hex8_dshape_gp(int g, double dN[8][3]) {
    const double x = 0.57735026918962576451;  // 1/sqrt(3)
    const double xi = hex_sign(g, 0) * x, eta = hex_sign(g, 1) * x, zeta = hex_sign(g, 2) * x;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double s0 = hex_sign(a, 0), s1 = hex_sign(a, 1), s2 = hex_sign(a, 2);
        dN[a][0] = s0 * (1.0 + s1 * eta) * (1.0 + s2 * zeta) * 0.125;
        dN[a][1] = s1 * (1.0 + s0 * xi) * (1.0 + s2 * zeta) * 0.125;
        dN[a][2] = s2 * (1.0 + s0 * xi) * (1.0 + s1 * eta) * 0.125;
    }
}

// dN_a/dxi of one node a at Gauss point g: the entries hex8_dshape_gp forms, by the same expressions
__device__ __forceinline__ void hex8_dshape_node(int a, int g, double d[3]) {
    const double x = 0.57735026918962576451;  // 1/sqrt(3)
    const double xi = hex_sign(g, 0) * x, eta = hex_sign(g, 1) * x, zeta = hex_sign(g, 2) * x;
    const double s0 = hex_sign(a, 0), s1 = hex_sign(a, 1), s2 = hex_sign(a, 2);
    d[0] = s0 * (1.0 + s1 * eta) * (1.0 + s2 * zeta) * 0.125;
    d[1] = s1 * (1.0 + s0 * xi) * (1.0 + s2 * zeta) * 0.125;
    d[2] = s2 * (1.0 + s0 * xi) * (1.0 + s1 * eta) * 0.125;
}

// gradient of one shape function from J^-1 and its natural derivatives: (J^-1 dN_a)_i
__device__ __forceinline__ void grad_from_jinv(const double Ji[9], const double d[3], double g[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] = Ji[3 * i] * d[0] + Ji[3 * i + 1] * d[1] + Ji[3 * i + 2] * d[2];
}

// Jacobian from nodal coordinates c[a][j]; returns det; fills grad[a][i] = (J^-1 dN_a)_i and,
// when Jout is given, J^-1 row-major
__device__ __forceinline__ double hex8_grad(const double dN[8][3], const double c[8][3],
                                            double grad[8][3], double* Jout = nullptr) {
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) s += dN[a][i] * c[a][j];
            J[i][j] = s;
        }
    const double d = det3(J);
    double Ji[3][3];
    inv3(J, d, Ji);
    double Jf[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Jf[3 * i + j] = Ji[i][j];
#pragma unroll
    for (int a = 0; a < 8; ++a) grad_from_jinv(Jf, dN[a], grad[a]);
    if (Jout) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Jout[k] = Jf[k];
    }
    return d;
}

// ------------------------------------------------------------------------------------------
// materials. props layout per model (as set through ceos_material_set):
//   1 GHL: E, nu   2 J2: E, nu, sigma_y, H   3 NH: C10, K   5 SVK: E, nu   7 CNH: mu, lambda
//   99 EXT: none (stress and tangent come per Gauss point from the caller, see mat_ext)
// ------------------------------------------------------------------------------------------
enum : int { kGHL = 1, kJ2 = 2, kNH = 3, kNHQ1 = 4, kSVK = 5, kCNH = 7, kVISCO = 12, kEXT = 99 };
constexpr int kExtPerGp = 27;  // external material point: S (6) + dS/dE upper (21)

__device__ __forceinline__ void elastic_up(double E, double nu, double D[21]) {
    // ModGeneralizedHookesLaw .f90:526-560, ModJ2Plasticity.f90:587-597
    const double cte = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
#pragma unroll
    for (int i = 0; i < 21; ++i) D[i] = 0.0;
    const double a = cte * (1.0 - nu), b = cte * nu, c = cte * (0.5 - nu);
    D[up_idx(0, 0)] = a; D[up_idx(0, 1)] = b; D[up_idx(0, 2)] = b;
    D[up_idx(1, 1)] = a; D[up_idx(1, 2)] = b;
    D[up_idx(2, 2)] = a;
    D[up_idx(3, 3)] = c; D[up_idx(4, 4)] = c; D[up_idx(5, 5)] = c;
}

__device__ __forceinline__ void small_strain(const double F[3][3], double e[6]) {
    e[0] = F[0][0] - 1.0;
    e[1] = F[1][1] - 1.0;
    e[2] = F[2][2] - 1.0;
    e[3] = F[0][1] + F[1][0];
    e[4] = F[1][2] + F[2][1];
    e[5] = F[0][2] + F[2][0];
}

// model 1, ModGeneralizedHookesLaw .f90:475-520: sigma = D eps
__device__ __forceinline__ void mat_ghl(const double* p, const double F[3][3], double s[6], double D[21]) {
    double e[6];
    small_strain(F, e);
    elastic_up(p[0], p[1], D);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc += D[i <= j ? up_idx(i, j) : up_idx(j, i)] * e[j];
        s[i] = acc;
    }
}

// model 7, ModCompressibleNeoHookean.f90:422-546
__device__ __forceinline__ void mat_cnh(const double* p, const double F[3][3], double s[6], double D[21]) {
    const double mu = p[0], lam = p[1];
    const double J = det3(F);
    const double lnJ = log(J);
    const double rJ = 1.0 / J;
    const double b00 = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2];
    const double b11 = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2];
    const double b22 = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2];
    const double b01 = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
    const double b12 = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
    const double b02 = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
    const double mJ = mu * rJ, vol = lam * rJ * lnJ;
    s[0] = mJ * (b00 - 1.0) + vol;
    s[1] = mJ * (b11 - 1.0) + vol;
    s[2] = mJ * (b22 - 1.0) + vol;
    s[3] = mJ * b01;
    s[4] = mJ * b12;
    s[5] = mJ * b02;
    const double a = lam * rJ, c = 2.0 * rJ * (mu - lam * lnJ);
#pragma unroll
    for (int i = 0; i < 21; ++i) D[i] = 0.0;
    D[up_idx(0, 0)] = a + c; D[up_idx(0, 1)] = a; D[up_idx(0, 2)] = a;
    D[up_idx(1, 1)] = a + c; D[up_idx(1, 2)] = a;
    D[up_idx(2, 2)] = a + c;
    D[up_idx(3, 3)] = 0.5 * c; D[up_idx(4, 4)] = 0.5 * c; D[up_idx(5, 5)] = 0.5 * c;
}

// Voigt "square" product T(A,A) of one symmetric tensor with itself, upper triangle only,
// Tools/MathTools/ModMathRoutines.f90:581-625
__device__ __forceinline__ void square_self_up(const double A[6], double T[21]) {
    T[up_idx(0, 0)] = A[0] * A[0];
    T[up_idx(0, 1)] = A[3] * A[3];
    T[up_idx(0, 2)] = A[5] * A[5];
    T[up_idx(0, 3)] = A[0] * A[3];
    T[up_idx(0, 4)] = A[3] * A[5];
    T[up_idx(0, 5)] = A[0] * A[5];
    T[up_idx(1, 1)] = A[1] * A[1];
    T[up_idx(1, 2)] = A[4] * A[4];
    T[up_idx(1, 3)] = A[3] * A[1];
    T[up_idx(1, 4)] = A[1] * A[4];
    T[up_idx(1, 5)] = A[3] * A[4];
    T[up_idx(2, 2)] = A[2] * A[2];
    T[up_idx(2, 3)] = A[5] * A[4];
    T[up_idx(2, 4)] = A[4] * A[2];
    T[up_idx(2, 5)] = A[5] * A[2];
    T[up_idx(3, 3)] = 0.5 * (A[0] * A[1] + A[3] * A[3]);
    T[up_idx(3, 4)] = 0.5 * (A[3] * A[4] + A[5] * A[1]);
    T[up_idx(3, 5)] = 0.5 * (A[0] * A[4] + A[5] * A[3]);
    T[up_idx(4, 4)] = 0.5 * (A[1] * A[2] + A[4] * A[4]);
    T[up_idx(4, 5)] = 0.5 * (A[3] * A[2] + A[5] * A[4]);
    T[up_idx(5, 5)] = 0.5 * (A[0] * A[2] + A[5] * A[5]);
}

// T = (1/J) A W A^T (W symmetric, upper), A built from F, Tools/MathTools/ModMathRoutines.f90:254-311
__device__ __forceinline__ void push_forward_up(const double W[21], const double F[3][3], double J,
                                                double T[21]) {
    // Ft(i,j) = F(j,i)
#define FT(i, j) F[(j)-1][(i)-1]
    double A[6][6];
    A[0][0] = FT(1, 1) * FT(1, 1); A[0][1] = FT(2, 1) * FT(2, 1); A[0][2] = FT(3, 1) * FT(3, 1);
    A[0][3] = 2.0 * FT(1, 1) * FT(2, 1); A[0][4] = 2.0 * FT(2, 1) * FT(3, 1); A[0][5] = 2.0 * FT(1, 1) * FT(3, 1);
    A[1][0] = FT(1, 2) * FT(1, 2); A[1][1] = FT(2, 2) * FT(2, 2); A[1][2] = FT(3, 2) * FT(3, 2);
    A[1][3] = 2.0 * FT(1, 2) * FT(2, 2); A[1][4] = 2.0 * FT(2, 2) * FT(3, 2); A[1][5] = 2.0 * FT(1, 2) * FT(3, 2);
    A[2][0] = FT(1, 3) * FT(1, 3); A[2][1] = FT(2, 3) * FT(2, 3); A[2][2] = FT(3, 3) * FT(3, 3);
    A[2][3] = 2.0 * FT(1, 3) * FT(2, 3); A[2][4] = 2.0 * FT(2, 3) * FT(3, 3); A[2][5] = 2.0 * FT(1, 3) * FT(3, 3);
    A[3][0] = FT(1, 1) * FT(1, 2); A[3][1] = FT(2, 1) * FT(2, 2); A[3][2] = FT(3, 1) * FT(3, 2);
    A[3][3] = FT(1, 1) * FT(2, 2) + FT(2, 1) * FT(1, 2);
    A[3][4] = FT(2, 1) * FT(3, 2) + FT(3, 1) * FT(2, 2);
    A[3][5] = FT(3, 1) * FT(1, 2) + FT(1, 1) * FT(3, 2);
    A[4][0] = FT(1, 2) * FT(1, 3); A[4][1] = FT(2, 2) * FT(2, 3); A[4][2] = FT(3, 2) * FT(3, 3);
    A[4][3] = FT(1, 2) * FT(2, 3) + FT(2, 2) * FT(1, 3);
    A[4][4] = FT(2, 2) * FT(3, 3) + FT(3, 2) * FT(2, 3);
    A[4][5] = FT(3, 2) * FT(1, 3) + FT(1, 2) * FT(3, 3);
    A[5][0] = FT(1, 1) * FT(1, 3); A[5][1] = FT(2, 1) * FT(2, 3); A[5][2] = FT(3, 1) * FT(3, 3);
    A[5][3] = FT(1, 3) * FT(2, 1) + FT(2, 3) * FT(1, 1);
    A[5][4] = FT(2, 3) * FT(3, 1) + FT(3, 3) * FT(2, 1);
    A[5][5] = FT(3, 3) * FT(1, 1) + FT(1, 3) * FT(3, 1);
#undef FT
    double AW[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += A[i][k] * W[k <= j ? up_idx(k, j) : up_idx(j, k)];
            AW[i][j] = s;
        }
    const double rJ = 1.0 / J;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += AW[i][k] * A[j][k];
            T[up_idx(i, j)] = s * rJ;
        }
}

// External material point (an FE^2 macro Gauss point whose response comes from an RVE solve or a
// surrogate): q = second Piola-Kirchhoff stress S (Voigt 11,22,33,12,23,13) followed by the
// upper triangle of its material tangent dS/dE (engineering shear).  The Cauchy stress
// sigma = F S F^T / J and the spatial tangent follow by the push forward the hyperelastic models
// use (ModNeoHookean.f90 calls Push_Forward_Voigt the same way).
__device__ __forceinline__ void mat_ext(const double* q, const double F[3][3], double s[6], double D[21]) {
    const double J = det3(F);
    const double S[3][3] = {{q[0], q[3], q[5]}, {q[3], q[1], q[4]}, {q[5], q[4], q[2]}};
    double FS[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FS[i][j] = F[i][0] * S[0][j] + F[i][1] * S[1][j] + F[i][2] * S[2][j];
    const double rJ = 1.0 / J;
    const int vi[6] = {0, 1, 2, 0, 1, 0}, vj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int a = 0; a < 6; ++a)
        s[a] = (FS[vi[a]][0] * F[vj[a]][0] + FS[vi[a]][1] * F[vj[a]][1] + FS[vi[a]][2] * F[vj[a]][2]) * rJ;
    push_forward_up(q + 6, F, J, D);
}

// model 3, ModNeoHookean.f90:409-615
// Neo-Hookean stress and tangent from the isochoric part at F and a given volumetric part:
// sigma = 2 C10 J^-5/3 (b - tr b/3 I) + pr I, Cbar = Ciso + cv1 Cinv (x) Cinv - cv2 Cinv (.) Cinv,
// D = push forward of Cbar.
__device__ __forceinline__ void nh_core(double C10, const double F[3][3], double J, double Jm23, double Jm53,
                                        double pr, double cv1, double cv2, double s[6], double D[21]) {
    const double b00 = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2];
    const double b11 = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2];
    const double b22 = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2];
    const double b01 = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
    const double b12 = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
    const double b02 = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
    const double trb = b00 + b11 + b22;
    const double g = 2.0 * C10 * Jm53;
    const double t3 = trb / 3.0;
    s[0] = g * (b00 - t3) + pr;
    s[1] = g * (b11 - t3) + pr;
    s[2] = g * (b22 - t3) + pr;
    s[3] = g * b01;
    s[4] = g * b12;
    s[5] = g * b02;
    // material tangent in Voigt form, then push forward
    double C[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    double Ci[3][3];
    inv3(C, det3(C), Ci);
    const double CV[6] = {C[0][0], C[1][1], C[2][2], C[0][1], C[1][2], C[0][2]};
    const double CiV[6] = {Ci[0][0], Ci[1][1], Ci[2][2], Ci[0][1], Ci[1][2], Ci[0][2]};
    const double sf = 2.0 * C10;
    const double ipSC = sf * CV[0] + sf * CV[1] + sf * CV[2];  // Inner_Product_Voigt(Sfric, C)
    double Siso[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) Siso[i] = Jm23 * ((i < 3 ? sf : 0.0) - (1.0 / 3.0) * ipSC * CiV[i]);
    double sq[21];
    square_self_up(CiV, sq);
    const double c1 = (2.0 / 3.0) * Jm23 * ipSC;
    double Cbar[21];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            const int k = up_idx(i, j);
            const double bl = CiV[i] * CiV[j];
            const double pm = sq[k] - (1.0 / 3.0) * bl;
            const double ciso = c1 * pm - (2.0 / 3.0) * (Siso[i] * CiV[j] + CiV[i] * Siso[j]);
            const double cvol = cv1 * bl - cv2 * sq[k];
            Cbar[k] = ciso + cvol;
        }
    push_forward_up(Cbar, F, J, D);
}

// model 3 (ModNeoHookean.f90:409-615): p = 3 K J^-2/3 (J^1/3 - 1), Cvol = J (p + J dp/dJ)
// Cinv (x) Cinv - 2 J p Cinv (.) Cinv
__device__ __forceinline__ void mat_nh(const double* p, const double F[3][3], double s[6], double D[21]) {
    const double C10 = p[0], K = p[1];
    const double J = det3(F);
    const double J13 = pow(J, 1.0 / 3.0);
    const double Jm23 = pow(J, -2.0 / 3.0);
    const double Jm53 = pow(J, -5.0 / 3.0);
    const double pr = 3.0 * K * Jm23 * (J13 - 1.0);
    const double d2 = -K * Jm53 * (J13 - 2.0);
    nh_core(C10, F, J, Jm23, Jm53, pr, J * (pr + J * d2), 2.0 * J * pr, s, D);
}

// model 4, the mean-dilatation (Q1P0) Neo-Hookean (ModNeoHookeanQ1P0.f90:429-484 stress,
// :540-613 tangent): the isochoric part at the Gauss point's J, the pressure of the element's
// Jbar = V/V0, pbar = 3 K (Jbar^1/3 - 1) / Jbar^2/3, Cp = J pbar (Cinv (x) Cinv - 2 Cinv (.) Cinv);
// the pressure's derivative enters the element stiffness through Bdiv (nhq1p0_d2psi)
__device__ __forceinline__ void mat_nhq1p0(const double* p, const double F[3][3], double jbar, double s[6],
                                           double D[21]) {
    const double C10 = p[0], K = p[1];
    const double J = det3(F);
    const double Jm23 = pow(J, -2.0 / 3.0);
    const double Jm53 = pow(J, -5.0 / 3.0);
    const double pbar = (3.0 * K * (pow(jbar, 1.0 / 3.0) - 1.0)) / pow(jbar, 2.0 / 3.0);
    nh_core(C10, F, J, Jm23, Jm53, pbar, J * pbar, 2.0 * J * pbar, s, D);
}

// SecondDerivativesOfPSI_Jbar_NeoHookeanQ1P0 (ModNeoHookeanQ1P0.f90:625-640)
__device__ __forceinline__ double nhq1p0_d2psi(const double* p, double jbar) {
    return -(p[1] * (pow(jbar, 1.0 / 3.0) - 2.0)) / pow(jbar, 5.0 / 3.0);
}

// Model 3 in the reference configuration: second Piola stress and material tangent dS/dE (upper
// Voigt, engineering shear) from C = F^T F and J = det F, the terms mat_nh forms before its push
// forward: S = 2 C10 J^-2/3 (I - tr C / 3 C^-1) + J p C^-1, p = 3 K J^-2/3 (J^1/3 - 1).
__device__ __forceinline__ void nh_material(double C10, double K, const double Cm[3][3], double J, double S[6],
                                            double Ct[21]) {
    double Ci[3][3];
    inv3(Cm, det3(Cm), Ci);
    const double CiV[6] = {Ci[0][0], Ci[1][1], Ci[2][2], Ci[0][1], Ci[1][2], Ci[0][2]};
    const double J13 = pow(J, 1.0 / 3.0);
    const double Jm23 = pow(J, -2.0 / 3.0);
    const double Jm53 = pow(J, -5.0 / 3.0);
    const double pr = 3.0 * K * Jm23 * (J13 - 1.0);
    const double sf = 2.0 * C10;
    const double ipSC = sf * (Cm[0][0] + Cm[1][1] + Cm[2][2]);
    double Siso[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        Siso[i] = Jm23 * ((i < 3 ? sf : 0.0) - (1.0 / 3.0) * ipSC * CiV[i]);
        S[i] = Siso[i] + J * pr * CiV[i];
    }
    double sq[21];
    square_self_up(CiV, sq);
    const double d2 = -K * Jm53 * (J13 - 2.0);
    const double c1 = (2.0 / 3.0) * Jm23 * ipSC;
    const double cv1 = J * (pr + J * d2), cv2 = 2.0 * J * pr;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            const int k = up_idx(i, j);
            const double bl = CiV[i] * CiV[j];
            const double pm = sq[k] - (1.0 / 3.0) * bl;
            Ct[k] = c1 * pm - (2.0 / 3.0) * (Siso[i] * CiV[j] + CiV[i] * Siso[j]) + cv1 * bl - cv2 * sq[k];
        }
}

// model 5, ModStVenantKirchhoff.f90:643-773
__device__ __forceinline__ void mat_svk(const double* p, const double F[3][3], double s[6], double D[21]) {
    const double E = p[0], nu = p[1];
    const double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E / (2.0 * (1.0 + nu));
    double S[3][3];
    {
        double Ee[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Ee[i][j] = 0.5 * (F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j] - (i == j ? 1.0 : 0.0));
        const double trE = Ee[0][0] + Ee[1][1] + Ee[2][2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) S[i][j] = (i == j ? lam * trE : 0.0) + 2.0 * mu * Ee[i][j];
    }
    const double J = det3(F);
    const double rJ = 1.0 / J;
    double FS[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FS[i][j] = F[i][0] * S[0][j] + F[i][1] * S[1][j] + F[i][2] * S[2][j];
    double sg[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) sg[i][j] = (FS[i][0] * F[j][0] + FS[i][1] * F[j][1] + FS[i][2] * F[j][2]) * rJ;
    s[0] = sg[0][0]; s[1] = sg[1][1]; s[2] = sg[2][2]; s[3] = sg[0][1]; s[4] = sg[1][2]; s[5] = sg[0][2];
    double b[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) b[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
    const int vi[6] = {0, 1, 2, 0, 1, 0};
    const int vj[6] = {0, 1, 2, 1, 2, 2};
    const double lJ = lam * rJ, mJ = mu * rJ;
#pragma unroll
    for (int I = 0; I < 6; ++I)
#pragma unroll
        for (int K = I; K < 6; ++K) {
            const int i = vi[I], j = vj[I], k = vi[K], l = vj[K];
            D[up_idx(I, K)] = lJ * b[i][j] * b[k][l] + mJ * (b[i][k] * b[j][l] + b[i][l] * b[j][k]);
        }
}

// model 2, ModJ2Plasticity.f90:385-479 (update) and :555-620 (tangent).
// old[0..5] committed plastic strain (engineering shear), old[6] accumulated plastic strain;
// nw receives the same 7 values after the update; returns the PlasticState flag.
__device__ __forceinline__ bool mat_j2(const double* p, const double F[3][3], const double old[7],
                                       double s[6], double D[21], double nw[7]) {
    const double E = p[0], nu = p[1], sy = p[2], H = p[3];
    double strain[6];
    small_strain(F, strain);
    const double K = E / (3.0 * (1.0 - 2.0 * nu));
    const double G = E / (2.0 * (1.0 + nu));
    double Eet[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) Eet[i] = strain[i] - old[i];
    const double evol = Eet[0] + Eet[1] + Eet[2];
    double Sdt[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double ed = Eet[i] - (i < 3 ? evol / 3.0 : 0.0);
        Sdt[i] = G * ed * (i < 3 ? 2.0 : 1.0);
    }
    const double pt = K * evol;
    const double nrm2 = Sdt[0] * Sdt[0] + Sdt[1] * Sdt[1] + Sdt[2] * Sdt[2] + 2.0 * Sdt[3] * Sdt[3] +
                        2.0 * Sdt[4] * Sdt[4] + 2.0 * Sdt[5] * Sdt[5];
    const double qt = sqrt(1.5 * nrm2);
    const double phi = qt - (sy + H * old[6]);
    if (phi <= 0.0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = Sdt[i] + (i < 3 ? pt : 0.0);
#pragma unroll
        for (int i = 0; i < 7; ++i) nw[i] = old[i];
        elastic_up(E, nu, D);
        return false;
    }
    const double dg = phi / (3.0 * G + H);
    const double f = 1.0 - dg * 3.0 * G / qt;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double sd = f * Sdt[i];
        s[i] = sd + (i < 3 ? pt : 0.0);
        const double ee = sd / (2.0 * G) + (i < 3 ? evol / 3.0 : 0.0);
        nw[i] = strain[i] - ee;
    }
    nw[6] = old[6] + dg;
    const double rn = 1.0 / sqrt(nrm2);
    double N[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) N[i] = Sdt[i] * rn;
    const double a = 2.0 * G * f;
    const double b = 6.0 * G * G * (dg / qt - 1.0 / (3.0 * G + H));
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            const double Is = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
            const double I1 = (i < 3 && j < 3) ? 1.0 : 0.0;
            D[up_idx(i, j)] = a * (Is - I1 / 3.0) + b * (N[i] * N[j]) + K * I1;
        }
    return true;
}

}  // namespace ceos
