// Viscoelastic matrix model on the device (model 12),
// ConstitutiveLibrary/ConstitutiveModelsLibrary/ModViscoelasticMatrix.f90:276-1028.
//
// Finite-strain variational viscoelasticity: S = S_inf(C) + Fv^-1 Se Fv^-T with a local Newton
// on X = (sym dv, gamma) (7 unknowns, tolerance 1e-5 on |R|_2, at most 10 iterations), and the
// algorithmic tangent from a 7x6 multi-right-hand-side solve, pushed forward to the spatial
// configuration.  Second-order tensors are column-major 9-vectors (p = i + 3j) and fourth-order
// tensors 9x9 arrays T[p][q] = T_ijkl, the layout of Tools/MathTools/ModVoigtNotation.f90.
// One thread evaluates one Gauss point; the 9x9 temporaries live in registers/scratch.
#pragma once
#include "device_math.hpp"

namespace ceos {

constexpr int kViscoState = 20;  // [0..8] Fv, [9..17] dv, [18] gamma, [19] time of the state

namespace vs {

__device__ __forceinline__ int swp(int p) { return (p % 3) * 3 + p / 3; }

__device__ inline void eye(double I[9]) {
#pragma unroll
    for (int q = 0; q < 9; ++q) I[q] = (q == 0 || q == 4 || q == 8) ? 1.0 : 0.0;
}
__device__ inline void tr(const double* A, double* T) {
#pragma unroll
    for (int q = 0; q < 9; ++q) T[q] = A[swp(q)];
}
__device__ inline double det(const double* A) {
    return A[0] * A[4] * A[8] - A[0] * A[7] * A[5] + A[1] * A[5] * A[6] - A[1] * A[3] * A[8] +
           A[2] * A[3] * A[7] - A[2] * A[4] * A[6];
}
__device__ inline void inv(const double* A, double* B) {
    const double r = 1.0 / det(A);
    B[0] = (A[4] * A[8] - A[7] * A[5]) * r;
    B[1] = (A[2] * A[7] - A[1] * A[8]) * r;
    B[2] = (A[1] * A[5] - A[2] * A[4]) * r;
    B[3] = (A[5] * A[6] - A[3] * A[8]) * r;
    B[4] = (A[0] * A[8] - A[2] * A[6]) * r;
    B[5] = (A[2] * A[3] - A[0] * A[5]) * r;
    B[6] = (A[3] * A[7] - A[4] * A[6]) * r;
    B[7] = (A[1] * A[6] - A[0] * A[7]) * r;
    B[8] = (A[0] * A[4] - A[1] * A[3]) * r;
}
__device__ inline void mul(const double* A, const double* B, double* C) {  // C = A.B (C may alias)
    double R[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) R[i + 3 * j] = A[i] * B[3 * j] + A[i + 3] * B[1 + 3 * j] + A[i + 6] * B[2 + 3 * j];
#pragma unroll
    for (int q = 0; q < 9; ++q) C[q] = R[q];
}
// T_ijkl = A_ik B_jl (SquareVoigt)
__device__ inline void square(const double* A, const double* B, double T[9][9]) {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) T[p][q] = A[(p % 3) + 3 * (q % 3)] * B[(p / 3) + 3 * (q / 3)];
}
__device__ inline void rsym(double T[9][9]) {  // (T_ijkl + T_ijlk)/2
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q)
            if (swp(q) > q) {
                const double v = 0.5 * T[p][q] + 0.5 * T[p][swp(q)];
                T[p][q] = v;
                T[p][swp(q)] = v;
            }
}
__device__ inline void lsym(double T[9][9]) {  // (T_ijkl + T_jikl)/2
#pragma unroll
    for (int p = 0; p < 9; ++p)
        if (swp(p) > p)
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const double v = 0.5 * T[p][q] + 0.5 * T[swp(p)][q];
                T[p][q] = v;
                T[swp(p)][q] = v;
            }
}
__device__ inline void mm(const double A[9][9], const double B[9][9], double C[9][9]) {  // C may alias A
    for (int p = 0; p < 9; ++p) {
        double row[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < 9; ++r) s += A[p][r] * B[r][q];
            row[q] = s;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) C[p][q] = row[q];
    }
}
// 9-vector -> 6 (Tensor2VoigtToTensor2VoigtSym: 11,22,33,12,23,31) and back (symmetric)
__device__ inline void to6(const double* A, double* T) {
    T[0] = A[0]; T[1] = A[4]; T[2] = A[8]; T[3] = A[3]; T[4] = A[7]; T[5] = A[2];
}
__device__ inline void from6(const double* A, double* T) {
    T[0] = A[0]; T[1] = A[3]; T[2] = A[5]; T[3] = A[3]; T[4] = A[1]; T[5] = A[4];
    T[6] = A[5]; T[7] = A[4]; T[8] = A[2];
}
__device__ __forceinline__ int s6(int a) {  // Tensor4VoigtToTensor4VoigtSym index map
    return a == 0 ? 0 : a == 1 ? 4 : a == 2 ? 8 : a == 3 ? 3 : a == 4 ? 7 : 6;
}

// partial-pivot Gaussian elimination, n = 7, nrhs right-hand sides (row-major B[7][nrhs])
template <int NRHS>
__device__ inline bool lu7(double A[7][7], double B[7][NRHS]) {
    for (int k = 0; k < 7; ++k) {
        int pm = k;
        for (int i = k + 1; i < 7; ++i)
            if (fabs(A[i][k]) > fabs(A[pm][k])) pm = i;
        if (A[pm][k] == 0.0) return false;
        if (pm != k) {
            for (int j = 0; j < 7; ++j) { const double t = A[k][j]; A[k][j] = A[pm][j]; A[pm][j] = t; }
            for (int r = 0; r < NRHS; ++r) { const double t = B[k][r]; B[k][r] = B[pm][r]; B[pm][r] = t; }
        }
        const double rp = 1.0 / A[k][k];
        for (int i = k + 1; i < 7; ++i) {
            const double f = A[i][k] * rp;
            for (int j = k + 1; j < 7; ++j) A[i][j] -= f * A[k][j];
            for (int r = 0; r < NRHS; ++r) B[i][r] -= f * B[k][r];
        }
    }
    for (int i = 6; i >= 0; --i) {
        const double rp = 1.0 / A[i][i];
        for (int r = 0; r < NRHS; ++r) {
            double x = B[i][r];
            for (int j = i + 1; j < 7; ++j) x -= A[i][j] * B[j][r];
            B[i][r] = x * rp;
        }
    }
    return true;
}

struct Kin {
    double Fv[9], FvFo[9], Jv, Ce[9], Cei[9], Se[9], Me[9];
};

__device__ inline void kin(const double* F, const double* Fo, const double* dv, double dt, double mue, Kin& k) {
    double I[9], A[9], Ai[9], t[9];
    eye(I);
#pragma unroll
    for (int q = 0; q < 9; ++q) A[q] = I[q] - dt * dv[q];
    inv(A, Ai);
    mul(Ai, Fo, k.Fv);
    k.Jv = det(k.Fv);
    inv(Fo, t);
    mul(k.Fv, t, k.FvFo);
    inv(k.Fv, t);
    double Fe[9], FeT[9];
    mul(F, t, Fe);
    tr(Fe, FeT);
    mul(FeT, Fe, k.Ce);
    inv(k.Ce, k.Cei);
#pragma unroll
    for (int q = 0; q < 9; ++q) k.Se[q] = mue * (I[q] - k.Cei[q]);
    mul(k.Ce, k.Se, k.Me);
}

// De = 2 mue rsym(Cei (x) Cei)
__device__ inline void dem(const Kin& k, double mue, double De[9][9]) {
    square(k.Cei, k.Cei, De);
    rsym(De);
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) De[p][q] *= 2.0 * mue;
}

// Residual_MATRIX (ModViscoelasticMatrix.f90:573-623)
__device__ inline void residual(const Kin& k, double gam, const double* dv, double dt, double niv, double R[7]) {
    double FT[9], Dp[9], DpT[9];
    tr(k.FvFo, FT);
    mul(FT, k.Me, Dp);
#pragma unroll
    for (int q = 0; q < 9; ++q) Dp[q] *= -dt;
    tr(Dp, DpT);
    double DL[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) DL[q] = (0.5 * (Dp[q] + DpT[q]) + dt * (niv * dv[q])) + gam * (dt * k.Jv * k.FvFo[q]);
    to6(DL, R);
    R[6] = k.Jv - 1.0;
}

// Jacobian_MATRIX (ModViscoelasticMatrix.f90:627-734); DSe = d Se / d dv
__device__ inline void jacobian(const Kin& k, const double De[9][9], const double* Fo, double gam, double dt,
                                double niv, double DSe[9][9], double Km[7][7]) {
    double I[9], CF[9];
    eye(I);
    mul(k.Ce, k.FvFo, CF);
    double B[9][9], DCe[9][9], T[9][9], DMe[9][9];
    square(CF, I, B);
    // DCe = -2 dt rlsym(B)
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) DCe[p][q] = B[p][q];
    rsym(DCe);
    lsym(DCe);
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) DCe[p][q] *= -2.0 * dt;
    // DSe = -dt rsym(De.B)
    mm(De, B, DSe);
    rsym(DSe);
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) DSe[p][q] *= -dt;
    // DMe = (I sq Se).DCe + (Ce sq I).DSe
    square(I, k.Se, T);
    mm(T, DCe, DMe);
    square(k.Ce, I, T);
    mm(T, DSe, B);  // B reused
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) DMe[p][q] += B[p][q];
    // D2He = -rlsym(dt^2 Aux1 sq Aux2 + dt (FvFo^T sq I).DMe)
    double Foi[9], A1[9], A2[9], t[9], FT[9];
    inv(Fo, Foi);
    tr(Foi, A1);
    mul(A1, k.FvFo, A1);
    mul(k.Fv, k.Me, t);
    tr(t, A2);
    tr(k.FvFo, FT);
    square(FT, I, T);
    mm(T, DMe, DCe);  // DCe reused as temporary
    square(A1, A2, T);
    const double dt2 = dt * dt;
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) T[p][q] = dt2 * T[p][q] + dt * DCe[p][q];
    rsym(T);
    lsym(T);
    // D2Jv = dt^2 Jv rlsym(FvFo ball FvFo + FvFo sq FvFo)
    square(k.FvFo, k.FvFo, B);
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) B[p][q] += k.FvFo[p] * k.FvFo[q];
    rsym(B);
    lsym(B);
    // D2L = -D2He' + dt niv Isym + gamma dt^2 Jv D2Jv'
    double D2L[9][9];
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const double isym = 0.5 * ((p == q ? 1.0 : 0.0) + (q == swp(p) ? 1.0 : 0.0));
            D2L[p][q] = (-T[p][q] + dt * (niv * isym)) + gam * (dt2 * k.Jv * B[p][q]);
        }
    double g9[9], g6[6];
#pragma unroll
    for (int q = 0; q < 9; ++q) g9[q] = dt * k.Jv * k.FvFo[q];
    to6(g9, g6);
    for (int a = 0; a < 6; ++a) {
#pragma unroll
        for (int b = 0; b < 6; ++b) Km[a][b] = D2L[s6(a)][s6(b)] * (b >= 3 ? 2.0 : 1.0);
        Km[a][6] = g6[a];
        Km[6][a] = g6[a] * (a >= 3 ? 2.0 : 1.0);
    }
    Km[6][6] = 0.0;
}

}  // namespace vs

// Update + tangent at one Gauss point.  props: K_inf, Mu_inf, Lambda_inf, K_e, Mu_e, Ni_v.
// F row-major 3x3; old/nw: kViscoState values.  Returns false if the local Newton failed.
__device__ inline bool mat_visco(const double* pr, const double F33[3][3], double time, const double* old,
                                 double s[6], double D[21], double* nw) {
    using namespace vs;
    const double muinf = pr[1], laminf = pr[2], mue = pr[4], niv = pr[5];
    double F[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) F[i + 3 * j] = F33[i][j];
    const double* Fo = old;
    const double dt = time - old[19];
    double dv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double gam = old[18];
    double X[7];
    to6(dv, X);
    X[6] = gam;
    Kin k;
    bool ok = false;
    double De[9][9], DSe[9][9], Km[7][7];
    for (int it = 0; it < 10; ++it) {
        kin(F, Fo, dv, dt, mue, k);
        double R[7];
        residual(k, gam, dv, dt, niv, R);
        double n2 = 0.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) n2 += R[q] * R[q];
        if (sqrt(n2) < 1.0e-5) { ok = true; break; }
        dem(k, mue, De);
        jacobian(k, De, Fo, gam, dt, niv, DSe, Km);
        double rhs[7][1];
#pragma unroll
        for (int q = 0; q < 7; ++q) rhs[q][0] = -R[q];
        if (!lu7<1>(Km, rhs)) break;
#pragma unroll
        for (int q = 0; q < 7; ++q) X[q] += rhs[q][0];
        from6(X, dv);
        gam = X[6];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) { nw[q] = k.Fv[q]; nw[9 + q] = dv[q]; }
    nw[18] = gam;
    nw[19] = time;
    // stress: S = S_inf + Fv^-1 Se Fv^-T; sigma = F S F^T / J
    double I[9], Ft[9], C[9], Ci[9], Fvi[9], FviT[9], t[9], S[9];
    eye(I);
    const double J = det(F);
    const double lnJ = log(J);
    tr(F, Ft);
    mul(Ft, F, C);
    inv(C, Ci);
    inv(k.Fv, Fvi);
    tr(Fvi, FviT);
    mul(Fvi, k.Se, t);
    mul(t, FviT, t);
#pragma unroll
    for (int q = 0; q < 9; ++q) S[q] = (muinf * (I[q] - Ci[q]) + laminf * lnJ * Ci[q]) + t[q];
    mul(F, S, t);
    mul(t, Ft, t);
    const double rJ = 1.0 / J;
    s[0] = t[0] * rJ; s[1] = t[4] * rJ; s[2] = t[8] * rJ; s[3] = t[3] * rJ; s[4] = t[7] * rJ; s[5] = t[6] * rJ;

    // ---- algorithmic tangent (GetTangentModulus_ViscoelasticMatrix_3D, :745-998) ----
    kin(F, Fo, dv, dt, mue, k);  // state of the returned solution (Fv recomputed from dv)
    dem(k, mue, De);
    jacobian(k, De, Fo, gam, dt, niv, DSe, Km);
    inv(k.Fv, Fvi);
    tr(Fvi, FviT);
    double A[9][9], Bm[9][9], W[9][9];
    // DSm/DC = (Dinf + rlsym(FiSq.De.FiTSq)) / 2
    square(Fvi, Fvi, A);          // FiSq
    square(FviT, FviT, Bm);       // FiTSq
    mm(A, De, W);
    mm(W, Bm, W);
    rsym(W);
    lsym(W);
    double DC6[6][6];
    {
        double sq[9][9];
        square(Ci, Ci, sq);
        rsym(sq);
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const int p = s6(a), q = s6(b);
                const double dinf = laminf * (Ci[p] * Ci[q]) + 2.0 * (muinf - laminf * lnJ) * sq[p][q];
                DC6[a][b] = 0.5 * (dinf + W[p][q]);
            }
    }
    // DSm/Ddv = rlsym(-2 dt (Fo^-1 sq Fv^-1 Se) + FiSq.DSe)
    double Foi[9], FviS[9];
    inv(Fo, Foi);
    mul(Fvi, k.Se, FviS);
    square(Foi, FviS, W);
    mm(A, DSe, A);  // A <- FiSq.DSe
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) W[p][q] = -2.0 * dt * W[p][q] + A[p][q];
    rsym(W);
    lsym(W);
    double Ds6[6][6];
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b) Ds6[a][b] = W[s6(a)][s6(b)];
    // D2L/DCDdv = -dt rlsym((FvFo^T sq I).DMe/DC)
    // (mm(A, B, C) may alias C with A but never with B)
    rsym(Bm);                 // DCe/DC = rsym(FiTSq)
    mm(De, Bm, A);            // DSe/DC = De.DCe/DC / 2
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) A[p][q] *= 0.5;
    square(I, k.Se, W);
    mm(W, Bm, W);             // W <- (I sq Se).DCe/DC
    square(k.Ce, I, Bm);
    mm(Bm, A, Bm);            // Bm <- (Ce sq I).DSe/DC
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) W[p][q] += Bm[p][q];   // DMe/DC
    double FT[9];
    tr(k.FvFo, FT);
    square(FT, I, A);
    mm(A, W, A);
    rsym(A);
    lsym(A);
    for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = 0; q < 9; ++q) W[p][q] = A[p][q];
    double rhs[7][6];
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b) rhs[a][b] = dt * W[s6(a)][s6(b)];  // -(-dt G)
#pragma unroll
    for (int b = 0; b < 6; ++b) rhs[6][b] = 0.0;
    lu7<6>(Km, rhs);
    double Dm[21];
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            if (b < a) continue;
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += Ds6[a][c] * (rhs[c][b] * (c >= 3 ? 2.0 : 1.0));
            Dm[up_idx(a, b)] = 2.0 * (DC6[a][b] + acc);
        }
    push_forward_up(Dm, F33, J, D);
    return ok;
}

}  // namespace ceos
