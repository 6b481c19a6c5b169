// FE^2 macro points on the GPU: the MLP surrogate of the homogenized RVE response and the
// conversion of RVE results into the external material points of the macro element kernel.
//
// A macro Gauss point p carries its deformation gradient F (column-major 9-vector) and the
// small-strain homogenized stiffness of its RVE, C_lin (upper Voigt, 21 values; the condensed
// tangent of the RVE at F = I, computed once per microstructure).  Its response is the second
// Piola-Kirchhoff stress S (Voigt 11,22,33,12,23,13) and the material tangent dS/dE (upper 21,
// engineering shear), written as one 27-double record per point ("external material point",
// model kEXT of the element kernel).
//
// Surrogate (no counterpart in the reference, which has no concurrent FE^2 loop):
//   U      = (F^T F)^(1/2) by a cyclic Jacobi eigen-decomposition of C = F^T F
//   x      = [(U - I)_Voigt / u_scale, (C_lin - d_mean) / d_std]          (n_in = 27, FP32)
//   y      = MLP(x)                                                        (FP32, ceos_mlp_forward)
//   S      = S_base + (|E|^2 / e_ref^2) (t_mean + t_std y),  E = (C - I) / 2
//   S_base = S_NH(C; K_V, mu_V) + (C_lin - C_iso(K_V, mu_V)) : E
//   dS/dE  = dS_base/dE + t (x) 2E / e_ref^2 + (|E|^2 / e_ref^2) t_std (dy/dx_U) dU/dE / u_scale
// with dU/dE from U dU + dU U = 2 dE solved in the eigenbasis of U.  The base is the Neo-Hookean
// model 3 with the Voigt-average bulk and shear moduli of C_lin plus the anisotropic rest of C_lin;
// it linearizes to C_lin exactly, so the surrogate is exact to first order at U = I, and it
// carries the volumetric nonlinearity (J-dependent pressure) of the Neo-Hookean phases, so the
// network only learns a residual that is small against both the volumetric and the shear stress.
// Training hull: the convex set {U : lam_lo <= eig(U) <= lam_hi} (the convex hull of the
// training stretches, which are drawn over exactly that set) times the box of the normalized
// descriptors seen in training.  Points outside are flagged for the RVE fallback.
//
// RVE results -> external points: S = F^-1 P and
//   dS_IJ/dE_KL = sum_ik Finv_Ii Finv_Kk (A_iJkL - delta_ik S_JL),  A_iJkL = dP_iJ/dF_kL
// (the conversion of First_Elasticity_Modulus_To_Second_Voigt, Tools/MathTools/
// ModMathRoutines.f90:315, used by GetTangentModulusBase, ConstitutiveLibrary/
// ModConstitutiveModel.f90:104-164), symmetrized over the minor and major index pairs.
#include <hip/hip_runtime.h>

#include <cstring>

#include "ceos_gpu.h"
#include "device_math.hpp"

using namespace ceos;

#define SUR_HIP_OK(x)                               \
    do {                                            \
        if ((x) != hipSuccess) return CEOS_ERR_HIP; \
    } while (0)

namespace {

constexpr int kNd = 21;            // descriptor components (C_lin upper Voigt)
constexpr int kNin = 6 + kNd;      // network inputs
constexpr int kNout = 6;
constexpr int kAux = 12;           // per point: eigenvalues of U (3), eigenvectors (9, column k = vector k)

struct SurCfg {
    double u_scale, lam_lo, lam_hi, e_ref;
    double d_mean[kNd], d_std[kNd], d_lo[kNd], d_hi[kNd];
    double t_mean[kNout], t_std[kNout];
};

__device__ __forceinline__ void voigt_pair(int a, int& i, int& j) {
    i = a < 3 ? a : (a == 3 ? 0 : (a == 4 ? 1 : 0));
    j = a < 3 ? a : (a == 3 ? 1 : 2);
}

// Cyclic Jacobi for a symmetric 3x3 matrix: on return a is (numerically) diagonal with the
// eigenvalues on its diagonal and Q holds the eigenvectors as columns (A = Q diag Q^T).
__host__ __device__ inline void jacobi_sym3(double a[3][3], double Q[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Q[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > 1.0e-40 * dia)) break;
        for (int r = 0; r < 3; ++r) {
            const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
            const double apq = a[p][q];
            if (apq == 0.0) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; ++k) {  // columns p, q of A P
                const double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {  // rows p, q of P^T (A P)
                const double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            for (int k = 0; k < 3; ++k) {
                const double qkp = Q[k][p], qkq = Q[k][q];
                Q[k][p] = c * qkp - s * qkq;
                Q[k][q] = s * qkp + c * qkq;
            }
        }
    }
}

__device__ __forceinline__ void load_F(const double* f, double F[3][3]) {  // column-major 9-vector
#pragma unroll
    for (int J = 0; J < 3; ++J)
#pragma unroll
        for (int i = 0; i < 3; ++i) F[i][J] = f[3 * J + i];
}

__device__ __forceinline__ void right_cauchy_green(const double F[3][3], double C[3][3]) {
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int J = 0; J < 3; ++J) C[I][J] = F[0][I] * F[0][J] + F[1][I] * F[1][J] + F[2][I] * F[2][J];
}

__global__ __launch_bounds__(256) void k_sur_features(int n, SurCfg cfg, const double* __restrict__ Fin,
                                                      const double* __restrict__ Clin, float* __restrict__ X,
                                                      int* __restrict__ flags, double* __restrict__ aux) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double F[3][3], C[3][3], Q[3][3];
    load_F(Fin + 9L * t, F);
    right_cauchy_green(F, C);
    jacobi_sym3(C, Q);
    double lam[3];
    bool inside = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lam[k] = sqrt(fmax(C[k][k], 0.0));
        inside = inside && lam[k] >= cfg.lam_lo && lam[k] <= cfg.lam_hi;
    }
    double U[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) U[i][j] = Q[i][0] * lam[0] * Q[j][0] + Q[i][1] * lam[1] * Q[j][1] + Q[i][2] * lam[2] * Q[j][2];
    float* x = X + (long)kNin * t;
    const double ru = 1.0 / cfg.u_scale;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        int i, j;
        voigt_pair(a, i, j);
        x[a] = (float)((U[i][j] - (i == j ? 1.0 : 0.0)) * ru);
    }
    const double* d = Clin + (long)kNd * t;
#pragma unroll
    for (int k = 0; k < kNd; ++k) {
        const double z = (d[k] - cfg.d_mean[k]) / cfg.d_std[k];
        x[6 + k] = (float)z;
        inside = inside && z >= cfg.d_lo[k] && z <= cfg.d_hi[k];
    }
    if (flags) flags[t] = inside ? 0 : 1;
    if (aux) {
        double* o = aux + (long)kAux * t;
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = lam[k];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) o[3 + 3 * i + j] = Q[i][j];
    }
}

__global__ __launch_bounds__(256) void k_sur_output(int n, SurCfg cfg, const double* __restrict__ Fin,
                                                    const double* __restrict__ Clin, const double* __restrict__ aux,
                                                    const float* __restrict__ Y, const float* __restrict__ Jx,
                                                    double* __restrict__ ext, double* __restrict__ Pout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double F[3][3], C[3][3];
    load_F(Fin + 9L * t, F);
    right_cauchy_green(F, C);
    double E[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) E[i][j] = 0.5 * (C[i][j] - (i == j ? 1.0 : 0.0));
    const double Ee[6] = {E[0][0], E[1][1], E[2][2], 2.0 * E[0][1], 2.0 * E[1][2], 2.0 * E[0][2]};
    const double re2 = 1.0 / (cfg.e_ref * cfg.e_ref);
    const double phi = (E[0][0] * E[0][0] + E[1][1] * E[1][1] + E[2][2] * E[2][2] +
                        2.0 * (E[0][1] * E[0][1] + E[1][2] * E[1][2] + E[0][2] * E[0][2])) * re2;
    double Cl[6][6];
    const double* cl = Clin + (long)kNd * t;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = a; b < 6; ++b) Cl[a][b] = Cl[b][a] = cl[up_idx(a, b)];
    // base: Neo-Hookean with the Voigt-average moduli of C_lin, plus C_lin's anisotropic rest : E
    const double Kv = (Cl[0][0] + Cl[1][1] + Cl[2][2] + 2.0 * (Cl[0][1] + Cl[0][2] + Cl[1][2])) / 9.0;
    const double mv = ((Cl[0][0] + Cl[1][1] + Cl[2][2]) - (Cl[0][1] + Cl[0][2] + Cl[1][2]) +
                       3.0 * (Cl[3][3] + Cl[4][4] + Cl[5][5])) / 15.0;
    const double lv = Kv - 2.0 * mv / 3.0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b)
            Cl[a][b] -= (a < 3 && b < 3 ? lv : 0.0) + (a == b ? (a < 3 ? 2.0 * mv : mv) : 0.0);
    double Snh[6], Cnh[21];
    nh_material(0.5 * mv, Kv, C, det3(F), Snh, Cnh);
    double tt[6], S[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        tt[a] = cfg.t_mean[a] + cfg.t_std[a] * (double)Y[(long)kNout * t + a];
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) acc += Cl[a][b] * Ee[b];
        S[a] = Snh[a] + acc + phi * tt[a];
    }
    const double* o = aux + (long)kAux * t;
    const double lam[3] = {o[0], o[1], o[2]};
    double Q[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Q[i][j] = o[3 + 3 * i + j];
    const float* J = Jx + (long)kNout * kNin * t;  // J[o][c] = dy_o / dx_c
    double Ct[6][6];
    const double ru = 1.0 / cfg.u_scale;
#pragma unroll 1
    for (int b = 0; b < 6; ++b) {
        int K, L;
        voigt_pair(b, K, L);
        // dE for a unit engineering strain component b, rotated into the eigenbasis of U
        double dEp[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                dEp[i][j] = K == L ? Q[K][i] * Q[K][j] : 0.5 * (Q[K][i] * Q[L][j] + Q[L][i] * Q[K][j]);
        double dUp[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) dUp[i][j] = 2.0 * dEp[i][j] / (lam[i] + lam[j]);
        double dx[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            int i, j;
            voigt_pair(c, i, j);
            double acc = 0.0;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) acc += Q[i][p] * dUp[p][q] * Q[j][q];
            dx[c] = acc * ru;
        }
        const double dphi = 2.0 * E[K][L] * re2;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += (double)J[kNin * a + c] * dx[c];
            Ct[a][b] = Cnh[a <= b ? up_idx(a, b) : up_idx(b, a)] + Cl[a][b] + tt[a] * dphi + phi * cfg.t_std[a] * acc;
        }
    }
    double* e = ext + (long)kExtPerGp * t;
#pragma unroll
    for (int a = 0; a < 6; ++a) e[a] = S[a];
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = a; b < 6; ++b) e[6 + up_idx(a, b)] = 0.5 * (Ct[a][b] + Ct[b][a]);
    if (Pout) {
        const double Sm[3][3] = {{S[0], S[3], S[5]}, {S[3], S[1], S[4]}, {S[5], S[4], S[2]}};
#pragma unroll
        for (int Jc = 0; Jc < 3; ++Jc)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                Pout[9L * t + 3 * Jc + i] = F[i][0] * Sm[0][Jc] + F[i][1] * Sm[1][Jc] + F[i][2] * Sm[2][Jc];
    }
}

__global__ __launch_bounds__(256) void k_rve_to_ext(int n, const int* __restrict__ sel, const double* __restrict__ Fin,
                                                    const double* __restrict__ P, const double* __restrict__ A,
                                                    double* __restrict__ ext) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double F[3][3], Fi[3][3];
    load_F(Fin + 9L * t, F);
    inv3(F, det3(F), Fi);
    const double* pp = P + 9L * t;
    double S[3][3];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int Jc = 0; Jc < 3; ++Jc) S[I][Jc] = Fi[I][0] * pp[3 * Jc] + Fi[I][1] * pp[3 * Jc + 1] + Fi[I][2] * pp[3 * Jc + 2];
    double* e = ext + (long)kExtPerGp * (sel ? sel[t] : t);
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        int I, Jc;
        voigt_pair(a, I, Jc);
        e[a] = 0.5 * (S[I][Jc] + S[Jc][I]);
    }
    if (!A) return;  // stress only: the point keeps the tangent it has
    const double* aa = A + 81L * t;  // aa[9 (3J + i) + 3L + k] = dP_iJ / dF_kL
    // C_IJKL = sum_ik Fi_Ii Fi_Kk (A_iJkL - delta_ik S_JL), the 21 upper Voigt entries after
    // averaging the 8 index permutations of the minor and major symmetries
    auto cijkl = [&](int I, int Jc, int K, int L) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                acc += Fi[I][i] * Fi[K][k] * (aa[9 * (3 * Jc + i) + 3 * L + k] - (i == k ? S[Jc][L] : 0.0));
        return acc;
    };
#pragma unroll 1
    for (int a = 0; a < 6; ++a) {
        int I, Jc;
        voigt_pair(a, I, Jc);
#pragma unroll 1
        for (int b = a; b < 6; ++b) {
            int K, L;
            voigt_pair(b, K, L);
            const double m1 = cijkl(I, Jc, K, L) + cijkl(Jc, I, K, L) + cijkl(I, Jc, L, K) + cijkl(Jc, I, L, K);
            const double m2 = cijkl(K, L, I, Jc) + cijkl(L, K, I, Jc) + cijkl(K, L, Jc, I) + cijkl(L, K, Jc, I);
            e[6 + up_idx(a, b)] = 0.125 * (m1 + m2);
        }
    }
}

}  // namespace

struct ceos_surrogate {
    ceos_mlp* mlp = nullptr;  // not owned
    SurCfg cfg;
    int max_points = 0;
    float *X = nullptr, *Y = nullptr, *J = nullptr;
    double* aux = nullptr;
};

namespace {

void sur_free(ceos_surrogate* s) {
    if (s->X) (void)hipFree(s->X);
    if (s->Y) (void)hipFree(s->Y);
    if (s->J) (void)hipFree(s->J);
    if (s->aux) (void)hipFree(s->aux);
    s->X = s->Y = s->J = nullptr;
    s->aux = nullptr;
}

}  // namespace

extern "C" {

int ceos_surrogate_create(ceos_mlp* mlp, const ceos_surrogate_cfg* cfg, int max_points, ceos_surrogate** out) {
    if (!out) return CEOS_ERR_ARG;
    *out = nullptr;
    if (!mlp || !cfg || max_points < 1) return CEOS_ERR_ARG;
    int n_in = 0, n_out = 0, max_batch = 0;
    if (ceos_mlp_shape(mlp, &n_in, nullptr, &n_out, &max_batch) != CEOS_OK) return CEOS_ERR_ARG;
    if (n_in != kNin || n_out != kNout || max_points > max_batch) return CEOS_ERR_ARG;
    if (!(cfg->u_scale > 0.0) || !(cfg->e_ref > 0.0) || !(cfg->lam_lo > 0.0) || !(cfg->lam_hi > cfg->lam_lo))
        return CEOS_ERR_ARG;
    for (int k = 0; k < kNd; ++k)
        if (!(cfg->d_std[k] > 0.0)) return CEOS_ERR_ARG;
    for (int k = 0; k < kNout; ++k)
        if (!(cfg->t_std[k] > 0.0)) return CEOS_ERR_ARG;
    ceos_surrogate* s = new ceos_surrogate();
    s->mlp = mlp;
    s->max_points = max_points;
    s->cfg.u_scale = cfg->u_scale;
    s->cfg.lam_lo = cfg->lam_lo;
    s->cfg.lam_hi = cfg->lam_hi;
    s->cfg.e_ref = cfg->e_ref;
    std::memcpy(s->cfg.d_mean, cfg->d_mean, sizeof(s->cfg.d_mean));
    std::memcpy(s->cfg.d_std, cfg->d_std, sizeof(s->cfg.d_std));
    std::memcpy(s->cfg.d_lo, cfg->d_lo, sizeof(s->cfg.d_lo));
    std::memcpy(s->cfg.d_hi, cfg->d_hi, sizeof(s->cfg.d_hi));
    std::memcpy(s->cfg.t_mean, cfg->t_mean, sizeof(s->cfg.t_mean));
    std::memcpy(s->cfg.t_std, cfg->t_std, sizeof(s->cfg.t_std));
    const long long m = max_points;
    if (hipMalloc((void**)&s->X, sizeof(float) * kNin * m) != hipSuccess ||
        hipMalloc((void**)&s->Y, sizeof(float) * kNout * m) != hipSuccess ||
        hipMalloc((void**)&s->J, sizeof(float) * kNout * kNin * m) != hipSuccess ||
        hipMalloc((void**)&s->aux, sizeof(double) * kAux * m) != hipSuccess) {
        sur_free(s);
        delete s;
        return CEOS_ERR_HIP;
    }
    *out = s;
    return CEOS_OK;
}

void ceos_surrogate_destroy(ceos_surrogate* s) {
    if (!s) return;
    (void)hipDeviceSynchronize();
    sur_free(s);
    delete s;
}

int ceos_surrogate_features_device(ceos_surrogate* s, int n, const double* dF, const double* dClin, float* dX,
                                   int* dflags, void* stream) {
    if (!s || n < 1 || !dF || !dClin || !dX) return CEOS_ERR_ARG;
    hipStream_t st = stream ? (hipStream_t)stream : hipStreamPerThread;
    hipLaunchKernelGGL(k_sur_features, dim3((n + 255) / 256), dim3(256), 0, st, n, s->cfg, dF, dClin, dX, dflags,
                       (double*)nullptr);
    SUR_HIP_OK(hipGetLastError());
    if (!stream) SUR_HIP_OK(hipStreamSynchronize(st));
    return CEOS_OK;
}

int ceos_surrogate_eval_device(ceos_surrogate* s, int n, const double* dF, const double* dClin, double* dext,
                               int* dflags, double* dP, void* stream) {
    if (!s || n < 1 || n > s->max_points || !dF || !dClin || !dext) return CEOS_ERR_ARG;
    // all three steps on one stream (the MLP would otherwise use its own)
    hipStream_t st = stream ? (hipStream_t)stream : hipStreamPerThread;
    hipLaunchKernelGGL(k_sur_features, dim3((n + 255) / 256), dim3(256), 0, st, n, s->cfg, dF, dClin, s->X, dflags,
                       s->aux);
    SUR_HIP_OK(hipGetLastError());
    const int rc = ceos_mlp_forward(s->mlp, s->X, n, s->Y, s->J, (void*)st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sur_output, dim3((n + 255) / 256), dim3(256), 0, st, n, s->cfg, dF, dClin, s->aux, s->Y,
                       s->J, dext, dP);
    SUR_HIP_OK(hipGetLastError());
    if (!stream) SUR_HIP_OK(hipStreamSynchronize(st));
    return CEOS_OK;
}

int ceos_rve_to_external_device(int n, const int* dsel, const double* dF, const double* dP, const double* dA,
                                double* dext, void* stream) {
    if (n < 0 || (n > 0 && (!dF || !dP || !dext))) return CEOS_ERR_ARG;
    if (n == 0) return CEOS_OK;
    hipStream_t st = stream ? (hipStream_t)stream : hipStreamPerThread;
    hipLaunchKernelGGL(k_rve_to_ext, dim3((n + 255) / 256), dim3(256), 0, st, n, dsel, dF, dP, dA, dext);
    SUR_HIP_OK(hipGetLastError());
    if (!stream) SUR_HIP_OK(hipStreamSynchronize(st));
    return CEOS_OK;
}

}  // extern "C"
