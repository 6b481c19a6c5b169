// Volume averages over the reference configuration (MultiScale/ModMultiscaleAnalysis.f90:140-441):
//   P_bar = (1/V0) sum_e sum_gp P w detJ_X,  P = J sigma F^-T (Tools/ModContinuumMechanics.f90:125-129)
//   F_bar = (1/V0) sum_e sum_gp F w detJ_X
// One thread per (element, Gauss point); each block writes 19 partial sums (P, F column-major,
// then the volume); the host adds the block partials in block order (deterministic).
#pragma once
#include "device_math.hpp"

namespace ceos {

constexpr int kHomBlock = 256;
constexpr int kHomVals = 19;

__global__ __launch_bounds__(256) void k_homogenize_partial(int ne, const int* __restrict__ conn,
                                                            const double* __restrict__ X,
                                                            const double* __restrict__ U,
                                                            const double* __restrict__ sig,
                                                            double* __restrict__ part) {
    __shared__ double red[kHomVals][kHomBlock];
    const int t = blockIdx.x * kHomBlock + threadIdx.x;
    double v[kHomVals];
#pragma unroll
    for (int k = 0; k < kHomVals; ++k) v[k] = 0.0;
    if (t < ne * 8) {
        const int s = t >> 3, g = t & 7;
        double dN[8][3], c[8][3], u[8][3], gX[8][3];
        hex8_dshape_gp(g, dN);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int n = conn[(long)s * 8 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                c[a][k] = X[3L * n + k];
                u[a][k] = U[3L * n + k];
            }
        }
        const double detX = hex8_grad(dN, c, gX);
        double F[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < 8; ++a) acc += gX[a][j] * u[a][i];
                F[i][j] = acc + (i == j ? 1.0 : 0.0);
            }
        const double J = det3(F);
        double Fi[3][3];
        inv3(F, J, Fi);
        double sg[3][3];
        const double* sp = sig;
        const long base = (long)g * ne + s;
        const long stride = 8L * ne;
        sg[0][0] = sp[base + 0 * stride]; sg[1][1] = sp[base + 1 * stride]; sg[2][2] = sp[base + 2 * stride];
        sg[0][1] = sg[1][0] = sp[base + 3 * stride];
        sg[1][2] = sg[2][1] = sp[base + 4 * stride];
        sg[0][2] = sg[2][0] = sp[base + 5 * stride];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                // P_ij = J sum_k sigma_ik (F^-1)_jk ; column-major index 3j + i
                const double Pij = J * (sg[i][0] * Fi[j][0] + sg[i][1] * Fi[j][1] + sg[i][2] * Fi[j][2]);
                v[3 * j + i] = Pij * detX;
                v[9 + 3 * j + i] = F[i][j] * detX;
            }
        v[18] = detX;
    }
#pragma unroll
    for (int k = 0; k < kHomVals; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kHomBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < kHomVals; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < kHomVals) part[(long)blockIdx.x * kHomVals + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_add(long n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] + b[i];
}

__global__ void k_zero_int(int* p) { p[0] = 0; p[1] = 0; }

// Deformation gradient at every Gauss point, F = I + sum_a u_a (x) grad_X N_a
// (ModElement.f90:1210-1262, the arithmetic of the element kernel's phase 1), in element order:
// F[(e 8 + g) 9 + 3J + i] = F_iJ (column-major 9-vectors, the reference's Voigt 9 layout).
// One thread per (slot, Gauss point); eid maps the slot to its element.
__global__ __launch_bounds__(256) void k_gp_deformation_gradient(int ne, const int* __restrict__ conn,
                                                                 const int* __restrict__ eid,
                                                                 const double* __restrict__ X,
                                                                 const double* __restrict__ U,
                                                                 double* __restrict__ Fout) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8L * ne) return;
    const int s = (int)(t >> 3), g = (int)(t & 7);
    double dN[8][3], c[8][3], u[8][3], gX[8][3];
    hex8_dshape_gp(g, dN);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int n = conn[(long)s * 8 + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[a][k] = X[3L * n + k];
            u[a][k] = U[3L * n + k];
        }
    }
    (void)hex8_grad(dN, c, gX);
    double* o = Fout + ((long)eid[s] * 8 + g) * 9;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) acc += gX[a][j] * u[a][i];
            o[3 * j + i] = acc + (i == j ? 1.0 : 0.0);
        }
}

}  // namespace ceos
