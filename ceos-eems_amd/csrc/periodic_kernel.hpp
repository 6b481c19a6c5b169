// Periodic reduction of a full hex8 system onto the periodic (torus) pattern of an RVE: the
// products with the periodicity matrix T of the reference's periodic system of equations
// (FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:77-222: K_red = T^T K T, R_red = T^T R,
// DeltaU = T DeltaU_red).  T maps every node to one reduced node (a slave to its master, the
// corners to the first corner), so T^T K T sums, for every entry of the reduced upper pattern,
// the full entries whose two DOFs map onto it, and T^T R sums the rows of the nodes that share a
// master.  Each output is one thread summing a fixed list in a fixed order: no atomics, the
// results are bitwise reproducible.
#pragma once
#include <stdint.h>

#include "system_kernel.hpp"

namespace ceos {

constexpr int kPerBlock = 256;
constexpr uint32_t kPerDoubled = 0x80000000u;  // contribution counted twice (both of a pair of
                                               // periodic images: K(i,j) and K(j,i) land on one entry)

// K_red[e] = sum of the contributing full entries (full slot in bits 0..30)
__global__ __launch_bounds__(kPerBlock) void k_periodic_reduce_k(long long nnz_red, const int* __restrict__ kptr,
                                                                 const uint32_t* __restrict__ kidx,
                                                                 const double* __restrict__ K,
                                                                 double* __restrict__ Kred) {
    const long long e = (long long)blockIdx.x * kPerBlock + threadIdx.x;
    if (e >= nnz_red) return;
    double s = 0.0;
    for (int c = kptr[e]; c < kptr[e + 1]; ++c) {
        const uint32_t w = kidx[c];
        const double v = K[w & ~kPerDoubled];
        s += (w & kPerDoubled) ? 2.0 * v : v;
    }
    Kred[e] = s;
}

// R_red[a] = sum over the nodes A of master a of F[A]; mask (the reduced system's DOF masks, or
// nullptr): DOFs with a mask bit set are zeroed (the constrained rows of a right-hand side)
__global__ __launch_bounds__(kPerBlock) void k_periodic_reduce_r(int nn_red, const int* __restrict__ rptr,
                                                                 const int* __restrict__ rnodes,
                                                                 const double* __restrict__ F,
                                                                 const uint8_t* __restrict__ mask,
                                                                 double* __restrict__ R) {
    const int a = blockIdx.x * kPerBlock + threadIdx.x;
    if (a >= nn_red) return;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    for (int k = rptr[a]; k < rptr[a + 1]; ++k) {
        const long A = rnodes[k];
        r0 += F[3 * A];
        r1 += F[3 * A + 1];
        r2 += F[3 * A + 2];
    }
    if (mask) {
        r0 = mask[3L * a] ? 0.0 : r0;
        r1 = mask[3L * a + 1] ? 0.0 : r1;
        r2 = mask[3L * a + 2] ? 0.0 : r2;
    }
    R[3L * a] = r0;
    R[3L * a + 1] = r1;
    R[3L * a + 2] = r2;
}

// U[A] += scale * w[map[A]]
__global__ __launch_bounds__(kPerBlock) void k_periodic_expand(int nn, const int* __restrict__ map,
                                                               const double* __restrict__ w, double scale,
                                                               double* __restrict__ U) {
    const int A = blockIdx.x * kPerBlock + threadIdx.x;
    if (A >= nn) return;
    const long a = map[A];
#pragma unroll
    for (int i = 0; i < 3; ++i) U[3L * A + i] += scale * w[3 * a + i];
}

// y = K x for the symmetric matrix stored as its upper CSR, one thread per node over the node's
// neighbour slots (the neighbours in slot order)
template <typename KT>
__global__ __launch_bounds__(kSysBlock) void k_sym_spmv(SysGeom g, const KT* __restrict__ K,
                                                        const double* __restrict__ x, double* __restrict__ y) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    if (A >= g.nn) return;
    double r[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < g.width; ++j) {
        const uint32_t src = g.slot_src[(long)j * g.nn + A];
        if (src == kSlotPad) continue;
        const long B = g.slot_nbr[(long)j * g.nn + A];
        double b[9];
        csr_block(K, g, A, src, b);
        const double x0 = x[3 * B], x1 = x[3 * B + 1], x2 = x[3 * B + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) r[i] += b[3 * i] * x0 + b[3 * i + 1] * x1 + b[3 * i + 2] * x2;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) y[3L * A + i] = r[i];
}

// the 9 sums sum_A F[A]_i X[A]_J (column-major, 3J + i) per workgroup, nodes in ascending order
// per thread, then a fixed-order tree over the threads
__global__ __launch_bounds__(kPerBlock) void k_periodic_moment_partial(int nn, const double* __restrict__ X,
                                                                       const double* __restrict__ F,
                                                                       double* __restrict__ part) {
    double s[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) s[k] = 0.0;
    for (int A = blockIdx.x * kPerBlock + threadIdx.x; A < nn; A += gridDim.x * kPerBlock) {
        const double f0 = F[3L * A], f1 = F[3L * A + 1], f2 = F[3L * A + 2];
#pragma unroll
        for (int J = 0; J < 3; ++J) {
            const double xj = X[3L * A + J];
            s[3 * J] += f0 * xj;
            s[3 * J + 1] += f1 * xj;
            s[3 * J + 2] += f2 * xj;
        }
    }
    __shared__ double ls[9][kPerBlock];
#pragma unroll
    for (int k = 0; k < 9; ++k) ls[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kPerBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 9; ++k) ls[k][threadIdx.x] += ls[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 9) part[9 * blockIdx.x + threadIdx.x] = ls[threadIdx.x][0];
}

// P[k] = (sum over the workgroups of part[.][k]) / V0, in workgroup order
__global__ void k_periodic_moment_final(int nblk, const double* __restrict__ part, double inv_v0,
                                        double* __restrict__ P) {
    const int k = threadIdx.x;
    if (k >= 9) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[9 * b + k];
    P[k] = s * inv_v0;
}

}  // namespace ceos
