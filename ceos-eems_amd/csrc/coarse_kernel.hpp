// Two-level preconditioner of the Newton system's PCG: block Jacobi plus a coarse correction on
// aggregates of nodes,
//   M^-1 r = D^-1 r + P A_c^-1 P^T r,   A_c = P^T K P,
// the additive two-level method with a rigid-body coarse space (the tentative prolongator of
// smoothed aggregation, Vanek, Mandel & Brezina 1996).  It replaces nothing in the reference, whose
// PARDISO call (Tools/MathTools/ModPardisoSolver.f90:343-380) is a direct solve: the block-Jacobi
// PCG it stands in for needs iterations that grow with the mesh (about 1,100 to 3,400 per solve on
// the 80^3 cube); the coarse space removes the low-energy error that makes them grow, so the count
// stays about constant at a fixed aggregate size.
//
// Layout (built on the host once per load case, coarse_setup in ceos_gpu.hip):
//   agg[nn]                aggregate of every node (boxes of the reference coordinates)
//   agg_ptr[na+1], agg_nodes[nn]   the nodes of every aggregate, ascending
//   basis[nn][18]          node A's rows of P (FP32): a 3x6 block, row-major (DOF k, mode m); the 6
//                          rigid body modes of the aggregate (translations, rotations about its
//                          centroid), zero on constrained DOFs, orthonormalized per aggregate
//   pair_ptr[npair+1], pair_ij[npair], pair_ent[]
//                          the coarse blocks (I, J) of the aggregate pairs that share a nonzero
//                          of K, and per pair the (node A of I, slot j of A) whose neighbour lies
//                          in J: A_c(I, J) = sum over them of B_A^T K_(A, nbr) B_nbr
//   Ac[nc][nc], Aci[nc][ld]  the coarse matrix (nc = 6 na, FP64) and its inverse (FP32, both
//                          triangles, rows padded to ld = nc rounded up to 4; dense: nc is a few
//                          thousand)
// FP32 for P and the inverse halves the preconditioner's traffic per iteration; the
// preconditioner only has to be symmetric positive definite and close to K^-1, and the Krylov
// vectors, dot products and the convergence test stay FP64.
// All reductions run in a fixed order, so the preconditioner, and the PCG iterates, are bitwise
// reproducible.
#pragma once
#include "system_kernel.hpp"

namespace ceos {

constexpr int kCoarseModes = 6;
constexpr int kCoarseBlock = 64;  // threads of the coarse assembly (one wavefront per block pair)

// A_c(I, J): one wavefront per aggregate pair; lanes take the pair's entries in turn and keep the
// 36 sums, which are then reduced across the wavefront in a fixed order
template <typename KT>
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_assemble(SysGeom g, const KT* __restrict__ K,
                                                                  const float* __restrict__ basis,
                                                                  const int* __restrict__ pair_ptr,
                                                                  const int2* __restrict__ pair_ij,
                                                                  const uint32_t* __restrict__ pair_ent, int nc,
                                                                  double* __restrict__ Ac) {
    const int pr = blockIdx.x;
    const int e0 = pair_ptr[pr], e1 = pair_ptr[pr + 1];
    double acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kCoarseBlock) {
        const uint32_t w = pair_ent[e];
        const int A = (int)(w & 0x1ffffffu), j = (int)(w >> 25);
        const long t = (long)j * g.nn + A;
        const int B = g.slot_nbr[t];
        double kb[9];
        csr_block(K, g, A, g.slot_src[t], kb);
        const float* ba = basis + 18L * A;
        const float* bb = basis + 18L * B;
        double T[3][6];  // K_(A,B) B_B
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int m = 0; m < 6; ++m)
                T[k][m] = kb[3 * k] * (double)bb[m] + kb[3 * k + 1] * (double)bb[6 + m] + kb[3 * k + 2] * (double)bb[12 + m];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const double b0 = ba[a], b1 = ba[6 + a], b2 = ba[12 + a];  // widened from FP32
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[6 * a + m] = fma(b0, T[0][m], fma(b1, T[1][m], fma(b2, T[2][m], acc[6 * a + m])));
        }
    }
#pragma unroll
    for (int k = 0; k < 36; ++k) {
        double v = acc[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        acc[k] = v;
    }
    if (threadIdx.x < 36) {
        const int2 ij = pair_ij[pr];
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 36; ++k) v = (k == (int)threadIdx.x) ? acc[k] : v;
        const int a = (int)threadIdx.x / 6, m = (int)threadIdx.x % 6;
        Ac[(long)(6 * ij.x + a) * nc + 6 * ij.y + m] = v;
    }
}

// coarse DOFs whose mode is empty on every node of the aggregate (all DOFs constrained): unit
// diagonal, so that A_c stays positive definite
__global__ void k_coarse_unit_diag(int n, const int* __restrict__ dead, int nc, double* __restrict__ Ac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Ac[(long)dead[i] * nc + dead[i]] = 1.0;
}

// the inverse from the Cholesky factorisation holds the upper triangle in column-major order
// (element (i, j), i <= j, at j nc + i): both triangles, row-major in FP32 with a row stride ld
// (nc rounded up to a multiple of 4, the padding zero) so that rows are read as float4
__global__ void k_coarse_inverse_f32(int nc, int ld, const double* __restrict__ A, float* __restrict__ out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nc * ld) return;
    const int i = (int)(t / ld), j = (int)(t - (long)i * ld);
    out[t] = j >= nc ? 0.0f : (float)(i <= j ? A[(long)j * nc + i] : A[(long)i * nc + j]);
}

// g = P^T r: one workgroup per aggregate, the nodes of the aggregate in ascending order per
// thread, then a fixed-order tree over the threads
__global__ __launch_bounds__(kSysBlock) void k_cg_restrict(const CgScalars* c, const int* __restrict__ agg_ptr,
                                                           const int* __restrict__ agg_nodes,
                                                           const float* __restrict__ basis,
                                                           const double* __restrict__ r, double* __restrict__ gc) {
    if (cg_stopped(c)) return;
    const int I = blockIdx.x;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int k = agg_ptr[I] + (int)threadIdx.x; k < agg_ptr[I + 1]; k += kSysBlock) {
        const int A = agg_nodes[k];
        const float* b = basis + 18L * A;
        const double r0 = r[3L * A], r1 = r[3L * A + 1], r2 = r[3L * A + 2];
#pragma unroll
        for (int m = 0; m < 6; ++m) s[m] += (double)b[m] * r0 + (double)b[6 + m] * r1 + (double)b[12 + m] * r2;
    }
    __shared__ double ls[6][kSysBlock];
#pragma unroll
    for (int m = 0; m < 6; ++m) ls[m][threadIdx.x] = s[m];
    __syncthreads();
    for (int w = kSysBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 6; ++m) ls[m][threadIdx.x] += ls[m][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) gc[6 * I + threadIdx.x] = ls[threadIdx.x][0];
}

// y = A_c^-1 g: one wavefront per row of the (symmetric, dense) inverse; every lane reads 4
// entries of the row at once (row stride ld, a multiple of 4; g holds ld entries, those past nc
// zero) into two partial sums, then a fixed-order reduction over the wavefront
__global__ __launch_bounds__(256) void k_cg_coarse_solve(const CgScalars* c, int nc, int ld,
                                                         const float* __restrict__ Aci,
                                                         const double* __restrict__ gc, double* __restrict__ y) {
    if (cg_stopped(c)) return;
    const int row = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= nc) return;
    const float4* a = reinterpret_cast<const float4*>(Aci + (long)row * ld);
    const double2* g2 = reinterpret_cast<const double2*>(gc);
    double s0 = 0.0, s1 = 0.0;
    for (int q = lane; q < (ld >> 2); q += 64) {
        const float4 v = a[q];
        const double2 ga = g2[2 * q], gb = g2[2 * q + 1];
        s0 = fma((double)v.x, ga.x, s0);
        s1 = fma((double)v.y, ga.y, s1);
        s0 = fma((double)v.z, gb.x, s0);
        s1 = fma((double)v.w, gb.y, s1);
    }
    double s = s0 + s1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) y[row] = s;
}

// z = D^-1 r + P y, partial sums of r.z and r.r (one per workgroup, grid-stride over the nodes
// like k_cg_update); at the start of PCG also p = z
__global__ __launch_bounds__(kSysBlock) void k_cg_prolong(int nn, const CgScalars* c, int start,
                                                          const double* __restrict__ Minv,
                                                          const int* __restrict__ agg,
                                                          const float* __restrict__ basis,
                                                          const double* __restrict__ y, const double* __restrict__ r,
                                                          double* __restrict__ z, double* __restrict__ p,
                                                          double* part_s, double* part_m) {
    if (cg_stopped(c)) return;
    const int stride = gridDim.x * kSysBlock;
    double rz = 0.0, rr = 0.0;
    for (int A = blockIdx.x * kSysBlock + threadIdx.x; A < nn; A += stride) {
        double rv[3], zv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) rv[i] = r[3L * A + i];
        precond(Minv, nn, A, rv, zv);
        const float* b = basis + 18L * A;
        const double* yv = y + 6L * agg[A];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) s = fma((double)b[6 * i + m], yv[m], s);
            zv[i] += s;
            z[3L * A + i] = zv[i];
            if (start) p[3L * A + i] = zv[i];
            rz += rv[i] * zv[i];
            rr += rv[i] * rv[i];
        }
    }
    block_sum2(rz, rr, part_s, part_m);
}

}  // namespace ceos
