// Gather assembly of the upper CSR tangent and of Fint from element records.
//
// Replaces the scatter of AssembleGlobalMatrixUpperTriangular / InternalForce
// (FEM/Subroutines/SubAssembleGlobalMatrixUpperTriangular.f90:13-69, SubInternalForce.f90:83)
// by its transpose: every CSR row segment owns its output and sums the element blocks that
// contribute to it, in element order, through a contribution map built once per mesh.
//
// Blocks whose contributions all come from one patch are written by the element kernel itself
// (they have no contribution here and are skipped), as are the Fint parts of their nodes.
//
// Work unit = one node block of the upper pattern (the diagonal block of node A, or the block
// (A, B) of its q-th upper neighbour B) that has boundary records, in CSR order.  The thread
// sums the contributing patch records and writes its 3 pieces of the rows 3A..3A+2;
// consecutive threads write neighbouring pieces of the same rows.
// Contribution word: bits 0..30 = offset of the 3x3 block (or of the 6-value diagonal block)
// in the record array, bit 31 = stored transposed (row-major values that are the columns of the
// CSR block).  The element kernel writes patch records (element_kernel.hpp) already in the CSR
// block's orientation, so the maps built for it never set the bit.
#pragma once
#include <stdint.h>

#include "element_kernel.hpp"

namespace ceos {

struct GatherParams {
    int nn;
    long long nblocks;
    long long blk0, blk1;    // the launch covers work-list entries [blk0, blk1)
    const int4* glist;       // work list (k_gather_list)
    const uint32_t* cword;   // [contributions]
    const int* fptr;         // [nn+1] residual-only pass: Fe records of node A at fword[fptr[A]..]
    const uint32_t* fword;   // offsets (doubles) of the Fe records
    const double* rec;       // patch records (F+K) or Fe parts (residual only)
    double* Fint;            // [3 nn]
    void* K;                 // [nnz]
};

// Contributions are taken kGatherBatch at a time: their words, then all their record pieces are
// requested before any is summed (a thread keeps up to 2 x 9 loads in flight instead of one
// word -> piece chain per contribution).  The sums still run in contribution order.  With patch
// records and 32-element patches a listed block has 2.3 contributions on average (80^3 cube):
// batches of 2 measured 0.237 ms per gather, 3 0.236 ms, 4 0.233 ms; non-temporal record loads
// 0.317 ms (the records the element kernel just wrote are still in the last-level cache).
constexpr int kGatherBatch = 4;

// Work list of the gather: one entry per node block that has boundary records, in CSR order
// (blocks written by the element kernel are absent, so no thread starts for them):
//   x = first contribution, y = contributions | (diagonal block) << 30,
//   z = K slot of the block's first entry in row 3A (diagonal: the row start; else the column 3B),
//   w = A << 7 | nup[A].
// Rows 3A+1 and 3A+2 of the block start at z + 3 nup + 3 and z + 6 nup + 5 (diagonal: at its
// (1,1) and (2,2) entries) or at z + 3 nup + 2 and z + 6 nup + 3 (off-diagonal).
template <typename KT>
__global__ __launch_bounds__(256) void k_gather_list(GatherParams p) {
    const long long i = p.blk0 + (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.blk1) return;
    const int4 d = p.glist[i];
    const int c0 = d.x, c1 = d.x + (d.y & 0x3fffffff);
    const long s = d.z;
    const int A = d.w >> 7, nu = d.w & 127;
    KT* K = (KT*)p.K;
    if ((d.y >> 30) & 1) {
        // diagonal block, stored upper 00 01 02 11 12 22, preceded by the lane's Fe
        double a[6] = {0, 0, 0, 0, 0, 0}, f[3] = {0, 0, 0};
        for (int c = c0; c < c1; c += kGatherBatch) {
            const double* dd[kGatherBatch];
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                dd[u] = (c + u < c1) ? p.rec + (long)(p.cword[c + u] & 0x7fffffffu) : nullptr;
            double v[kGatherBatch][9];
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                if (dd[u]) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) v[u][k] = dd[u][k - 3];
                }
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                if (dd[u]) {
#pragma unroll
                    for (int k = 0; k < 6; ++k) a[k] += v[u][3 + k];
#pragma unroll
                    for (int k = 0; k < 3; ++k) f[k] += v[u][k];
                }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) p.Fint[3L * A + k] = f[k];
        const long r1 = s + 3 + 3 * nu, r2 = r1 + 2 + 3 * nu;
        K[s] = (KT)a[0]; K[s + 1] = (KT)a[1]; K[s + 2] = (KT)a[2];
        K[r1] = (KT)a[3]; K[r1 + 1] = (KT)a[4];
        K[r2] = (KT)a[5];
        return;
    }
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = c0; c < c1; c += kGatherBatch) {
        uint32_t w[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) w[u] = (c + u < c1) ? p.cword[c + u] : 0xffffffffu;
        double v[kGatherBatch][9];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u)
            if (w[u] != 0xffffffffu) {
                const double* dd = p.rec + (long)(w[u] & 0x7fffffffu);
#pragma unroll
                for (int k = 0; k < 9; ++k) v[u][k] = dd[k];
            }
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            if (w[u] == 0xffffffffu) continue;
            if (w[u] >> 31) {
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int k = 0; k < 3; ++k) a[3 * r + k] += v[u][3 * k + r];
            } else {
#pragma unroll
                for (int k = 0; k < 9; ++k) a[k] += v[u][k];
            }
        }
    }
    const long s1 = s + 3 * nu + 2, s2 = s + 6 * nu + 3;
    K[s] = (KT)a[0]; K[s + 1] = (KT)a[1]; K[s + 2] = (KT)a[2];
    K[s1] = (KT)a[3]; K[s1 + 1] = (KT)a[4]; K[s1 + 2] = (KT)a[5];
    K[s2] = (KT)a[6]; K[s2 + 1] = (KT)a[7]; K[s2 + 2] = (KT)a[8];
}

// residual-only gather: Fint of node A from the Fe records of its patches, in patch order (the
// element kernel writes one 3-double Fe record per boundary diagonal patch record in this pass;
// fword holds their offsets).  One thread per node, the offsets and then their Fe records requested
// kGatherBatch at a time.
__global__ __launch_bounds__(256) void k_gather_fint(GatherParams p) {
    const int A = blockIdx.x * 256 + threadIdx.x;
    if (A >= p.nn) return;
    const int c0 = p.fptr[A], c1 = p.fptr[A + 1];
    if (c0 == c1) return;  // the node's diagonal block is complete: its patch wrote Fint
    double f[3] = {0.0, 0.0, 0.0};
    for (int c = c0; c < c1; c += kGatherBatch) {
        uint32_t w[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) w[u] = (c + u < c1) ? p.fword[c + u] : 0u;
        double v[kGatherBatch][3];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u)
            if (c + u < c1) {
                const double* d = p.rec + w[u];
#pragma unroll
                for (int k = 0; k < 3; ++k) v[u][k] = d[k];
            }
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u)
            if (c + u < c1) {
#pragma unroll
                for (int k = 0; k < 3; ++k) f[k] += v[u][k];
            }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) p.Fint[3L * A + k] = f[k];
}

// nodes that belong to no element: Fint and the diagonal K block are zero (the reference zeroes
// Kg%Val and Fint before assembling)
template <typename KT>
__global__ __launch_bounds__(256) void k_zero_free_nodes(int n, const int* nodes, const int* rowstart, const int* nup,
                                                         double* Fint, void* Kv) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int A = nodes[i];
    for (int k = 0; k < 3; ++k) Fint[3L * A + k] = 0.0;
    if (!Kv) return;
    KT* K = (KT*)Kv;
    const long r0 = rowstart[A], r1 = r0 + 3 + 3 * nup[A], r2 = r1 + 2 + 3 * nup[A];
    K[r0] = (KT)0; K[r0 + 1] = (KT)0; K[r0 + 2] = (KT)0;
    K[r1] = (KT)0; K[r1 + 1] = (KT)0;
    K[r2] = (KT)0;
}

}  // namespace ceos
