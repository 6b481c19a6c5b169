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
// (A, B) of its q-th upper neighbour B), enumerated in CSR order: block blk belongs to node
// blknode[blk] and is its q = blk - blkstart[A]-th block.  The thread sums the contributing
// element blocks and writes its 3 pieces of the rows 3A..3A+2; consecutive threads write
// consecutive pieces of the same rows.
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
    long long blk0, blk1;    // the launch covers blocks [blk0, blk1)
    const int* blknode;      // [nblocks] node A of block
    const int* rowstart;     // [nn] slot of (3A, 3A)
    const int* nup;          // [nn]
    const int* blkstart;     // [nn] global block index of the diagonal block of A
    const int* cptr;         // [nblocks+1] contribution ranges per block
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
// records a node block has 1.5 contributions on average (80^3 cube): batches of 2 measured
// 0.397 ms per gather, 4 0.402 ms, 8 0.477 ms.
constexpr int kGatherBatch = 2;

template <typename KT>
__global__ __launch_bounds__(256) void k_gather_blocks(GatherParams p) {
    const long long blk = p.blk0 + (long long)blockIdx.x * 256 + threadIdx.x;
    if (blk >= p.blk1) return;
    const int c0 = p.cptr[blk], c1 = p.cptr[blk + 1];
    if (c0 == c1) return;  // a complete block: its patch's workgroup wrote it (and Fint)
    const int A = p.blknode[blk];
    const int q = (int)(blk - p.blkstart[A]);
    const int nu = p.nup[A];
    const long r0 = p.rowstart[A], r1 = r0 + 3 + 3 * nu, r2 = r1 + 2 + 3 * nu;
    KT* K = (KT*)p.K;
    if (q == 0) {
        // diagonal block, stored upper 00 01 02 11 12 22, preceded by the lane's Fe
        double a[6] = {0, 0, 0, 0, 0, 0}, f[3] = {0, 0, 0};
        for (int c = c0; c < c1; c += kGatherBatch) {
            const double* d[kGatherBatch];
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                d[u] = (c + u < c1) ? p.rec + (long)(p.cword[c + u] & 0x7fffffffu) : nullptr;
            double v[kGatherBatch][9];
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                if (d[u]) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) v[u][k] = d[u][k - 3];
                }
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                if (d[u]) {
#pragma unroll
                    for (int k = 0; k < 6; ++k) a[k] += v[u][3 + k];
#pragma unroll
                    for (int k = 0; k < 3; ++k) f[k] += v[u][k];
                }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) p.Fint[3L * A + k] = f[k];
        if (K) {
            K[r0] = (KT)a[0]; K[r0 + 1] = (KT)a[1]; K[r0 + 2] = (KT)a[2];
            K[r1] = (KT)a[3]; K[r1 + 1] = (KT)a[4];
            K[r2] = (KT)a[5];
        }
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
                const double* d = p.rec + (long)(w[u] & 0x7fffffffu);
#pragma unroll
                for (int k = 0; k < 9; ++k) v[u][k] = d[k];
            }
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            if (w[u] == 0xffffffffu) continue;
            if (w[u] >> 31) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int k = 0; k < 3; ++k) a[3 * i + k] += v[u][3 * k + i];
            } else {
#pragma unroll
                for (int k = 0; k < 9; ++k) a[k] += v[u][k];
            }
        }
    }
    const long o = 3L * (q - 1);
    K[r0 + 3 + o] = (KT)a[0]; K[r0 + 4 + o] = (KT)a[1]; K[r0 + 5 + o] = (KT)a[2];
    K[r1 + 2 + o] = (KT)a[3]; K[r1 + 3 + o] = (KT)a[4]; K[r1 + 4 + o] = (KT)a[5];
    K[r2 + 1 + o] = (KT)a[6]; K[r2 + 2 + o] = (KT)a[7]; K[r2 + 3 + o] = (KT)a[8];
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
