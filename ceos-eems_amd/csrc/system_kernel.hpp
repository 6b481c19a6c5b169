// The macro Newton step around the element path: Dirichlet boundary conditions on R and K, the
// residual norm, and a block-Jacobi preconditioned conjugate-gradient solve of K DX = -R.
//
//   residual + BC   EvaluateR / EvaluateKt (FEM/Modules/ModFEMSystemOfEquations.f90:40-93) with
//                   ApplyBoundaryConditionsNEW (FEM/BoundaryConditions/ModBoundaryConditions.f90:246-340):
//                   R = Fint - Fext; at it = 0, R += K (Ubar - U) over the prescribed DOFs (the
//                   product uses K before its rows are cleared); R(prescribed) = -(Ubar - U) at
//                   it = 0 and 0 afterwards; R(fixed) = 0; rows and columns of constrained DOFs
//                   of K cleared, their diagonal set to 1.
//   constrained-entry map: AllocatePrescDispSparseMapping / AllocateFixedSupportSparseMapping
//                   (ModBoundaryConditions.f90:344-536) search all nnz entries for every
//                   constrained DOF; here a per-DOF mask is tested while each row is streamed.
//   norm            NewtonRaphsonFull_Solve (Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:109-117):
//                   1 = L2, 2 = max |R|.
//   linear solve    the PARDISO call of the Newton step (Tools/MathTools/LinearSolver/ModPardisoSolver.f90:343-380)
//                   is replaced by PCG on the device.
//
// Symmetric matrix access.  K is stored as its upper CSR triangle in node blocks (rows 3A..3A+2:
// the diagonal block's upper part, then one 3x3 block per upper neighbour B > A, in rank order).
// For every node the neighbour slots (all B adjacent to A, ascending, A itself included) are
// described by `src` words: (owner << 7) | rank, owner = the node whose rows store the block
// (A for B >= A, B for B < A, stored transposed), rank 0 = the diagonal block, rank q >= 1 the
// q-th upper block of the owner; kSlotPad marks an unused slot.
//
// SELL copy for the PCG: nodes in chunks of 64 (one wavefront); chunk c stores its `width`
// slots, 9 entries each, 64 lanes innermost: value ((c*width + j)*9 + 3a + b)*64 + lane holds
// entry (3A + a, 3 nbr + b) of slot j of node A = 64 c + lane.  A wavefront reads 64
// consecutive values per (slot, entry) and a workgroup streams 4 contiguous chunk records;
// neighbour ids likewise, (c*width + j)*64 + lane.  Built once per Newton iteration from the
// BC-applied CSR, FP64 or FP32 (vectors, products and dot products stay FP64).
#pragma once
#include <stdint.h>

namespace ceos {

constexpr uint8_t kBcPresc = 1, kBcFixed = 2;
constexpr uint32_t kSlotPad = 0xffffffffu;
constexpr int kSysBlock = 256;
constexpr int kCgSlots = 3;    // neighbour slots per load batch of the FP32 product (1, 2, 3, 9, 27
                               // measured 0.163, 0.156, 0.154, 0.180, 0.262 ms per 80^3 PCG iteration)
constexpr int kCgSlots64 = 3;  // and of the FP64 one
// workgroups of k_cg_update and k_cg_pupdate (grid-stride over the nodes): every workgroup of the
// next kernel sums all their partials, so fewer partials cut that re-read traffic
constexpr int kCgUpdBlocks = 1024;

struct SysGeom {
    int nn;
    int width;                 // neighbour slots per node
    const int* rowstart;       // [nn] CSR offset of row 3A
    const int* nup;            // [nn] upper neighbours of A
    const int* up_ptr;         // [nn + 1]
    const int* up_nodes;       // upper neighbours, ascending per node
    const int* slot_nbr;       // [width][nn] neighbour node of every slot (A for padding)
    const uint32_t* slot_src;  // [width][nn]
    const uint8_t* dof_mask;   // [3 nn] kBcPresc | kBcFixed
    const uint8_t* node_mask;  // [nn] OR of the node's three DOF masks
    const int* sell_nbr;       // [chunks][width][64] neighbour ids in the SELL order
};

constexpr int kSellChunk = 64;

__device__ inline long sell_val(int width, int A, int j, int m) {
    return (((long)(A >> 6) * width + j) * 9 + m) * kSellChunk + (A & 63);
}

// block (rows: DOFs of A, columns: DOFs of the slot's neighbour) of the symmetric matrix
template <typename KT>
__device__ inline void csr_block(const KT* __restrict__ K, const SysGeom& g, int A, uint32_t src, double b[9]) {
    const int owner = (int)(src >> 7), q = (int)(src & 127u);
    const long r0 = g.rowstart[owner];
    const long nu = g.nup[owner];
    const long r1 = r0 + 3 + 3 * nu, r2 = r1 + 2 + 3 * nu;
    if (q == 0) {
        b[0] = (double)K[r0]; b[1] = (double)K[r0 + 1]; b[2] = (double)K[r0 + 2];
        b[4] = (double)K[r1]; b[5] = (double)K[r1 + 1];
        b[8] = (double)K[r2];
        b[3] = b[1]; b[6] = b[2]; b[7] = b[5];
        return;
    }
    const long o = 3L * (q - 1);
    double t[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t[k] = (double)K[r0 + 3 + o + k];
        t[3 + k] = (double)K[r1 + 2 + o + k];
        t[6 + k] = (double)K[r2 + 1 + o + k];
    }
    if (owner == A) {
#pragma unroll
        for (int k = 0; k < 9; ++k) b[k] = t[k];
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) b[3 * i + k] = t[3 * k + i];
    }
}

// block-level reduction of (sum, max) pairs in a fixed order; thread 0 writes part[blockIdx]
__device__ inline void block_sum_max(double s, double m, double* part_s, double* part_m) {
    __shared__ double ls[kSysBlock], lm[kSysBlock];
    ls[threadIdx.x] = s;
    lm[threadIdx.x] = m;
    __syncthreads();
    for (int w = kSysBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ls[threadIdx.x] += ls[threadIdx.x + w];
            lm[threadIdx.x] = fmax(lm[threadIdx.x], lm[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (part_s) part_s[blockIdx.x] = ls[0];
        if (part_m) part_m[blockIdx.x] = lm[0];
    }
}

// block-level reduction of two sums in a fixed order; thread 0 writes pa/pb[blockIdx]
__device__ inline void block_sum2(double a, double b, double* pa, double* pb) {
    __shared__ double la[kSysBlock], lb[kSysBlock];
    la[threadIdx.x] = a;
    lb[threadIdx.x] = b;
    __syncthreads();
    for (int w = kSysBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            la[threadIdx.x] += la[threadIdx.x + w];
            lb[threadIdx.x] += lb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pa[blockIdx.x] = la[0];
        pb[blockIdx.x] = lb[0];
    }
}

// sums (and maxima) of n partials by one workgroup, fixed order: thread t takes t, t+256, ...
__device__ inline void reduce_partials(const double* part_s, const double* part_m, int n, double* out_s,
                                       double* out_m) {
    __shared__ double ls[kSysBlock], lm[kSysBlock];
    double s = 0.0, m = 0.0;
    for (int i = threadIdx.x; i < n; i += kSysBlock) {
        if (part_s) s += part_s[i];
        if (part_m) m = fmax(m, part_m[i]);
    }
    ls[threadIdx.x] = s;
    lm[threadIdx.x] = m;
    __syncthreads();
    for (int w = kSysBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ls[threadIdx.x] += ls[threadIdx.x + w];
            lm[threadIdx.x] = fmax(lm[threadIdx.x], lm[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (out_s) *out_s = ls[0];
    if (out_m) *out_m = lm[0];
}

// R of the Newton step with the boundary conditions, one thread per node; partial sums of R^2
// and maxima of |R| per workgroup.  rmod != 0: add K (Ubar - U) over the prescribed DOFs.
template <typename KT>
__global__ __launch_bounds__(kSysBlock) void k_bc_residual(SysGeom g, const KT* __restrict__ K, int rmod,
                                                           const double* __restrict__ Fint,
                                                           const double* __restrict__ Fext,
                                                           const double* __restrict__ Ubar,
                                                           const double* __restrict__ U, double* __restrict__ R,
                                                           double* part_s, double* part_m) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    double ss = 0.0, mm = 0.0;
    if (A < g.nn) {
        double r[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) r[i] = Fint[3L * A + i] - (Fext ? Fext[3L * A + i] : 0.0);
        if (rmod) {
            for (int j = 0; j < g.width; ++j) {
                const uint32_t src = g.slot_src[(long)j * g.nn + A];
                if (src == kSlotPad) continue;
                const int B = g.slot_nbr[(long)j * g.nn + A];
                if (!(g.node_mask[B] & kBcPresc)) continue;
                double ud[3];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    ud[k] = (g.dof_mask[3L * B + k] & kBcPresc) ? Ubar[3L * B + k] - U[3L * B + k] : 0.0;
                double b[9];
                csr_block(K, g, A, src, b);
#pragma unroll
                for (int i = 0; i < 3; ++i) r[i] += b[3 * i] * ud[0] + b[3 * i + 1] * ud[1] + b[3 * i + 2] * ud[2];
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const uint8_t m = g.dof_mask[3L * A + i];
            if (m & kBcFixed) r[i] = 0.0;
            else if (m & kBcPresc) r[i] = rmod ? -(Ubar[3L * A + i] - U[3L * A + i]) : 0.0;
            R[3L * A + i] = r[i];
            ss += r[i] * r[i];
            mm = fmax(mm, fabs(r[i]));
        }
    }
    block_sum_max(ss, mm, part_s, part_m);
}

// norm of R: out[0] = L2, out[1] = max |R|, out[2] = out[norm_type - 1]
__global__ __launch_bounds__(kSysBlock) void k_norm_final(const double* part_s, const double* part_m, int n,
                                                          int norm_type, double* out) {
    double s, m;
    reduce_partials(part_s, part_m, n, &s, &m);
    if (threadIdx.x == 0) {
        out[0] = sqrt(s);
        out[1] = m;
        out[2] = norm_type == 1 ? out[0] : out[1];
    }
}

// rows and columns of constrained DOFs of the upper CSR: 0, diagonal 1; one thread per node
template <typename KT>
__global__ __launch_bounds__(kSysBlock) void k_bc_matrix(SysGeom g, KT* __restrict__ K) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    if (A >= g.nn) return;
    const long r0 = g.rowstart[A];
    const long nu = g.nup[A];
    const long rs[3] = {r0, r0 + 3 + 3 * nu, r0 + 3 + 3 * nu + 2 + 3 * nu};
    const uint8_t ma[3] = {g.dof_mask[3L * A], g.dof_mask[3L * A + 1], g.dof_mask[3L * A + 2]};
    // diagonal block: row i, columns i..2
#pragma unroll
    for (int i = 0; i < 3; ++i)
        for (int k = i; k < 3; ++k)
            if (ma[i] | ma[k]) K[rs[i] + (k - i)] = (KT)((k == i) ? 1.0 : 0.0);
    for (int q = 0; q < nu; ++q) {
        const int B = g.up_nodes[g.up_ptr[A] + q];
        const uint8_t nb = g.node_mask[B];
        if (!(nb | ma[0] | ma[1] | ma[2])) continue;
        uint8_t mb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) mb[k] = g.dof_mask[3L * B + k];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const long base = rs[i] + (3 - i) + 3L * q;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (ma[i] | mb[k]) K[base + k] = (KT)0.0;
        }
    }
}

// SELL copy of the symmetric matrix, one thread per (slot, node)
template <typename KT, typename VT>
__global__ __launch_bounds__(kSysBlock) void k_build_ell(SysGeom g, const KT* __restrict__ K, VT* __restrict__ val) {
    const long t = (long)blockIdx.x * kSysBlock + threadIdx.x;
    if (t >= (long)g.width * g.nn) return;
    const int j = (int)(t / g.nn), A = (int)(t - (long)j * g.nn);
    const uint32_t src = g.slot_src[t];
    double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (src != kSlotPad) csr_block(K, g, A, src, b);
#pragma unroll
    for (int m = 0; m < 9; ++m) val[sell_val(g.width, A, j, m)] = (VT)b[m];
}

// inverse of every node's diagonal block (symmetric, stored 00 01 02 11 12 22 as [6][nn]).
// The preconditioner must be positive definite (MINRES needs it; PCG breaks down without it):
// a block that is not (leading minors not all positive) is replaced by the inverse of its
// absolute diagonal, a zero diagonal entry by 1.
template <typename KT>
__global__ __launch_bounds__(kSysBlock) void k_block_jacobi(SysGeom g, const KT* __restrict__ K, double* __restrict__ Minv) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    if (A >= g.nn) return;
    double b[9];
    csr_block(K, g, A, (uint32_t)A << 7, b);
    const double c00 = b[4] * b[8] - b[5] * b[7];
    const double c01 = b[2] * b[7] - b[1] * b[8];
    const double c02 = b[1] * b[5] - b[2] * b[4];
    const double det = b[0] * c00 + b[3] * c01 + b[6] * c02;
    const long nn = g.nn;
    const double m2 = b[0] * b[4] - b[1] * b[3];
    if (!(b[0] > 0.0 && m2 > 0.0 && det > 0.0) || !isfinite(det)) {
        const double d0 = fabs(b[0]), d1 = fabs(b[4]), d2 = fabs(b[8]);
        Minv[A] = d0 > 0.0 ? 1.0 / d0 : 1.0; Minv[nn + A] = 0.0; Minv[2 * nn + A] = 0.0;
        Minv[3 * nn + A] = d1 > 0.0 ? 1.0 / d1 : 1.0; Minv[4 * nn + A] = 0.0;
        Minv[5 * nn + A] = d2 > 0.0 ? 1.0 / d2 : 1.0;
        return;
    }
    const double id = 1.0 / det;
    Minv[A] = c00 * id;
    Minv[nn + A] = c01 * id;
    Minv[2 * nn + A] = c02 * id;
    Minv[3 * nn + A] = (b[0] * b[8] - b[2] * b[6]) * id;
    Minv[4 * nn + A] = (b[2] * b[3] - b[0] * b[5]) * id;
    Minv[5 * nn + A] = (b[0] * b[4] - b[1] * b[3]) * id;
}

// PCG scalars on the device.  r.z of iteration k lives in rz[k & 1]: the kernel that forms
// r.z of iteration k + 1 reads rz[k & 1] in every workgroup while workgroup 0 writes rz[(k+1) & 1].
struct CgScalars {
    double rz[2];
    double alpha, beta;
    double rr, bb;   // r . r, b . b
    double tol2;     // rtol^2
    int done;        // 1 = converged, 2 = breakdown (p.Kp <= 0), 3 = iteration limit
    int iters;
    int maxit;
    int pad;
};

__device__ inline void precond(const double* __restrict__ Minv, long nn, int A, const double r[3], double z[3]) {
    const double m00 = Minv[A], m01 = Minv[nn + A], m02 = Minv[2 * nn + A];
    const double m11 = Minv[3 * nn + A], m12 = Minv[4 * nn + A], m22 = Minv[5 * nn + A];
    z[0] = m00 * r[0] + m01 * r[1] + m02 * r[2];
    z[1] = m01 * r[0] + m11 * r[1] + m12 * r[2];
    z[2] = m02 * r[0] + m12 * r[1] + m22 * r[2];
}

// the PCG state as one value for the whole workgroup.  Workgroup 0 of k_cg_update and
// k_cg_pupdate may set `done` while other workgroups of the same launch start: a workgroup then
// sees either value, never a mix, and both are consistent (the skipped work is unused).
__device__ inline bool cg_stopped(const CgScalars* c) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = __atomic_load_n(&c->done, __ATOMIC_RELAXED);
    __syncthreads();
    return s_done != 0;
}

// two sums of n partials by one workgroup in a fixed order (every workgroup that calls it gets
// bitwise the same values)
__device__ inline void reduce2(const double* pa, const double* pb, int n, double& a, double& b) {
    __shared__ double la[kSysBlock], lb[kSysBlock];
    double s = 0.0, q = 0.0;
    for (int i = threadIdx.x; i < n; i += kSysBlock) {
        s += pa[i];
        if (pb) q += pb[i];
    }
    la[threadIdx.x] = s;
    lb[threadIdx.x] = q;
    __syncthreads();
    for (int w = kSysBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            la[threadIdx.x] += la[threadIdx.x + w];
            lb[threadIdx.x] += lb[threadIdx.x + w];
        }
        __syncthreads();
    }
    a = la[0];
    b = lb[0];
}

// x = 0, r = b, z = p = M^-1 b; partials of r.z and r.r.  TWOLEVEL: x = 0, r = b only, and the
// stop flag cleared (the coarse correction, coarse_kernel.hpp, forms z, p and the partials)
template <bool TWOLEVEL = false>
__global__ __launch_bounds__(kSysBlock) void k_cg_init(int nn, const double* __restrict__ Minv, const double* __restrict__ b,
                                                       double* x, double* r, double* z, double* p,
                                                       double* part_s, double* part_m, CgScalars* c = nullptr) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    if constexpr (TWOLEVEL) {
        if (A == 0) c->done = 0;
        if (A < nn) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                x[3L * A + i] = 0.0;
                r[3L * A + i] = b[3L * A + i];
            }
        }
        return;
    }
    double rz = 0.0, rr = 0.0;
    if (A < nn) {
        double rv[3], zv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) rv[i] = b[3L * A + i];
        precond(Minv, nn, A, rv, zv);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[3L * A + i] = 0.0;
            r[3L * A + i] = rv[i];
            z[3L * A + i] = zv[i];
            p[3L * A + i] = zv[i];
            rz += rv[i] * zv[i];
            rr += rv[i] * rv[i];
        }
    }
    block_sum2(rz, rr, part_s, part_m);
}

__global__ __launch_bounds__(kSysBlock) void k_cg_init_final(const double* part_s, const double* part_m, int n,
                                                             double tol2, int maxit, CgScalars* c) {
    double rz, rr;
    reduce2(part_s, part_m, n, rz, rr);
    if (threadIdx.x == 0) {
        c->tol2 = tol2;
        c->maxit = maxit;
        c->rz[0] = rz;
        c->rr = c->bb = rr;
        c->beta = 0.0;
        c->alpha = 0.0;
        c->iters = 0;
        c->done = (rr == 0.0) ? 1 : 0;
        if (!c->done && c->maxit <= 0) c->done = 3;
    }
}

// the outcome of a PCG run for the host: {iterations, relative residual |r| / |b|, flag}
__global__ void k_cg_report(const CgScalars* c, double* out) {
    if (threadIdx.x == 0) {
        out[0] = (double)c->iters;
        out[1] = c->bb > 0.0 ? sqrt(c->rr / c->bb) : 0.0;
        out[2] = (double)c->done;
    }
}

// q = K p, partial sums of p.q.  The neighbour slots are taken G at a time: their ids, the
// G x 9 matrix values and the gathered p are all requested before any is used, so a thread
// keeps G x 10 + G x 3 loads in flight instead of one dependent chain per slot.
template <typename VT, int G>
__global__ __launch_bounds__(kSysBlock) void k_cg_spmv(SysGeom g, const VT* __restrict__ val, const CgScalars* c,
                                                       const double* __restrict__ p, double* __restrict__ q,
                                                       double* part_s) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    // the first slot batch is requested before the stop-flag read (a thread past nn reads node
    // 0's slots and discards them)
    const int Al = A < g.nn ? A : 0;
    const int* nb = g.sell_nbr + (long)(Al >> 6) * g.width * kSellChunk + (Al & 63);
    const VT* vb = val + sell_val(g.width, Al, 0, 0);
    int B[G];
    VT v[G][9];
    auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < G; ++u) B[u] = nb[(j0 + u) * kSellChunk];
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int m = 0; m < 9; ++m) v[u][m] = __builtin_nontemporal_load(vb + ((j0 + u) * 9 + m) * kSellChunk);
    };
    if (G <= g.width) load(0);
    if (cg_stopped(c)) return;
    double pq = 0.0;
    if (A < g.nn) {
        double acc[3] = {0.0, 0.0, 0.0};
        int j0 = 0;
        for (; j0 + G <= g.width; j0 += G) {
            if (j0 > 0) load(j0);
            double pb[G][3];
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int k = 0; k < 3; ++k) pb[u][k] = p[3L * B[u] + k];
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    acc[i] += (double)v[u][3 * i] * pb[u][0] + (double)v[u][3 * i + 1] * pb[u][1] +
                              (double)v[u][3 * i + 2] * pb[u][2];
        }
        for (int j = j0; j < g.width; ++j) {
            const int Bj = nb[j * kSellChunk];
            const VT* vj = vb + j * 9 * kSellChunk;
#pragma unroll
            for (int i = 0; i < 3; ++i)
                acc[i] += (double)vj[(3 * i) * kSellChunk] * p[3L * Bj] + (double)vj[(3 * i + 1) * kSellChunk] * p[3L * Bj + 1] +
                          (double)vj[(3 * i + 2) * kSellChunk] * p[3L * Bj + 2];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            q[3L * A + i] = acc[i];
            pq += p[3L * A + i] * acc[i];
        }
    }
    block_sum_max(pq, 0.0, part_s, nullptr);
}

// alpha = r.z / p.q (p.q reduced by every workgroup alike); x += alpha p, r -= alpha q,
// z = M^-1 r; partials of r.z and r.r.  TWOLEVEL: x and r only (the coarse correction forms z and
// the partials)
template <bool TWOLEVEL = false>
__global__ __launch_bounds__(kSysBlock) void k_cg_update(int nn, int par, const double* __restrict__ Minv,
                                                         CgScalars* c, const double* pq_part, int nblk,
                                                         const double* __restrict__ p, const double* __restrict__ q,
                                                         double* x, double* r, double* z, double* part_s,
                                                         double* part_m) {
    // the first node's operands are requested first: their latency overlaps the stop-flag read
    // and the reduction of p.q (a thread past nn loads node 0 and discards it)
    const int stride = gridDim.x * kSysBlock;
    int A = blockIdx.x * kSysBlock + threadIdx.x;
    double pv[3], qv[3], xv[3], rv0[3], mi[6];
    auto load = [&](int n) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            pv[i] = p[3L * n + i];
            qv[i] = q[3L * n + i];
            xv[i] = x[3L * n + i];
            rv0[i] = r[3L * n + i];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) mi[k] = Minv[(long)k * nn + n];
    };
    load(A < nn ? A : 0);
    if (cg_stopped(c)) return;
    double pq, unused;
    reduce2(pq_part, nullptr, nblk, pq, unused);
    if (!(pq > 0.0)) {  // the same verdict in every workgroup
        if (blockIdx.x == 0 && threadIdx.x == 0) __atomic_store_n(&c->done, 2, __ATOMIC_RELAXED);
        return;
    }
    const double alpha = c->rz[par] / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->alpha = alpha;
    double rz = 0.0, rr = 0.0;
    for (bool first = true; A < nn; A += stride, first = false) {
        if (!first) load(A);
        double rv[3], zv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[3L * A + i] = xv[i] + alpha * pv[i];
            rv[i] = rv0[i] - alpha * qv[i];
            r[3L * A + i] = rv[i];
        }
        if constexpr (TWOLEVEL) continue;
        zv[0] = mi[0] * rv[0] + mi[1] * rv[1] + mi[2] * rv[2];  // precond() on the loaded inverse
        zv[1] = mi[1] * rv[0] + mi[3] * rv[1] + mi[4] * rv[2];
        zv[2] = mi[2] * rv[0] + mi[4] * rv[1] + mi[5] * rv[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            z[3L * A + i] = zv[i];
            rz += rv[i] * zv[i];
            rr += rv[i] * rv[i];
        }
    }
    if constexpr (!TWOLEVEL) block_sum2(rz, rr, part_s, part_m);
}

// beta = r.z (new) / r.z (old), the convergence test, p = z + beta p; workgroup 0 records the
// scalars.  A workgroup that starts after workgroup 0 has set `done` skips the p update, which
// is then no longer used.
__global__ __launch_bounds__(kSysBlock) void k_cg_pupdate(int nn, int par, CgScalars* c, const double* rz_part,
                                                          const double* rr_part, int nblk,
                                                          const double* __restrict__ z, double* __restrict__ p) {
    // z and p of the first node are requested before the stop-flag read and the reductions
    const int stride = gridDim.x * kSysBlock;
    int A = blockIdx.x * kSysBlock + threadIdx.x;
    double zv[3], pv[3];
    auto load = [&](int n) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            zv[i] = z[3L * n + i];
            pv[i] = p[3L * n + i];
        }
    };
    load(A < nn ? A : 0);
    if (cg_stopped(c)) return;
    double rz_new, rr;
    reduce2(rz_part, rr_part, nblk, rz_new, rr);
    const double beta = rz_new / c->rz[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->rz[par ^ 1] = rz_new;
        c->beta = beta;
        c->rr = rr;
        c->iters += 1;
        if (rr <= c->tol2 * c->bb) __atomic_store_n(&c->done, 1, __ATOMIC_RELAXED);
        else if (c->iters >= c->maxit) __atomic_store_n(&c->done, 3, __ATOMIC_RELAXED);
    }
    for (bool first = true; A < nn; A += stride, first = false) {
        if (!first) load(A);
#pragma unroll
        for (int i = 0; i < 3; ++i) p[3L * A + i] = zv[i] + beta * pv[i];
    }
}

// ---------------------------------------------------------------------------------------------
// Preconditioned MINRES (Paige & Saunders 1975, in the form of the recurrences below) for the
// symmetric indefinite tangents on which PCG breaks down.  M is the block-Jacobi operator above
// (positive definite).  Per iteration k, with z = M^-1 r2 and beta = sqrt(r2 . z) from before:
//   k_mr_spmv:    v = z / beta, y = K v - (beta / beta_old) r1, partials of v . y (= alpha)
//   k_mr_lanczos: r_new = y - (alpha / beta) r2, z = M^-1 r_new, partials of r_new . z
//   k_mr_update:  beta_new = sqrt(r_new . z), the next Givens rotation, w = (v - eps w_{k-2} -
//                 delta w_{k-1}) / gamma, x += phi w; stop when phibar <= rtol beta_1 (the
//                 residual in the M^-1 norm).
// r1, r2, r_new rotate through 3 buffers; w_k overwrites w_{k-2}, so w takes 2.  The rotation
// state of iteration k is st[k & 1]: every workgroup of k_mr_update reads st[k & 1] and
// workgroup 0 writes st[(k + 1) & 1].  k_mr_update updates x, so it must not stop itself
// half-way: it leaves its verdict in `verdict`, and the next k_mr_spmv (which writes no
// solution) turns it into `done`.
// ---------------------------------------------------------------------------------------------
struct MrState {
    double beta, oldb;            // beta_k, beta_{k-1}
    double cs, sn, dbar, epsln;   // rotation carried into iteration k
    double phibar;
    double pad;
};

struct MrScalars {
    MrState st[2];
    double alpha;      // alpha of the current iteration (k_mr_lanczos, read by k_mr_update)
    double beta1;      // sqrt(b . M^-1 b)
    double rtol;
    int done;          // 1 converged, 2 preconditioner not positive (r . z < 0), 3 iteration limit
    int verdict;       // the same codes, set by k_mr_update, effective from the next k_mr_spmv
    int iters;
    int maxit;
};

__device__ inline bool mr_stopped(const MrScalars* c) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = __atomic_load_n(&c->done, __ATOMIC_RELAXED);
    __syncthreads();
    return s_done != 0;
}

// x = 0, r2 = b, z = M^-1 b, w's = 0; partials of b . z
__global__ __launch_bounds__(kSysBlock) void k_mr_init(int nn, const double* __restrict__ Minv, const double* __restrict__ b,
                                                       double* x, double* r1, double* r2, double* z, double* w1,
                                                       double* w2, double* part) {
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    double bz = 0.0;
    if (A < nn) {
        double rv[3], zv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) rv[i] = b[3L * A + i];
        precond(Minv, nn, A, rv, zv);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[3L * A + i] = 0.0;
            r1[3L * A + i] = rv[i];  // unused at k = 0 (beta / beta_old = 0)
            r2[3L * A + i] = rv[i];
            z[3L * A + i] = zv[i];
            w1[3L * A + i] = 0.0;
            w2[3L * A + i] = 0.0;
            bz += rv[i] * zv[i];
        }
    }
    block_sum_max(bz, 0.0, part, nullptr);
}

__global__ __launch_bounds__(kSysBlock) void k_mr_init_final(const double* part, int n, MrScalars* c) {
    double bz, unused;
    reduce2(part, nullptr, n, bz, unused);
    if (threadIdx.x == 0) {
        MrState& s0 = c->st[0];
        c->iters = 0;
        c->verdict = 0;
        if (!(bz >= 0.0)) {
            c->done = 2;
            return;
        }
        const double b1 = sqrt(bz);
        c->beta1 = b1;
        s0.beta = b1;
        s0.oldb = 0.0;
        s0.cs = -1.0;
        s0.sn = 0.0;
        s0.dbar = 0.0;
        s0.epsln = 0.0;
        s0.phibar = b1;
        c->done = (b1 == 0.0) ? 1 : 0;
        if (!c->done && c->maxit <= 0) c->done = 3;
    }
}

template <typename VT, int G>
__global__ __launch_bounds__(kSysBlock) void k_mr_spmv(SysGeom g, const VT* __restrict__ val, MrScalars* c, int par,
                                                       const double* __restrict__ z, const double* __restrict__ r1,
                                                       double* __restrict__ v, double* __restrict__ y, double* part) {
    if (mr_stopped(c)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && c->verdict) __atomic_store_n(&c->done, c->verdict, __ATOMIC_RELAXED);
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    const MrState& s = c->st[par];
    const double ib = 1.0 / s.beta;
    const double f = s.oldb > 0.0 ? s.beta / s.oldb : 0.0;
    double vy = 0.0;
    if (A < g.nn) {
        double acc[3] = {0.0, 0.0, 0.0};
        const int* nb = g.sell_nbr + (long)(A >> 6) * g.width * kSellChunk + (A & 63);
        const VT* vb = val + sell_val(g.width, A, 0, 0);
        int j0 = 0;
        for (; j0 + G <= g.width; j0 += G) {
            int B[G];
            VT m[G][9];
#pragma unroll
            for (int u = 0; u < G; ++u) B[u] = nb[(j0 + u) * kSellChunk];
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int q = 0; q < 9; ++q) m[u][q] = __builtin_nontemporal_load(vb + ((j0 + u) * 9 + q) * kSellChunk);
            double pb[G][3];
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int k = 0; k < 3; ++k) pb[u][k] = z[3L * B[u] + k];
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    acc[i] += (double)m[u][3 * i] * pb[u][0] + (double)m[u][3 * i + 1] * pb[u][1] +
                              (double)m[u][3 * i + 2] * pb[u][2];
        }
        for (int j = j0; j < g.width; ++j) {
            const int B = nb[j * kSellChunk];
            const VT* m = vb + j * 9 * kSellChunk;
#pragma unroll
            for (int i = 0; i < 3; ++i)
                acc[i] += (double)m[(3 * i) * kSellChunk] * z[3L * B] + (double)m[(3 * i + 1) * kSellChunk] * z[3L * B + 1] +
                          (double)m[(3 * i + 2) * kSellChunk] * z[3L * B + 2];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double vi = z[3L * A + i] * ib;
            const double yi = acc[i] * ib - f * r1[3L * A + i];
            v[3L * A + i] = vi;
            y[3L * A + i] = yi;
            vy += vi * yi;
        }
    }
    block_sum_max(vy, 0.0, part, nullptr);
}

__global__ __launch_bounds__(kSysBlock) void k_mr_lanczos(int nn, int par, const double* __restrict__ Minv, MrScalars* c,
                                                          const double* vy_part, int nblk, const double* __restrict__ r2,
                                                          double* __restrict__ y, double* __restrict__ z, double* part) {
    if (mr_stopped(c)) return;
    double alpha, unused;
    reduce2(vy_part, nullptr, nblk, alpha, unused);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->alpha = alpha;
    const double f = alpha / c->st[par].beta;
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    double rz = 0.0;
    if (A < nn) {
        double rv[3], zv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            rv[i] = y[3L * A + i] - f * r2[3L * A + i];
            y[3L * A + i] = rv[i];
        }
        precond(Minv, nn, A, rv, zv);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            z[3L * A + i] = zv[i];
            rz += rv[i] * zv[i];
        }
    }
    block_sum_max(rz, 0.0, part, nullptr);
}

// the rotation and the solution update; w_new overwrites the buffer of w_{k-2}
__global__ __launch_bounds__(kSysBlock) void k_mr_update(int nn, int par, MrScalars* c, const double* rz_part, int nblk,
                                                         const double* __restrict__ v, const double* __restrict__ w_km1,
                                                         double* __restrict__ w_km2, double* __restrict__ x) {
    if (mr_stopped(c)) return;
    double rz, unused;
    reduce2(rz_part, nullptr, nblk, rz, unused);
    const MrState s = c->st[par];
    const double alpha = c->alpha;
    const double beta_new = rz >= 0.0 ? sqrt(rz) : 0.0;
    const double oldeps = s.epsln;
    const double delta = s.cs * s.dbar + s.sn * alpha;
    const double gbar = s.sn * s.dbar - s.cs * alpha;
    const double epsln = s.sn * beta_new;
    const double dbar = -s.cs * beta_new;
    double gamma = sqrt(gbar * gbar + beta_new * beta_new);
    gamma = fmax(gamma, 1e-300);
    const double cs = gbar / gamma, sn = beta_new / gamma;
    const double phi = cs * s.phibar, phibar = sn * s.phibar;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        MrState& t = c->st[par ^ 1];
        t.beta = beta_new;
        t.oldb = s.beta;
        t.cs = cs;
        t.sn = sn;
        t.dbar = dbar;
        t.epsln = epsln;
        t.phibar = phibar;
        c->iters += 1;
        if (!(rz >= 0.0)) c->verdict = 2;
        else if (phibar <= c->rtol * c->beta1 || beta_new == 0.0) c->verdict = 1;
        else if (c->iters >= c->maxit) c->verdict = 3;
    }
    const int A = blockIdx.x * kSysBlock + threadIdx.x;
    if (A < nn) {
        const double ig = 1.0 / gamma;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double w = (v[3L * A + i] - oldeps * w_km2[3L * A + i] - delta * w_km1[3L * A + i]) * ig;
            w_km2[3L * A + i] = w;
            x[3L * A + i] += phi * w;
        }
    }
}

// u += s d; if x != null, x = X + u (UpdateMeshCoordinates)
__global__ __launch_bounds__(kSysBlock) void k_newton_update(long n, double s, const double* __restrict__ d, double* u,
                                                             const double* __restrict__ X, double* x) {
    const long t = (long)blockIdx.x * kSysBlock + threadIdx.x;
    if (t >= n) return;
    const double v = u[t] + s * d[t];
    u[t] = v;
    if (x) x[t] = X[t] + v;
}

}  // namespace ceos
