// Batched periodic RVE solver: every RVE of a batch solved to equilibrium under its macro
// deformation gradient in ONE persistent launch, one workgroup per RVE at a time.
//
// Work of the reference it replaces (file:line relative to SourceCode/):
//   periodic system of equations, K_red = T^T K T, R_red = T^T R, expansion of the fluctuation
//                      FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:77-222
//   affine part U = (F - I) Y and the fixed corner
//                      MultiScale/ModMultiscaleBoundaryConditions.f90:200-289, 632-725
//   Newton (max-abs norm, relative test)  Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:74-163
//   golden-ratio cut-back of the load      FEM/ModFEMAnalysis.f90:1270-1340
//   homogenized first Piola stress        MultiScale/ModMultiscaleAnalysis.f90:140-329
//   the PARDISO solve is replaced by preconditioned CG (block Jacobi plus a coarse level of
//   aggregate translations).
//
// Formulation on the torus.  The periodic voxel mesh of side n has n^3 nodes (the master nodes
// of the reference's T): node (i,j,k) = i + n (j + n k); element (i,j,k) joins the nodes
// ((i,j,k) + delta_a) mod n, delta_a the hex8 corner offsets.  Unknown: the fluctuation w of
// every torus node, the displacement of element corner a being U_a = (F - I) Y_a + w(node_a)
// with Y_a the unwrapped corner position (i,j,k) + delta_a times h = 1/n.  Assembling the
// element blocks on the torus is T^T K T; node 0 (the 8 cube corners) has w = 0 (identity rows).
//
// Storage of the reduced tangent, per RVE being solved: the 27 neighbour blocks of every node,
// M[(o*9 + 3a + b) * nn + i] = K(3i+a, 3 nbr(i,o)+b), o = (dx+1) + 3(dy+1) + 9(dz+1).
// The assembly colours the voxels by coordinate parity (8 colours, n even), so the elements of
// one colour share no node and their read-add-write of M and R needs no atomics.
//
// Workgroup of 256 threads = 4 waves.  Assembly: 32 elements per round, 8 lanes each, the same
// three phases as the element kernel (nodes -> LDS; lane = Gauss point; lane = node).
// CG: p lives in LDS (n <= 16: 98 KB), everything else in the workgroup's slice of HBM scratch;
// thread t handles nodes t, t+256, ...; dot products reduce over the workgroup.
#pragma once
#include "element_kernel.hpp"

namespace ceos {

constexpr int kRveThreads = 256;
constexpr int kRveWaves = kRveThreads / 64;
constexpr int kRveElemPerRound = 32;          // assembly: threads 0..255, 8 lanes per element
constexpr int kRveAsmThreads = 8 * kRveElemPerRound;
constexpr int kRveMaxSide = 16;
// Coarse level of the CG preconditioner: the three translations of every aggregate of
// kRveAgg^3 torus nodes (the last aggregate along an axis is shorter when kRveAgg does not
// divide n), node 0 left out.  At most 4^3 aggregates: 192 coarse unknowns.
constexpr int kRveAgg = 4;
constexpr int kRveMaxAggSide = (kRveMaxSide + kRveAgg - 1) / kRveAgg;
constexpr int kRveMaxCoarse = 3 * kRveMaxAggSide * kRveMaxAggSide * kRveMaxAggSide;
constexpr int kRveCoarsePacked = kRveMaxCoarse * (kRveMaxCoarse + 1) / 2;  // 18528 doubles
constexpr int kRveMax3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
// the assembly's element slices, the coarse matrix while it is inverted (packed upper triangle,
// FP64) and, in the CG, p of the largest RVE share the LDS: 148 KB
constexpr int kRveLds = kRveMax3(kRveElemPerRound * kLdsPerElem, 3 * kRveMaxSide * kRveMaxSide * kRveMaxSide,
                                 kRveCoarsePacked);
// w 3, R 3, x 3, r 3, q 3, z 3, Minv 6, B 27, then x, r, q, z of the condensation's three right-hand
// sides (9 each, interleaved per node: [9 i + 3 s + k]; rve_vec3, right after the RVE's B)
constexpr int kRveVecPerNode = 87;
static_assert(kRveLds >= 3 * kRveMaxSide * kRveMaxSide * kRveMaxSide, "p must fit in LDS");
static_assert(kRveMaxCoarse <= kRveThreads, "one thread per coarse unknown");

struct RveParams {
    int n_rve;
    const int* order;             // RVE ids in queue order (largest first)
    const int* side;              // [n_rve]
    const long long* node_off;    // [n_rve] first node of each RVE in w
    const long long* phase_off;   // [n_rve] first voxel of each RVE in phase
    const uint8_t* phase;         // voxel phase (0 matrix, 1 inclusion), x fastest
    const double* props;          // [2][8] material constants per phase
    const double* Fbar;           // [n_rve][9] macro F, column-major (3J + i)
    double* Fconv;                // [n_rve][9] last converged macro F
    double* w;                    // [nodes][3] converged fluctuation
    double* P;                    // [n_rve][9] homogenized first Piola stress, column-major
    double* A;                    // [n_rve][81] dP/dF ([3J+i][3L+k]) or nullptr
    int* info;                    // [n_rve][4] status, Newton iterations, CG iterations, cut-backs
    void* Mbuf;                   // [slots][243 * max_nodes] of MT (double-sized slots)
    double* Vbuf;                 // [slots][kRveVecPerNode * max_nodes]
    long long max_nodes;
    int* queue;                   // next queue position (zeroed before the launch)
    double newton_rtol, newton_atol, pcg_rtol;
    double tangent_rtol;          // CG tolerance of the condensation solves
    int newton_maxit, pcg_maxit, max_cutbacks;
    int probe;                    // 1: k_rve_tangent only runs one condensation solve per RVE with the
                                  // Newton steps' CG (the tile probe); A and info are left as they are
    int inexact;                  // 1: CG tolerance follows the Newton residual (forcing term)
    int nt_tangent;               // 1: launch the kernels whose CG reads the reduced tangent with
                                  // non-temporal loads (NT = true)
    long long* prof;              // [n_rve][2] clock ticks in the Newton assemblies, in the Newton linear
                                  // solves (preconditioner set-up included), or nullptr
    float* Cbuf;                  // [slots][kRveMaxCoarse^2] inverse of the coarse matrix
    int coarse;                   // 1: two-level preconditioner (block Jacobi + aggregate translations)
    float* V;                     // [nodes][27] the last condensation's solutions v_m (m = 3L + k), per node
    int* vvalid;                  // [n_rve] 1: V holds the RVE's solutions of an earlier condensation
    int warm;                     // 1: a condensation starts from V where it is valid
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// workgroup sums of N values; every thread gets the totals, summed in a fixed order
// (red: >= kRveWaves N doubles of LDS)
template <int N>
__device__ __forceinline__ void wg_sum(double (&v)[N], double* red) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();  // the previous reduction's readers are done with red
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) red[(threadIdx.x >> 6) * N + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kRveWaves; ++w) t += red[w * N + k];
        v[k] = t;
    }
}

__device__ __forceinline__ double wg_max(double v, double* red) {
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int w = 1; w < kRveWaves; ++w) t = fmax(t, red[w]);
    return t;
}

// hex8 corner offsets (---,+--,++-,-+-,--+,+-+,+++,-++)
__device__ __forceinline__ int corner_dx(int a) { return ((a + 1) >> 1) & 1; }
__device__ __forceinline__ int corner_dy(int a) { return (a >> 1) & 1; }
__device__ __forceinline__ int corner_dz(int a) { return (a >> 2) & 1; }
__device__ __forceinline__ int nbr_offset(int a, int b) {  // o of node b seen from node a
    return (corner_dx(b) - corner_dx(a) + 1) + 3 * (corner_dy(b) - corner_dy(a) + 1) +
           9 * (corner_dz(b) - corner_dz(a) + 1);
}

template <typename MT>
struct RveCtx {
    int n, nn;
    double h;
    const uint8_t* phase;
    double* wconv;  // converged fluctuation of this RVE (global state)
    MT* M;          // reduced tangent: FP64, or FP32 storage read into FP64 products
    double *wc, *R, *x, *r, *q, *z, *Minv, *B;
    int na, nc;     // aggregates per axis, coarse unknowns (3 na^3)
    float* C;       // [nc][nc] inverse of the coarse matrix (symmetric)
};

// Element pass over the whole RVE at the macro gradient Fa: zero and assemble M (if WANT_K) and
// R.  Returns true if some element has detJ <= 1e-13.
template <int MODEL, bool WANT_K, typename MT>
__device__ bool rve_assemble(const RveCtx<MT>& c, const RveParams& p, const double* Fa, double* lds,
                             int (*snode)[8]) {
    const int tid = threadIdx.x;
    const long long nm = 243LL * c.nn;
    if constexpr (WANT_K) {
        for (long long k = tid; k < nm; k += kRveThreads) c.M[k] = 0.0;
    }
    for (int k = tid; k < 3 * c.nn; k += kRveThreads) c.R[k] = 0.0;
    bool bad = false;
    __syncthreads();
    const int n = c.n, half = n >> 1, ncol = half * half * half;
    const int lane = tid & 7, le = tid >> 3;
    double* sm = lds + le * kLdsPerElem;
    for (int color = 0; color < 8; ++color) {
        const int cx = color & 1, cy = (color >> 1) & 1, cz = color >> 2;
        for (int base = 0; base < ncol; base += kRveElemPerRound) {
            const bool worker = tid < kRveAsmThreads;  // the other threads only pass the barriers
            const int el = base + le;
            const bool active = worker && el < ncol;
            const int ex = 2 * (el % half) + cx, ey = 2 * ((el / half) % half) + cy, ez = 2 * (el / (half * half)) + cz;
            // phase 0: corner a = lane
            if (worker) {
                const int gx = ex + corner_dx(lane), gy = ey + corner_dy(lane), gz = ez + corner_dz(lane);
                const int node = active ? ((gx % n) + n * ((gy % n) + n * (gz % n))) : 0;
                snode[le][lane] = node;
                const double Y[3] = {gx * c.h, gy * c.h, gz * c.h};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double u = active ? c.wc[3 * node + i] : 0.0;
#pragma unroll
                    for (int J = 0; J < 3; ++J) u += (Fa[3 * i + J] - (i == J ? 1.0 : 0.0)) * Y[J];
                    sm[lane * 9 + i] = Y[i];
                    sm[lane * 9 + 3 + i] = Y[i] + u;
                    sm[lane * 9 + 6 + i] = u;
                }
            }
            __syncthreads();
            // phase 1: Gauss point g = lane
            const int g = lane;
            double gxs[8][3];
            double jix[9];  // J^-1 of the current configuration (phase 2 recomputes grad_x N)
            double s[6], D[21];
            double wdet = 0.0;
            if (worker) {
                double dN[8][3];
                hex8_dshape_gp(g, dN);
                double cc[8][3];
#pragma unroll
                for (int a = 0; a < 8; ++a)
#pragma unroll
                    for (int k = 0; k < 3; ++k) cc[a][k] = sm[a * 9 + k];
                double gX[8][3];
                const double detX = hex8_grad(dN, cc, gX);
                double F[3][3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double acc = 0.0;
#pragma unroll
                        for (int a = 0; a < 8; ++a) acc += gX[a][j] * sm[a * 9 + 6 + i];
                        F[i][j] = acc + (i == j ? 1.0 : 0.0);
                    }
#pragma unroll
                for (int a = 0; a < 8; ++a)
#pragma unroll
                    for (int k = 0; k < 3; ++k) cc[a][k] = sm[a * 9 + 3 + k];
                const double detx = hex8_grad(dN, cc, gxs, jix);
                bad |= active && !(detx > kDetTol && detX > kDetTol);
                wdet = detx;
                const int ph = active ? c.phase[ex + n * (ey + n * ez)] : 0;
                double nw[1];
                material<MODEL>(p.props + 8 * ph, F, 0.0, nullptr, s, D, nw);
                if (!active) {
#pragma unroll
                    for (int k = 0; k < 6; ++k) s[k] = 0.0;
#pragma unroll
                    for (int k = 0; k < 21; ++k) D[k] = 0.0;
                    wdet = 0.0;
                }
            }
            __syncthreads();  // node data no longer needed
            if (worker) {
#pragma unroll
                for (int k = 0; k < 9; ++k) sm[g * 9 + k] = jix[k];
#pragma unroll
                for (int k = 0; k < 6; ++k) sm[kOffSig + g * 6 + k] = s[k] * wdet;
                if constexpr (WANT_K) {
#pragma unroll
                    for (int k = 0; k < 21; ++k) sm[kOffD + g * 21 + k] = D[k] * wdet;
                }
            }
            __syncthreads();
            // phase 2: node b = lane; read-add-write of R and M (no other element of the colour
            // shares a node; each block slot is written by exactly one lane of the element)
            if (worker) {
                const int b = lane;
                const int npart = (b < 4) ? 4 : 3;
                double fe0, fe1, fe2, kd[6], ko[4][9];
                node_blocks<true, WANT_K>(sm, b, fe0, fe1, fe2, kd, ko);
                if (active) {
                    const int nb = snode[le][b];
                    double r0 = c.R[3 * nb], r1 = c.R[3 * nb + 1], r2 = c.R[3 * nb + 2];
                    c.R[3 * nb] = r0 + fe0;
                    c.R[3 * nb + 1] = r1 + fe1;
                    c.R[3 * nb + 2] = r2 + fe2;
                    if constexpr (WANT_K) {
                        const int nn = c.nn;
                        MT* M = c.M;
                        {
                            const double full[9] = {kd[0], kd[1], kd[2], kd[1], kd[3], kd[4], kd[2], kd[4], kd[5]};
                            double old[9];
#pragma unroll
                            for (int k = 0; k < 9; ++k) old[k] = M[(13 * 9 + k) * nn + nb];
#pragma unroll
                            for (int k = 0; k < 9; ++k) M[(13 * 9 + k) * nn + nb] = old[k] + full[k];
                        }
                        // pair (c, b): K_cb into the row of c, K_bc = K_cb^T into the row of b
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            if (j < npart) {
                                const int cn = (b + j + 1) & 7;
                                const int nc = snode[le][cn];
                                const int oc = nbr_offset(cn, b) * 9 * nn + nc;
                                const int ob = nbr_offset(b, cn) * 9 * nn + nb;
                                double o1[9], o2[9];
#pragma unroll
                                for (int k = 0; k < 9; ++k) {
                                    o1[k] = M[oc + k * nn];
                                    o2[k] = M[ob + k * nn];
                                }
#pragma unroll
                                for (int r = 0; r < 3; ++r)
#pragma unroll
                                    for (int t = 0; t < 3; ++t) {
                                        M[oc + (3 * r + t) * nn] = o1[3 * r + t] + ko[j][3 * r + t];
                                        M[ob + (3 * r + t) * nn] = o2[3 * r + t] + ko[j][3 * t + r];
                                    }
                            }
                        }
                    }
                }
            }
            __syncthreads();  // LDS and this round's writes before the next round
        }
    }
    return bad;
}

// y = M v for node i (27 neighbour blocks; v in LDS); node 0 is the fixed corner (identity row,
// its column never multiplies a nonzero since v_0 = 0).  The 81 tangent entries of one z-plane of
// neighbours are loaded before any of them is used (the scheduling barrier keeps the loads
// together), so a row costs three HBM round trips instead of one per block.
// NT: the tangent rows are read with non-temporal loads (a full batch streams far more tangents
// than the caches hold; a compile-time choice, as a runtime branch here costs the condensation
// kernel registers)
template <typename MT, bool NT>
__device__ __forceinline__ void rve_row(const RveCtx<MT>& c, int i, const double* v, double y[3]) {
    const int n = c.n, nn = c.nn;
    const int ix = i % n, iy = (i / n) % n, iz = i / (n * n);
    const int xs[3] = {ix == 0 ? n - 1 : ix - 1, ix, ix == n - 1 ? 0 : ix + 1};
    const int ys[3] = {iy == 0 ? n - 1 : iy - 1, iy, iy == n - 1 ? 0 : iy + 1};
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
        const int zz = iz + dz < 0 ? n - 1 : (iz + dz == n ? 0 : iz + dz);
        // entries (o 9 + k), o = (dx+1) + 3 (dy+1) + 9 (dz+1): this plane is q = 9 ob + k, q < 81
        const MT* mp = c.M + (long long)(81 * (dz + 1)) * nn + i;
        double m[81];
#pragma unroll
        for (int q = 0; q < 81; ++q) {
            if constexpr (NT) m[q] = (double)__builtin_nontemporal_load(mp + (long long)q * nn);
            else m[q] = (double)mp[(long long)q * nn];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int by = 0; by < 3; ++by) {
#pragma unroll
            for (int bx = 0; bx < 3; ++bx) {
                const int j = xs[bx] + n * (ys[by] + n * zz);
                const int b = bx + 3 * by;
                const double v0 = v[3 * j], v1 = v[3 * j + 1], v2 = v[3 * j + 2];
                a0 += m[9 * b] * v0 + m[9 * b + 1] * v1 + m[9 * b + 2] * v2;
                a1 += m[9 * b + 3] * v0 + m[9 * b + 4] * v1 + m[9 * b + 5] * v2;
                a2 += m[9 * b + 6] * v0 + m[9 * b + 7] * v1 + m[9 * b + 8] * v2;
            }
        }
    }
    y[0] = a0;
    y[1] = a1;
    y[2] = a2;
}

// block-Jacobi: inverse of every node's diagonal block (symmetric, upper 6 values); node 0 -> 0
template <typename MT>
__device__ void rve_precond(const RveCtx<MT>& c) {
    const long long nn = c.nn;
    for (int i = threadIdx.x; i < c.nn; i += kRveThreads) {
        double Bm[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) Bm[a][b] = c.M[(13LL * 9 + 3 * a + b) * nn + i];
        double Bi[3][3];
        inv3(Bm, det3(Bm), Bi);
        const bool fixed = i == 0;
        double* mi = c.Minv + 6 * i;
        mi[0] = fixed ? 0.0 : Bi[0][0];
        mi[1] = fixed ? 0.0 : 0.5 * (Bi[0][1] + Bi[1][0]);
        mi[2] = fixed ? 0.0 : 0.5 * (Bi[0][2] + Bi[2][0]);
        mi[3] = fixed ? 0.0 : Bi[1][1];
        mi[4] = fixed ? 0.0 : 0.5 * (Bi[1][2] + Bi[2][1]);
        mi[5] = fixed ? 0.0 : Bi[2][2];
    }
}

__device__ __forceinline__ void apply_minv(const double* mi, const double r[3], double z[3]) {
    z[0] = mi[0] * r[0] + mi[1] * r[1] + mi[2] * r[2];
    z[1] = mi[1] * r[0] + mi[3] * r[1] + mi[4] * r[2];
    z[2] = mi[2] * r[0] + mi[4] * r[1] + mi[5] * r[2];
}

// packed row-major upper triangle of an nc x nc matrix: entry (i, j), i <= j
__device__ __forceinline__ int coarse_upk(int i, int j, int nc) { return i * nc - (i * (i - 1)) / 2 + (j - i); }

// Coarse matrix A_c = P^T K P of the aggregate translations (P maps the coarse unknowns of an
// aggregate to the three unknowns of each of its nodes; node 0 and its column left out), then
// inverted.  One wave per aggregate, lane = node (lane & 3, lane >> 2 & 3, lane >> 4) of its
// block: the lane reads the 27 neighbour blocks of its node once (the same rows as the CG product)
// and sorts them into 8 classes by the axes along which the step leaves the aggregate; the wave
// then sums, for every neighbour aggregate b >= a, the classes of its lanes that land on b (in
// the fixed butterfly order of wave_sum; one aggregate is one wave's, so the adds need no
// atomics).  The packed upper triangle (FP64, Ap in LDS) is
// inverted by the symmetric sweep operator (pivot k: a_kk <- -1 / a_kk, a_ik <- a_ik / a_kk,
// a_ij <- a_ij - a_ik a_kj / a_kk), which leaves -A_c^-1 (A_c is SPD: no pivoting).  The inverse
// goes to c.C in FP32, both triangles from the same value, so the preconditioner stays exactly
// symmetric.  rk: nc doubles of LDS.  Returns false (no coarse level for this tangent) if a
// pivot is not positive.
template <typename MT>
__device__ bool rve_coarse_setup(const RveCtx<MT>& c, double* Ap, double* rk) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = c.n, na = c.na, nc = c.nc;
    const long long nn = c.nn;
    const int np = nc * (nc + 1) / 2;
    for (int e = tid; e < np; e += kRveThreads) Ap[e] = 0.0;
    __syncthreads();
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const int cnt = na >= 3 ? 3 : na, d0 = na >= 3 ? -1 : 0;  // distinct neighbour aggregates per axis
    for (int a = wave; a < na * na * na; a += kRveWaves) {
        const int ax = a % na, ay = (a / na) % na, az = a / (na * na);
        const int x0 = kRveAgg * ax, y0 = kRveAgg * ay, z0 = kRveAgg * az;
        const int wx = min(kRveAgg, n - x0), wy = min(kRveAgg, n - y0), wz = min(kRveAgg, n - z0);  // >= 2
        const bool live = lx < wx && ly < wy && lz < wz;
        const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
        const int i = live ? x + n * (y + n * z) : 0;
        // two passes, by whether the step leaves the aggregate along z (4 classes of 9 sums each
        // in registers); the second pass adds to what the first wrote
#pragma unroll 1
        for (int zp = 0; zp < 2; ++zp) {
            double acc[4][9];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int k = 0; k < 9; ++k) acc[q][k] = 0.0;
            if (i != 0) {
                // one z-plane of neighbours (9 offsets, 81 entries) requested at a time; a plane
                // belongs to this pass if the step along z stays in (zp = 0) or leaves (zp = 1)
                // the aggregate
#pragma unroll 1
                for (int dz = -1; dz <= 1; ++dz) {
                    if ((lz + dz < 0 || lz + dz >= wz ? 1 : 0) != zp) continue;
                    const int zz = (z + dz + n) % n;
                    const MT* mp = c.M + (long long)(81 * (dz + 1)) * nn + i;
                    double m[81];
#pragma unroll
                    for (int q = 0; q < 81; ++q) m[q] = (double)mp[(long long)q * nn];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int ob = 0; ob < 9; ++ob) {
                        const int dx = ob % 3 - 1, dy = ob / 3 - 1;
                        const int xx = (x + dx + n) % n, yy = (y + dy + n) % n;
                        const bool col0 = xx + n * (yy + n * zz) == 0;  // node 0's column
                        const int cls = (lx + dx < 0 || lx + dx >= wx ? 1 : 0) + (ly + dy < 0 || ly + dy >= wy ? 2 : 0);
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (q == cls && !col0) {
#pragma unroll
                                for (int k = 0; k < 9; ++k) acc[q][k] += m[9 * ob + k];
                            }
                    }
                }
            }
            // the aggregate each class lands on: the neighbour along every axis whose bit is set,
            // on the side where the lane sits (a block is at least 2 nodes wide: never both sides)
            int bt[4];
            const int bz = (az + (zp ? (lz == 0 ? -1 : 1) : 0) + na) % na;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int bx = (ax + ((q & 1) ? (lx == 0 ? -1 : 1) : 0) + na) % na;
                const int by = (ay + ((q & 2) ? (ly == 0 ? -1 : 1) : 0) + na) % na;
                bt[q] = bx + na * (by + na * bz);
            }
            for (int t = 0; t < cnt * cnt * cnt; ++t) {
                const int tx = (ax + d0 + t % cnt + na) % na, ty = (ay + d0 + (t / cnt) % cnt + na) % na,
                          tz = (az + d0 + t / (cnt * cnt) + na) % na;
                const int b = tx + na * (ty + na * tz);
                if (b < a) continue;  // the lower triangle is the transpose
                double mine = 0.0;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    double v = 0.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) v += bt[q] == b ? acc[q][k] : 0.0;
                    v = wave_sum(v);
                    if (lane == k) mine = v;
                }
                if (lane < 9) {
                    const int r = 3 * a + lane / 3, col = 3 * b + lane % 3;
                    if (r <= col) Ap[coarse_upk(r, col, nc)] += mine;
                }
            }
        }
    }
    // symmetric sweep, one thread per column j of the upper triangle: entries (i, j), i <= j, at
    // packed offsets j, j + (nc - 1), j + (nc - 1) + (nc - 2), ...; at row i the lanes of a wave
    // touch consecutive doubles
    for (int k = 0; k < nc; ++k) {
        __syncthreads();  // the previous pivot's updates (or the assembly) complete
        for (int t = tid; t < nc; t += kRveThreads) rk[t] = Ap[t <= k ? coarse_upk(t, k, nc) : coarse_upk(k, t, nc)];
        __syncthreads();
        const double piv = rk[k];
        if (!(piv > 0.0)) return false;  // the same value for every thread
        const double pinv = 1.0 / piv;
        if (tid < nc) {
            const int j = tid;
            const double rj = rk[j] * pinv;
            int e = j;
            for (int i = 0; i <= j; ++i) {
                const double v = Ap[e];
                Ap[e] = i == k ? (j == k ? -pinv : v * pinv) : (j == k ? v * pinv : v - rk[i] * rj);
                e += nc - 1 - i;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < nc * nc; e += kRveThreads) {
        const int r = e / nc, col = e - r * nc;
        c.C[e] = (float)(-Ap[r <= col ? coarse_upk(r, col, nc) : coarse_upk(col, r, nc)]);
    }
    __syncthreads();
    return true;
}

// Restriction rc = P^T r into cv[0, nc) and the coarse correction yc = A_c^-1 rc into
// cv[nc, 2 nc); returns this thread's part of rc . yc (threads t < nc: unknown t = 3 a + comp).
// r must be complete (a barrier before the call); cv[nc, 2 nc) may be read after the caller's
// next barrier.  r is zero at node 0, so the restriction needs no mask.  Both steps are chains of
// dependent global loads inside one workgroup, so each requests a batch of operands before
// summing any: the restriction one z-layer of the aggregate (16 nodes) at a time, the coarse
// product kRveCoarseBatch entries at a time (larger batches cost the CG loop registers).
constexpr int kRveCoarseBatch = 16;

template <typename MT>
__device__ double rve_coarse_apply(const RveCtx<MT>& c, double* cv) {
    const int tid = threadIdx.x, n = c.n, na = c.na, nc = c.nc;
    if (tid < nc) {
        const int a = tid / 3, comp = tid - 3 * a;
        const int x0 = kRveAgg * (a % na), y0 = kRveAgg * ((a / na) % na), z0 = kRveAgg * (a / (na * na));
        const int wx = min(kRveAgg, n - x0), wy = min(kRveAgg, n - y0), wz = min(kRveAgg, n - z0);
        double s = 0.0;
#pragma unroll 1
        for (int lz = 0; lz < wz; ++lz) {
            const double* r0 = c.r + 3LL * (x0 + n * (y0 + n * (z0 + lz))) + comp;
            // clamped addresses, masked values: the loads stay unconditional (one batch)
            double v[kRveAgg][kRveAgg];
#pragma unroll
            for (int ly = 0; ly < kRveAgg; ++ly)
#pragma unroll
                for (int lx = 0; lx < kRveAgg; ++lx) v[ly][lx] = r0[3 * (min(lx, wx - 1) + n * min(ly, wy - 1))];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int ly = 0; ly < kRveAgg; ++ly)
#pragma unroll
                for (int lx = 0; lx < kRveAgg; ++lx) s += (lx < wx && ly < wy) ? v[ly][lx] : 0.0;
        }
        cv[tid] = s;
    }
    __syncthreads();
    double part = 0.0;
    if (tid < nc) {
        const float* Cp = c.C + tid;
        double y = 0.0;
#pragma unroll 1
        for (int k0 = 0; k0 < nc; k0 += kRveCoarseBatch) {
            float cb[kRveCoarseBatch];
#pragma unroll
            for (int k = 0; k < kRveCoarseBatch; ++k) cb[k] = Cp[(long long)min(k0 + k, nc - 1) * nc];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < kRveCoarseBatch; ++k)
                if (k0 + k < nc) y += (double)cb[k] * cv[k0 + k];
        }
        cv[nc + tid] = y;
        part = cv[tid] * y;
    }
    return part;
}

// coarse unknown index 3 a of node i's aggregate
__device__ __forceinline__ int rve_agg3(int i, int n, int na) {
    const int ix = i % n, iy = (i / n) % n, iz = i / (n * n);
    return 3 * (ix / kRveAgg + na * (iy / kRveAgg + na * (iz / kRveAgg)));
}

// preconditioned CG on the reduced system: M x = sign * rhs (node 0 fixed) to
// ||r|| <= rtol ||b||.  p lives in LDS, the other vectors in the workgroup's scratch.  The vector
// sweeps take G nodes per thread at a time and load all their operands before computing (one
// round trip per group); G is chosen among kCgGroups by ceos_rve_autotune.  Preconditioner:
// block Jacobi, plus (coarse) the aggregate translations, z = D^-1 r + P A_c^-1 P^T r (additive
// two-level, SPD; r . z gains rc . yc).  cv: 2 kRveMaxCoarse doubles of LDS.  Returns the
// iterations used.
constexpr int kCgGroups[3] = {2, 4, 8};
constexpr int kCgGroupDefault = 4;

template <typename MT, int G, bool NT>
__device__ int rve_pcg(const RveCtx<MT>& c, const RveParams& p, const double* rhs, double sign, double rtol,
                       double* pl, double* red, double* cv, bool coarse) {
    const int tid = threadIdx.x;
    const int nn = c.nn;
    constexpr int S = kRveThreads * G;
    double acc[2] = {0.0, 0.0};
    for (int i0 = tid; i0 < nn; i0 += S) {
        double r[G][3], mi[G][6];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int i = i0 + g * kRveThreads;
            const int ii = i < nn ? i : 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) r[g][k] = rhs[3 * ii + k];
#pragma unroll
            for (int k = 0; k < 6; ++k) mi[g][k] = c.Minv[6 * ii + k];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int i = i0 + g * kRveThreads;
            if (i >= nn) continue;
            double rr[3], z[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) rr[k] = i == 0 ? 0.0 : sign * r[g][k];
            apply_minv(mi[g], rr, z);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                c.x[3 * i + k] = 0.0;
                c.r[3 * i + k] = rr[k];
                pl[3 * i + k] = z[k];
                acc[0] += rr[k] * z[k];
                acc[1] += rr[k] * rr[k];
            }
        }
    }
    double sums[3] = {acc[0], acc[1], 0.0};
    if (coarse) {
        __syncthreads();  // r complete
        sums[2] = rve_coarse_apply(c, cv);
    }
    wg_sum<3>(sums, red);
    double rz = sums[0] + sums[2];
    const double tol2 = rtol * rtol * sums[1];
    if (!(sums[1] > 0.0)) return 0;
    if (coarse) {
        for (int i = tid; i < nn; i += kRveThreads) {
            if (i == 0) continue;
            const int a3 = rve_agg3(i, c.n, c.na);
#pragma unroll
            for (int k = 0; k < 3; ++k) pl[3 * i + k] += cv[c.nc + a3 + k];
        }
    }
    int it = 0;
    while (it < p.pcg_maxit) {
        ++it;
        __syncthreads();  // p complete
        double pq[1] = {0.0};
        for (int i = tid; i < nn; i += kRveThreads) {
            double y[3];
            if (i == 0) {
                y[0] = y[1] = y[2] = 0.0;
            } else {
                rve_row<MT, NT>(c, i, pl, y);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                c.q[3 * i + k] = y[k];
                pq[0] += pl[3 * i + k] * y[k];
            }
        }
        wg_sum<1>(pq, red);
        if (!(pq[0] > 0.0)) break;  // not positive definite along p
        const double alpha = rz / pq[0];
        double s2[2] = {0.0, 0.0};
        for (int i0 = tid; i0 < nn; i0 += S) {
            double xv[G][3], rv[G][3], qv[G][3], mi[G][6];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int i = i0 + g * kRveThreads;
                const int ii = i < nn ? i : 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    xv[g][k] = c.x[3 * ii + k];
                    rv[g][k] = c.r[3 * ii + k];
                    qv[g][k] = c.q[3 * ii + k];
                }
#pragma unroll
                for (int k = 0; k < 6; ++k) mi[g][k] = c.Minv[6 * ii + k];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int i = i0 + g * kRveThreads;
                if (i >= nn) continue;
                double r[3], z[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c.x[3 * i + k] = xv[g][k] + alpha * pl[3 * i + k];
                    r[k] = rv[g][k] - alpha * qv[g][k];
                    c.r[3 * i + k] = r[k];
                }
                apply_minv(mi[g], r, z);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c.z[3 * i + k] = z[k];
                    s2[0] += r[k] * r[k];
                    s2[1] += r[k] * z[k];
                }
            }
        }
        wg_sum<2>(s2, red);
        if (s2[0] <= tol2) break;
        if (coarse) {
            double cz[1] = {rve_coarse_apply(c, cv)};  // r complete: wg_sum ended on a barrier
            wg_sum<1>(cz, red);
            s2[1] += cz[0];
        }
        const double beta = s2[1] / rz;
        rz = s2[1];
        for (int i0 = tid; i0 < nn; i0 += S) {
            double zv[G][3];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int i = i0 + g * kRveThreads;
                const int ii = i < nn ? i : 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) zv[g][k] = c.z[3 * ii + k];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int i = i0 + g * kRveThreads;
                if (i >= nn) continue;
                double yc[3] = {0.0, 0.0, 0.0};
                if (coarse && i != 0) {
                    const int a3 = rve_agg3(i, c.n, c.na);
#pragma unroll
                    for (int k = 0; k < 3; ++k) yc[k] = cv[c.nc + a3 + k];
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) pl[3 * i + k] = zv[g][k] + yc[k] + beta * pl[3 * i + k];
            }
        }
    }
    __syncthreads();
    return it;
}

// x, r, q, z (k = 0..3) of the condensation's three right-hand sides: 9 nn doubles each after the
// RVE's 27 nn values of B (within the slot: 24 + 27 + 36 <= kRveVecPerNode doubles per node);
// derived here rather than kept in RveCtx, whose every pointer costs both kernels registers
template <typename MT>
__device__ __forceinline__ double* rve_vec3(const RveCtx<MT>& c, int k) {
    return c.B + (27LL + 9LL * k) * c.nn;
}

// y_s = M p_s for node i and the three right-hand sides s of the condensation (p in LDS, FP32,
// [9 j + 3 s + k]): the tangent rows are read once for the three products
template <typename MT, bool NT>
__device__ __forceinline__ void rve_row3(const RveCtx<MT>& c, int i, const float* v3, double y[9]) {
    const int n = c.n, nn = c.nn;
    const int ix = i % n, iy = (i / n) % n, iz = i / (n * n);
    const int xs[3] = {ix == 0 ? n - 1 : ix - 1, ix, ix == n - 1 ? 0 : ix + 1};
    const int ys[3] = {iy == 0 ? n - 1 : iy - 1, iy, iy == n - 1 ? 0 : iy + 1};
#pragma unroll
    for (int t = 0; t < 9; ++t) y[t] = 0.0;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
        const int zz = iz + dz < 0 ? n - 1 : (iz + dz == n ? 0 : iz + dz);
        const MT* mp = c.M + (long long)(81 * (dz + 1)) * nn + i;
        double m[81];
#pragma unroll
        for (int q = 0; q < 81; ++q) {
            if constexpr (NT) m[q] = (double)__builtin_nontemporal_load(mp + (long long)q * nn);
            else m[q] = (double)mp[(long long)q * nn];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int by = 0; by < 3; ++by) {
#pragma unroll
            for (int bx = 0; bx < 3; ++bx) {
                const int j = xs[bx] + n * (ys[by] + n * zz);
                const int b = bx + 3 * by;
                const float* vj = v3 + 9 * j;
                double v[9];
#pragma unroll
                for (int t = 0; t < 9; ++t) v[t] = (double)vj[t];
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        y[3 * q + a] += m[9 * b + 3 * a] * v[3 * q] + m[9 * b + 3 * a + 1] * v[3 * q + 1] +
                                        m[9 * b + 3 * a + 2] * v[3 * q + 2];
            }
        }
    }
}

// rve_coarse_apply for the three right-hand sides: rc_s = P^T r_s into cv[s nc + u] and
// yc_s = A_c^-1 rc_s into cv[3 nc + s nc + u]; part[s] = this thread's part of rc_s . yc_s.  The
// coarse inverse is read once for the three.  Restriction: two rows of an aggregate layer (8
// nodes, 24 values) per round trip, clamped addresses and masked values.
template <typename MT>
__device__ void rve_coarse_apply3(const RveCtx<MT>& c, double* cv, double part[3]) {
    const int tid = threadIdx.x, n = c.n, na = c.na, nc = c.nc;
    if (tid < nc) {
        const int a = tid / 3, comp = tid - 3 * a;
        const int x0 = kRveAgg * (a % na), y0 = kRveAgg * ((a / na) % na), z0 = kRveAgg * (a / (na * na));
        const int wx = min(kRveAgg, n - x0), wy = min(kRveAgg, n - y0), wz = min(kRveAgg, n - z0);
        double s3[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
        for (int lz = 0; lz < wz; ++lz) {
            const double* r0 = rve_vec3(c, 1) + 9LL * (x0 + n * (y0 + n * (z0 + lz))) + comp;
#pragma unroll 1
            for (int ly0 = 0; ly0 < kRveAgg; ly0 += 2) {
                double v[2][kRveAgg][3];
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int lx = 0; lx < kRveAgg; ++lx)
#pragma unroll
                        for (int q = 0; q < 3; ++q)
                            v[dy][lx][q] = r0[9 * (min(lx, wx - 1) + n * min(ly0 + dy, wy - 1)) + 3 * q];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int lx = 0; lx < kRveAgg; ++lx)
#pragma unroll
                        for (int q = 0; q < 3; ++q) s3[q] += (lx < wx && ly0 + dy < wy) ? v[dy][lx][q] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) cv[q * nc + tid] = s3[q];
    }
    __syncthreads();
    part[0] = part[1] = part[2] = 0.0;
    if (tid < nc) {
        const float* Cp = c.C + tid;
        double y3[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
        for (int k0 = 0; k0 < nc; k0 += kRveCoarseBatch) {
            float cb[kRveCoarseBatch];
#pragma unroll
            for (int k = 0; k < kRveCoarseBatch; ++k) cb[k] = Cp[(long long)min(k0 + k, nc - 1) * nc];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < kRveCoarseBatch; ++k)
                if (k0 + k < nc) {
                    const double cf = (double)cb[k];
#pragma unroll
                    for (int q = 0; q < 3; ++q) y3[q] += cf * cv[q * nc + k0 + k];
                }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            cv[3 * nc + q * nc + tid] = y3[q];
            part[q] = cv[q * nc + tid] * y3[q];
        }
    }
}

// Three condensation solves at once: K_red v_s = -b_(m0+s), s = 0, 1, 2, each a PCG (block Jacobi
// plus the coarse level) to ||r_s|| <= rtol ||b_s||.  Every iteration streams the tangent rows once
// for the three solves, where three single solves stream them three times: the tangent is the
// dominant traffic of a CG iteration (972 of about 1,250 bytes per node in FP32 storage).  The
// three directions p_s live in LDS in FP32 (they fit where one FP64 p did); x_s and r_s are updated
// with the same rounded p_s that q_s = K p_s is formed from, so r_s stays the residual of x_s and
// the solves reach the FP64 tolerances (the rounded directions cost about 1 % more iterations).
// A converged right-hand side is frozen (alpha_s = 0) while the others go on.  x_s ends in
// rve_vec3(c, 0);
// its[s] = iterations of solve s.
template <typename MT, bool NT>
__device__ void rve_pcg3(const RveCtx<MT>& c, const RveParams& p, int m0, double rtol, float* pl, double* red,
                         double* cv, bool coarse, int its[3], const float* x0) {
    const int tid = threadIdx.x;
    const int nn = c.nn, nc = c.nc;
    const long long nn3 = 3LL * nn;
    double* x3 = rve_vec3(c, 0);
    double* r3 = rve_vec3(c, 1);
    double* q3 = rve_vec3(c, 2);
    double* z3 = rve_vec3(c, 3);
    // warm start (x0, FP32 per node [27 i + 3 m + k]): x = x0 exactly, also as the direction in LDS
    // for r0 = -b - K x0; the stopping test stays relative to ||b||
    if (x0) {
        for (int i = tid; i < nn; i += kRveThreads) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float v = i == 0 ? 0.0f : x0[27LL * i + 3 * (m0 + q) + k];
                    pl[9 * i + 3 * q + k] = v;
                    x3[9 * i + 3 * q + k] = (double)v;
                }
        }
        __syncthreads();
    }
    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // r_s . z_s, r_s . r_s, b_s . b_s
    for (int i = tid; i < nn; i += kRveThreads) {
        double b[9], mi[6], y[9];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k) b[3 * q + k] = c.B[(m0 + q) * nn3 + 3 * i + k];
#pragma unroll
        for (int k = 0; k < 6; ++k) mi[k] = c.Minv[6 * i + k];
        if (x0 && i != 0) {
            rve_row3<MT, NT>(c, i, pl, y);
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) y[t] = 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double rr[3], z[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) rr[k] = i == 0 ? 0.0 : -b[3 * q + k] - y[3 * q + k];
            apply_minv(mi, rr, z);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (!x0) {
                    x3[9 * i + 3 * q + k] = 0.0;
                    pl[9 * i + 3 * q + k] = (float)z[k];
                } else {
                    z3[9 * i + 3 * q + k] = z[k];  // p is still x0 for the other rows' products
                }
                r3[9 * i + 3 * q + k] = rr[k];
                acc[q] += rr[k] * z[k];
                acc[3 + q] += rr[k] * rr[k];
                const double bk = i == 0 ? 0.0 : b[3 * q + k];
                acc[6 + q] += bk * bk;
            }
        }
    }
    if (x0) {
        __syncthreads();  // every row's product with x0 is done
        for (int i = tid; i < nn; i += kRveThreads) {
#pragma unroll
            for (int t = 0; t < 9; ++t) pl[9 * i + t] = (float)z3[9 * i + t];
        }
    }
    double sums[12] = {acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], 0.0, 0.0, 0.0, acc[6], acc[7], acc[8]};
    if (coarse) {
        __syncthreads();  // r complete
        double pc[3];
        rve_coarse_apply3(c, cv, pc);
        sums[6] = pc[0];
        sums[7] = pc[1];
        sums[8] = pc[2];
    }
    wg_sum<12>(sums, red);
    double rz[3], tol2[3];
    bool act[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        rz[q] = sums[q] + sums[6 + q];
        tol2[q] = rtol * rtol * sums[9 + q];
        act[q] = sums[3 + q] > tol2[q];
        its[q] = 0;
    }
    if (coarse) {
        for (int i = tid; i < nn; i += kRveThreads) {
            if (i == 0) continue;
            const int a3 = rve_agg3(i, c.n, c.na);
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    pl[9 * i + 3 * q + k] = (float)((double)pl[9 * i + 3 * q + k] + cv[3 * nc + q * nc + a3 + k]);
        }
    }
    int it = 0;
    while ((act[0] || act[1] || act[2]) && it < p.pcg_maxit) {
        ++it;
        __syncthreads();  // p complete
        double pq[3] = {0.0, 0.0, 0.0};
        for (int i = tid; i < nn; i += kRveThreads) {
            double y[9];
            if (i == 0) {
#pragma unroll
                for (int t = 0; t < 9; ++t) y[t] = 0.0;
            } else {
                rve_row3<MT, NT>(c, i, pl, y);
            }
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    q3[9 * i + 3 * q + k] = y[3 * q + k];
                    pq[q] += (double)pl[9 * i + 3 * q + k] * y[3 * q + k];
                }
        }
        wg_sum<3>(pq, red);
        double alpha[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (act[q] && !(pq[q] > 0.0)) {  // not positive definite along p_s: this solve stops
                act[q] = false;
                its[q] = it;
            }
            alpha[q] = act[q] ? rz[q] / pq[q] : 0.0;
        }
        double s2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // r_s . r_s, r_s . z_s
        for (int i = tid; i < nn; i += kRveThreads) {
            double xv[9], rv[9], qv[9], mi[6];
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                xv[t] = x3[9 * i + t];
                rv[t] = r3[9 * i + t];
                qv[t] = q3[9 * i + t];
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) mi[k] = c.Minv[6 * i + k];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double r[3], z[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    x3[9 * i + 3 * q + k] = xv[3 * q + k] + alpha[q] * (double)pl[9 * i + 3 * q + k];
                    r[k] = rv[3 * q + k] - alpha[q] * qv[3 * q + k];
                    r3[9 * i + 3 * q + k] = r[k];
                }
                apply_minv(mi, r, z);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    z3[9 * i + 3 * q + k] = z[k];
                    s2[q] += r[k] * r[k];
                    s2[3 + q] += r[k] * z[k];
                }
            }
        }
        wg_sum<6>(s2, red);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (act[q] && s2[q] <= tol2[q]) {
                act[q] = false;
                its[q] = it;
            }
        if (!(act[0] || act[1] || act[2])) break;
        double cz[3] = {0.0, 0.0, 0.0};
        if (coarse) {
            rve_coarse_apply3(c, cv, cz);  // r complete: wg_sum ended on a barrier
            wg_sum<3>(cz, red);
        }
        double beta[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double rzn = s2[3 + q] + cz[q];
            beta[q] = act[q] ? rzn / rz[q] : 0.0;
            if (act[q]) rz[q] = rzn;
        }
        for (int i = tid; i < nn; i += kRveThreads) {
            double zv[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) zv[t] = z3[9 * i + t];
            __builtin_amdgcn_sched_barrier(0);
            const int a3 = (coarse && i != 0) ? rve_agg3(i, c.n, c.na) : 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (!act[q]) continue;  // frozen: its direction no longer moves x_s
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double yc = (coarse && i != 0) ? cv[3 * nc + q * nc + a3 + k] : 0.0;
                    pl[9 * i + 3 * q + k] = (float)(zv[3 * q + k] + yc + beta[q] * (double)pl[9 * i + 3 * q + k]);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
        if (act[q]) its[q] = it;  // the iteration limit
    __syncthreads();
}

// engineering strain (11,22,33,12,23,13) of grad_x Phi_I = e_i (x) g_J, I = 3J + i
__device__ __forceinline__ void mode_strain(const double Fi[3][3], int I, double e[6]) {
    const int J = I / 3, i = I % 3;
    const double g0 = Fi[J][0], g1 = Fi[J][1], g2 = Fi[J][2];
    e[0] = i == 0 ? g0 : 0.0;
    e[1] = i == 1 ? g1 : 0.0;
    e[2] = i == 2 ? g2 : 0.0;
    e[3] = (i == 0 ? g1 : 0.0) + (i == 1 ? g0 : 0.0);
    e[4] = (i == 1 ? g2 : 0.0) + (i == 2 ? g1 : 0.0);
    e[5] = (i == 0 ? g2 : 0.0) + (i == 2 ? g0 : 0.0);
}

// Gauss-point sums at the converged state, per thread:
//   Pacc: P = J sigma F^-T times the reference volume weight (column-major) and the reference
//         volume (HomogenizeStress, MultiScale/ModMultiscaleAnalysis.f90:140-329);
//   T1 (WANT_T): the homogenized tangent part Phi^T K Phi for the affine modes
//         Phi_(iJ) = e_i Y_J: grad_x Phi_(iJ) = e_i (x) g_J with g_J row J of F^-1, so
//         (Phi_I^T K Phi_K) = sum_gp w detJ [eps(I)^T D eps(K) + delta_ik g_J . sigma g_L],
//         45 upper entries (I <= K, I = 3J + i).
template <int MODEL, bool WANT_T, typename MT>
__device__ void rve_gp_pass(const RveCtx<MT>& c, const RveParams& p, const double* Fa, double* lds, double Pacc[10],
                            double T1[45]) {
    const int tid = threadIdx.x;
    const int n = c.n, nel = n * n * n;
    const int lane = tid & 7, le = tid >> 3;
    double* sm = lds + le * kLdsPerElem;
#pragma unroll
    for (int k = 0; k < 45; ++k) T1[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 10; ++k) Pacc[k] = 0.0;
    for (int base = 0; base < nel; base += kRveElemPerRound) {
        const bool worker = tid < kRveAsmThreads;
        const int el = base + le;
        const bool active = worker && el < nel;
        const int ex = el % n, ey = (el / n) % n, ez = el / (n * n);
        if (worker) {
            const int gx = ex + corner_dx(lane), gy = ey + corner_dy(lane), gz = ez + corner_dz(lane);
            const int node = active ? ((gx % n) + n * ((gy % n) + n * (gz % n))) : 0;
            const double Y[3] = {gx * c.h, gy * c.h, gz * c.h};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double u = active ? c.wc[3 * node + i] : 0.0;
#pragma unroll
                for (int J = 0; J < 3; ++J) u += (Fa[3 * i + J] - (i == J ? 1.0 : 0.0)) * Y[J];
                sm[lane * 9 + i] = Y[i];
                sm[lane * 9 + 3 + i] = Y[i] + u;
                sm[lane * 9 + 6 + i] = u;
            }
        }
        __syncthreads();
        if (active) {
            const int g = lane;
            double dN[8][3];
            hex8_dshape_gp(g, dN);
            double cc[8][3], gX[8][3], gx8[8][3];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) cc[a][k] = sm[a * 9 + k];
            const double detX = hex8_grad(dN, cc, gX);
            double F[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a < 8; ++a) acc += gX[a][j] * sm[a * 9 + 6 + i];
                    F[i][j] = acc + (i == j ? 1.0 : 0.0);
                }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) cc[a][k] = sm[a * 9 + 3 + k];
            const double detx = hex8_grad(dN, cc, gx8);
            const int ph = c.phase[ex + n * (ey + n * ez)];
            double s[6], D[21], nw[1];
            material<MODEL>(p.props + 8 * ph, F, 0.0, nullptr, s, D, nw);
            {
                double P3[3][3];
                first_piola(s, F, P3);
#pragma unroll
                for (int J = 0; J < 3; ++J)
#pragma unroll
                    for (int i = 0; i < 3; ++i) Pacc[3 * J + i] += P3[i][J] * detX;
                Pacc[9] += detX;
            }
            if constexpr (WANT_T) {
            double Fi[3][3];
            inv3(F, det3(F), Fi);
            double gsg[3][3];  // g_J . sigma g_L
            {
                const double sg[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
#pragma unroll
                for (int J = 0; J < 3; ++J)
#pragma unroll
                    for (int L = 0; L < 3; ++L) {
                        double acc = 0.0;
#pragma unroll
                        for (int a = 0; a < 3; ++a)
#pragma unroll
                            for (int b2 = 0; b2 < 3; ++b2) acc += Fi[J][a] * sg[a][b2] * Fi[L][b2];
                        gsg[J][L] = acc;
                    }
            }
            int k45 = 0;
#pragma unroll
            for (int I = 0; I < 9; ++I) {
                double eI[6], De[6];
                mode_strain(Fi, I, eI);
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    double acc = 0.0;
#pragma unroll
                    for (int b2 = 0; b2 < 6; ++b2) acc += D[a <= b2 ? up_idx(a, b2) : up_idx(b2, a)] * eI[b2];
                    De[a] = acc;
                }
#pragma unroll
                for (int K = I; K < 9; ++K) {
                    double eK[6];
                    mode_strain(Fi, K, eK);
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a < 6; ++a) acc += De[a] * eK[a];
                    if ((I % 3) == (K % 3)) acc += gsg[I / 3][K / 3];
                    T1[k45++] += acc * detx;
                }
            }
            }
        }
        __syncthreads();
    }
}

// b_(kL) = T^T K Phi_(kL) for the 9 affine modes: row i of M against the mode's values at the
// neighbour offsets (translation invariance of K removes the absolute position); node 0 -> 0.
// B[(3L + k)][i][a] = sum_o M[o][a][k](i) d_L(o).
template <typename MT>
__device__ void rve_affine_rhs(const RveCtx<MT>& c) {
    const long long nn = c.nn;
    for (int i = threadIdx.x; i < c.nn; i += kRveThreads) {
        double bb[9][3];
#pragma unroll
        for (int m = 0; m < 9; ++m) bb[m][0] = bb[m][1] = bb[m][2] = 0.0;
        if (i != 0) {
            for (int o = 0; o < 27; ++o) {
                const double d[3] = {((o % 3) - 1) * c.h, (((o / 3) % 3) - 1) * c.h, ((o / 9) - 1) * c.h};
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const double m = c.M[((long long)o * 9 + 3 * a + k) * nn + i];
#pragma unroll
                        for (int L = 0; L < 3; ++L) bb[3 * L + k][a] += m * d[L];
                    }
            }
        }
#pragma unroll
        for (int m = 0; m < 9; ++m)
#pragma unroll
            for (int a = 0; a < 3; ++a) c.B[(long long)m * 3 * nn + 3 * i + a] = bb[m][a];
    }
}

// the queue position's RVE and the workgroup's scratch slice
template <typename MT>
__device__ __forceinline__ RveCtx<MT> rve_ctx(const RveParams& p, int rid, long long slot) {
    RveCtx<MT> c;
    c.n = p.side[rid];
    c.nn = c.n * c.n * c.n;
    c.h = 1.0 / c.n;
    c.phase = p.phase + p.phase_off[rid];
    c.wconv = p.w + 3 * p.node_off[rid];
    c.M = reinterpret_cast<MT*>(p.Mbuf) + slot * 243 * p.max_nodes;
    double* V = p.Vbuf + slot * kRveVecPerNode * p.max_nodes;
    const long long mn = p.max_nodes;
    c.wc = V;
    c.R = V + 3 * mn;
    c.x = V + 6 * mn;
    c.r = V + 9 * mn;
    c.q = V + 12 * mn;
    c.z = V + 15 * mn;
    c.Minv = V + 18 * mn;
    c.B = V + 24 * mn;
    c.na = (c.n + kRveAgg - 1) / kRveAgg;
    c.nc = 3 * c.na * c.na * c.na;
    c.C = p.Cbuf + slot * (long long)(kRveMaxCoarse * kRveMaxCoarse);
    return c;
}

template <int MODEL, typename MT, int G, bool NT>
__global__ __launch_bounds__(kRveThreads) void k_rve_solve(RveParams p) {
    __shared__ double lds[kRveLds];
    __shared__ double red[kRveWaves * 45];
    __shared__ double cv[2 * kRveMaxCoarse];  // coarse residual and correction (setup: a pivot row)
    __shared__ int snode[kRveElemPerRound][8];
    __shared__ int sq;
    __shared__ double sFa[9];
    const int tid = threadIdx.x;
    const long long slot = blockIdx.x;
    const double gr = 1.6180339887498948482;  // (1 + sqrt 5) / 2, FEM/ModFEMAnalysis.f90:1172
    for (;;) {
        if (tid == 0) sq = atomicAdd(p.queue, 1);
        __syncthreads();
        const int qi = sq;
        __syncthreads();
        if (qi >= p.n_rve) break;
        const int rid = p.order[qi];
        RveCtx<MT> c = rve_ctx<MT>(p, rid, slot);
        for (int k = tid; k < 3 * c.nn; k += kRveThreads) c.wc[k] = c.wconv[k];
        if (p.warm && p.vvalid[rid]) {
            // first-order predictor for the first attempt: the last condensation's solutions are
            // dw/dF at the converged state, w += sum_m v_m (F - F_conv)_m (a cut-back restarts
            // from w_conv); Newton then corrects to the same test
            const float* Vr = p.V + 27LL * p.node_off[rid];
            double dF[9];
#pragma unroll
            for (int m = 0; m < 9; ++m) dF[m] = p.Fbar[9 * rid + m] - p.Fconv[9 * rid + m];
            for (int i = tid; i < c.nn; i += kRveThreads) {
                if (i == 0) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double dw = 0.0;
#pragma unroll
                    for (int m = 0; m < 9; ++m) dw += (double)Vr[27LL * i + 3 * m + k] * dF[m];
                    c.wc[3 * i + k] += dw;
                }
            }
        }
        __syncthreads();

        double alpha_min = 0.0, alpha = 1.0;
        int cut = 0, newton_total = 0, pcg_total = 0;
        bool done = false, failed = false;
        for (int guard = 0; guard < 64 && !done && !failed; ++guard) {
            // macro F of the substep, row-major in LDS (the converged F of the RVE is only
            // updated after the loop)
            if (tid < 9) {
                const int i = tid / 3, J = tid % 3;
                const double f0 = p.Fconv[9 * rid + 3 * J + i], ft = p.Fbar[9 * rid + 3 * J + i];
                sFa[tid] = alpha == 1.0 ? ft : f0 + alpha * (ft - f0);
            }
            __syncthreads();
            // Newton on the reduced system
            bool ok = false;
            bool coarse = false;  // coarse inverse of this substep's first tangent, kept for the others
            double nr0 = 0.0;
            for (int it = 0; it <= p.newton_maxit; ++it) {
                const long long t0 = wall_clock64();
                const bool bad = rve_assemble<MODEL, true>(c, p, sFa, lds, snode);
                if (p.prof && tid == 0) p.prof[2 * rid] += wall_clock64() - t0;
                double rmax = 0.0;
                for (int i = tid; i < c.nn; i += kRveThreads) {
                    if (i == 0) continue;
#pragma unroll
                    for (int k = 0; k < 3; ++k) rmax = fmax(rmax, fabs(c.R[3 * i + k]));
                }
                double flags[1] = {bad ? 1.0 : 0.0};
                wg_sum<1>(flags, red);
                const double nr = wg_max(rmax, red);
                if (flags[0] > 0.0 || !(nr <= 1.0e300)) break;
                if (it == 0) nr0 = nr;
                newton_total += it == 0 ? 0 : 1;
                if (nr <= fmax(p.newton_rtol * nr0, p.newton_atol)) {
                    ok = true;
                    break;
                }
                if (it == p.newton_maxit) break;
                const long long t1 = wall_clock64();
                rve_precond(c);
                // the first step's coarse inverse serves the substep's later steps too: the
                // preconditioner stays SPD, and the tangent moves little between Newton steps
                if (it == 0) coarse = p.coarse && rve_coarse_setup(c, lds, cv);
                __syncthreads();
                // inexact Newton: the linear solve only needs to beat the nonlinear residual
                // reduction still to come (forcing 0.1 ||R|| / ||R_0||), and never more than half the
                // way to the Newton test itself (0.5 target / ||R||: the last step is not oversolved);
                // within [pcg_rtol, 1e-2]
                const double target = fmax(p.newton_rtol * nr0, p.newton_atol);
                const double eta =
                    p.inexact ? fmin(1e-2, fmax(p.pcg_rtol, fmax(0.1 * nr / nr0, 0.5 * target / nr))) : p.pcg_rtol;
                pcg_total += rve_pcg<MT, G, NT>(c, p, c.R, -1.0, eta, lds, red, cv, coarse);
                if (p.prof && tid == 0) p.prof[2 * rid + 1] += wall_clock64() - t1;
                for (int k = tid; k < 3 * c.nn; k += kRveThreads) c.wc[k] += c.x[k];
                __syncthreads();
            }
            if (!ok) {
                for (int k = tid; k < 3 * c.nn; k += kRveThreads) c.wc[k] = c.wconv[k];
                __syncthreads();
                alpha = alpha_min + (1.0 - 1.0 / gr) * (alpha - alpha_min);
                if (++cut > p.max_cutbacks) failed = true;
                continue;
            }
            for (int k = tid; k < 3 * c.nn; k += kRveThreads) c.wconv[k] = c.wc[k];
            __syncthreads();
            if (alpha == 1.0) {
                done = true;
            } else {
                const double aux = alpha_min;
                alpha_min = alpha;
                alpha = fmin(alpha + gr * (alpha - aux), 1.0);
            }
        }
        if (done) {
            double Pbar[10], T1[45];
            rve_gp_pass<MODEL, false>(c, p, sFa, lds, Pbar, T1);
            wg_sum<10>(Pbar, red);
            if (tid < 9) {
                p.P[9 * rid + tid] = Pbar[tid] / Pbar[9];
                p.Fconv[9 * rid + tid] = p.Fbar[9 * rid + tid];
            }
        }
        if (tid == 0) {
            p.info[4 * rid] = done ? 0 : -1;
            p.info[4 * rid + 1] = newton_total;
            p.info[4 * rid + 2] = pcg_total;
            p.info[4 * rid + 3] = cut;
        }
        __syncthreads();
    }
}


// Homogenized tangent dP/dF of every converged RVE by static condensation (second persistent
// launch, after k_rve_solve): the reduced tangent at the converged state is assembled again,
//   A = (Phi^T K Phi + sum_m b_m^T v_m) / V,  K_red v_m = -b_m,  b_m = T^T K Phi_m,
// nine CG solves per RVE.  RVEs whose solve failed are skipped.
template <int MODEL, typename MT, int G, bool NT>
__global__ __launch_bounds__(kRveThreads) void k_rve_tangent(RveParams p) {
    __shared__ double lds[kRveLds];
    __shared__ double red[kRveWaves * 45];
    __shared__ double cv[6 * kRveMaxCoarse];  // coarse residuals and corrections of three solves
    __shared__ int snode[kRveElemPerRound][8];
    __shared__ int sq;
    __shared__ double sFa[9];
    // Phi^T K Phi (45 upper), V0, condensed parts (81): kept in LDS, so that no register holds them
    // through the nine CG solves (their 90 VGPRs would push the CG's batched loads into spills)
    __shared__ double sT[45 + 1 + 81];
    const int tid = threadIdx.x;
    const long long slot = blockIdx.x;
    for (;;) {
        if (tid == 0) sq = atomicAdd(p.queue, 1);
        __syncthreads();
        const int qi = sq;
        __syncthreads();
        if (qi >= p.n_rve) break;
        const int rid = p.order[qi];
        if (!p.probe && p.info[4 * rid] != 0) continue;
        RveCtx<MT> c = rve_ctx<MT>(p, rid, slot);
        for (int k = tid; k < 3 * c.nn; k += kRveThreads) c.wc[k] = c.wconv[k];
        if (tid < 9) sFa[tid] = p.Fconv[9 * rid + 3 * (tid % 3) + tid / 3];
        __syncthreads();
        rve_assemble<MODEL, true>(c, p, sFa, lds, snode);
        {
            double Pbar[10], T1[45];
            rve_gp_pass<MODEL, true>(c, p, sFa, lds, Pbar, T1);
            wg_sum<10>(Pbar, red);
            wg_sum<45>(T1, red);
            if (tid == 0) {
#pragma unroll
                for (int k = 0; k < 45; ++k) sT[k] = T1[k];
                sT[45] = Pbar[9];
            }
        }
        rve_affine_rhs(c);
        rve_precond(c);
        const bool coarse = p.coarse && rve_coarse_setup(c, lds, cv);
        __syncthreads();
        const long long nn = c.nn;
        int pcg_total = 0;
        // condensed part b_k . v_m, v_m = K_red^-1 (-b_m): the nine solves three at a time
        double* Aout = p.A + 81LL * rid;
        if (p.probe) {
            // tile probe: the first affine mode's solve with the single-right-hand-side CG of the
            // Newton steps, whose vector sweeps the tile G shapes
            rve_pcg<MT, G, NT>(c, p, c.B, -1.0, p.tangent_rtol, lds, red, cv, coarse);
            continue;
        }
        // the last condensation's solutions of this RVE start the solves (a nearby state: the
        // initial residual is a fraction of b); this one's are kept for the next
        float* Vr = p.V + 27LL * p.node_off[rid];
        const bool warm = p.warm && p.vvalid[rid];
        for (int m0 = 0; m0 < 9; m0 += 3) {
            int its[3];
            rve_pcg3<MT, NT>(c, p, m0, p.tangent_rtol, reinterpret_cast<float*>(lds), red, cv, coarse, its,
                             warm ? Vr : nullptr);
            pcg_total += its[0] + its[1] + its[2];
            const double* x3 = rve_vec3(c, 0);
            for (int i = tid; i < c.nn; i += kRveThreads) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int k = 0; k < 3; ++k) Vr[27LL * i + 3 * (m0 + q) + k] = (float)x3[9 * i + 3 * q + k];
            }
            double dots[27];  // [9 s + k] = b_k . v_(m0+s)
#pragma unroll
            for (int k = 0; k < 27; ++k) dots[k] = 0.0;
            for (int i = tid; i < c.nn; i += kRveThreads) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    double bk[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) bk[k] = c.B[(long long)k * 3 * nn + 3 * i + a];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const double v = x3[9 * i + 3 * q + a];
#pragma unroll
                        for (int k = 0; k < 9; ++k) dots[9 * q + k] += bk[k] * v;
                    }
                }
            }
            wg_sum<27>(dots, red);
            if (tid == 0) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int k = 0; k < 9; ++k) sT[46 + 9 * k + m0 + q] = dots[9 * q + k];
            }
        }
        __syncthreads();
        if (tid < 81) {
            const int I = tid / 9, K = tid % 9;
            const int lo = I < K ? I : K, hi = I < K ? K : I;
            const int idx = lo * 9 - (lo * (lo - 1)) / 2 + (hi - lo);
            Aout[tid] = (sT[idx] + sT[46 + tid]) / sT[45];
        }
        if (tid == 0) {
            p.info[4 * rid + 2] += pcg_total;
            p.vvalid[rid] = 1;
        }
        __syncthreads();
    }
}

}  // namespace ceos
