// Fused hex8 element kernel: constitutive update + internal force + tangent stiffness + CSR
// scatter, one pass over the elements of one colour.
//
// Work of the reference loops it replaces (file:line relative to SourceCode/):
//   SolveConstitutiveModel   FEM/Subroutines/SubSolveConstitutiveModel.f90:56-136
//   InternalForce            FEM/Subroutines/SubInternalForce.f90:49-97 (+ ModElement.f90:484-562)
//   TangentStiffnessMatrix   FEM/Subroutines/SubTangentStiffnessMatrix.f90:63,294-310
//                            (+ ModElement.f90:200-348, SubAssembleGlobalMatrixUpperTriangular.f90)
//
// Mapping (wave64): 8 lanes per element, 8 elements per wavefront.
//   phase 0: lane a gathers node a (X, x, u) into LDS.
//   phase 1: lane g evaluates Gauss point g: grad_X N, F, grad_x N, detJ, sigma, D (upper),
//            writes sigma (and the material state) to HBM and (grad_x N, w detJ sigma,
//            w detJ D) to LDS.
//   phase 2: lane b owns node b: Fe_b = sum_g B_b^T sigma w detJ, the diagonal 3x3 block
//            (b,b) and the off-diagonal blocks (c,b), c = b+1..b+3 (and b+4 for b < 4) mod 8,
//            K_cb = sum_g w detJ (B_c^T D B_b + (gradN_c . sigma gradN_b) I).
//            All 36 node blocks of the element are computed exactly once.
// Scatter is atomic-free: elements of one colour share no node, colours run in stream order,
// and the first colour that touches a node block writes instead of adding (so neither K nor
// Fint needs zeroing).  CSR slots come from the node-blocked structure of the pattern:
//   row 3A+i starts at rowstart[A] + {0, 3+3 nup[A], 5+6 nup[A]}[i]; inside it the diagonal
//   block columns come first, then 3 columns per upper neighbour in ascending node order, so
//   slot(3A+i, 3B+k) = rowptr(A,i) + (3-i) + 3 rank_A(B) + k for A < B.
// rank and the first-touch bit of each off-diagonal block are 1 byte per (element, lane, j).
#pragma once
#include <stdint.h>

#include "device_math.hpp"
#include "visco.hpp"

namespace ceos {

constexpr int kEPB = 32;              // elements per workgroup (256 threads = 4 waves)
// doubles per element in LDS between the phases: J^-1 of the current configuration 8x9, w detJ
// sigma 8x6, w detJ D (upper) 8x21 = 288.  Phase 2 recomputes grad_x N from J^-1 and the natural
// derivatives (the 192 doubles of grad_x N itself would take 52.7 KB per workgroup, 3 workgroups
// per CU).  The slice stride is padded to 296: the 8 elements of a wavefront read the same offset
// of their slices at once, and a stride of 288 doubles (576 dwords, a multiple of the bank count)
// puts all 8 on the same banks; 296 (592 dwords) moves each slice by 16 banks and keeps the
// slices 64-byte aligned for 128-bit LDS reads.  37.9 KB per workgroup, plus the 2 KB table of
// dN/dxi phase 2 reads (39.9 KB): still 4 workgroups per CU, the 8 waves the register count
// allows.
constexpr int kLdsPerElem = 296;
constexpr int kOffSig = 72;
constexpr int kOffD = 120;
constexpr int kJ2State = 8;           // [0..5] plastic strain, [6] acc. plastic strain, [7] flag
constexpr int kStateStride = 20;      // state rows per Gauss point in the device arrays (max of models)
// Pieces of the gather assembly: lane b computes piece k = 0, [Fe_b (3), diagonal block (b,b)
// upper 00 01 02 11 12 22], and pieces k = 1 + j, the block (c,b) of its partner
// c = (b + j + 1) mod 8 (j < 4 for b < 4, j < 3 otherwise), rows = dofs of c, row-major; 36 pieces
// of 9 doubles per element, numbered q = kPieceBase[b] + k.
//
// Patches: the 16 elements of one workgroup (consecutive slots, chosen spatially compact on the
// host) share node blocks.  Their pieces are summed in LDS, block by block in slot order, and the
// workgroup writes one 9-double patch record per distinct node block of the patch (on a 4x2x2
// brick of a structured mesh 341 instead of 576 pieces).  Every off-diagonal patch record holds its
// CSR block (A,B), A < B, with rows = dofs of A: a piece whose partner node has the larger global
// id is staged transposed.  Tables built once per mesh (ceos_gpu.hip):
//   ptab[slot][36]  staging entry of each piece: bits 0..9 position in its half's staging area,
//                   bit 10 half, bit 11 transposed (a half's staging area holds the first
//                   contribution of its record r at r - r0, the further ones after all firsts);
//   pinfo[patch]    {offset of the patch's records (doubles), offset of its slot table, number of
//                   records, first record of the second half};
//   pslot[]         per record: bits 0..9 staging position of its second contribution, bits
//                   10..13 contributions;
//   pinfo2[patch]   {boundary records, boundary diagonal records, offset of the patch's Fe
//                   records (residual-only pass), complete diagonal records};
//   pkslot[]        per complete record: {K slot of its row 0, node A << 7 | nup[A]}.
// A record is complete when no other patch contributes to its node block: the workgroup writes
// it straight into K (a diagonal one also into Fint).  The boundary records, which the gather
// sums over patches, come first in a patch (diagonal ones first), then the complete ones
// (diagonal ones first).  The residual-only pass stages only the Fe parts: complete diagonal
// records go to Fint, boundary diagonal ones to 3-double Fe records, summed in the same order as
// in the F+K pass.
// The staging area of one half holds at most kStageMax pieces.
constexpr int kPieces = 36;
constexpr int kStageMax = (kEPB * kLdsPerElem) / 9;  // 1052 pieces of 9 doubles in the LDS slices
// staging positions in the tables: ptab bits 0..11 position, bit 12 half, bit 13 transposed;
// pslot bits 0..11 position of the second contribution, bits 12..15 contributions (at most 8)
constexpr int kPosBits = 12;
constexpr unsigned kPosMask = (1u << kPosBits) - 1u;
static_assert(kStageMax <= (1 << kPosBits), "staging positions need more bits");
__host__ __device__ constexpr int piece_base(int b) { return b < 4 ? 5 * b : 20 + 4 * (b - 4); }

struct ElemParams {
    int ne;                        // elements (device slot order)
    const int* conn;               // [ne][8] node ids, 0-based
    const int* mat;                // [ne] dense material index
    const double* props;           // [nmat][8]
    const double* X;               // [nn][3] reference coordinates
    const double* x;               // [nn][3] current coordinates
    const double* U;               // [nn][3] total displacement
    const uint32_t* pairs;         // [ne][8]: byte j-1 = rank | first<<7 of block (b+j, b)
    const uint8_t* diag_first;     // [ne]: bit b = first writer of node b (diag block, Fint)
    const int* rowstart;           // [nn] slot of (3A, 3A)
    const int* nup;                // [nn] upper neighbour count
    double* Fint;                  // [3 nn]
    void* K;                       // [nnz] double or float
    double* sig;                   // [6][8][ne]
    double* scratch;               // patch records (gather assembly); [ne][8][3] Fe parts in the
                                   // residual-only pass
    const uint16_t* ptab;          // [ne][36] staging entries of the pieces (gather assembly)
    const int4* pinfo;             // [patches] record offset, slot-table offset, records, split
    const uint16_t* pslot;         // per patch record: position of its 2nd contribution | contributions << 12
    const int4* pinfo2;            // [patches] boundary / boundary diagonal records, Fe offset, complete diagonal
    const int2* pkslot;            // [patch records] complete records: K slot of row 0, node << 7 | nup
    const double* st_old;          // [nstate][8][ne] committed state
    double* st_new;                // [nstate][8][ne] state of the last update
    uint8_t* elem_err;             // [ne]: bit 0 detJ <= 1e-13, bit 1 local Newton failed
    int* err_count;                // [0] elements with detJ failures, [1] local Newton failures
    double time;                   // analysis time of the evaluation (viscous time step)
    const int* eid;                // [ne] slot -> element id (0-based)
    const double* ext;             // [ne_elements][8][27] external material points (model kEXT)
};

__device__ __forceinline__ long rowptr(int rs, int nu, int i) {
    return (long)rs + (i == 0 ? 0 : (i == 1 ? 3 + 3 * nu : 5 + 6 * nu));
}

template <int MODEL>
constexpr int model_nstate() {
    return MODEL == kJ2 ? 7 : (MODEL == kVISCO ? kViscoState : (MODEL == kEXT ? kExtPerGp : 1));
}

// returns the J2 PlasticState flag, or for the viscoelastic model whether its local Newton
// converged; the other models return false.  jbar: the element's V/V0 (model 4 only).
template <int MODEL>
__device__ __forceinline__ bool material(const double* p, const double F[3][3], double time, const double* old,
                                         double s[6], double D[21], double* nw, double jbar = 1.0) {
    if constexpr (MODEL == kGHL) { mat_ghl(p, F, s, D); return false; }
    else if constexpr (MODEL == kNHQ1) { mat_nhq1p0(p, F, jbar, s, D); return false; }
    else if constexpr (MODEL == kCNH) { mat_cnh(p, F, s, D); return false; }
    else if constexpr (MODEL == kNH) { mat_nh(p, F, s, D); return false; }
    else if constexpr (MODEL == kSVK) { mat_svk(p, F, s, D); return false; }
    else if constexpr (MODEL == kVISCO) { return mat_visco(p, F, time, old, s, D, nw); }
    else if constexpr (MODEL == kEXT) { mat_ext(old, F, s, D); return false; }
    else { return mat_j2(p, F, old, s, D, nw); }
}

// GetTangentModulusBase (ConstitutiveLibrary/ModConstitutiveModel.f90:104-164): central
// differences (h = 1e-8) of the first Piola stress P = J sigma F^-T over the 9 components of F,
// A[(i,J)][(k,L)] = dP_kL/dF_iJ (column-major pairs), converted to the material modulus
// C_IJKL = sum_ik Finv_Ii Finv_Kk (A[(i,J)][(k,L)] - delta_ik S_JL)
// (First_Elasticity_Modulus_To_Second_Voigt, Tools/MathTools/ModMathRoutines.f90:315) and
// pushed forward.  The model is evaluated 18 times from the same committed state.
__device__ __forceinline__ void first_piola(const double s[6], const double F[3][3], double P[3][3]) {
    const double J = det3(F);
    double Fi[3][3];
    inv3(F, J, Fi);
    const double sg[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[i][j] = J * (sg[i][0] * Fi[j][0] + sg[i][1] * Fi[j][1] + sg[i][2] * Fi[j][2]);
}

template <int MODEL>
__device__ __noinline__ void tangent_numerical(const double* pr, const double F0[3][3], double time,
                                               const double* old, const double s0[6], double D[21],
                                               double jbar) {
    constexpr int NS = model_nstate<MODEL>();
    const double h = 1.0e-8;
    double A[9][9];
    double F[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = F0[i][j];
#pragma unroll 1
    for (int ij = 0; ij < 9; ++ij) {
        const int i = ij % 3, j = ij / 3;
        double P[2][3][3];
#pragma unroll 1
        for (int sgn = 0; sgn < 2; ++sgn) {
            F[i][j] = F0[i][j] + (sgn == 0 ? h : -h);
            double s[6], Dt[21], nw[NS];
            material<MODEL>(pr, F, time, old, s, Dt, nw, jbar);
            first_piola(s, F, P[sgn]);
        }
        F[i][j] = F0[i][j];
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int k = 0; k < 3; ++k) A[ij][k + 3 * l] = (P[0][k][l] - P[1][k][l]) / (2.0 * h);
    }
    double P0[3][3], V[3][3], S[3][3];
    first_piola(s0, F0, P0);
    inv3(F0, det3(F0), V);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) S[a][b] = V[a][0] * P0[0][b] + V[a][1] * P0[1][b] + V[a][2] * P0[2][b];
    const int vi[6] = {0, 1, 2, 0, 1, 0}, vj[6] = {0, 1, 2, 1, 2, 2};
    double C[21];
#pragma unroll 1
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) {
            const int I = vi[a], J = vj[a], K = vi[b], L = vj[b];
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    acc += V[I][i] * V[K][k] * (A[i + 3 * J][k + 3 * L] - (i == k ? S[J][L] : 0.0));
            C[up_idx(a, b)] = acc;
        }
    push_forward_up(C, F0, det3(F0), D);
}

// Phase 2 of the element evaluation, lane = node b, from the Gauss-point data in LDS (sm: J^-1
// [8][9], w detJ sigma [8][6] at kOffSig, w detJ D (upper, 21) [8] at kOffD; grad_x N of a node
// at a Gauss point is recomputed as J^-1 dN/dxi, the expression hex8_grad uses):
//   Fe_b = sum_g B_b^T sigma, the upper diagonal block K_bb (00 01 02 11 12 22) and the blocks
//   K_cb (rows = dofs of c, cols = dofs of b) for c = (b + j + 1) mod 8, j < 4 (b < 4) or j < 3,
//   K_cb = sum_g B_c^T D B_b + (gradN_c . sigma gradN_b) I (the geometric part for NL).
// TABLE: dN/dxi of node a at Gauss point g read from dnt[(8 g + a) * 4 + k] (an LDS table)
// instead of evaluated (9 FP64 instructions per node and Gauss point)
template <bool NL, bool WANT_K, bool TABLE = false>
__device__ __forceinline__ void node_blocks(const double* sm, int b, double& fe0, double& fe1, double& fe2,
                                            double kd[6], double ko[4][9], const double* dnt = nullptr) {
    const int npart = (b < 4) ? 4 : 3;
    fe0 = 0.0;
    fe1 = 0.0;
    fe2 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) kd[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 9; ++i) ko[j][i] = 0.0;

#pragma unroll 1
    for (int gp = 0; gp < 8; ++gp) {
        double Ji[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Ji[k] = sm[gp * 9 + k];
        const double* Sp = sm + kOffSig + gp * 6;
        double gb[3];
        {
            double d[3];
            if constexpr (TABLE) {
#pragma unroll
                for (int k = 0; k < 3; ++k) d[k] = dnt[(gp * 8 + b) * 4 + k];
            } else {
                hex8_dshape_node(b, gp, d);
            }
            grad_from_jinv(Ji, d, gb);
        }
        const double b0 = gb[0], b1 = gb[1], b2 = gb[2];
        const double S0 = Sp[0], S1 = Sp[1], S2 = Sp[2], S3 = Sp[3], S4 = Sp[4], S5 = Sp[5];
        const double sg0 = S0 * b0 + S3 * b1 + S5 * b2;
        const double sg1 = S3 * b0 + S1 * b1 + S4 * b2;
        const double sg2 = S5 * b0 + S4 * b1 + S2 * b2;
        fe0 += sg0;
        fe1 += sg1;
        fe2 += sg2;
        if constexpr (WANT_K) {
            const double* Dp = sm + kOffD + gp * 21;
            double Dl[21];
#pragma unroll
            for (int k = 0; k < 21; ++k) Dl[k] = Dp[k];
#define DD(I, J) Dl[(I) <= (J) ? up_idx(I, J) : up_idx(J, I)]
            double DB[6][3];
#pragma unroll
            for (int I = 0; I < 6; ++I) {
                DB[I][0] = DD(I, 0) * b0 + DD(I, 3) * b1 + DD(I, 5) * b2;
                DB[I][1] = DD(I, 1) * b1 + DD(I, 3) * b0 + DD(I, 4) * b2;
                DB[I][2] = DD(I, 2) * b2 + DD(I, 4) * b1 + DD(I, 5) * b0;
            }
#undef DD
            // The block entries accumulate through FMA chains straight into the accumulators
            // (acc = fma(c, DB, acc) three times): a sum of three products added to acc would
            // cost a MUL, two FMAs and an ADD per entry.
            // diagonal block, upper part (k >= i)
            {
                kd[0] = fma(b0, DB[0][0], kd[0]); kd[0] = fma(b1, DB[3][0], kd[0]); kd[0] = fma(b2, DB[5][0], kd[0]);
                kd[1] = fma(b0, DB[0][1], kd[1]); kd[1] = fma(b1, DB[3][1], kd[1]); kd[1] = fma(b2, DB[5][1], kd[1]);
                kd[2] = fma(b0, DB[0][2], kd[2]); kd[2] = fma(b1, DB[3][2], kd[2]); kd[2] = fma(b2, DB[5][2], kd[2]);
                kd[3] = fma(b1, DB[1][1], kd[3]); kd[3] = fma(b0, DB[3][1], kd[3]); kd[3] = fma(b2, DB[4][1], kd[3]);
                kd[4] = fma(b1, DB[1][2], kd[4]); kd[4] = fma(b0, DB[3][2], kd[4]); kd[4] = fma(b2, DB[4][2], kd[4]);
                kd[5] = fma(b2, DB[2][2], kd[5]); kd[5] = fma(b1, DB[4][2], kd[5]); kd[5] = fma(b0, DB[5][2], kd[5]);
                if constexpr (NL) {
                    const double geo = b0 * sg0 + b1 * sg1 + b2 * sg2;
                    kd[0] += geo;
                    kd[3] += geo;
                    kd[5] += geo;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j < npart) {
                    const int c = (b + j + 1) & 7;
                    double gc[3], d[3];
                    if constexpr (TABLE) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) d[k] = dnt[(gp * 8 + c) * 4 + k];
                    } else {
                        hex8_dshape_node(c, gp, d);
                    }
                    grad_from_jinv(Ji, d, gc);
                    const double c0 = gc[0], c1 = gc[1], c2 = gc[2];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        double& r0 = ko[j][0 * 3 + k];
                        double& r1 = ko[j][1 * 3 + k];
                        double& r2 = ko[j][2 * 3 + k];
                        r0 = fma(c0, DB[0][k], r0); r0 = fma(c1, DB[3][k], r0); r0 = fma(c2, DB[5][k], r0);
                        r1 = fma(c1, DB[1][k], r1); r1 = fma(c0, DB[3][k], r1); r1 = fma(c2, DB[4][k], r1);
                        r2 = fma(c2, DB[2][k], r2); r2 = fma(c1, DB[4][k], r2); r2 = fma(c0, DB[5][k], r2);
                    }
                    if constexpr (NL) {
                        // geometric part (gradN_c . sigma gradN_b) on the block's diagonal
                        const double geo = c0 * sg0 + c1 * sg1 + c2 * sg2;
                        ko[j][0] += geo;
                        ko[j][4] += geo;
                        ko[j][8] += geo;
                    }
                }
            }
        }
    }
}

// NUMT: tangent modulus by central differences (a separate instantiation, so that the default
// kernel carries no call and keeps its register allocation)
template <int MODEL, bool NL, bool WANT_K, typename KT, bool GATHER, bool NUMT = false>
__global__ __launch_bounds__(kEPB * 8) void k_element(ElemParams p, int e_begin, int e_end, int patch_begin) {
    __shared__ double smem[kEPB * kLdsPerElem];
    __shared__ int snode[GATHER ? 1 : kEPB][8];  // partner node ids (scatter assembly only)
    __shared__ double sdn[8 * 8 * 4];             // dN/dxi of node a at Gauss point g: [g][a][3 + pad]
    for (int i = threadIdx.x; i < 64; i += kEPB * 8) {
        double d[3];
        hex8_dshape_node(i & 7, i >> 3, d);
        sdn[4 * i] = d[0];
        sdn[4 * i + 1] = d[1];
        sdn[4 * i + 2] = d[2];
        sdn[4 * i + 3] = 0.0;
    }
    const int lane = threadIdx.x & 7;
    const int le = threadIdx.x >> 3;
    const int e = e_begin + blockIdx.x * kEPB + le;
    const bool active = e < e_end;
    double* sm = smem + le * kLdsPerElem;

    // ---------------- phase 0: gather node lane -----------------
    int node = 0;
    if (active) node = p.conn[(long)e * 8 + lane];
    if constexpr (!GATHER) snode[le][lane] = node;
    // patch header and this lane's staging entries, requested with the node data
    int4 pi = make_int4(0, 0, 0, 0);
    int4 pi2 = make_int4(0, 0, 0, 0);
    unsigned ptv[5] = {0u, 0u, 0u, 0u, 0u};
    {
        double v[9];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = active ? p.X[3L * node + k] : 0.0;
            v[3 + k] = active ? p.x[3L * node + k] : 0.0;
            v[6 + k] = active ? p.U[3L * node + k] : 0.0;
        }
        if constexpr (GATHER) {
            pi = p.pinfo[patch_begin + blockIdx.x];
            pi2 = p.pinfo2[patch_begin + blockIdx.x];
            const uint16_t* tab = p.ptab + (long)(active ? e : e_begin) * kPieces + piece_base(lane);
#pragma unroll
            for (int k = 0; k < 5; ++k) ptv[k] = (WANT_K || k == 0) && (k < 4 || lane < 4) ? tab[k] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) sm[lane * 9 + k] = v[k];
    }
    __syncthreads();

    // ---------------- phase 1: Gauss point g = lane -----------------
    const int g = lane;
    double gx[8][3];
    double jix[9];  // J^-1 of the current configuration at this Gauss point
    double s[6], D[21];
    double wdet;
    bool bad;
    bool nrfail = false;
    double jbar = 1.0, vol0 = 1.0;           // model 4: element volume ratio, reference volume
    double kdiv = 0.0;                       // model 4: d2Psi/dJbar2 / V0
    double bdiv[MODEL == kNHQ1 ? 8 : 1][3];  // model 4: integral of grad_x N over the element
    {
        double dN[8][3];
        hex8_dshape_gp(g, dN);
        double c[8][3];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) c[a][k] = sm[a * 9 + k];
        double gX[8][3];
        const double detX = hex8_grad(dN, c, gX);
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
            for (int k = 0; k < 3; ++k) c[a][k] = sm[a * 9 + 3 + k];
        const double detx = hex8_grad(dN, c, gx, jix);
        bad = active && !(detx > kDetTol && detX > kDetTol);
        wdet = detx;  // Gauss weight 1 (ModElementHexa8.f90:326)
        if constexpr (MODEL == kNHQ1) {
            // ElementVolume (ModElement.f90:1313-1427): V and V0 over the element's 8 lanes
            // (butterfly sums: every lane gets the same bits), Jbar = V/V0 at every Gauss point
            // (SubSolveConstitutiveModel.f90:87-100); Bdiv = sum_g grad_x N w detJ, the integral
            // of div N over the current element (ModElement.f90:459)
            double v = detx, vx = detX;
#pragma unroll
            for (int o = 1; o < 8; o <<= 1) {
                v += __shfl_xor(v, o);
                vx += __shfl_xor(vx, o);
            }
            jbar = v / vx;
            vol0 = vx;
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double t = gx[a][k] * detx;
#pragma unroll
                    for (int o = 1; o < 8; o <<= 1) t += __shfl_xor(t, o);
                    bdiv[a][k] = t;
                }
        }

        constexpr int NS = model_nstate<MODEL>();
        double old[NS], nw[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) old[v] = 0.0;
        if constexpr (MODEL == kVISCO) { old[0] = old[4] = old[8] = 1.0; }
        const int mi = active ? p.mat[e] : 0;
        const double* pr = p.props + 8 * mi;
        if constexpr (MODEL == kJ2 || MODEL == kVISCO) {
            if (active) {
#pragma unroll
                for (int v = 0; v < NS; ++v) old[v] = p.st_old[((long)v * 8 + g) * p.ne + e];
            }
        }
        if constexpr (MODEL == kEXT) {
            if (active) {
                const double* q = p.ext + ((long)p.eid[e] * 8 + g) * kExtPerGp;
#pragma unroll
                for (int v = 0; v < NS; ++v) old[v] = q[v];
            }
        }
        const bool flag = material<MODEL>(pr, F, p.time, old, s, D, nw, jbar);
        if constexpr (WANT_K && NUMT && MODEL != kEXT) tangent_numerical<MODEL>(pr, F, p.time, old, s, D, jbar);
        if constexpr (MODEL == kNHQ1) kdiv = nhq1p0_d2psi(pr, jbar) / vol0;
        if (active) {
#pragma unroll
            for (int k = 0; k < 6; ++k) p.sig[((long)k * 8 + g) * p.ne + e] = s[k];
            if constexpr (MODEL == kJ2 || MODEL == kVISCO) {
#pragma unroll
                for (int v = 0; v < NS; ++v) p.st_new[((long)v * 8 + g) * p.ne + e] = nw[v];
            }
            if constexpr (MODEL == kJ2) p.st_new[((long)7 * 8 + g) * p.ne + e] = flag ? 1.0 : 0.0;
        }
        if constexpr (MODEL == kVISCO) nrfail = active && !flag;
    }
    {
        // per-element OR of the 8 Gauss-point checks (8 lanes of one element)
        const unsigned long long m = __ballot(bad ? 1 : 0);
        const unsigned long long mn = __ballot(nrfail ? 1 : 0);
        const unsigned bits = (unsigned)(m >> (threadIdx.x & 56)) & 0xffu;  // this element's 8 lanes
        const unsigned nbits = (unsigned)(mn >> (threadIdx.x & 56)) & 0xffu;
        if (active && lane == 0) {
            p.elem_err[e] = (bits ? 1 : 0) | (nbits ? 2 : 0);
            if (bits) atomicAdd(p.err_count, 1);
            if (nbits) atomicAdd(p.err_count + 1, 1);
        }
    }
    __syncthreads();  // node data no longer needed: reuse the LDS slice for GP data
#pragma unroll
    for (int k = 0; k < 9; ++k) sm[g * 9 + k] = jix[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) sm[kOffSig + g * 6 + k] = s[k] * wdet;
    if constexpr (WANT_K) {
#pragma unroll
        for (int k = 0; k < 21; ++k) sm[kOffD + g * 21 + k] = D[k] * wdet;
    }
    __syncthreads();

    // ---------------- phase 2: node b = lane -----------------
    // words of the patch records this thread sums (records tid + 128 i; a patch has at most
    // 36 kEPB = 576), requested before the block computation that hides their latency
    constexpr int kRecPerThread = (kPieces * kEPB + kEPB * 8 - 1) / (kEPB * 8);
    unsigned psw[kRecPerThread];
    int2 pks[kRecPerThread];
    if constexpr (GATHER) {
#pragma unroll
        for (int i = 0; i < kRecPerThread; ++i) {
            const int r = (int)threadIdx.x + kEPB * 8 * i;
            psw[i] = r < pi.z ? p.pslot[pi.y + r] : 0u;
            pks[i] = (r < pi.z && r >= pi2.x) ? p.pkslot[pi.y + r] : make_int2(0, 0);
        }
    }
    const int b = lane;
    const int npart = (b < 4) ? 4 : 3;
    double fe0, fe1, fe2;
    double kd[6];
    double ko[4][9];
    node_blocks<NL, WANT_K, true>(sm, b, fe0, fe1, fe2, kd, ko, sdn);
    if constexpr (MODEL == kNHQ1 && NL && WANT_K) {
        // Ke += d2Psi/dJbar2 / V0 Bdiv^T Bdiv (ModElement.f90:464-467): block (c,b) gains
        // kdiv Bdiv_c (x) Bdiv_b.  The lane-dependent rows are selected with unrolled compares
        // (a runtime index into bdiv would move it to scratch memory).
        double bb[3] = {0.0, 0.0, 0.0}, bc[4][3];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) bb[k] = (a == b) ? bdiv[a][k] : bb[k];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = (b + j + 1) & 7;
#pragma unroll
            for (int k = 0; k < 3; ++k) bc[j][k] = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) bc[j][k] = (a == c) ? bdiv[a][k] : bc[j][k];
        }
        kd[0] += kdiv * bb[0] * bb[0];
        kd[1] += kdiv * bb[0] * bb[1];
        kd[2] += kdiv * bb[0] * bb[2];
        kd[3] += kdiv * bb[1] * bb[1];
        kd[4] += kdiv * bb[1] * bb[2];
        kd[5] += kdiv * bb[2] * bb[2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j < npart) {
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int t = 0; t < 3; ++t) ko[j][3 * r + t] += kdiv * bc[j][r] * bb[t];
            }
        }
    }
    if constexpr (GATHER && !WANT_K) {
        // residual only: the Fe parts through the patch's diagonal records (the F+K pass's sums)
        const unsigned t = ptv[0];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            __syncthreads();
            if (active && ((t >> kPosBits) & 1u) == (unsigned)half) {
                double* dst = smem + 9 * (t & kPosMask);
                dst[0] = fe0;
                dst[1] = fe1;
                dst[2] = fe2;
            }
            __syncthreads();
            const int r0 = half ? pi.w : 0, r1 = half ? pi.z : pi.w;
            const int nbd = pi2.y, cd0 = pi2.x, cd1 = pi2.x + pi2.w;  // diagonal records
#pragma unroll
            for (int i = 0; i < kRecPerThread; ++i) {
                const int r = (int)threadIdx.x + kEPB * 8 * i;
                if (r < r0 || r >= r1 || !(r < nbd || (r >= cd0 && r < cd1))) continue;
                const unsigned w = psw[i];
                double* dst = smem + 9 * (r - r0);
                const double* ext = smem + 9 * (w & kPosMask);
                const int cnt = (int)(w >> kPosBits);
                double a0 = dst[0], a1 = dst[1], a2 = dst[2];
#pragma unroll
                for (int c = 1; c < 8; ++c) {
                    if (c < cnt) {
                        a0 += ext[9 * (c - 1)];
                        a1 += ext[9 * (c - 1) + 1];
                        a2 += ext[9 * (c - 1) + 2];
                    }
                }
                if (r >= cd0) {
                    // complete: the node's whole Fint
                    const int A = pks[i].y >> 7;
                    p.Fint[3L * A] = a0;
                    p.Fint[3L * A + 1] = a1;
                    p.Fint[3L * A + 2] = a2;
                } else {
                    dst[0] = a0;
                    dst[1] = a1;
                    dst[2] = a2;
                }
            }
            __syncthreads();
            const int n3 = 3 * max(min(r1, nbd) - r0, 0);
            for (int o = (int)threadIdx.x; o < n3; o += kEPB * 8)
                p.scratch[(long)pi2.z + 3L * r0 + o] = smem[9 * (o / 3) + o % 3];
        }
        return;
    }
    if constexpr (GATHER) {
        // patch assembly: in two halves, stage the pieces whose patch record falls in the half
        // (each at its position, transposed where the record's orientation asks), then sum every
        // record of the half over its contributions in slot order and write it (unit-stride
        // across the workgroup's threads)
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            __syncthreads();  // the Gauss-point data (half 0) or the staged pieces (half 1) are read
            if (active) {
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    if (k <= npart) {
                        const unsigned t = ptv[k];
                        if (((t >> kPosBits) & 1u) == (unsigned)half) {
                            double* dst = smem + 9 * (t & kPosMask);
                            if (k == 0) {
                                dst[0] = fe0;
                                dst[1] = fe1;
                                dst[2] = fe2;
#pragma unroll
                                for (int q = 0; q < 6; ++q) dst[3 + q] = kd[q];
                            } else if ((t >> (kPosBits + 1)) & 1u) {
#pragma unroll
                                for (int r = 0; r < 3; ++r)
#pragma unroll
                                    for (int c = 0; c < 3; ++c) dst[3 * r + c] = ko[k - 1][3 * c + r];
                            } else {
#pragma unroll
                                for (int q = 0; q < 9; ++q) dst[q] = ko[k - 1][q];
                            }
                        }
                    }
                }
            }
            __syncthreads();
            const int r0 = half ? pi.w : 0, r1 = half ? pi.z : pi.w;
            // sum every record in place: its first contribution sits at r - r0, the others at
            // the record's extra positions; the half's records are then one contiguous block
#pragma unroll
            for (int i = 0; i < kRecPerThread; ++i) {
                const int r = (int)threadIdx.x + kEPB * 8 * i;
                if (r < r0 || r >= r1) continue;
                const unsigned w = psw[i];
                double* dst = smem + 9 * (r - r0);
                const double* ext = smem + 9 * (w & kPosMask);
                const int cnt = (int)(w >> kPosBits);
                double acc[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) acc[q] = dst[q];
#pragma unroll
                for (int c = 1; c < 8; ++c) {
                    if (c < cnt) {
#pragma unroll
                        for (int q = 0; q < 9; ++q) acc[q] += ext[9 * (c - 1) + q];
                    }
                }
#pragma unroll
                for (int q = 0; q < 9; ++q) dst[q] = acc[q];
            }
            __syncthreads();
            // boundary records: non-temporal, unit stride across the workgroup (they are read back
            // once, by the gather, long after this workgroup ends, so they need not hold L2 lines)
            const int nbnd = pi2.x;
            double* out = p.scratch + (long)pi.x + 9L * r0;
            const int n9 = 9 * max(min(r1, nbnd) - r0, 0);
            for (int o = (int)threadIdx.x; o < n9; o += kEPB * 8) __builtin_nontemporal_store(smem[o], out + o);
            // complete records: straight into K (rows 3A..3A+2 of the CSR block) and Fint,
            // non-temporal like the records (the gather that follows then keeps L2 for its reads)
            KT* Kg = (KT*)p.K;
#pragma unroll
            for (int i = 0; i < kRecPerThread; ++i) {
                const int r = (int)threadIdx.x + kEPB * 8 * i;
                if (r < r0 || r >= r1 || r < nbnd) continue;
                const double* v = smem + 9 * (r - r0);
                const long row0 = pks[i].x;
                const int A = pks[i].y >> 7, nu = pks[i].y & 127;
                if (r < nbnd + pi2.w) {
                    __builtin_nontemporal_store(v[0], p.Fint + (3L * A));
                    __builtin_nontemporal_store(v[1], p.Fint + (3L * A + 1));
                    __builtin_nontemporal_store(v[2], p.Fint + (3L * A + 2));
                    const long row1 = row0 + 3 + 3 * nu, row2 = row1 + 2 + 3 * nu;
                    __builtin_nontemporal_store((KT)v[3], Kg + (row0)); __builtin_nontemporal_store((KT)v[4], Kg + (row0 + 1)); __builtin_nontemporal_store((KT)v[5], Kg + (row0 + 2));
                    __builtin_nontemporal_store((KT)v[6], Kg + (row1)); __builtin_nontemporal_store((KT)v[7], Kg + (row1 + 1));
                    __builtin_nontemporal_store((KT)v[8], Kg + (row2));
                } else {
                    const long row1 = row0 + 2 + 3 * nu, row2 = row1 + 1 + 3 * nu;
                    __builtin_nontemporal_store((KT)v[0], Kg + (row0)); __builtin_nontemporal_store((KT)v[1], Kg + (row0 + 1)); __builtin_nontemporal_store((KT)v[2], Kg + (row0 + 2));
                    __builtin_nontemporal_store((KT)v[3], Kg + (row1)); __builtin_nontemporal_store((KT)v[4], Kg + (row1 + 1)); __builtin_nontemporal_store((KT)v[5], Kg + (row1 + 2));
                    __builtin_nontemporal_store((KT)v[6], Kg + (row2)); __builtin_nontemporal_store((KT)v[7], Kg + (row2 + 1)); __builtin_nontemporal_store((KT)v[8], Kg + (row2 + 2));
                }
            }
        }
        return;
    }
    if (!active) return;

    // ---------------- scatter -----------------
    // Every read of K / Fint is issued before any write: the slots of one lane are distinct,
    // but the compiler cannot prove it, so interleaved read-add-write sequences would wait on
    // one memory round trip each.
    const int B = node;
    const bool dfirst = (p.diag_first[e] >> b) & 1;
    {
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        if (!dfirst) {
            f0 = p.Fint[3L * B + 0];
            f1 = p.Fint[3L * B + 1];
            f2 = p.Fint[3L * B + 2];
        }
        p.Fint[3L * B + 0] = f0 + fe0;
        p.Fint[3L * B + 1] = f1 + fe1;
        p.Fint[3L * B + 2] = f2 + fe2;
    }
    if constexpr (WANT_K) {
        KT* K = (KT*)p.K;
        const int rsB = p.rowstart[B], nuB = p.nup[B];
        const uint32_t pw = p.pairs[(long)e * 8 + b];
        int slot[42];
        double val[42];
        bool rd[42];   // slot needs the old value (not the first writer)
        bool wr[42];   // slot is written by this lane
        {
            const int r0 = (int)rowptr(rsB, nuB, 0), r1 = (int)rowptr(rsB, nuB, 1), r2 = (int)rowptr(rsB, nuB, 2);
            const int ds[6] = {r0, r0 + 1, r0 + 2, r1, r1 + 1, r2};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                slot[q] = ds[q];
                val[q] = kd[q];
                rd[q] = !dfirst;
                wr[q] = true;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool valid = j < npart;
            const int c = (b + j + 1) & 7;
            const int C = snode[le][c];
            const uint32_t byte = (pw >> (8 * j)) & 0xffu;
            const int rank = (int)(byte & 0x7fu);
            const bool first = (byte >> 7) & 1u;
            const bool lowC = C < B;
            const int rsC = lowC ? p.rowstart[C] : rsB;
            const int nuC = lowC ? p.nup[C] : nuB;
            // C < B: rows of C, K(3C+i, 3B+k) = K_cb(i,k); else rows of B, K(3B+k, 3C+i) = K_cb(i,k)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int base = (int)rowptr(rsC, nuC, r) + (3 - r) + 3 * rank;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int q = 6 + 9 * j + 3 * r + t;
                    slot[q] = base + t;
                    val[q] = lowC ? ko[j][r * 3 + t] : ko[j][t * 3 + r];
                    rd[q] = valid && !first;
                    wr[q] = valid;
                }
            }
        }
        double old[42];
#pragma unroll
        for (int q = 0; q < 42; ++q) old[q] = rd[q] ? (double)K[slot[q]] : 0.0;
#pragma unroll
        for (int q = 0; q < 42; ++q)
            if (wr[q]) K[slot[q]] = (KT)(old[q] + val[q]);
    }
}

}  // namespace ceos
