// libceos_gpu.so -- host runtime and C ABI of the MI355X element hot path (see include/ceos_gpu.h).
//
// Setup (once per mesh, host C++):
//   * node-blocked CSR pattern of the element cliques (the reference builds the same pattern
//     with linked lists, FEM/ModFEMAnalysis.f90:209-273 and ModSparseMatrixRoutines.f90);
//   * greedy element colouring (elements of one colour share no node), elements of a colour
//     grouped by material model and kept in id order; device arrays use this slot order;
//   * per element: the CSR rank byte and first-touch bit of its 28 off-diagonal node blocks
//     and the first-touch bits of its 8 nodes (see element_kernel.hpp).
// Evaluation: one launch of the fused element kernel per (colour, model) range, in stream order.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/ceos_gpu.h"
#include "element_kernel.hpp"
#include "homogenize_kernel.hpp"
#include "assembly_kernel.hpp"
#include "system_kernel.hpp"
#include "coarse_kernel.hpp"
#include "periodic_kernel.hpp"

using namespace ceos;

#define HIP_OK(expr)                                                                           \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            fprintf(stderr, "ceos_gpu: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                                       \
            return CEOS_ERR_HIP;                                                               \
        }                                                                                      \
    } while (0)

// Element records of ceos_evaluate_device_static: a fixed buffer in the code object, one copy per
// device, held until the process ends.  Host code reaches it only through the per-device address
// that static_scratch() obtains with hipGetSymbolAddress (the symbol itself is a host-side
// stand-in, not a device address).  Nothing in it has to start at zero: an evaluation writes
// every record it later reads.
constexpr size_t kStaticScratchBytes = size_t(128) << 20;
constexpr size_t kStaticScratchDoubles = kStaticScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {nullptr};
bool g_scratch_tried[kMaxDevices] = {false};

// device address of `scratch` on `device` (which must be current), looked up once per device
double* static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_tried[device]) {
        g_scratch_tried[device] = true;
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            fprintf(stderr, "ceos_gpu: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
                    hipGetErrorString(e));
            ptr = nullptr;
        }
        g_scratch_ptr[device] = (double*)ptr;
    }
    return g_scratch_ptr[device];
}

constexpr int kMaxMat = 64;
constexpr int kMaxProps = 8;

struct MatDef {
    int model = 0;
    double p[kMaxProps] = {0};
    bool set = false;
};

struct LaunchRange {
    int begin, end, model;
    int patch0 = 0;  // gather assembly: patch of the range's first slot (patches = kEPB-slot chunks)
};

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <typename T>
hipError_t dalloc(T*& p, size_t n) {
    dfree(p);
    return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
}

// buffers whose size is fixed for the life of a handle (it depends on nn, ne or the pattern only):
// allocated once and kept across table rebuilds, so their addresses never change
template <typename T>
hipError_t dalloc_once(T*& p, size_t n) {
    if (p) return hipSuccess;
    return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
}

// buffers whose size depends on the slot order: reallocated only when they must grow
template <typename T>
hipError_t dalloc_fit(T*& p, size_t& cap, size_t n) {
    if (p && cap >= n) return hipSuccess;
    dfree(p);
    cap = 0;
    const hipError_t e = hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) cap = n;
    return e;
}

void set_status(ceos_status* st, int code, int nfail, const char* msg) {
    if (!st) return;
    st->error = code != 0;
    st->error_id = code;
    st->n_failed = nfail;
    snprintf(st->description, sizeof(st->description), "%s", msg ? msg : "");
}

bool model_supported(int m) {
    return m == kGHL || m == kJ2 || m == kNH || m == kNHQ1 || m == kSVK || m == kCNH || m == kVISCO || m == kEXT;
}

int model_nprops(int m) {
    switch (m) {
        case kJ2: return 4;
        case kVISCO: return 6;
        case kEXT: return 0;
        default: return 2;
    }
}

}  // namespace

struct ceos_handle {
    int nn = 0, ne = 0;
    int nl = 1, elem_tech = CEOS_ELEM_FULL_INTEGRATION, precision = CEOS_K_FP64;
    bool fused = true;
    int device = 0;
    hipStream_t stream = nullptr;

    std::vector<int> conn0;      // [ne][8] original order, 0-based
    std::vector<int> mat_of;     // [ne] dense material index
    std::vector<int> mat_ids;    // dense index -> user material id
    MatDef mats[kMaxMat];
    std::vector<double> X;       // [nn][3]

    // pattern (host)
    std::vector<int> up_ptr, up_nodes;  // upper neighbours per node (sorted)
    std::vector<int> rowstart, nup;
    long long nnz = 0;
    std::vector<int> color;      // [ne]
    int ncolor = 0;

    // slot order
    bool finalized = false;
    std::vector<int> perm;       // slot -> element
    std::vector<int> iperm;      // element -> slot
    std::vector<LaunchRange> ranges;
    bool has_state = false;
    int slabs = 0;           // 0 = automatic
    int nslabs_used = 1;
    int assembly = CEOS_ASSEMBLY_GATHER;
    int numerical_tangent = 0;
    long long nblocks = 0;   // node blocks of the upper pattern (gather assembly)
    // pipelined evaluation (gather assembly): element launches in chunks of the slot order on the
    // caller's stream, each followed on a second stream by the gather of the CSR prefix whose
    // contributing elements are all done
    int pipe_req = -1;                   // requested chunks (-1 = automatic)
    int pipe_chunks = 1;                 // chunks in use (1 = no pipeline)
    std::vector<int> blk_last_slot;      // [nblocks] last (largest) slot contributing to the block
    std::vector<long long> pipe_slot;    // [chunks + 1] slot boundaries
    std::vector<long long> pipe_blk;     // [chunks + 1] block boundaries of the gather groups
    hipStream_t stream2 = nullptr;
    std::vector<hipEvent_t> pipe_ev;     // [chunks + 1]

    // device
    double *dX = nullptr, *dx = nullptr, *dU = nullptr, *dFint = nullptr, *dK = nullptr;
    float* dK32 = nullptr;
    double *dsig = nullptr, *dst_old = nullptr, *dst_new = nullptr, *dprops = nullptr;
    int *dconn = nullptr, *dmat = nullptr, *drowstart = nullptr, *dnup = nullptr;
    int* dperm = nullptr;               // slot -> element (device copy of perm)
    const double* dext = nullptr;       // caller's external material points (model kEXT)
    bool has_ext = false;
    uint32_t* dpairs = nullptr;
    uint8_t *ddiag_first = nullptr, *delem_err = nullptr;
    int* derr = nullptr;
    double* dhpart = nullptr;  // homogenisation partial sums
    double* dscratch = nullptr;  // patch records (gather assembly), allocated by ensure_records
    size_t scratch_cap = 0;
    long long rec_doubles = 0;   // doubles of the patch records of one F+K evaluation
    uint16_t* dptab = nullptr;   // [ne][36] staging entries of the pieces
    int4* dpinfo = nullptr;      // [patches]
    int4* dpinfo2 = nullptr;     // [patches] boundary records, boundary diagonal records, Fe offset,
                                 // complete diagonal records
    int2* dpkslot = nullptr;     // [patch records] complete records: K slot of row 0, node << 7 | nup
    long long n_complete_records = 0;
    long long n_complete_diag = 0;     // complete diagonal records (node's whole Fint in the element kernel)
    long long n_boundary_diag = 0;     // boundary diagonal records
    long long n_gather_blocks = 0;     // node blocks with at least one boundary record
    long long n_gather_diag = 0;       // of which diagonal
    int n_free_nodes = 0;        // nodes in no element (their Fint and diagonal K are zeroed)
    int* dfree_nodes = nullptr;
    long long fe_doubles = 0;    // doubles of the Fe records of one residual-only evaluation
    uint16_t* dpslot = nullptr;  // [patch records]
    int* dfptr = nullptr;        // [nn+1] residual-only Fe map
    uint32_t* dfword = nullptr;  // [8 ne]
    size_t pinfo_cap = 0, pslot_cap = 0, cword_cap = 0, pinfo2_cap = 0, pkslot_cap = 0, fword_cap = 0;
    // ordering of the handle's own stream after device entries that ran on a caller's stream
    hipEvent_t ev_caller = nullptr;
    bool caller_pending = false;
    int4* dglist = nullptr;      // gather work list (assembly_kernel.hpp)
    long long n_glist = 0;
    std::vector<long long> glist_pos;  // [nblocks + 1] work-list entries before every block
    uint32_t* dcword = nullptr;
    int hpart_n = 0;

    bool f_valid = false, k_valid = false, have_u = false;
    bool k_bc = false;  // the handle's K carries the boundary conditions (ceos_residual_bc)
    double time = 0.0;

    // Newton system: boundary conditions, residual norm, PCG (system_kernel.hpp)
    bool sys_ready = false;
    int sys_width = 0;
    int *dup_ptr = nullptr, *dup_nodes = nullptr, *dslot_nbr = nullptr, *dsell_nbr = nullptr;
    uint32_t* dslot_src = nullptr;
    uint8_t *ddof_mask = nullptr, *dnode_mask = nullptr;
    int n_presc = 0, n_fixed = 0;
    double *dR = nullptr, *dpart = nullptr, *dpart2 = nullptr, *dpart3 = nullptr, *dnorm = nullptr;
    double *dFext = nullptr, *dUbar = nullptr;  // staging of the host residual call
    int solver_precision = -1;                  // ELL values: -1 = the handle's precision
    void* dell = nullptr;
    bool ell_f32 = false;
    double *dMinv = nullptr, *dcg_r = nullptr, *dcg_z = nullptr, *dcg_q = nullptr;
    double* dcg_p = nullptr;
    double *dcg_b = nullptr, *dcg_x = nullptr;  // staging of the host solve call
    CgScalars* dcg = nullptr;
    double *dmr[3] = {nullptr, nullptr, nullptr};  // MINRES: the third r buffer, v, and the second w
    MrScalars* dmr_c = nullptr;
    int solver_method = CEOS_SOLVER_AUTO;
    int cg_check = 32;                          // iterations between host checks of the Krylov state
    std::vector<int> slot_nbr_host;             // [width][nn] neighbour of every slot (host copy)
    std::vector<uint32_t> slot_src_host;        // [width][nn]
    std::vector<uint8_t> dof_mask_host;         // [3 nn] constrained DOFs of the load case
    // two-level preconditioner (coarse_kernel.hpp)
    int prec_kind = CEOS_PREC_BLOCK_JACOBI;
    double agg_size = 8.0;                      // aggregate box edge, in mean element edge lengths
    bool coarse_ready = false;                  // aggregates, basis and maps of the current load case
    bool coarse_ok = false;                     // the last coarse factorisation succeeded
    int n_agg = 0, n_coarse = 0, n_pair = 0, n_dead = 0;
    int *dagg = nullptr, *dagg_ptr = nullptr, *dagg_nodes = nullptr, *dpair_ptr = nullptr, *ddead = nullptr;
    int2* dpair_ij = nullptr;
    uint32_t* dpair_ent = nullptr;
    float *dbasis = nullptr, *dAci32 = nullptr;
    double *dAc = nullptr, *dAci = nullptr, *dgc = nullptr, *dyc = nullptr;
    int coarse_reuse = 0;                       // solves a coarse inverse serves after the one that built it
    int coarse_uses = 0;                        // solves since the last build
    bool coarse_built = false;
    rocblas_handle blas = nullptr;
    rocblas_int* dpotrf_info = nullptr;

    // phase timing: 3 events per evaluation (start, element kernels done, assembly done)
    std::vector<hipEvent_t> tev;
    int timing_cap = 0, timing_n = 0;
    bool timing_on = false;
    long long timing_launches = 0;
};

// ------------------------------------------------------------------------------------------
// setup helpers
// ------------------------------------------------------------------------------------------
static int build_pattern(ceos_handle* h) {
    const int nn = h->nn, ne = h->ne;
    std::vector<int> deg(nn + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a) deg[h->conn0[8 * e + a] + 1]++;
    for (int i = 0; i < nn; ++i) deg[i + 1] += deg[i];
    std::vector<int> ne_list(deg[nn]);
    {
        std::vector<int> fill(deg.begin(), deg.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int a = 0; a < 8; ++a) ne_list[fill[h->conn0[8 * e + a]]++] = e;
    }
    h->up_ptr.assign(nn + 1, 0);
    h->up_nodes.clear();
    h->up_nodes.reserve((size_t)nn * 14);
    std::vector<int> nb;
    for (int A = 0; A < nn; ++A) {
        if (deg[A + 1] == deg[A]) return CEOS_ERR_ARG;  // node in no element
        nb.clear();
        for (int q = deg[A]; q < deg[A + 1]; ++q) {
            const int e = ne_list[q];
            for (int a = 0; a < 8; ++a) {
                const int B = h->conn0[8 * e + a];
                if (B > A) nb.push_back(B);
            }
        }
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        if (nb.size() > 127) return CEOS_ERR_ARG;  // rank must fit 7 bits
        h->up_nodes.insert(h->up_nodes.end(), nb.begin(), nb.end());
        h->up_ptr[A + 1] = (int)h->up_nodes.size();
    }
    h->rowstart.assign(nn, 0);
    h->nup.assign(nn, 0);
    long long acc = 0;
    for (int A = 0; A < nn; ++A) {
        const int nu = h->up_ptr[A + 1] - h->up_ptr[A];
        h->nup[A] = nu;
        h->rowstart[A] = (int)acc;
        acc += 6 + 9LL * nu;
        if (acc > 2147483000LL) return CEOS_ERR_ARG;  // int32 CSR as in the reference
    }
    h->nnz = acc;
    return CEOS_OK;
}

static int build_colors(ceos_handle* h) {
    const int ne = h->ne;
    std::vector<uint64_t> used(h->nn, 0);
    h->color.assign(ne, 0);
    int nc = 0;
    for (int e = 0; e < ne; ++e) {
        uint64_t u = 0;
        for (int a = 0; a < 8; ++a) u |= used[h->conn0[8 * e + a]];
        int c = 0;
        while (c < 64 && ((u >> c) & 1ull)) ++c;
        if (c == 64) return CEOS_ERR_ARG;
        h->color[e] = c;
        nc = std::max(nc, c + 1);
        for (int a = 0; a < 8; ++a) used[h->conn0[8 * e + a]] |= (1ull << c);
    }
    h->ncolor = nc;
    return CEOS_OK;
}

static inline int rank_of(const ceos_handle* h, int A, int B) {  // A < B
    const int* b = h->up_nodes.data() + h->up_ptr[A];
    const int* e = h->up_nodes.data() + h->up_ptr[A + 1];
    const int* p = std::lower_bound(b, e, B);
    if (p == e || *p != B) return -1;
    return (int)(p - b);
}

static void free_device(ceos_handle* h) {
    dfree(h->dX); dfree(h->dx); dfree(h->dU); dfree(h->dFint); dfree(h->dK); dfree(h->dK32);
    dfree(h->dsig); dfree(h->dst_old); dfree(h->dst_new); dfree(h->dprops);
    dfree(h->dconn); dfree(h->dmat); dfree(h->drowstart); dfree(h->dnup); dfree(h->dperm);
    dfree(h->dpairs); dfree(h->ddiag_first); dfree(h->delem_err); dfree(h->derr); dfree(h->dhpart);
    dfree(h->dscratch); dfree(h->dglist); dfree(h->dcword);
    dfree(h->dptab); dfree(h->dpinfo); dfree(h->dpslot); dfree(h->dfptr); dfree(h->dfword); dfree(h->dpinfo2);
    dfree(h->dpkslot); dfree(h->dfree_nodes);
    dfree(h->dup_ptr); dfree(h->dup_nodes); dfree(h->dslot_nbr); dfree(h->dslot_src); dfree(h->dsell_nbr);
    dfree(h->ddof_mask); dfree(h->dnode_mask); dfree(h->dR); dfree(h->dpart); dfree(h->dpart2);
    dfree(h->dnorm); dfree(h->dFext); dfree(h->dUbar);
    if (h->dell) (void)hipFree(h->dell);
    h->dell = nullptr;
    dfree(h->dMinv); dfree(h->dcg_r); dfree(h->dcg_z); dfree(h->dcg_q); dfree(h->dcg_p);
    dfree(h->dcg_b); dfree(h->dcg_x); dfree(h->dcg); dfree(h->dpart3);
    dfree(h->dmr[0]); dfree(h->dmr[1]); dfree(h->dmr[2]); dfree(h->dmr_c);
    dfree(h->dagg); dfree(h->dagg_ptr); dfree(h->dagg_nodes); dfree(h->dpair_ptr); dfree(h->ddead);
    dfree(h->dpair_ij); dfree(h->dpair_ent); dfree(h->dbasis); dfree(h->dAc); dfree(h->dAci); dfree(h->dgc);
    dfree(h->dAci32);
    dfree(h->dyc); dfree(h->dpotrf_info);
    if (h->blas) (void)rocblas_destroy_handle(h->blas);
    h->blas = nullptr;
}

static int upload_props(ceos_handle* h) {
    std::vector<double> props((size_t)kMaxMat * kMaxProps, 0.0);
    for (size_t i = 0; i < h->mat_ids.size(); ++i)
        for (int k = 0; k < kMaxProps; ++k) props[i * kMaxProps + k] = h->mats[i].p[k];
    HIP_OK(hipMemcpyAsync(h->dprops, props.data(), sizeof(double) * props.size(), hipMemcpyHostToDevice,
                          h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));  // props is a temporary
    return CEOS_OK;
}

// Spatially compact order of one launch group's elements, whose consecutive runs of kEPB slots
// are the patches of the gather assembly: elements sorted by the tile of their reference centroid
// (tiles of 4h x 2h x 2h, h the mean element size from the body diagonals), tiles in z, y, x
// order, element ids inside a tile.  On a structured mesh every patch is a 4x2x2 brick; on any
// other mesh the order only decides how many node blocks a patch shares (never correctness).
static void patch_order(const ceos_handle* h, std::vector<int>& elems) {
    const size_t n = elems.size();
    if (n < 2) return;
    double lo[3] = {1e300, 1e300, 1e300}, hsum = 0.0;
    std::vector<double> cen(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const int* cn = &h->conn0[(size_t)8 * elems[i]];
        double c[3] = {0, 0, 0};
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) {
                const double x = h->X[3L * cn[a] + k];
                c[k] += 0.125 * x;
                lo[k] = std::min(lo[k], x);
            }
        for (int k = 0; k < 3; ++k) cen[3 * i + k] = c[k];
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = h->X[3L * cn[6] + k] - h->X[3L * cn[0] + k];
            d2 += d * d;
        }
        hsum += std::sqrt(d2 / 3.0);
    }
    const double hh = hsum / (double)n;
    if (!(hh > 0.0)) return;
    // 4 x 4 x 2 elements for 32-element patches (4 x 2 x 2 for 16, 4 x 4 x 4 for 64)
    const double tile[3] = {4.0 * hh, (kEPB >= 32 ? 4.0 : 2.0) * hh, (kEPB >= 64 ? 4.0 : 2.0) * hh};
    std::vector<std::pair<std::array<long long, 3>, int>> key(n);
    for (size_t i = 0; i < n; ++i) {
        std::array<long long, 3> t;
        for (int k = 0; k < 3; ++k) t[2 - k] = (long long)std::floor((cen[3 * i + k] - lo[k]) / tile[k]);
        key[i] = {t, elems[i]};
    }
    std::sort(key.begin(), key.end());
    for (size_t i = 0; i < n; ++i) elems[i] = key[i].second;
}

// Tables of the patch assembly (element_kernel.hpp, assembly_kernel.hpp).  For every patch (kEPB
// consecutive slots of a launch range): the distinct node blocks its pieces contribute to, each a
// record summed over its contributions in slot order.  A record is complete when no other patch
// contributes to its block: the element kernel writes it straight into K (and a diagonal one into
// Fint).  The other, boundary records go to the record buffer and through the gather.  Record
// order inside a patch: boundary diagonal, boundary off-diagonal, complete diagonal, complete
// off-diagonal, block order inside each group.  Also: the staging place of every piece, the
// contribution map of the gather over boundary records (patch order inside every block), and the
// map of the residual-only pass over the Fe parts of boundary diagonal records (the same sums, so
// both passes give the same Fint bit for bit).
static int build_gather_map(ceos_handle* h) {
    const int nn = h->nn, ne = h->ne;
    if ((long long)ne * 9 * kPieces >= (1LL << 31) || nn >= (1 << 24)) return CEOS_ERR_ARG;
    std::vector<int> blkstart(nn);
    long long nb = 0;
    for (int A = 0; A < nn; ++A) {
        blkstart[A] = (int)nb;
        nb += 1 + h->nup[A];
    }
    if (nb >= (1LL << 31)) return CEOS_ERR_ARG;
    std::vector<int> blknode(nb);
    for (int A = 0; A < nn; ++A)
        for (int q = 0; q <= h->nup[A]; ++q) blknode[blkstart[A] + q] = A;
    h->nblocks = nb;
    // patches: kEPB-slot chunks of every launch range (the ranges tile [0, ne) in order)
    std::vector<int> pstart;
    for (auto& r : h->ranges) {
        r.patch0 = (int)pstart.size();
        for (int sl = r.begin; sl < r.end; sl += kEPB) pstart.push_back(sl);
    }
    const int np = (int)pstart.size();
    pstart.push_back(ne);
    struct Piece {
        int blk, slot, q;
        bool tr;
    };
    std::vector<Piece> pc;
    auto pieces_of = [&](int p) -> int {
        pc.clear();
        for (int sl = pstart[p]; sl < pstart[p + 1]; ++sl) {
            const int* cn = &h->conn0[(size_t)8 * h->perm[sl]];
            for (int b = 0; b < 8; ++b) {
                const int B = cn[b];
                const int npt = b < 4 ? 4 : 3;
                pc.push_back({blkstart[B], sl, piece_base(b), false});
                for (int j = 0; j < npt; ++j) {
                    const int C = cn[(b + j + 1) & 7];
                    const int lo = std::min(B, C), hi = std::max(B, C);
                    const int r = rank_of(h, lo, hi);
                    if (r < 0) return CEOS_ERR_SLOT;
                    pc.push_back({blkstart[lo] + 1 + r, sl, piece_base(b) + 1 + j, C > B});
                }
            }
        }
        std::sort(pc.begin(), pc.end(), [](const Piece& a, const Piece& b) {
            return a.blk != b.blk ? a.blk < b.blk : a.slot < b.slot;
        });
        return CEOS_OK;
    };
    // pass 1: how many patches contribute to every block
    std::vector<uint8_t> npatch(nb, 0);
    for (int p = 0; p < np; ++p) {
        if (pieces_of(p)) return CEOS_ERR_SLOT;
        for (size_t i = 0; i < pc.size(); ++i)
            if (i == 0 || pc[i].blk != pc[i - 1].blk) npatch[pc[i].blk] = (uint8_t)std::min(255, npatch[pc[i].blk] + 1);
    }
    // pass 2: records, staging, tables
    std::vector<uint16_t> ptab((size_t)ne * kPieces);
    std::vector<int4> pinfo(np), pinfo2(np);
    std::vector<uint16_t> pslot;
    std::vector<int2> pkslot;  // per record: complete records: K slot of row 0, node A << 7 | nup[A]
    std::vector<int> rec_blk;  // block of every boundary record
    std::vector<int> rec_patch;
    std::vector<int> fe_node;  // node of every boundary diagonal record (residual-only Fe records)
    pslot.reserve((size_t)ne * 24);
    pkslot.reserve((size_t)ne * 24);
    long long rec_total = 0;
    long long ncomplete = 0;
    h->n_complete_diag = 0;
    for (int p = 0; p < np; ++p) {
        if (pieces_of(p)) return CEOS_ERR_SLOT;
        // runs of equal blocks, grouped: boundary diagonal, boundary off-diagonal, complete
        // diagonal, complete off-diagonal (block order inside each group)
        std::vector<std::pair<int, int>> runs[4];  // (first piece, count)
        for (size_t i = 0; i < pc.size();) {
            size_t j = i;
            while (j < pc.size() && pc[j].blk == pc[i].blk) ++j;
            const int blk = pc[i].blk;
            const bool diag = blk == blkstart[blknode[blk]];
            const bool complete = npatch[blk] == 1;
            runs[(complete ? 2 : 0) + (diag ? 0 : 1)].push_back({(int)i, (int)(j - i)});
            i = j;
        }
        std::vector<std::pair<int, int>> rec;
        for (auto& g : runs) rec.insert(rec.end(), g.begin(), g.end());
        const int nrec = (int)rec.size();
        const int nbd = (int)runs[0].size(), nbnd = nbd + (int)runs[1].size(), ncd = (int)runs[2].size();
        std::vector<int> cum(nrec + 1, 0);
        for (int r = 0; r < nrec; ++r) cum[r + 1] = cum[r] + rec[r].second;
        int split = nrec;
        for (int r = 0; r < nrec; ++r)
            if (2 * cum[r] >= cum[nrec]) { split = r; break; }
        if (cum[split] > kStageMax || cum[nrec] - cum[split] > kStageMax) return CEOS_ERR_ARG;
        if ((rec_total + nbnd) * 9 >= (1LL << 31)) return CEOS_ERR_ARG;
        pinfo[p] = make_int4((int)(rec_total * 9), (int)pslot.size(), nrec, split);
        pinfo2[p] = make_int4(nbnd, nbd, (int)(3 * fe_node.size()), ncd);
        // staging of a half: the first contribution of record r at r - r0, the further ones of
        // every record after all the firsts, in record order
        int extra = 0;
        for (int r = 0; r < nrec; ++r) {
            const int half = r >= split ? 1 : 0;
            const int r0 = half ? split : 0;
            if (r == r0) extra = (half ? nrec : split) - r0;
            const int first = rec[r].first, cnt = rec[r].second;
            pslot.push_back((uint16_t)(extra | (cnt << kPosBits)));
            for (int i = first; i < first + cnt; ++i) {
                const Piece& q = pc[i];
                const int pos = i == first ? r - r0 : extra++;
                ptab[(size_t)q.slot * kPieces + q.q] = (uint16_t)(pos | (half << kPosBits) | ((q.tr ? 1 : 0) << (kPosBits + 1)));
            }
            const int blk = pc[first].blk;
            const int A = blknode[blk];
            if (r < nbnd) {
                pkslot.push_back(make_int2(0, 0));
                rec_blk.push_back(blk);
                rec_patch.push_back(p);
                if (r < nbd) fe_node.push_back(A);
            } else {
                const int q = blk - blkstart[A];  // 0 = diagonal, else 1 + rank of the neighbour
                const long long row0 = (long long)h->rowstart[A] + (q ? 3 + 3LL * (q - 1) : 0);
                pkslot.push_back(make_int2((int)row0, (A << 7) | h->nup[A]));
                ++ncomplete;
                if (q == 0) ++h->n_complete_diag;
            }
        }
        rec_total += nbnd;
    }
    h->rec_doubles = rec_total * 9;
    // contribution map: block -> boundary records (patch order); complete blocks have none
    std::vector<int> cptr(nb + 1, 0);
    for (int b : rec_blk) cptr[b + 1]++;
    h->n_gather_blocks = h->n_gather_diag = 0;
    for (long long b = 0; b < nb; ++b)
        if (cptr[b + 1]) {
            ++h->n_gather_blocks;
            if (b == blkstart[blknode[b]]) ++h->n_gather_diag;
        }
    for (long long b = 0; b < nb; ++b) cptr[b + 1] += cptr[b];
    std::vector<uint32_t> cword(rec_total);
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    h->blk_last_slot.assign(nb, -1);
    for (long long i = 0; i < rec_total; ++i) {
        const int b = rec_blk[i], p = rec_patch[i];
        const bool diag = b == blkstart[blknode[b]];
        cword[fill[b]++] = (uint32_t)(9 * i + (diag ? 3 : 0));
        h->blk_last_slot[b] = std::max(h->blk_last_slot[b], pstart[p + 1] - 1);
    }
    // residual-only pass: node -> Fe records of boundary diagonal records (3 doubles each),
    // patch order; a node whose diagonal block is complete gets its Fint from the element kernel
    h->fe_doubles = 3LL * (long long)fe_node.size();
    h->n_boundary_diag = (long long)fe_node.size();
    std::vector<int> fptr(nn + 1, 0);
    for (int A : fe_node) fptr[A + 1]++;
    for (int A = 0; A < nn; ++A) fptr[A + 1] += fptr[A];
    std::vector<uint32_t> fword(fe_node.size());
    {
        std::vector<int> ff(fptr.begin(), fptr.end() - 1);
        for (size_t i = 0; i < fe_node.size(); ++i) fword[ff[fe_node[i]]++] = (uint32_t)(3 * i);
    }
    // the work list of the gather: the blocks with boundary records, in CSR order
    std::vector<int4> glist;
    glist.reserve((size_t)h->n_gather_blocks);
    h->glist_pos.assign(nb + 1, 0);
    for (long long b = 0; b < nb; ++b) {
        h->glist_pos[b] = (long long)glist.size();
        const int cnt = cptr[b + 1] - cptr[b];
        if (!cnt) continue;
        const int A = blknode[b];
        const int q = (int)(b - blkstart[A]);
        const long long slot = (long long)h->rowstart[A] + (q ? 3 + 3LL * (q - 1) : 0);
        if (cnt >= (1 << 30) || slot >= (1LL << 31)) return CEOS_ERR_ARG;
        glist.push_back(make_int4(cptr[b], cnt | (q == 0 ? (1 << 30) : 0), (int)slot, (A << 7) | h->nup[A]));
    }
    h->glist_pos[nb] = (long long)glist.size();
    h->n_glist = (long long)glist.size();
    HIP_OK(dalloc(h->dglist, glist.size()));
    HIP_OK(dalloc_fit(h->dcword, h->cword_cap, cword.size()));
    HIP_OK(dalloc_once(h->dptab, ptab.size()));
    HIP_OK(dalloc_fit(h->dpinfo, h->pinfo_cap, pinfo.size()));
    HIP_OK(dalloc_fit(h->dpinfo2, h->pinfo2_cap, pinfo2.size()));
    HIP_OK(dalloc_fit(h->dpslot, h->pslot_cap, pslot.size()));
    HIP_OK(dalloc_fit(h->dpkslot, h->pkslot_cap, pkslot.size()));
    HIP_OK(dalloc_once(h->dfptr, nn + 1));
    HIP_OK(dalloc_fit(h->dfword, h->fword_cap, fword.size()));
    if (!glist.empty())
        HIP_OK(hipMemcpy(h->dglist, glist.data(), sizeof(int4) * glist.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dcword, cword.data(), sizeof(uint32_t) * cword.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dptab, ptab.data(), sizeof(uint16_t) * ptab.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpinfo, pinfo.data(), sizeof(int4) * pinfo.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpinfo2, pinfo2.data(), sizeof(int4) * pinfo2.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpslot, pslot.data(), sizeof(uint16_t) * pslot.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpkslot, pkslot.data(), sizeof(int2) * pkslot.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dfptr, fptr.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dfword, fword.data(), sizeof(uint32_t) * fword.size(), hipMemcpyHostToDevice));
    h->n_complete_records = ncomplete;
    std::vector<int> freen;
    for (int A = 0; A < nn; ++A)
        if (npatch[blkstart[A]] == 0) freen.push_back(A);
    h->n_free_nodes = (int)freen.size();
    if (!freen.empty()) {
        HIP_OK(dalloc(h->dfree_nodes, freen.size()));
        HIP_OK(hipMemcpy(h->dfree_nodes, freen.data(), sizeof(int) * freen.size(), hipMemcpyHostToDevice));
    }
    return CEOS_OK;
}

// Chunks of the pipelined evaluation.  Chunk k covers slots [pipe_slot[k], pipe_slot[k+1]); the
// gather group after it covers blocks [pipe_blk[k], pipe_blk[k+1]), the longest CSR prefix whose
// blocks only have contributors in chunks <= k (the prefix keeps every gather launch a contiguous
// run of rows; on meshes numbered like their geometry the groups are balanced).  Off by default:
// on the 80^3 cube the element kernel moves 1.8 GB per evaluation itself, so the two kernels
// contend for HBM (4 chunks: 365 against 373 M elem/s).
static int plan_pipeline(ceos_handle* h) {
    int S = h->pipe_req < 0 ? 1 : std::max(1, h->pipe_req);
    S = std::min(S, std::max(1, h->ne / 1024));
    h->pipe_chunks = S;
    h->pipe_slot.assign(S + 1, 0);
    // chunk boundaries at patch starts (a patch's records come from one workgroup)
    std::vector<int> starts;
    for (const auto& r : h->ranges)
        for (int sl = r.begin; sl < r.end; sl += kEPB) starts.push_back(sl);
    for (int k = 0; k <= S; ++k) {
        const long long want = (long long)h->ne * k / S;
        const auto it = std::lower_bound(starts.begin(), starts.end(), (int)std::min<long long>(want, h->ne));
        h->pipe_slot[k] = (k == S || it == starts.end()) ? h->ne : *it;
    }
    h->pipe_slot[0] = 0;
    h->pipe_blk.assign(S + 1, h->nblocks);
    h->pipe_blk[0] = 0;
    int chunk_prefix = 0;
    long long b = 0;
    for (int k = 0; k < S; ++k) {
        // advance over blocks whose contributors all lie in chunks <= k
        while (b < h->nblocks) {
            const long long ls = h->blk_last_slot[b];
            int ck = 0;
            while (ck < S - 1 && ls >= h->pipe_slot[ck + 1]) ++ck;
            chunk_prefix = std::max(chunk_prefix, ck);
            if (chunk_prefix > k) break;
            ++b;
        }
        h->pipe_blk[k + 1] = k == S - 1 ? h->nblocks : b;
    }
    if (S > 1) {
        if (!h->stream2) HIP_OK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
        while ((int)h->pipe_ev.size() < S + 1) {
            hipEvent_t e;
            HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            h->pipe_ev.push_back(e);
        }
    }
    return CEOS_OK;
}

// slot order, launch ranges, rank/first-touch tables, per-element device arrays
static int finalize(ceos_handle* h) {
    if (h->finalized) return CEOS_OK;
    for (size_t i = 0; i < h->mat_ids.size(); ++i) {
        if (!h->mats[i].set) return CEOS_ERR_ARG;
        // the reference pairs the mean-dilatation technology with its Q1P0 models only
        // (ModConstitutiveModelLibrary.f90:492-556): model 4 <=> CEOS_ELEM_MEAN_DILATATION
        if ((h->mats[i].model == kNHQ1) != (h->elem_tech == CEOS_ELEM_MEAN_DILATATION)) return CEOS_ERR_ARG;
    }
    const int ne = h->ne;
    // J2 state of a previous slot order, kept in element order across the rebuild
    std::vector<double> keep_old, keep_new;
    if (h->dst_old && !h->perm.empty()) {
        const size_t n = (size_t)ne * 8 * kStateStride;
        std::vector<double> a(n), b(n);
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipMemcpy(a.data(), h->dst_old, sizeof(double) * n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(b.data(), h->dst_new, sizeof(double) * n, hipMemcpyDeviceToHost));
        keep_old.resize(n);
        keep_new.resize(n);
        for (int sl = 0; sl < ne; ++sl) {
            const int e = h->perm[sl];
            for (int q = 0; q < 8 * kStateStride; ++q) {
                keep_old[(size_t)q * ne + e] = a[(size_t)q * ne + sl];
                keep_new[(size_t)q * ne + e] = b[(size_t)q * ne + sl];
            }
        }
    }
    // group by (colour, model), id order inside a group
    std::vector<int> models;
    for (size_t i = 0; i < h->mat_ids.size(); ++i) models.push_back(h->mats[i].model);
    std::sort(models.begin(), models.end());
    models.erase(std::unique(models.begin(), models.end()), models.end());
    h->has_state = std::find(models.begin(), models.end(), (int)kJ2) != models.end() ||
                   std::find(models.begin(), models.end(), (int)kVISCO) != models.end();
    h->has_ext = std::find(models.begin(), models.end(), (int)kEXT) != models.end();
    h->perm.clear();
    h->ranges.clear();
    {
        std::vector<std::vector<int>> bucket((size_t)h->ncolor * models.size());
        for (int e = 0; e < ne; ++e) {
            const int m = h->mats[h->mat_of[e]].model;
            const int mi = (int)(std::lower_bound(models.begin(), models.end(), m) - models.begin());
            bucket[(size_t)h->color[e] * models.size() + mi].push_back(e);
        }
        if (h->assembly == CEOS_ASSEMBLY_GATHER) {
            // no write conflicts: one launch per model, elements in patch order
            for (auto& v : bucket) v.clear();
            for (int e = 0; e < ne; ++e) {
                const int m = h->mats[h->mat_of[e]].model;
                const int mi = (int)(std::lower_bound(models.begin(), models.end(), m) - models.begin());
                bucket[mi].push_back(e);
            }
            for (size_t mi = 0; mi < models.size(); ++mi) patch_order(h, bucket[mi]);
        }
        const int ncol = h->assembly == CEOS_ASSEMBLY_GATHER ? 1 : h->ncolor;
        // execution order: for each chunk of the id-ordered colour lists, every colour (and
        // model) in turn; one chunk (the default) is plain colour order
        int S = h->slabs;
        if (S <= 0) S = 1;
        h->nslabs_used = S;
        for (int sl = 0; sl < S; ++sl)
            for (int c = 0; c < ncol; ++c)
                for (size_t mi = 0; mi < models.size(); ++mi) {
                    auto& v = bucket[(size_t)c * models.size() + mi];
                    const size_t b0 = v.size() * sl / S, b1 = v.size() * (sl + 1) / S;
                    if (b1 <= b0) continue;
                    LaunchRange r{(int)h->perm.size(), 0, models[mi]};
                    h->perm.insert(h->perm.end(), v.begin() + b0, v.begin() + b1);
                    r.end = (int)h->perm.size();
                    h->ranges.push_back(r);
                }
    }
    h->iperm.assign(ne, 0);
    for (int s = 0; s < ne; ++s) h->iperm[h->perm[s]] = s;

    if (h->assembly == CEOS_ASSEMBLY_GATHER) {
        int rc = build_gather_map(h);
        if (rc) return rc;
        rc = plan_pipeline(h);
        if (rc) return rc;
    }

    // per slot tables in execution order; first-touch over node blocks
    std::vector<int> conn_s((size_t)ne * 8), mat_s(ne);
    const bool colored = h->assembly == CEOS_ASSEMBLY_COLORED;
    std::vector<uint32_t> pairs(colored ? (size_t)ne * 8 : 0, 0);
    std::vector<uint8_t> dfirst(colored ? ne : 0, 0);
    std::vector<uint64_t> touched(colored ? (h->nnz + 63) / 64 : 0, 0);
    auto test_and_set = [&](long long id) {
        const bool was = (touched[id >> 6] >> (id & 63)) & 1ull;
        touched[id >> 6] |= (1ull << (id & 63));
        return !was;
    };
    for (int s = 0; s < ne; ++s) {
        const int e = h->perm[s];
        const int* cn = &h->conn0[(size_t)8 * e];
        for (int a = 0; a < 8; ++a) conn_s[(size_t)8 * s + a] = cn[a];
        mat_s[s] = h->mat_of[e];
        if (!colored) continue;
        uint8_t df = 0;
        for (int b = 0; b < 8; ++b)
            if (test_and_set(h->rowstart[cn[b]])) df |= (uint8_t)(1u << b);
        dfirst[s] = df;
        for (int b = 0; b < 8; ++b) {
            const int np = b < 4 ? 4 : 3;
            uint32_t w = 0;
            for (int j = 0; j < np; ++j) {
                const int c = (b + j + 1) & 7;
                const int A = std::min(cn[b], cn[c]), B = std::max(cn[b], cn[c]);
                const int r = rank_of(h, A, B);
                if (r < 0 || A == B) return CEOS_ERR_SLOT;
                const long long id = (long long)h->rowstart[A] + 3 + 3LL * r;
                const uint32_t first = test_and_set(id) ? 1u : 0u;
                w |= ((uint32_t)r | (first << 7)) << (8 * j);
            }
            pairs[(size_t)8 * s + b] = w;
        }
    }

    HIP_OK(dalloc_once(h->dconn, (size_t)ne * 8));
    HIP_OK(dalloc_once(h->dmat, ne));
    HIP_OK(dalloc_once(h->dperm, ne));
    HIP_OK(hipMemcpyAsync(h->dperm, h->perm.data(), sizeof(int) * ne, hipMemcpyHostToDevice, h->stream));
    if (colored) {
        HIP_OK(dalloc_once(h->dpairs, (size_t)ne * 8));
        HIP_OK(dalloc_once(h->ddiag_first, ne));
    } else {
        dfree(h->dpairs);
        dfree(h->ddiag_first);
    }
    HIP_OK(dalloc_once(h->delem_err, ne));
    HIP_OK(dalloc_once(h->dsig, (size_t)ne * 48));
    HIP_OK(hipMemcpyAsync(h->dconn, conn_s.data(), sizeof(int) * conn_s.size(), hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(h->dmat, mat_s.data(), sizeof(int) * ne, hipMemcpyHostToDevice, h->stream));
    if (colored) {
        HIP_OK(hipMemcpyAsync(h->dpairs, pairs.data(), sizeof(uint32_t) * pairs.size(), hipMemcpyHostToDevice,
                              h->stream));
        HIP_OK(hipMemcpyAsync(h->ddiag_first, dfirst.data(), ne, hipMemcpyHostToDevice, h->stream));
    }
    HIP_OK(hipMemsetAsync(h->delem_err, 0, ne, h->stream));
    HIP_OK(hipMemsetAsync(h->dsig, 0, sizeof(double) * (size_t)ne * 48, h->stream));
    if (h->has_state) {
        const size_t n = (size_t)ne * 8 * kStateStride;
        HIP_OK(dalloc_once(h->dst_old, n));
        HIP_OK(dalloc_once(h->dst_new, n));
        if (!keep_old.empty()) {
            std::vector<double> a(n), b(n);
            for (int sl = 0; sl < ne; ++sl) {
                const int e = h->perm[sl];
                for (int q = 0; q < 8 * kStateStride; ++q) {
                    a[(size_t)q * ne + sl] = keep_old[(size_t)q * ne + e];
                    b[(size_t)q * ne + sl] = keep_new[(size_t)q * ne + e];
                }
            }
            HIP_OK(hipMemcpy(h->dst_old, a.data(), sizeof(double) * n, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(h->dst_new, b.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        } else {
            // initial state: zero, viscous deformation gradient Fv = I
            // (ModJ2Plasticity.f90:151-154, ModViscoelasticMatrix.f90:120-133)
            std::vector<double> a(n, 0.0);
            for (int sl = 0; sl < ne; ++sl)
                if (h->mats[h->mat_of[h->perm[sl]]].model == kVISCO)
                    for (int g = 0; g < 8; ++g)
                        for (int v : {0, 4, 8}) a[((size_t)v * 8 + g) * ne + sl] = 1.0;
            HIP_OK(hipMemcpy(h->dst_old, a.data(), sizeof(double) * n, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(h->dst_new, a.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        }
    } else {
        dfree(h->dst_old);
        dfree(h->dst_new);
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    h->finalized = true;
    return CEOS_OK;
}

// element record buffer of the gather assembly (records of every element, ne * kRecord doubles)
static int ensure_records(ceos_handle* h) {
    if (h->assembly != CEOS_ASSEMBLY_GATHER) return CEOS_OK;
    HIP_OK(dalloc_fit(h->dscratch, h->scratch_cap, (size_t)h->rec_doubles));
    return CEOS_OK;
}

static bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) return false;
    return cs != hipStreamCaptureStatusNone;
}

// A device entry that queued work on a caller's stream: the entries that run on the handle's
// own stream (state commit, status, state and stress reads, host-array calls) wait for it.
// Work queued while the caller's stream is being captured is ordered by whoever launches the
// graph, not by the library.
static int note_caller_stream(ceos_handle* h, hipStream_t s) {
    if (s == h->stream || stream_capturing(s)) return CEOS_OK;
    HIP_OK(hipEventRecord(h->ev_caller, s));
    h->caller_pending = true;
    return CEOS_OK;
}

static int after_caller_stream(ceos_handle* h) {
    if (!h->caller_pending) return CEOS_OK;
    HIP_OK(hipStreamWaitEvent(h->stream, h->ev_caller, 0));
    h->caller_pending = false;
    return CEOS_OK;
}

// tables of the evaluation (finalize) and the record buffer; refuses to build them on a stream
// under capture (building synchronises and allocates)
static int prepare_for(ceos_handle* h, hipStream_t s) {
    if (h->finalized && (h->assembly != CEOS_ASSEMBLY_GATHER || (h->dscratch && h->scratch_cap >= (size_t)h->rec_doubles)))
        return CEOS_OK;
    if (stream_capturing(s)) return CEOS_ERR_ARG;
    int rc = finalize(h);
    if (rc) return rc;
    return ensure_records(h);
}

// ------------------------------------------------------------------------------------------
// kernel dispatch
// ------------------------------------------------------------------------------------------
template <int MODEL, bool NL, bool WANT_K, typename KT, bool G, bool NUMT = false>
static void launch_t(const ElemParams& p, int b, int e, int patch0, hipStream_t s) {
    const int n = e - b;
    const int grid = (n + kEPB - 1) / kEPB;
    hipLaunchKernelGGL((k_element<MODEL, NL, WANT_K, KT, G, NUMT>), dim3(grid), dim3(kEPB * 8), 0, s, p, b, e,
                       patch0);
}

template <int MODEL, typename KT, bool G>
static void launch_m(const ElemParams& p, int b, int e, int pc, bool nl, bool wk, bool numt, hipStream_t s) {
    if (nl) {
        if (wk && numt) launch_t<MODEL, true, true, KT, G, true>(p, b, e, pc, s);
        else if (wk) launch_t<MODEL, true, true, KT, G>(p, b, e, pc, s);
        else launch_t<MODEL, true, false, KT, G>(p, b, e, pc, s);
    } else {
        if (wk && numt) launch_t<MODEL, false, true, KT, G, true>(p, b, e, pc, s);
        else if (wk) launch_t<MODEL, false, true, KT, G>(p, b, e, pc, s);
        else launch_t<MODEL, false, false, KT, G>(p, b, e, pc, s);
    }
}

template <typename KT, bool G>
static void launch_range(const ElemParams& p, const LaunchRange& r, bool nl, bool wk, bool numt, hipStream_t s) {
    switch (r.model) {
        case kGHL: launch_m<kGHL, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kJ2: launch_m<kJ2, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kNH: launch_m<kNH, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kNHQ1: launch_m<kNHQ1, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kSVK: launch_m<kSVK, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kCNH: launch_m<kCNH, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kVISCO: launch_m<kVISCO, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, numt, s); break;
        case kEXT: launch_m<kEXT, KT, G>(p, r.begin, r.end, r.patch0, nl, wk, false, s); break;  // tangent is given
        default: break;
    }
}

// one evaluation: sigma/state for every GP, Fint, and (if K != null) the tangent.  rec = the
// element-record buffer; null = the handle's own (prepared, and allocated if missing)
static int evaluate(ceos_handle* h, const double* dU, double* dFint, void* dK, hipStream_t s,
                    double* rec = nullptr) {
    int st = CEOS_OK;
    if (!rec) {
        st = prepare_for(h, s);
        if (st) return st;
        rec = h->dscratch;
    } else if (!h->finalized) {
        return CEOS_ERR_ARG;
    }
    if (s == h->stream) {
        st = after_caller_stream(h);
        if (st) return st;
    }
    if (h->has_ext && !h->dext) return CEOS_ERR_ARG;  // ceos_set_external_gp was not called
    ElemParams p;
    p.ne = h->ne;
    p.eid = h->dperm;
    p.ext = h->dext;
    p.conn = h->dconn;
    p.mat = h->dmat;
    p.props = h->dprops;
    p.X = h->dX;
    p.x = h->dx;
    p.U = dU;
    p.pairs = h->dpairs;
    p.diag_first = h->ddiag_first;
    p.rowstart = h->drowstart;
    p.nup = h->dnup;
    p.Fint = dFint;
    p.K = dK;
    p.sig = h->dsig;
    p.st_old = h->dst_old;
    p.st_new = h->dst_new;
    p.elem_err = h->delem_err;
    p.err_count = h->derr;
    p.time = h->time;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, h->derr);
    p.scratch = rec;
    p.ptab = h->dptab;
    p.pinfo = h->dpinfo;
    p.pslot = h->dpslot;
    p.pinfo2 = h->dpinfo2;
    p.pkslot = h->dpkslot;
    hipEvent_t* ev = (h->timing_on && h->timing_n < h->timing_cap) ? &h->tev[3 * h->timing_n] : nullptr;
    if (ev) HIP_OK(hipEventRecord(ev[0], s));
    const bool wk = dK != nullptr;
    const bool f32 = h->precision == CEOS_K_FP32_STORED;
    const bool gather = h->assembly == CEOS_ASSEMBLY_GATHER;
    const bool numt = h->numerical_tangent != 0;
    GatherParams g;
    g.nn = h->nn;
    g.nblocks = h->nblocks;
    g.blk0 = 0;
    g.blk1 = h->n_glist;
    g.glist = h->dglist;
    g.cword = h->dcword;
    g.fptr = h->dfptr;
    g.fword = h->dfword;
    g.rec = rec;
    g.Fint = dFint;
    g.K = dK;
    if (gather && wk && h->pipe_chunks > 1) {
        // element chunk k on s, then on stream2 (after chunk k's event) the gather of group k;
        // the caller's stream waits for stream2 at the end (fork / join, capturable)
        const int S = h->pipe_chunks;
        hipStream_t s2 = h->stream2;
        for (int k = 0; k < S; ++k) {
            for (const auto& r : h->ranges) {
                LaunchRange q = r;
                q.begin = (int)std::max<long long>(r.begin, h->pipe_slot[k]);
                q.end = (int)std::min<long long>(r.end, h->pipe_slot[k + 1]);
                q.patch0 = r.patch0 + (q.begin - r.begin) / kEPB;  // chunks start at patch starts
                if (q.end > q.begin) {
                    if (f32) launch_range<float, true>(p, q, h->nl != 0, wk, numt, s);
                    else launch_range<double, true>(p, q, h->nl != 0, wk, numt, s);
                }
            }
            HIP_OK(hipEventRecord(h->pipe_ev[k], s));
            HIP_OK(hipStreamWaitEvent(s2, h->pipe_ev[k], 0));
            g.blk0 = h->glist_pos[h->pipe_blk[k]];
            g.blk1 = h->glist_pos[h->pipe_blk[k + 1]];
            if (g.blk1 > g.blk0) {
                const unsigned nbk = (unsigned)((g.blk1 - g.blk0 + 255) / 256);
                if (f32) hipLaunchKernelGGL(k_gather_list<float>, dim3(nbk), dim3(256), 0, s2, g);
                else hipLaunchKernelGGL(k_gather_list<double>, dim3(nbk), dim3(256), 0, s2, g);
            }
        }
        if (ev) HIP_OK(hipEventRecord(ev[1], s));
        HIP_OK(hipEventRecord(h->pipe_ev[S], s2));
        HIP_OK(hipStreamWaitEvent(s, h->pipe_ev[S], 0));
        if (h->n_free_nodes) {
            const unsigned nbk = (unsigned)((h->n_free_nodes + 255) / 256);
            if (f32) hipLaunchKernelGGL(k_zero_free_nodes<float>, dim3(nbk), dim3(256), 0, s, h->n_free_nodes,
                                        h->dfree_nodes, h->drowstart, h->dnup, dFint, dK);
            else hipLaunchKernelGGL(k_zero_free_nodes<double>, dim3(nbk), dim3(256), 0, s, h->n_free_nodes,
                                    h->dfree_nodes, h->drowstart, h->dnup, dFint, dK);
        }
        if (ev) {
            HIP_OK(hipEventRecord(ev[2], s));
            h->timing_n++;
            h->timing_launches += (long long)h->ranges.size() * S;
        }
        HIP_OK(hipGetLastError());
        return CEOS_OK;
    }
    for (const auto& r : h->ranges) {
        if (gather) {
            // records are FP64 either way; KT is the type of the complete blocks written to K
            if (f32) launch_range<float, true>(p, r, h->nl != 0, wk, numt, s);
            else launch_range<double, true>(p, r, h->nl != 0, wk, numt, s);
        } else {
            if (f32) launch_range<float, false>(p, r, h->nl != 0, wk, numt, s);
            else launch_range<double, false>(p, r, h->nl != 0, wk, numt, s);
        }
    }
    if (ev) HIP_OK(hipEventRecord(ev[1], s));
    if (gather) {
        if (wk) {
            if (h->n_glist) {
                const unsigned nbk = (unsigned)((h->n_glist + 255) / 256);
                if (f32) hipLaunchKernelGGL(k_gather_list<float>, dim3(nbk), dim3(256), 0, s, g);
                else hipLaunchKernelGGL(k_gather_list<double>, dim3(nbk), dim3(256), 0, s, g);
            }
        } else {
            hipLaunchKernelGGL(k_gather_fint, dim3((h->nn + 255) / 256), dim3(256), 0, s, g);
        }
        if (h->n_free_nodes) {
            const unsigned nbk = (unsigned)((h->n_free_nodes + 255) / 256);
            if (f32) hipLaunchKernelGGL(k_zero_free_nodes<float>, dim3(nbk), dim3(256), 0, s, h->n_free_nodes,
                                        h->dfree_nodes, h->drowstart, h->dnup, dFint, dK);
            else hipLaunchKernelGGL(k_zero_free_nodes<double>, dim3(nbk), dim3(256), 0, s, h->n_free_nodes,
                                    h->dfree_nodes, h->drowstart, h->dnup, dFint, dK);
        }
    }
    if (ev) {
        HIP_OK(hipEventRecord(ev[2], s));
        h->timing_n++;
        h->timing_launches += (long long)h->ranges.size();
    }
    HIP_OK(hipGetLastError());
    return CEOS_OK;
}

static int read_status(ceos_handle* h, ceos_status* st) {
    int nerr[2] = {0, 0};
    int rc = after_caller_stream(h);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(nerr, h->derr, sizeof(nerr), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (nerr[0] || nerr[1]) {
        std::vector<uint8_t> flags(h->ne);
        HIP_OK(hipMemcpy(flags.data(), h->delem_err, h->ne, hipMemcpyDeviceToHost));
        int n = 0;
        for (int s = 0; s < h->ne; ++s) n += flags[s] != 0;
        if (nerr[0]) {
            set_status(st, CEOS_ERR_DETJ, n,
                       "ElementVolume / DeformationGradient: Determinant of the Jacobian Matrix <= 0.0d0");
            return CEOS_ERR_DETJ;
        }
        set_status(st, CEOS_ERR_LOCAL_NR, n, "Max Iteration in Local Newton-Raphson - Viscoelastic Matrix Model");
        return CEOS_ERR_LOCAL_NR;
    }
    set_status(st, CEOS_OK, 0, "");
    return CEOS_OK;
}

static int ensure_k_buffers(ceos_handle* h) {
    if (h->precision == CEOS_K_FP32_STORED) {
        if (!h->dK32) HIP_OK(dalloc(h->dK32, (size_t)h->nnz));
    } else {
        if (!h->dK) HIP_OK(dalloc(h->dK, (size_t)h->nnz));
    }
    return CEOS_OK;
}

static void* kbuf(ceos_handle* h) {
    return h->precision == CEOS_K_FP32_STORED ? (void*)h->dK32 : (void*)h->dK;
}

// ------------------------------------------------------------------------------------------
// Newton system around the element path (system_kernel.hpp)
// ------------------------------------------------------------------------------------------

// neighbour slots of every node (lower neighbours ascending, the node, upper neighbours
// ascending), the upper lists and the per-DOF mask, on the device; once per mesh
static int ensure_system(ceos_handle* h) {
    if (h->sys_ready) return CEOS_OK;
    const int nn = h->nn;
    if (nn >= (1 << 25)) return CEOS_ERR_ARG;  // owner field of the slot words
    std::vector<int> lptr(nn + 1, 0);
    for (int B = 0; B < nn; ++B)
        for (int q = h->up_ptr[B]; q < h->up_ptr[B + 1]; ++q) lptr[h->up_nodes[q] + 1]++;
    for (int A = 0; A < nn; ++A) lptr[A + 1] += lptr[A];
    std::vector<uint32_t> lsrc(lptr[nn]);
    {
        std::vector<int> fill(lptr.begin(), lptr.end() - 1);
        for (int B = 0; B < nn; ++B)  // ascending B: every lower list comes out sorted
            for (int q = h->up_ptr[B]; q < h->up_ptr[B + 1]; ++q)
                lsrc[fill[h->up_nodes[q]]++] = ((uint32_t)B << 7) | (uint32_t)(q - h->up_ptr[B] + 1);
    }
    int width = 0;
    for (int A = 0; A < nn; ++A) width = std::max(width, lptr[A + 1] - lptr[A] + 1 + h->nup[A]);
    std::vector<int> nbr((size_t)width * nn);
    std::vector<uint32_t> src((size_t)width * nn, kSlotPad);
    for (int A = 0; A < nn; ++A) {
        int j = 0;
        auto put = [&](int B, uint32_t w) {
            nbr[(size_t)j * nn + A] = B;
            src[(size_t)j * nn + A] = w;
            ++j;
        };
        for (int k = lptr[A]; k < lptr[A + 1]; ++k) put((int)(lsrc[k] >> 7), lsrc[k]);
        put(A, (uint32_t)A << 7);
        for (int q = h->up_ptr[A]; q < h->up_ptr[A + 1]; ++q)
            put(h->up_nodes[q], ((uint32_t)A << 7) | (uint32_t)(q - h->up_ptr[A] + 1));
        for (; j < width; ++j) nbr[(size_t)j * nn + A] = A;
    }
    // neighbour ids in the SELL order of the PCG product (padding lanes of the last chunk: node 0)
    const long nchunk = (nn + kSellChunk - 1) / kSellChunk;
    std::vector<int> sell((size_t)nchunk * width * kSellChunk, 0);
    for (int A = 0; A < nn; ++A)
        for (int j = 0; j < width; ++j)
            sell[((size_t)(A / kSellChunk) * width + j) * kSellChunk + (A % kSellChunk)] = nbr[(size_t)j * nn + A];
    const int nblk = (nn + kSysBlock - 1) / kSysBlock;
    HIP_OK(dalloc(h->dsell_nbr, sell.size()));
    HIP_OK(hipMemcpy(h->dsell_nbr, sell.data(), sizeof(int) * sell.size(), hipMemcpyHostToDevice));
    HIP_OK(dalloc(h->dup_ptr, nn + 1));
    HIP_OK(dalloc(h->dup_nodes, h->up_nodes.size()));
    HIP_OK(dalloc(h->dslot_nbr, nbr.size()));
    HIP_OK(dalloc(h->dslot_src, src.size()));
    HIP_OK(dalloc(h->ddof_mask, 3L * nn));
    HIP_OK(dalloc(h->dnode_mask, nn));
    HIP_OK(dalloc(h->dR, 3L * nn));
    HIP_OK(dalloc(h->dpart, nblk));
    HIP_OK(dalloc(h->dpart2, nblk));
    HIP_OK(dalloc(h->dpart3, nblk));
    HIP_OK(dalloc(h->dnorm, 3));
    HIP_OK(hipMemcpy(h->dup_ptr, h->up_ptr.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dup_nodes, h->up_nodes.data(), sizeof(int) * h->up_nodes.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dslot_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dslot_src, src.data(), sizeof(uint32_t) * src.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(h->ddof_mask, 0, 3L * nn));
    HIP_OK(hipMemset(h->dnode_mask, 0, nn));
    h->sys_width = width;
    h->slot_nbr_host.swap(nbr);
    h->slot_src_host.swap(src);
    h->sys_ready = true;
    return CEOS_OK;
}

static SysGeom sys_geom(const ceos_handle* h) {
    SysGeom g;
    g.nn = h->nn;
    g.width = h->sys_width;
    g.rowstart = h->drowstart;
    g.nup = h->dnup;
    g.up_ptr = h->dup_ptr;
    g.up_nodes = h->dup_nodes;
    g.slot_nbr = h->dslot_nbr;
    g.slot_src = h->dslot_src;
    g.dof_mask = h->ddof_mask;
    g.node_mask = h->dnode_mask;
    g.sell_nbr = h->dsell_nbr;
    return g;
}

// residual with the boundary conditions, the BC on K and the norm (all on stream s)
static int residual_bc(ceos_handle* h, int it, const double* dU, const double* dFint, void* dK,
                       const double* dFext, const double* dUbar, double* dR, int norm_type, double* dnorm,
                       hipStream_t s) {
    if (!h->sys_ready) return CEOS_ERR_ARG;
    if (norm_type != 1 && norm_type != 2) return CEOS_ERR_ARG;
    const bool rmod = it == 0 && h->n_presc > 0;
    if (rmod && (!dUbar || !dU || !dK)) return CEOS_ERR_ARG;
    const SysGeom g = sys_geom(h);
    const int nblk = (h->nn + kSysBlock - 1) / kSysBlock;
    const bool f32 = h->precision == CEOS_K_FP32_STORED;
    if (f32)
        hipLaunchKernelGGL(k_bc_residual<float>, dim3(nblk), dim3(kSysBlock), 0, s, g, (const float*)dK, (int)rmod,
                           dFint, dFext, dUbar, dU, dR, h->dpart, h->dpart2);
    else
        hipLaunchKernelGGL(k_bc_residual<double>, dim3(nblk), dim3(kSysBlock), 0, s, g, (const double*)dK,
                           (int)rmod, dFint, dFext, dUbar, dU, dR, h->dpart, h->dpart2);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(kSysBlock), 0, s, (const double*)h->dpart,
                       (const double*)h->dpart2, nblk, norm_type, dnorm);
    if (dK && h->n_presc + h->n_fixed > 0) {
        if (f32) hipLaunchKernelGGL(k_bc_matrix<float>, dim3(nblk), dim3(kSysBlock), 0, s, g, (float*)dK);
        else hipLaunchKernelGGL(k_bc_matrix<double>, dim3(nblk), dim3(kSysBlock), 0, s, g, (double*)dK);
    }
    HIP_OK(hipGetLastError());
    return CEOS_OK;
}

static int ensure_solver(ceos_handle* h) {
    const int want_prec = h->solver_precision < 0 ? h->precision : h->solver_precision;
    const bool f32 = want_prec == CEOS_K_FP32_STORED;
    const size_t nval = (size_t)h->sys_width * 9 * kSellChunk * ((h->nn + kSellChunk - 1) / kSellChunk);
    if (!h->dell || h->ell_f32 != f32) {
        if (h->dell) (void)hipFree(h->dell);
        h->dell = nullptr;
        HIP_OK(hipMalloc(&h->dell, nval * (f32 ? sizeof(float) : sizeof(double))));
        h->ell_f32 = f32;
    }
    const long n = 3L * h->nn;
    if (!h->dMinv) {
        HIP_OK(dalloc(h->dMinv, 6L * h->nn));
        HIP_OK(dalloc(h->dcg_r, n));
        HIP_OK(dalloc(h->dcg_z, n));
        HIP_OK(dalloc(h->dcg_q, n));
        HIP_OK(dalloc(h->dcg_p, n));
        HIP_OK(dalloc(h->dcg, 1));
    }
    return CEOS_OK;
}

static int ensure_minres(ceos_handle* h) {
    if (h->dmr_c) return CEOS_OK;
    const long n = 3L * h->nn;
    for (int i = 0; i < 3; ++i) HIP_OK(dalloc(h->dmr[i], n));
    HIP_OK(dalloc(h->dmr_c, 1));
    return CEOS_OK;
}

// ------------------------------------------------------------------------------------------
// two-level preconditioner: aggregates, coarse basis and coarse-matrix maps (coarse_kernel.hpp)
// ------------------------------------------------------------------------------------------
constexpr int kMaxCoarse = 12288;  // coarse DOFs: the dense A_c and its inverse take 2 x 1.2 GB at most

// Boxes of the reference coordinates with an edge of agg_size mean element edges (grown until the
// coarse space fits kMaxCoarse); the 6 rigid-body modes of every aggregate about its centroid,
// zero on constrained DOFs, orthonormalized (modified Gram-Schmidt, twice); the entries of every
// coarse block pair.  Built on the host once per load case (it depends on the BC mask).
static int coarse_setup(ceos_handle* h) {
    const int nn = h->nn, width = h->sys_width;
    if (width >= 128 || nn >= (1 << 25)) return CEOS_ERR_ARG;  // entry word: node 25 bits, slot 7 bits
    const double* X = h->X.data();
    double hmean = 0.0;
    for (int e = 0; e < h->ne; ++e) {
        const int a = h->conn0[8L * e], b = h->conn0[8L * e + 1];
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) d2 += (X[3L * b + k] - X[3L * a + k]) * (X[3L * b + k] - X[3L * a + k]);
        hmean += std::sqrt(d2);
    }
    hmean /= std::max(h->ne, 1);
    double lo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
        double mn = X[k], mx = X[k];
        for (int A = 1; A < nn; ++A) mn = std::min(mn, X[3L * A + k]), mx = std::max(mx, X[3L * A + k]);
        lo[k] = mn;
        ext[k] = mx - mn;
    }
    std::vector<int> agg(nn);
    int na = 0;
    for (double H = h->agg_size * hmean;; H *= 1.25) {
        int nb[3];
        for (int k = 0; k < 3; ++k) nb[k] = std::max(1, (int)std::floor(ext[k] / std::max(H, 1e-300) + 0.5));
        std::vector<int> bin(nn);
        for (int A = 0; A < nn; ++A) {
            int c[3];
            for (int k = 0; k < 3; ++k)
                c[k] = ext[k] > 0.0 ? std::min(nb[k] - 1, (int)((X[3L * A + k] - lo[k]) / ext[k] * nb[k])) : 0;
            bin[A] = c[0] + nb[0] * (c[1] + nb[1] * c[2]);
        }
        std::vector<int> id((size_t)nb[0] * nb[1] * nb[2], -1);
        na = 0;
        for (int A = 0; A < nn; ++A)
            if (id[bin[A]] < 0) id[bin[A]] = 0;
        for (auto& v : id)
            if (v == 0) v = na++;
        if (6L * na <= kMaxCoarse) {
            for (int A = 0; A < nn; ++A) agg[A] = id[bin[A]];
            break;
        }
    }
    const int nc = 6 * na;
    std::vector<int> aptr(na + 1, 0), anodes(nn);
    for (int A = 0; A < nn; ++A) aptr[agg[A] + 1]++;
    for (int I = 0; I < na; ++I) aptr[I + 1] += aptr[I];
    {
        std::vector<int> f(aptr.begin(), aptr.end() - 1);
        for (int A = 0; A < nn; ++A) anodes[f[agg[A]]++] = A;
    }
    // coarse basis
    std::vector<double> basis(18L * nn, 0.0);
    std::vector<int> dead;
    const std::vector<uint8_t>& dm = h->dof_mask_host;
    for (int I = 0; I < na; ++I) {
        const int n0 = aptr[I], n1 = aptr[I + 1], m = n1 - n0;
        double c[3] = {0, 0, 0};
        for (int k = n0; k < n1; ++k)
            for (int d = 0; d < 3; ++d) c[d] += X[3L * anodes[k] + d] / m;
        std::vector<double> M(18L * m);  // [node][dof][mode]
        for (int k = 0; k < m; ++k) {
            const int A = anodes[n0 + k];
            const double dx = X[3L * A] - c[0], dy = X[3L * A + 1] - c[1], dz = X[3L * A + 2] - c[2];
            const double v[3][6] = {{1, 0, 0, 0, dz, -dy}, {0, 1, 0, -dz, 0, dx}, {0, 0, 1, dy, -dx, 0}};
            for (int i = 0; i < 3; ++i) {
                const bool free_dof = dm.empty() || dm[3L * A + i] == 0;
                for (int q = 0; q < 6; ++q) M[18L * k + 6 * i + q] = free_dof ? v[i][q] : 0.0;
            }
        }
        for (int q = 0; q < 6; ++q) {
            double n0q = 0.0;
            for (long r = 0; r < 3L * m; ++r) n0q += M[6 * r + q] * M[6 * r + q];
            n0q = std::sqrt(n0q);
            for (int pass = 0; pass < 2; ++pass)
                for (int l = 0; l < q; ++l) {
                    double d = 0.0;
                    for (long r = 0; r < 3L * m; ++r) d += M[6 * r + l] * M[6 * r + q];
                    for (long r = 0; r < 3L * m; ++r) M[6 * r + q] -= d * M[6 * r + l];
                }
            double nq = 0.0;
            for (long r = 0; r < 3L * m; ++r) nq += M[6 * r + q] * M[6 * r + q];
            nq = std::sqrt(nq);
            const bool alive = n0q > 0.0 && nq > 1e-8 * n0q;
            for (long r = 0; r < 3L * m; ++r) M[6 * r + q] = alive ? M[6 * r + q] / nq : 0.0;
            if (!alive) dead.push_back(6 * I + q);
        }
        for (int k = 0; k < m; ++k)
            for (int t = 0; t < 18; ++t) basis[18L * anodes[n0 + k] + t] = M[18L * k + t];
    }
    // pair lists: per aggregate I, the (node, slot) entries of its nodes grouped by the neighbour's
    // aggregate J (ascending J, then node, then slot)
    std::vector<int> pptr(1, 0);
    std::vector<int2> pij;
    std::vector<uint32_t> pent;
    pent.reserve((size_t)width * nn);
    std::vector<std::pair<int, uint32_t>> ent;
    for (int I = 0; I < na; ++I) {
        ent.clear();
        for (int k = aptr[I]; k < aptr[I + 1]; ++k) {
            const int A = anodes[k];
            for (int j = 0; j < width; ++j) {
                const size_t t = (size_t)j * nn + A;
                if (h->slot_src_host[t] == kSlotPad) continue;
                ent.push_back({agg[h->slot_nbr_host[t]], (uint32_t)A | ((uint32_t)j << 25)});
            }
        }
        std::stable_sort(ent.begin(), ent.end(), [](const std::pair<int, uint32_t>& a,
                                                    const std::pair<int, uint32_t>& b) { return a.first < b.first; });
        for (size_t i = 0; i < ent.size(); ++i) {
            if (i == 0 || ent[i].first != ent[i - 1].first) {
                if (i) pptr.push_back((int)pent.size());
                pij.push_back(make_int2(I, ent[i].first));
            }
            pent.push_back(ent[i].second);
        }
        if (!ent.empty()) pptr.push_back((int)pent.size());
    }
    if (pent.size() >= (size_t)INT32_MAX) return CEOS_ERR_ARG;
    HIP_OK(dalloc(h->dagg, nn));
    HIP_OK(dalloc(h->dagg_ptr, na + 1));
    HIP_OK(dalloc(h->dagg_nodes, nn));
    std::vector<float> basis32(basis.begin(), basis.end());
    HIP_OK(dalloc(h->dbasis, basis32.size()));
    HIP_OK(dalloc(h->dpair_ptr, pptr.size()));
    HIP_OK(dalloc(h->dpair_ij, pij.size()));
    HIP_OK(dalloc(h->dpair_ent, pent.size()));
    HIP_OK(dalloc(h->ddead, dead.size()));
    HIP_OK(dalloc(h->dAc, (size_t)nc * nc));
    HIP_OK(dalloc(h->dAci, (size_t)nc * nc));
    const int ld = (nc + 3) & ~3;
    HIP_OK(dalloc(h->dAci32, (size_t)nc * ld));
    HIP_OK(dalloc(h->dgc, ld));
    {
        const std::vector<double> zeros(ld, 0.0);  // the entries past nc stay zero
        HIP_OK(hipMemcpy(h->dgc, zeros.data(), sizeof(double) * ld, hipMemcpyHostToDevice));
    }
    HIP_OK(dalloc(h->dyc, nc));
    HIP_OK(dalloc(h->dpotrf_info, 1));
    HIP_OK(hipMemcpy(h->dagg, agg.data(), sizeof(int) * nn, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dagg_ptr, aptr.data(), sizeof(int) * (na + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dagg_nodes, anodes.data(), sizeof(int) * nn, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dbasis, basis32.data(), sizeof(float) * basis32.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpair_ptr, pptr.data(), sizeof(int) * pptr.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpair_ij, pij.data(), sizeof(int2) * pij.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dpair_ent, pent.data(), sizeof(uint32_t) * pent.size(), hipMemcpyHostToDevice));
    if (!dead.empty()) HIP_OK(hipMemcpy(h->ddead, dead.data(), sizeof(int) * dead.size(), hipMemcpyHostToDevice));
    if (!h->blas && rocblas_create_handle(&h->blas) != rocblas_status_success) return CEOS_ERR_HIP;
    h->n_agg = na;
    h->n_coarse = nc;
    h->n_pair = (int)pij.size();
    h->n_dead = (int)dead.size();
    h->coarse_ready = true;
    h->coarse_built = false;
    return CEOS_OK;
}

// A_c = P^T K P of the BC-applied K and its inverse (Cholesky, rocSOLVER); on a factorisation
// that fails (A_c not positive definite) the solve falls back to block Jacobi.  Synchronises s.
static int coarse_operator(ceos_handle* h, const void* dK, hipStream_t s) {
    const int nc = h->n_coarse;
    const SysGeom g = sys_geom(h);
    // blocks of aggregate pairs that share no nonzero of K are zero and written by no pair
    HIP_OK(hipMemsetAsync(h->dAc, 0, sizeof(double) * nc * nc, s));
    if (h->precision == CEOS_K_FP32_STORED)
        hipLaunchKernelGGL(k_coarse_assemble<float>, dim3(h->n_pair), dim3(kCoarseBlock), 0, s, g, (const float*)dK,
                           (const float*)h->dbasis, (const int*)h->dpair_ptr, (const int2*)h->dpair_ij,
                           (const uint32_t*)h->dpair_ent, nc, h->dAc);
    else
        hipLaunchKernelGGL(k_coarse_assemble<double>, dim3(h->n_pair), dim3(kCoarseBlock), 0, s, g, (const double*)dK,
                           (const float*)h->dbasis, (const int*)h->dpair_ptr, (const int2*)h->dpair_ij,
                           (const uint32_t*)h->dpair_ent, nc, h->dAc);
    if (h->n_dead)
        hipLaunchKernelGGL(k_coarse_unit_diag, dim3((h->n_dead + 255) / 256), dim3(256), 0, s, h->n_dead,
                           (const int*)h->ddead, nc, h->dAc);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(h->dAci, h->dAc, sizeof(double) * nc * nc, hipMemcpyDeviceToDevice, s));
    if (rocblas_set_stream(h->blas, s) != rocblas_status_success) return CEOS_ERR_HIP;
    rocblas_int info = 0;
    bool ok = rocsolver_dpotrf(h->blas, rocblas_fill_upper, nc, h->dAci, nc, h->dpotrf_info) == rocblas_status_success;
    if (ok) {
        HIP_OK(hipMemcpyAsync(&info, h->dpotrf_info, sizeof(info), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        ok = info == 0;
    }
    if (ok) {
        ok = rocsolver_dpotri(h->blas, rocblas_fill_upper, nc, h->dAci, nc, h->dpotrf_info) == rocblas_status_success;
        if (ok) {
            HIP_OK(hipMemcpyAsync(&info, h->dpotrf_info, sizeof(info), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            ok = info == 0;
        }
    }
    if (ok) {
        const int ld = (nc + 3) & ~3;
        const long nt = (long)nc * ld;
        hipLaunchKernelGGL(k_coarse_inverse_f32, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, nc, ld,
                           (const double*)h->dAci, h->dAci32);
        HIP_OK(hipGetLastError());
    }
    h->coarse_ok = ok;
    return CEOS_OK;
}

// x = K^-1 b by block-Jacobi PCG from x = 0; K (upper CSR, the handle's precision) after the
// boundary conditions.  Synchronises s every cg_check iterations to read the state.
// the SELL copy of the BC-applied K and the block-Jacobi inverses
static void build_operator(ceos_handle* h, const void* dK, hipStream_t s) {
    const SysGeom g = sys_geom(h);
    const int nblk = (h->nn + kSysBlock - 1) / kSysBlock;
    const long nt = (long)h->sys_width * h->nn;
    const unsigned nbe = (unsigned)((nt + kSysBlock - 1) / kSysBlock);
    if (h->precision == CEOS_K_FP32_STORED) {
        if (h->ell_f32) hipLaunchKernelGGL((k_build_ell<float, float>), dim3(nbe), dim3(kSysBlock), 0, s, g, (const float*)dK, (float*)h->dell);
        else hipLaunchKernelGGL((k_build_ell<float, double>), dim3(nbe), dim3(kSysBlock), 0, s, g, (const float*)dK, (double*)h->dell);
        hipLaunchKernelGGL(k_block_jacobi<float>, dim3(nblk), dim3(kSysBlock), 0, s, g, (const float*)dK, h->dMinv);
    } else {
        if (h->ell_f32) hipLaunchKernelGGL((k_build_ell<double, float>), dim3(nbe), dim3(kSysBlock), 0, s, g, (const double*)dK, (float*)h->dell);
        else hipLaunchKernelGGL((k_build_ell<double, double>), dim3(nbe), dim3(kSysBlock), 0, s, g, (const double*)dK, (double*)h->dell);
        hipLaunchKernelGGL(k_block_jacobi<double>, dim3(nblk), dim3(kSysBlock), 0, s, g, (const double*)dK, h->dMinv);
    }
}

// the coarse correction of the two-level preconditioner: z = D^-1 r + P A_c^-1 P^T r and the
// partials of r.z, r.r (one per update workgroup); at the start also p = z
static void coarse_apply(ceos_handle* h, int start, hipStream_t s) {
    const int nupd = std::min((h->nn + kSysBlock - 1) / kSysBlock, kCgUpdBlocks);
    hipLaunchKernelGGL(k_cg_restrict, dim3(h->n_agg), dim3(kSysBlock), 0, s, (const CgScalars*)h->dcg,
                       (const int*)h->dagg_ptr, (const int*)h->dagg_nodes, (const float*)h->dbasis,
                       (const double*)h->dcg_r, h->dgc);
    hipLaunchKernelGGL(k_cg_coarse_solve, dim3((h->n_coarse + 3) / 4), dim3(256), 0, s, (const CgScalars*)h->dcg,
                       h->n_coarse, (h->n_coarse + 3) & ~3, (const float*)h->dAci32, (const double*)h->dgc, h->dyc);
    hipLaunchKernelGGL(k_cg_prolong, dim3(nupd), dim3(kSysBlock), 0, s, h->nn, (const CgScalars*)h->dcg, start,
                       (const double*)h->dMinv, (const int*)h->dagg, (const float*)h->dbasis,
                       (const double*)h->dyc, (const double*)h->dcg_r, h->dcg_z, h->dcg_p, h->dpart, h->dpart2);
}

static bool two_level_on(const ceos_handle* h) {
    return h->prec_kind == CEOS_PREC_TWO_LEVEL && h->coarse_ready && h->coarse_built && h->coarse_ok;
}

// PCG start from x = 0: state scalars (tolerance, iteration limit) set by a kernel, so the
// sequence holds no host copy and may be captured
static void pcg_start(ceos_handle* h, const double* db, double* dx, double rtol, int maxit, hipStream_t s,
                      bool two_level = false) {
    const int nn = h->nn;
    const int nblk = (nn + kSysBlock - 1) / kSysBlock;
    if (two_level) {
        hipLaunchKernelGGL(k_cg_init<true>, dim3(nblk), dim3(kSysBlock), 0, s, nn, (const double*)h->dMinv, db, dx,
                           h->dcg_r, h->dcg_z, h->dcg_p, h->dpart, h->dpart2, h->dcg);
        coarse_apply(h, 1, s);
        hipLaunchKernelGGL(k_cg_init_final, dim3(1), dim3(kSysBlock), 0, s, (const double*)h->dpart,
                           (const double*)h->dpart2, std::min(nblk, kCgUpdBlocks), rtol * rtol, maxit, h->dcg);
        return;
    }
    hipLaunchKernelGGL(k_cg_init<false>, dim3(nblk), dim3(kSysBlock), 0, s, nn, (const double*)h->dMinv, db, dx, h->dcg_r,
                       h->dcg_z, h->dcg_p, h->dpart, h->dpart2, (CgScalars*)nullptr);
    hipLaunchKernelGGL(k_cg_init_final, dim3(1), dim3(kSysBlock), 0, s, (const double*)h->dpart,
                       (const double*)h->dpart2, nblk, rtol * rtol, maxit, h->dcg);
}

// iterations k0 .. k0 + n - 1 of PCG.  Per iteration: q = K p (partials of p.q in dpart3, one per
// node block); alpha, x, r, z (partials of r.z and r.r in dpart, dpart2, one per update
// workgroup); beta, test, p.  r.z of iteration k sits in rz[k & 1].  After the test has set the
// stop flag the kernels return at once, so iterations past convergence change nothing.
static void pcg_iterations(ceos_handle* h, double* dx, int k0, int n, hipStream_t s, bool two_level = false) {
    const SysGeom g = sys_geom(h);
    const int nn = h->nn;
    const int nblk = (nn + kSysBlock - 1) / kSysBlock;
    const int nupd = std::min(nblk, kCgUpdBlocks);
    for (int k = k0; k < k0 + n; ++k) {
        const int par = k & 1;
        if (h->ell_f32)
            hipLaunchKernelGGL((k_cg_spmv<float, kCgSlots>), dim3(nblk), dim3(kSysBlock), 0, s, g,
                               (const float*)h->dell, (const CgScalars*)h->dcg, (const double*)h->dcg_p, h->dcg_q,
                               h->dpart3);
        else
            hipLaunchKernelGGL((k_cg_spmv<double, kCgSlots64>), dim3(nblk), dim3(kSysBlock), 0, s, g,
                               (const double*)h->dell, (const CgScalars*)h->dcg, (const double*)h->dcg_p, h->dcg_q,
                               h->dpart3);
        if (two_level) {
            hipLaunchKernelGGL(k_cg_update<true>, dim3(nupd), dim3(kSysBlock), 0, s, nn, par, (const double*)h->dMinv,
                               h->dcg, (const double*)h->dpart3, nblk, (const double*)h->dcg_p,
                               (const double*)h->dcg_q, dx, h->dcg_r, h->dcg_z, h->dpart, h->dpart2);
            coarse_apply(h, 0, s);
        } else {
            hipLaunchKernelGGL(k_cg_update<false>, dim3(nupd), dim3(kSysBlock), 0, s, nn, par, (const double*)h->dMinv,
                               h->dcg, (const double*)h->dpart3, nblk, (const double*)h->dcg_p,
                               (const double*)h->dcg_q, dx, h->dcg_r, h->dcg_z, h->dpart, h->dpart2);
        }
        hipLaunchKernelGGL(k_cg_pupdate, dim3(nupd), dim3(kSysBlock), 0, s, nn, par, h->dcg, (const double*)h->dpart,
                           (const double*)h->dpart2, nupd, (const double*)h->dcg_z, h->dcg_p);
    }
}

// block-Jacobi PCG from x = 0.  Synchronises s every cg_check iterations to read the state.
static int pcg_run(ceos_handle* h, const double* db, double* dx, double rtol, int maxit, hipStream_t s, int* iters,
                   double* relres, int* flag) {
    const bool tl = two_level_on(h);
    pcg_start(h, db, dx, rtol, maxit, s, tl);
    CgScalars c;
    int k = 0;
    for (;;) {
        HIP_OK(hipMemcpyAsync(&c, h->dcg, sizeof(c), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        if (c.done) break;
        pcg_iterations(h, dx, k, h->cg_check, s, tl);
        k += h->cg_check;
        HIP_OK(hipGetLastError());
    }
    if (iters) *iters = c.iters;
    if (relres) *relres = c.bb > 0.0 ? sqrt(c.rr / c.bb) : 0.0;
    if (flag) *flag = c.done;
    return CEOS_OK;
}

// block-Jacobi preconditioned MINRES from x = 0 (symmetric indefinite K).  r1, r2, r_new rotate
// through dcg_r, dcg_q, dmr[0]; z = dcg_z; v = dmr[1]; w through dcg_p and dmr[2].
static int minres_run(ceos_handle* h, const double* db, double* dx, double rtol, int maxit, hipStream_t s,
                      int* iters, double* relres, int* flag) {
    int rc = ensure_minres(h);
    if (rc) return rc;
    const SysGeom g = sys_geom(h);
    const int nn = h->nn;
    const int nblk = (nn + kSysBlock - 1) / kSysBlock;
    double* R[3] = {h->dcg_r, h->dcg_q, h->dmr[0]};
    double* W[2] = {h->dcg_p, h->dmr[2]};
    double* z = h->dcg_z;
    double* v = h->dmr[1];
    MrScalars c0;
    memset(&c0, 0, sizeof(c0));
    c0.rtol = rtol;
    c0.maxit = maxit;
    HIP_OK(hipMemcpyAsync(h->dmr_c, &c0, sizeof(c0), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mr_init, dim3(nblk), dim3(kSysBlock), 0, s, nn, (const double*)h->dMinv, db, dx, R[0], R[1], z,
                       W[0], W[1], h->dpart);
    hipLaunchKernelGGL(k_mr_init_final, dim3(1), dim3(kSysBlock), 0, s, (const double*)h->dpart, nblk, h->dmr_c);
    MrScalars c;
    int k = 0;
    for (;;) {
        HIP_OK(hipMemcpyAsync(&c, h->dmr_c, sizeof(c), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        if (c.done || c.verdict) break;
        for (int i = 0; i < h->cg_check; ++i, ++k) {
            const int par = k & 1;
            double* r1 = R[k % 3];
            double* r2 = R[(k + 1) % 3];
            double* rn = R[(k + 2) % 3];
            double* wkm2 = W[k & 1];  // w_{k-2}; w_k replaces it
            double* wkm1 = W[(k + 1) & 1];
            if (h->ell_f32)
                hipLaunchKernelGGL((k_mr_spmv<float, kCgSlots>), dim3(nblk), dim3(kSysBlock), 0, s, g,
                                   (const float*)h->dell, h->dmr_c, par, (const double*)z,
                                   (const double*)r1, v, rn, h->dpart3);
            else
                hipLaunchKernelGGL((k_mr_spmv<double, kCgSlots64>), dim3(nblk), dim3(kSysBlock), 0, s, g,
                                   (const double*)h->dell, h->dmr_c, par, (const double*)z,
                                   (const double*)r1, v, rn, h->dpart3);
            hipLaunchKernelGGL(k_mr_lanczos, dim3(nblk), dim3(kSysBlock), 0, s, nn, par, (const double*)h->dMinv,
                               h->dmr_c, (const double*)h->dpart3, nblk, (const double*)r2, rn, z, h->dpart);
            hipLaunchKernelGGL(k_mr_update, dim3(nblk), dim3(kSysBlock), 0, s, nn, par, h->dmr_c,
                               (const double*)h->dpart, nblk, (const double*)v, (const double*)wkm1, wkm2, dx);
        }
        HIP_OK(hipGetLastError());
    }
    const int verdict = c.done ? c.done : c.verdict;
    if (iters) *iters = c.iters;
    if (relres) *relres = c.beta1 > 0.0 ? c.st[c.iters & 1].phibar / c.beta1 : 0.0;
    if (flag) *flag = verdict == 1 ? 4 : verdict;
    return CEOS_OK;
}

// x = K^-1 b from x = 0; K (upper CSR, the handle's precision) after the boundary conditions.
// CEOS_SOLVER_PCG: PCG only (flag 2 on an indefinite K); CEOS_SOLVER_MINRES: MINRES only;
// CEOS_SOLVER_AUTO: PCG, and MINRES from x = 0 when PCG breaks down.
static int linear_solve(ceos_handle* h, const void* dK, const double* db, double* dx, double rtol, int maxit,
                        hipStream_t s, int* iters, double* relres, int* flag) {
    if (!h->sys_ready || !dK || !db || !dx || !(rtol > 0.0) || maxit < 0) return CEOS_ERR_ARG;
    int rc = ensure_solver(h);
    if (rc) return rc;
    build_operator(h, dK, s);
    if (h->prec_kind == CEOS_PREC_TWO_LEVEL && h->solver_method != CEOS_SOLVER_MINRES) {
        if (!h->coarse_ready) {
            rc = coarse_setup(h);
            if (rc) return rc;
        }
        if (!h->coarse_built || h->coarse_uses > h->coarse_reuse) {
            rc = coarse_operator(h, dK, s);
            if (rc) return rc;
            h->coarse_built = true;
            h->coarse_uses = 0;
        }
        h->coarse_uses++;
    }
    int it = 0, fl = 0;
    double rr = 0.0;
    if (h->solver_method != CEOS_SOLVER_MINRES) {
        rc = pcg_run(h, db, dx, rtol, maxit, s, &it, &rr, &fl);
        if (rc) return rc;
    }
    if (h->solver_method == CEOS_SOLVER_MINRES || (h->solver_method == CEOS_SOLVER_AUTO && fl == 2)) {
        int it2 = 0;
        rc = minres_run(h, db, dx, rtol, maxit, s, &it2, &rr, &fl);
        if (rc) return rc;
        it += it2;
    }
    if (iters) *iters = it;
    if (relres) *relres = rr;
    if (flag) *flag = fl;
    return CEOS_OK;
}

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
extern "C" {

int ceos_mesh_create(int nn, const double* X, int ne, int npe, const int* conn, const int* mat_id,
                     ceos_handle** out) {
    if (!out) return CEOS_ERR_ARG;
    *out = nullptr;
    if (nn <= 0 || ne <= 0 || npe != 8 || !X || !conn) return CEOS_ERR_ARG;
    ceos_handle* h = new ceos_handle();
    h->nn = nn;
    h->ne = ne;
    h->X.assign(X, X + 3L * nn);
    h->conn0.resize((size_t)8 * ne);
    for (long i = 0; i < 8L * ne; ++i) {
        const int n = conn[i] - 1;
        if (n < 0 || n >= nn) { delete h; return CEOS_ERR_ARG; }
        h->conn0[i] = n;
    }
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a)
            for (int b = a + 1; b < 8; ++b)
                if (h->conn0[8 * e + a] == h->conn0[8 * e + b]) { delete h; return CEOS_ERR_ARG; }
    {
        std::map<int, int> dense;
        for (int e = 0; e < ne; ++e) dense[mat_id ? mat_id[e] : 1] = 0;
        if ((int)dense.size() > kMaxMat) { delete h; return CEOS_ERR_ARG; }
        for (auto& kv : dense) { kv.second = (int)h->mat_ids.size(); h->mat_ids.push_back(kv.first); }
        h->mat_of.resize(ne);
        for (int e = 0; e < ne; ++e) h->mat_of[e] = dense[mat_id ? mat_id[e] : 1];
    }
    int st = build_pattern(h);
    if (!st) st = build_colors(h);
    if (st) { delete h; return st; }
    if (hipGetDevice(&h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return CEOS_ERR_HIP;
    }
    (void)static_scratch(h->device);  // once per device: the address ceos_evaluate_device_static uses
    if (hipEventCreateWithFlags(&h->ev_caller, hipEventDisableTiming) != hipSuccess) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return CEOS_ERR_HIP;
    }
    auto fail = [&](int code) {
        free_device(h);
        (void)hipEventDestroy(h->ev_caller);
        (void)hipStreamDestroy(h->stream);
        delete h;
        return code;
    };
    if (dalloc(h->dX, 3L * nn) || dalloc(h->dx, 3L * nn) || dalloc(h->dU, 3L * nn) ||
        dalloc(h->dFint, 3L * nn) || dalloc(h->drowstart, nn) || dalloc(h->dnup, nn) ||
        dalloc(h->dprops, (size_t)kMaxMat * kMaxProps) || dalloc(h->derr, 2))
        return fail(CEOS_ERR_HIP);
    if (hipMemcpy(h->dX, X, sizeof(double) * 3L * nn, hipMemcpyHostToDevice) ||
        hipMemcpy(h->dx, X, sizeof(double) * 3L * nn, hipMemcpyHostToDevice) ||
        hipMemset(h->dU, 0, sizeof(double) * 3L * nn) ||
        hipMemcpy(h->drowstart, h->rowstart.data(), sizeof(int) * nn, hipMemcpyHostToDevice) ||
        hipMemcpy(h->dnup, h->nup.data(), sizeof(int) * nn, hipMemcpyHostToDevice) ||
        hipMemset(h->dprops, 0, sizeof(double) * kMaxMat * kMaxProps) ||
        hipMemset(h->derr, 0, 2 * sizeof(int)))
        return fail(CEOS_ERR_HIP);
    *out = h;
    return CEOS_OK;
}

static void timing_free(ceos_handle* h) {
    for (hipEvent_t e : h->tev) (void)hipEventDestroy(e);
    h->tev.clear();
    h->timing_cap = h->timing_n = 0;
    h->timing_on = false;
    h->timing_launches = 0;
}

void ceos_mesh_destroy(ceos_handle* h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    timing_free(h);
    free_device(h);
    for (hipEvent_t e : h->pipe_ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(h->ev_caller);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    (void)hipStreamDestroy(h->stream);
    delete h;
}

int ceos_set_pipeline(ceos_handle* h, int chunks) {
    if (!h || chunks < 0 || chunks > 64) return CEOS_ERR_ARG;
    h->pipe_req = chunks;
    h->finalized = false;  // replanned at the next evaluation
    return CEOS_OK;
}

int ceos_material_set(ceos_handle* h, int mat_id, int model, const double* props, int nprops) {
    if (!h || (!props && nprops > 0) || !model_supported(model) || nprops < model_nprops(model) ||
        nprops > kMaxProps)
        return CEOS_ERR_ARG;
    auto it = std::find(h->mat_ids.begin(), h->mat_ids.end(), mat_id);
    if (it == h->mat_ids.end()) return CEOS_ERR_ARG;
    const int i = (int)(it - h->mat_ids.begin());
    MatDef& m = h->mats[i];
    if (m.set && m.model != model) h->finalized = false;  // launch grouping changes
    if (!m.set) h->finalized = false;
    m.model = model;
    for (int k = 0; k < kMaxProps; ++k) m.p[k] = k < nprops && props ? props[k] : 0.0;
    m.set = true;
    h->f_valid = h->k_valid = false;
    return upload_props(h);
}

int ceos_set_analysis(ceos_handle* h, int nl_analysis, int elem_tech) {
    if (!h || (elem_tech != CEOS_ELEM_FULL_INTEGRATION && elem_tech != CEOS_ELEM_MEAN_DILATATION)) return CEOS_ERR_ARG;
    h->nl = nl_analysis ? 1 : 0;
    if (h->elem_tech != elem_tech) h->finalized = false;  // the material check runs again
    h->elem_tech = elem_tech;
    h->f_valid = h->k_valid = false;
    return CEOS_OK;
}

int ceos_set_precision(ceos_handle* h, int mode) {
    if (!h || (mode != CEOS_K_FP64 && mode != CEOS_K_FP32_STORED)) return CEOS_ERR_ARG;
    h->precision = mode;
    h->k_valid = false;
    return CEOS_OK;
}

int ceos_set_fused(ceos_handle* h, int fused) {
    if (!h) return CEOS_ERR_ARG;
    h->fused = fused != 0;
    return CEOS_OK;
}

int ceos_set_tangent_mode(ceos_handle* h, int mode) {
    if (!h || (mode != CEOS_TANGENT_ANALYTIC && mode != CEOS_TANGENT_NUMERICAL)) return CEOS_ERR_ARG;
    h->numerical_tangent = mode == CEOS_TANGENT_NUMERICAL;
    h->k_valid = false;
    return CEOS_OK;
}

int ceos_set_assembly(ceos_handle* h, int mode) {
    if (!h || (mode != CEOS_ASSEMBLY_GATHER && mode != CEOS_ASSEMBLY_COLORED)) return CEOS_ERR_ARG;
    if (mode != h->assembly) {
        h->assembly = mode;
        h->finalized = false;
        h->f_valid = h->k_valid = false;
    }
    return CEOS_OK;
}

int ceos_set_schedule(ceos_handle* h, int slabs) {
    if (!h || slabs < 0 || slabs > 4096) return CEOS_ERR_ARG;
    h->slabs = slabs;
    h->finalized = false;
    h->f_valid = h->k_valid = false;
    return CEOS_OK;
}

int ceos_mesh_info(ceos_handle* h, int* n_colors, long long* nnz_upper) {
    if (!h) return CEOS_ERR_ARG;
    if (n_colors) *n_colors = h->ncolor;
    if (nnz_upper) *nnz_upper = h->nnz;
    return CEOS_OK;
}

int ceos_assembly_info(ceos_handle* h, long long info[8]) {
    if (!h || !info) return CEOS_ERR_ARG;
    if (!h->finalized) return CEOS_ERR_ARG;
    long long patches = 0;
    for (auto& r : h->ranges) patches += (r.end - r.begin + kEPB - 1) / kEPB;
    info[0] = patches;
    info[1] = h->rec_doubles / 9;       // boundary records (one contribution word each)
    info[2] = h->n_boundary_diag;
    info[3] = h->n_complete_records;
    info[4] = h->n_complete_diag;
    info[5] = h->nblocks;
    info[6] = h->n_gather_blocks;
    info[7] = h->n_gather_diag;
    return CEOS_OK;
}

int ceos_csr_pattern(ceos_handle* h, int upper, int* n, long long* nnz, int* rowmap, int* col) {
    if (!h) return CEOS_ERR_ARG;
    const int nn = h->nn, ndof = 3 * nn;
    if (upper) {
        if (n) *n = ndof;
        if (nnz) *nnz = h->nnz;
        if (!rowmap && !col) return CEOS_OK;
        if (!rowmap || !col) return CEOS_ERR_ARG;
        long long k = 0;
        for (int A = 0; A < nn; ++A) {
            for (int i = 0; i < 3; ++i) {
                const int r = 3 * A + i;
                rowmap[r] = (int)(k + 1);
                for (int c = r; c < 3 * A + 3; ++c) col[k++] = c + 1;
                for (int q = h->up_ptr[A]; q < h->up_ptr[A + 1]; ++q) {
                    const int B = h->up_nodes[q];
                    for (int kk = 0; kk < 3; ++kk) col[k++] = 3 * B + kk + 1;
                }
            }
        }
        rowmap[ndof] = (int)(k + 1);
        return k == h->nnz ? CEOS_OK : CEOS_ERR_SLOT;
    }
    // full pattern: every neighbour (lower and upper) of the node, 3 columns each
    std::vector<int> lo_cnt(nn, 0);
    for (int A = 0; A < nn; ++A)
        for (int q = h->up_ptr[A]; q < h->up_ptr[A + 1]; ++q) lo_cnt[h->up_nodes[q]]++;
    long long total = 0;
    for (int A = 0; A < nn; ++A) total += 3LL * 3 * (1 + lo_cnt[A] + h->nup[A]);
    if (n) *n = ndof;
    if (nnz) *nnz = total;
    if (!rowmap && !col) return CEOS_OK;
    if (!rowmap || !col) return CEOS_ERR_ARG;
    std::vector<std::vector<int>> lower(nn);
    for (int A = 0; A < nn; ++A)
        for (int q = h->up_ptr[A]; q < h->up_ptr[A + 1]; ++q) lower[h->up_nodes[q]].push_back(A);
    long long k = 0;
    for (int A = 0; A < nn; ++A) {
        std::vector<int> nbs = lower[A];
        std::sort(nbs.begin(), nbs.end());
        nbs.push_back(A);
        for (int q = h->up_ptr[A]; q < h->up_ptr[A + 1]; ++q) nbs.push_back(h->up_nodes[q]);
        for (int i = 0; i < 3; ++i) {
            rowmap[3 * A + i] = (int)(k + 1);
            for (int B : nbs)
                for (int kk = 0; kk < 3; ++kk) col[k++] = 3 * B + kk + 1;
        }
    }
    rowmap[ndof] = (int)(k + 1);
    return CEOS_OK;
}

int ceos_solve_constitutive(ceos_handle* h, double time, const double* U, ceos_status* st) {
    if (!h || !U) return CEOS_ERR_ARG;
    int rc = prepare_for(h, h->stream);
    if (rc) { set_status(st, rc, 0, "material not set or mesh setup failed"); return rc; }
    rc = after_caller_stream(h);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(h->dU, U, sizeof(double) * 3L * h->nn, hipMemcpyHostToDevice, h->stream));
    h->have_u = true;
    h->time = time;
    void* K = nullptr;
    if (h->fused) {
        rc = ensure_k_buffers(h);
        if (rc) return rc;
        K = kbuf(h);
    }
    rc = evaluate(h, h->dU, h->dFint, K, h->stream);
    if (rc) { set_status(st, rc, 0, "launch failed"); return rc; }
    h->f_valid = true;
    h->k_valid = K != nullptr;
    h->k_bc = false;
    return read_status(h, st);
}

int ceos_internal_force(ceos_handle* h, double* Fint, ceos_status* st) {
    if (!h || !Fint || !h->have_u) return CEOS_ERR_ARG;
    if (!h->f_valid) {
        int rc = evaluate(h, h->dU, h->dFint, nullptr, h->stream);
        if (rc) return rc;
        h->f_valid = true;
    }
    HIP_OK(hipMemcpyAsync(Fint, h->dFint, sizeof(double) * 3L * h->nn, hipMemcpyDeviceToHost, h->stream));
    return read_status(h, st);
}

static int tangent_common(ceos_handle* h) {
    if (!h->have_u) return CEOS_ERR_ARG;
    if (!h->k_valid) {
        int rc = ensure_k_buffers(h);
        if (rc) return rc;
        rc = evaluate(h, h->dU, h->dFint, kbuf(h), h->stream);
        if (rc) return rc;
        h->k_valid = h->f_valid = true;
        h->k_bc = false;
    }
    return CEOS_OK;
}

int ceos_tangent(ceos_handle* h, double* Kval) {
    if (!h || !Kval) return CEOS_ERR_ARG;
    if (h->precision != CEOS_K_FP64) return CEOS_ERR_ARG;
    int rc = tangent_common(h);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(Kval, h->dK, sizeof(double) * h->nnz, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return CEOS_OK;
}

int ceos_tangent_f32(ceos_handle* h, float* Kval) {
    if (!h || !Kval) return CEOS_ERR_ARG;
    if (h->precision != CEOS_K_FP32_STORED) return CEOS_ERR_ARG;
    int rc = tangent_common(h);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(Kval, h->dK32, sizeof(float) * h->nnz, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return CEOS_OK;
}

int ceos_update_mesh(ceos_handle* h, const double* U) {
    if (!h || !U) return CEOS_ERR_ARG;
    std::vector<double> x(3L * h->nn);
    for (long i = 0; i < 3L * h->nn; ++i) x[i] = h->X[i] + U[i];
    int rc = after_caller_stream(h);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(h->dx, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    h->f_valid = h->k_valid = false;
    return CEOS_OK;
}

int ceos_switch_converged_state(ceos_handle* h) {
    if (!h) return CEOS_ERR_ARG;
    int rc = after_caller_stream(h);
    if (rc) return rc;
    if (h->finalized && h->has_state) {
        HIP_OK(hipMemcpyAsync(h->dst_old, h->dst_new, sizeof(double) * (size_t)h->ne * 8 * kStateStride,
                              hipMemcpyDeviceToDevice, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    h->f_valid = h->k_valid = false;
    return CEOS_OK;
}

int ceos_failed_elements(ceos_handle* h, int* ids, int cap) {
    if (!h) return CEOS_ERR_ARG;
    if (after_caller_stream(h)) return CEOS_ERR_HIP;
    HIP_OK(hipStreamSynchronize(h->stream));
    if (!h->finalized) return 0;
    std::vector<uint8_t> flags(h->ne);
    if (hipMemcpy(flags.data(), h->delem_err, h->ne, hipMemcpyDeviceToHost) != hipSuccess) return CEOS_ERR_HIP;
    std::vector<int> out;
    for (int s = 0; s < h->ne; ++s)
        if (flags[s]) out.push_back(h->perm[s] + 1);
    std::sort(out.begin(), out.end());
    for (int i = 0; i < (int)out.size() && i < cap; ++i) ids[i] = out[i];
    return (int)out.size();
}

int ceos_get_gp_stress(ceos_handle* h, double* sig) {
    if (!h || !sig || !h->finalized) return CEOS_ERR_ARG;
    if (after_caller_stream(h)) return CEOS_ERR_HIP;
    HIP_OK(hipStreamSynchronize(h->stream));
    const int ne = h->ne;
    std::vector<double> d((size_t)ne * 48);
    HIP_OK(hipMemcpy(d.data(), h->dsig, sizeof(double) * d.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < ne; ++s) {
        const int e = h->perm[s];
        for (int g = 0; g < 8; ++g)
            for (int k = 0; k < 6; ++k) sig[((size_t)e * 8 + g) * 6 + k] = d[((size_t)k * 8 + g) * ne + s];
    }
    return CEOS_OK;
}

int ceos_get_state(ceos_handle* h, int which, double* state) {
    if (!h || !state) return CEOS_ERR_ARG;
    if (after_caller_stream(h)) return CEOS_ERR_HIP;
    HIP_OK(hipStreamSynchronize(h->stream));
    int rc = finalize(h);
    if (rc) return rc;
    const int ne = h->ne;
    if (!h->has_state) {
        memset(state, 0, sizeof(double) * (size_t)ne * 8 * kStateStride);
        return CEOS_OK;
    }
    std::vector<double> d((size_t)ne * 8 * kStateStride);
    HIP_OK(hipMemcpy(d.data(), which ? h->dst_new : h->dst_old, sizeof(double) * d.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < ne; ++s) {
        const int e = h->perm[s];
        for (int g = 0; g < 8; ++g)
            for (int v = 0; v < kStateStride; ++v)
                state[((size_t)e * 8 + g) * kStateStride + v] = d[((size_t)v * 8 + g) * ne + s];
    }
    return CEOS_OK;
}

int ceos_set_state(ceos_handle* h, int which, const double* state) {
    if (!h || !state) return CEOS_ERR_ARG;
    if (after_caller_stream(h)) return CEOS_ERR_HIP;
    HIP_OK(hipStreamSynchronize(h->stream));
    int rc = finalize(h);
    if (rc) return rc;
    if (!h->has_state) return CEOS_OK;
    const int ne = h->ne;
    std::vector<double> d((size_t)ne * 8 * kStateStride);
    for (int s = 0; s < ne; ++s) {
        const int e = h->perm[s];
        for (int g = 0; g < 8; ++g)
            for (int v = 0; v < kStateStride; ++v)
                d[((size_t)v * 8 + g) * ne + s] = state[((size_t)e * 8 + g) * kStateStride + v];
    }
    HIP_OK(hipMemcpy(which ? h->dst_new : h->dst_old, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice));
    h->f_valid = h->k_valid = false;
    return CEOS_OK;
}

int ceos_homogenize(ceos_handle* h, double P[9], double F[9]) {
    if (!h || !P || !F || !h->have_u || !h->finalized) return CEOS_ERR_ARG;
    if (after_caller_stream(h)) return CEOS_ERR_HIP;
    const int nblk = (h->ne * 8 + kHomBlock - 1) / kHomBlock;
    if (h->hpart_n < nblk) {
        HIP_OK(dalloc(h->dhpart, (size_t)nblk * kHomVals));
        h->hpart_n = nblk;
    }
    hipLaunchKernelGGL(k_homogenize_partial, dim3(nblk), dim3(kHomBlock), 0, h->stream, h->ne, h->dconn, h->dX,
                       h->dU, h->dsig, h->dhpart);
    HIP_OK(hipGetLastError());
    std::vector<double> part((size_t)nblk * kHomVals);
    HIP_OK(hipMemcpyAsync(part.data(), h->dhpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    double acc[kHomVals] = {0};
    for (int b = 0; b < nblk; ++b)
        for (int k = 0; k < kHomVals; ++k) acc[k] += part[(size_t)b * kHomVals + k];
    const double V = acc[18];
    for (int k = 0; k < 9; ++k) {
        P[k] = acc[k] / V;
        F[k] = acc[9 + k] / V;
    }
    return CEOS_OK;
}

int ceos_state_save(ceos_handle* h, const char* path) {
    if (!h || !path) return CEOS_ERR_ARG;
    int rc = finalize(h);
    if (rc) return rc;
    rc = after_caller_stream(h);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(h->stream));
    FILE* f = fopen(path, "wb");
    if (!f) return CEOS_ERR_ARG;
    const int hdr[4] = {0x43454f53, h->nn, h->ne, h->has_state ? 1 : 0};
    std::vector<double> x(3L * h->nn);
    bool ok = hipMemcpy(x.data(), h->dx, sizeof(double) * x.size(), hipMemcpyDeviceToHost) == hipSuccess;
    ok = ok && fwrite(hdr, sizeof(hdr), 1, f) == 1 && fwrite(x.data(), sizeof(double), x.size(), f) == x.size();
    if (ok && h->has_state) {
        std::vector<double> s((size_t)h->ne * 8 * kStateStride);
        for (int which = 0; which < 2 && ok; ++which) {
            ok = ceos_get_state(h, which, s.data()) == CEOS_OK &&
                 fwrite(s.data(), sizeof(double), s.size(), f) == s.size();
        }
    }
    fclose(f);
    return ok ? CEOS_OK : CEOS_ERR_ARG;
}

int ceos_state_load(ceos_handle* h, const char* path) {
    if (!h || !path) return CEOS_ERR_ARG;
    int rc = finalize(h);
    if (rc) return rc;
    rc = after_caller_stream(h);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(h->stream));
    FILE* f = fopen(path, "rb");
    if (!f) return CEOS_ERR_ARG;
    int hdr[4];
    bool ok = fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == 0x43454f53 && hdr[1] == h->nn && hdr[2] == h->ne &&
              hdr[3] == (h->has_state ? 1 : 0);
    std::vector<double> x(3L * h->nn);
    ok = ok && fread(x.data(), sizeof(double), x.size(), f) == x.size();
    if (ok) ok = hipMemcpy(h->dx, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && h->has_state) {
        std::vector<double> s((size_t)h->ne * 8 * kStateStride);
        for (int which = 0; which < 2 && ok; ++which)
            ok = fread(s.data(), sizeof(double), s.size(), f) == s.size() && ceos_set_state(h, which, s.data()) == CEOS_OK;
    }
    fclose(f);
    h->f_valid = h->k_valid = false;
    return ok ? CEOS_OK : CEOS_ERR_ARG;
}

int ceos_evaluate_device(ceos_handle* h, double time, const double* dU, double* dFint, void* dK, void* stream) {
    if (!h || !dU || !dFint) return CEOS_ERR_ARG;
    h->time = time;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    int rc = evaluate(h, dU, dFint, dK, s);
    if (rc) return rc;
    return note_caller_stream(h, s);
}

int ceos_evaluate_device_static(ceos_handle* h, double time, const double* dU, double* dFint, void* dK,
                                void* stream) {
    if (!h || !dU || !dFint) return CEOS_ERR_ARG;
    if (!h->finalized) {
        fprintf(stderr, "ceos_evaluate_device_static: the handle is not prepared (call ceos_mesh_prepare)\n");
        return CEOS_ERR_ARG;
    }
    const size_t need = dK ? (size_t)h->rec_doubles : (size_t)h->fe_doubles;
    if (h->assembly == CEOS_ASSEMBLY_GATHER && need > kStaticScratchDoubles) {
        fprintf(stderr,
                "ceos_evaluate_device_static: %d elements need %.1f MiB of element records (%s), the static "
                "buffer holds %zu MiB\n",
                h->ne, need * sizeof(double) / 1048576.0, dK ? "F+K" : "residual only", kStaticScratchBytes >> 20);
        return CEOS_ERR_ARG;
    }
    double* buf = static_scratch(h->device);
    if (!buf) return CEOS_ERR_HIP;
    h->time = time;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    int rc = evaluate(h, dU, dFint, dK, s, buf);
    if (rc) return rc;
    return note_caller_stream(h, s);
}

int ceos_mesh_prepare(ceos_handle* h, int records) {
    if (!h) return CEOS_ERR_ARG;
    int rc = finalize(h);
    if (rc) return rc;
    return records ? ensure_records(h) : CEOS_OK;
}

int ceos_update_mesh_device(ceos_handle* h, const double* dU, void* stream) {
    if (!h || !dU) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    if (s == h->stream && after_caller_stream(h)) return CEOS_ERR_HIP;
    const long n = 3L * h->nn;
    hipLaunchKernelGGL(k_add, dim3((n + 255) / 256), dim3(256), 0, s, n, h->dX, dU, h->dx);
    HIP_OK(hipGetLastError());
    h->f_valid = h->k_valid = false;
    return note_caller_stream(h, s);
}

int ceos_timing_start(ceos_handle* h, int max_evaluations) {
    if (!h || max_evaluations <= 0 || max_evaluations > (1 << 20)) return CEOS_ERR_ARG;
    HIP_OK(hipStreamSynchronize(h->stream));
    timing_free(h);
    h->tev.resize(3L * max_evaluations, nullptr);
    for (auto& e : h->tev) HIP_OK(hipEventCreate(&e));
    h->timing_cap = max_evaluations;
    h->timing_on = true;
    return CEOS_OK;
}

int ceos_timing_stop(ceos_handle* h, double* element_ms, double* assembly_ms, int* n_evaluations,
                     long long* element_launches) {
    if (!h || !h->timing_on) return CEOS_ERR_ARG;
    double te = 0.0, ta = 0.0;
    for (int i = 0; i < h->timing_n; ++i) {
        float a = 0.f, b = 0.f;
        HIP_OK(hipEventSynchronize(h->tev[3 * i + 2]));
        HIP_OK(hipEventElapsedTime(&a, h->tev[3 * i], h->tev[3 * i + 1]));
        HIP_OK(hipEventElapsedTime(&b, h->tev[3 * i + 1], h->tev[3 * i + 2]));
        te += a;
        ta += b;
    }
    if (element_ms) *element_ms = te;
    if (assembly_ms) *assembly_ms = ta;
    if (n_evaluations) *n_evaluations = h->timing_n;
    if (element_launches) *element_launches = h->timing_launches;
    timing_free(h);
    return CEOS_OK;
}

int ceos_device_status(ceos_handle* h, ceos_status* st) {
    if (!h) return CEOS_ERR_ARG;
    return read_status(h, st);
}

int ceos_set_external_gp(ceos_handle* h, const double* dext) {
    if (!h) return CEOS_ERR_ARG;
    h->dext = dext;
    h->f_valid = h->k_valid = false;
    return CEOS_OK;
}

int ceos_gp_deformation_gradient_device(ceos_handle* h, const double* dU, double* dF, void* stream) {
    if (!h || !dU || !dF) return CEOS_ERR_ARG;
    int rc = finalize(h);
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    const long n = 8L * h->ne;
    hipLaunchKernelGGL(k_gp_deformation_gradient, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, h->ne,
                       h->dconn, h->dperm, h->dX, dU, dF);
    HIP_OK(hipGetLastError());
    return note_caller_stream(h, s);
}

int ceos_gp_deformation_gradient(ceos_handle* h, const double* U, double* F) {
    if (!h || !U || !F) return CEOS_ERR_ARG;
    int rc = finalize(h);
    if (rc) return rc;
    double *dUt = nullptr, *dF = nullptr;  // temporaries: the handle's own U stays untouched
    if (hipMalloc((void**)&dUt, sizeof(double) * 3L * h->nn) != hipSuccess) return CEOS_ERR_HIP;
    if (hipMalloc((void**)&dF, sizeof(double) * 72L * h->ne) != hipSuccess) {
        (void)hipFree(dUt);
        return CEOS_ERR_HIP;
    }
    rc = CEOS_ERR_HIP;
    if (hipMemcpyAsync(dUt, U, sizeof(double) * 3L * h->nn, hipMemcpyHostToDevice, h->stream) == hipSuccess &&
        ceos_gp_deformation_gradient_device(h, dUt, dF, h->stream) == CEOS_OK &&
        hipMemcpyAsync(F, dF, sizeof(double) * 72L * h->ne, hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
        hipStreamSynchronize(h->stream) == hipSuccess)
        rc = CEOS_OK;
    (void)hipFree(dUt);
    (void)hipFree(dF);
    return rc;
}

int ceos_bc_set(ceos_handle* h, int n_presc, const int* presc_dof, int n_fixed, const int* fixed_dof) {
    if (!h || n_presc < 0 || n_fixed < 0 || (n_presc && !presc_dof) || (n_fixed && !fixed_dof)) return CEOS_ERR_ARG;
    int rc = ensure_system(h);
    if (rc) return rc;
    const long ndof = 3L * h->nn;
    std::vector<uint8_t> dm(ndof, 0), nm(h->nn, 0);
    for (int i = 0; i < n_presc; ++i) {
        const long d = (long)presc_dof[i] - 1;
        if (d < 0 || d >= ndof) return CEOS_ERR_ARG;
        dm[d] |= kBcPresc;
    }
    for (int i = 0; i < n_fixed; ++i) {
        const long d = (long)fixed_dof[i] - 1;
        if (d < 0 || d >= ndof) return CEOS_ERR_ARG;
        dm[d] |= kBcFixed;
    }
    for (long d = 0; d < ndof; ++d) nm[d / 3] |= dm[d];
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipMemcpy(h->ddof_mask, dm.data(), ndof, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->dnode_mask, nm.data(), h->nn, hipMemcpyHostToDevice));
    h->n_presc = n_presc;
    h->n_fixed = n_fixed;
    if (dm != h->dof_mask_host) h->coarse_ready = h->coarse_built = false;  // P vanishes on constrained DOFs
    h->dof_mask_host.swap(dm);
    // the linear solver's buffers too, so that the solve of a Newton step may be captured
    return ensure_solver(h);
}

int ceos_residual_bc_device(ceos_handle* h, int it, const double* dU, const double* dFint, void* dK,
                            const double* dFext, const double* dUbar, double* dR, int norm_type, double* dnorm,
                            void* stream) {
    if (!h || !dFint || !dR || !dnorm || it < 0) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    int rc = residual_bc(h, it, dU, dFint, dK, dFext, dUbar, dR, norm_type, dnorm, s);
    if (rc) return rc;
    return note_caller_stream(h, s);
}

int ceos_residual_bc(ceos_handle* h, int it, const double* Fext, const double* Ubar, double* R, int norm_type,
                     double* normR) {
    if (!h || it < 0 || !h->sys_ready) return CEOS_ERR_ARG;
    if (it == 0 && h->n_presc > 0 && !Ubar) return CEOS_ERR_ARG;
    if (h->k_bc) h->k_valid = false;  // EvaluateKt assembles a fresh tangent every call
    int rc = tangent_common(h);
    if (rc) return rc;
    const long n = 3L * h->nn;
    if (Fext) {
        if (!h->dFext) HIP_OK(dalloc(h->dFext, n));
        HIP_OK(hipMemcpyAsync(h->dFext, Fext, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    }
    if (Ubar) {
        if (!h->dUbar) HIP_OK(dalloc(h->dUbar, n));
        HIP_OK(hipMemcpyAsync(h->dUbar, Ubar, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    }
    rc = residual_bc(h, it, h->dU, h->dFint, kbuf(h), Fext ? h->dFext : nullptr, Ubar ? h->dUbar : nullptr, h->dR,
                     norm_type, h->dnorm, h->stream);
    if (rc) return rc;
    h->k_bc = true;
    double nrm[3];
    HIP_OK(hipMemcpyAsync(nrm, h->dnorm, sizeof(nrm), hipMemcpyDeviceToHost, h->stream));
    if (R) HIP_OK(hipMemcpyAsync(R, h->dR, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (normR) *normR = nrm[2];
    return CEOS_OK;
}

int ceos_set_linear_solver(ceos_handle* h, int method) {
    if (!h || (method != CEOS_SOLVER_AUTO && method != CEOS_SOLVER_PCG && method != CEOS_SOLVER_MINRES))
        return CEOS_ERR_ARG;
    h->solver_method = method;
    return CEOS_OK;
}

int ceos_set_preconditioner(ceos_handle* h, int kind, double aggregate_size, int coarse_reuse) {
    if (!h || (kind != CEOS_PREC_BLOCK_JACOBI && kind != CEOS_PREC_TWO_LEVEL) || coarse_reuse < 0)
        return CEOS_ERR_ARG;
    const double a = aggregate_size > 0.0 ? aggregate_size : 8.0;
    if (a != h->agg_size) h->coarse_ready = false;
    h->prec_kind = kind;
    h->agg_size = a;
    h->coarse_reuse = coarse_reuse;
    h->coarse_built = false;
    return CEOS_OK;
}

int ceos_preconditioner_info(ceos_handle* h, int* n_aggregates, int* n_coarse) {
    if (!h) return CEOS_ERR_ARG;
    if (n_aggregates) *n_aggregates = h->coarse_ready ? h->n_agg : 0;
    if (n_coarse) *n_coarse = h->coarse_ready ? h->n_coarse : 0;
    return CEOS_OK;
}

int ceos_coarse_space(ceos_handle* h, int* agg, double* basis, double* Ac) {
    if (!h || !h->coarse_ready) return CEOS_ERR_ARG;
    HIP_OK(hipDeviceSynchronize());
    if (agg) HIP_OK(hipMemcpy(agg, h->dagg, sizeof(int) * h->nn, hipMemcpyDeviceToHost));
    if (basis) {
        std::vector<float> b32(18L * h->nn);
        HIP_OK(hipMemcpy(b32.data(), h->dbasis, sizeof(float) * b32.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < b32.size(); ++i) basis[i] = b32[i];
    }
    if (Ac) HIP_OK(hipMemcpy(Ac, h->dAc, sizeof(double) * (size_t)h->n_coarse * h->n_coarse, hipMemcpyDeviceToHost));
    return CEOS_OK;
}

int ceos_set_solver_precision(ceos_handle* h, int mode) {
    if (!h || (mode != -1 && mode != CEOS_K_FP64 && mode != CEOS_K_FP32_STORED)) return CEOS_ERR_ARG;
    h->solver_precision = mode;
    return CEOS_OK;
}

int ceos_linear_solve_device(ceos_handle* h, const void* dK, const double* db, double* dx, double rtol, int maxit,
                             int* iters, double* relres, int* flag, void* stream) {
    if (!h) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    int rc = linear_solve(h, dK, db, dx, rtol, maxit, s, iters, relres, flag);
    if (rc) return rc;
    return note_caller_stream(h, s);
}

int ceos_linear_solve_device_async(ceos_handle* h, const void* dK, const double* db, double* dx, double rtol,
                                   int maxit, double* dresult, void* stream) {
    if (!h || !h->sys_ready || !dK || !db || !dx || !dresult || !(rtol > 0.0) || maxit < 0) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    const int want_prec = h->solver_precision < 0 ? h->precision : h->solver_precision;
    const bool ready = h->dell && h->dMinv && h->ell_f32 == (want_prec == CEOS_K_FP32_STORED);
    if (!ready) {
        if (stream_capturing(s)) return CEOS_ERR_ARG;  // the buffers are allocated outside a capture
        int rc = ensure_solver(h);
        if (rc) return rc;
    }
    build_operator(h, dK, s);
    pcg_start(h, db, dx, rtol, maxit, s);
    pcg_iterations(h, dx, 0, maxit, s);
    hipLaunchKernelGGL(k_cg_report, dim3(1), dim3(64), 0, s, (const CgScalars*)h->dcg, dresult);
    HIP_OK(hipGetLastError());
    return note_caller_stream(h, s);
}

// ------------------------------------------------------------------------------------------
// Periodic reduction (general-mesh RVEs): T^T K T, T^T R and T w between a full hex8 handle and
// the handle of its periodic (torus) mesh; see periodic_kernel.hpp.
// ------------------------------------------------------------------------------------------
struct ceos_periodic {
    ceos_handle* full = nullptr;
    ceos_handle* red = nullptr;
    int nn = 0, nn_red = 0;
    long long nnz_red = 0;
    double v0 = 0.0;
    int nblk = 0;
    int *dmap = nullptr, *dkptr = nullptr, *drptr = nullptr, *drnodes = nullptr;
    uint32_t* dkidx = nullptr;
    double* dpart = nullptr;
};

// reference volume of a hex8 element by the 2x2x2 Gauss rule (exact for the trilinear map)
static double hex8_volume_host(const double* X, const int* c) {
    const double x = 0.57735026918962576451;
    auto sg = [](int a, int k) {
        const int s = k == 0 ? (((a + 1) >> 1) & 1) : (k == 1 ? ((a >> 1) & 1) : ((a >> 2) & 1));
        return s ? 1.0 : -1.0;
    };
    double v = 0.0;
    for (int g = 0; g < 8; ++g) {
        const double xi = sg(g, 0) * x, eta = sg(g, 1) * x, zeta = sg(g, 2) * x;
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 8; ++a) {
            const double s0 = sg(a, 0), s1 = sg(a, 1), s2 = sg(a, 2);
            const double d[3] = {s0 * (1.0 + s1 * eta) * (1.0 + s2 * zeta) * 0.125,
                                 s1 * (1.0 + s0 * xi) * (1.0 + s2 * zeta) * 0.125,
                                 s2 * (1.0 + s0 * xi) * (1.0 + s1 * eta) * 0.125};
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) J[i][k] += X[3L * c[a] + i] * d[k];
        }
        v += J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
    return v;
}

// 0-based CSR slot of the entry (3A+i, 3B+k), A <= B (i <= k when A == B), of a handle's upper
// pattern, or -1 when B is not an upper neighbour of A
static long long upper_slot(const ceos_handle* h, const std::vector<long long>& rs, int A, int B, int i, int k) {
    const long long nu = h->nup[A];
    const long long row = rs[A] + (i == 0 ? 0 : (i == 1 ? 3 + 3 * nu : 5 + 6 * nu));
    if (A == B) return row + (k - i);
    const int* b = h->up_nodes.data() + h->up_ptr[A];
    const int* e = h->up_nodes.data() + h->up_ptr[A + 1];
    const int* f = std::lower_bound(b, e, B);
    if (f == e || *f != B) return -1;
    return row + (3 - i) + 3LL * (f - b) + k;
}

void ceos_periodic_destroy(ceos_periodic* p) {
    if (!p) return;
    dfree(p->dmap); dfree(p->dkptr); dfree(p->drptr); dfree(p->drnodes); dfree(p->dkidx); dfree(p->dpart);
    delete p;
}

int ceos_periodic_create(ceos_handle* full, ceos_handle* red, const int* node_map, ceos_periodic** out) {
    if (!full || !red || !node_map || !out || full == red) return CEOS_ERR_ARG;
    *out = nullptr;
    if (full->precision != CEOS_K_FP64 || red->precision != CEOS_K_FP64) return CEOS_ERR_ARG;
    const int nn = full->nn, nr = red->nn;
    std::vector<int> cnt(nr, 0);
    for (int A = 0; A < nn; ++A) {
        if (node_map[A] < 0 || node_map[A] >= nr) return CEOS_ERR_ARG;
        cnt[node_map[A]]++;
    }
    for (int a = 0; a < nr; ++a)
        if (!cnt[a]) return CEOS_ERR_ARG;  // every reduced node is the image of some node
    auto row_starts = [](const ceos_handle* h) {
        std::vector<long long> rs(h->nn + 1, 0);
        for (int A = 0; A < h->nn; ++A) rs[A + 1] = rs[A] + 6 + 9LL * h->nup[A];
        return rs;
    };
    const std::vector<long long> rsf = row_starts(full), rsr = row_starts(red);
    if (rsf[nn] != full->nnz || rsr[nr] != red->nnz || full->nnz >= (1LL << 31)) return CEOS_ERR_ARG;
    // every stored full entry, in full slot order, onto its reduced upper entry
    std::vector<std::pair<long long, uint32_t>> con;  // (reduced slot, full slot | doubled)
    con.reserve((size_t)full->nnz);
    for (int A = 0; A < nn; ++A) {
        const int a = node_map[A];
        for (int i = 0; i < 3; ++i)
            for (int k = i; k < 3; ++k) {
                const long long fs = upper_slot(full, rsf, A, A, i, k);
                con.emplace_back(upper_slot(red, rsr, a, a, i, k), (uint32_t)fs);
            }
        for (int q = full->up_ptr[A]; q < full->up_ptr[A + 1]; ++q) {
            const int B = full->up_nodes[q], b = node_map[B];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) {
                    const long long fs = upper_slot(full, rsf, A, B, i, k);
                    long long rs;
                    uint32_t w = (uint32_t)fs;
                    if (a < b) rs = upper_slot(red, rsr, a, b, i, k);
                    else if (a > b) rs = upper_slot(red, rsr, b, a, k, i);
                    else {
                        // two periodic images in one element: K(i,k) and K(k,i) of the pair both
                        // land on the reduced diagonal block
                        rs = upper_slot(red, rsr, a, a, std::min(i, k), std::max(i, k));
                        if (i == k) w |= kPerDoubled;
                    }
                    if (rs < 0) return CEOS_ERR_SLOT;
                    con.emplace_back(rs, w);
                }
        }
    }
    std::vector<int> kptr((size_t)red->nnz + 1, 0);
    for (const auto& c : con) kptr[c.first + 1]++;
    for (long long e = 0; e < red->nnz; ++e) {
        if (!kptr[e + 1]) return CEOS_ERR_SLOT;  // a reduced entry no full entry reaches
        kptr[e + 1] += kptr[e];
    }
    std::vector<uint32_t> kidx(con.size());
    {
        std::vector<int> fill(kptr.begin(), kptr.end() - 1);
        for (const auto& c : con) kidx[fill[c.first]++] = c.second;  // stable: full slot order
    }
    std::vector<int> rptr(nr + 1, 0), rnodes(nn);
    for (int a = 0; a < nr; ++a) rptr[a + 1] = rptr[a] + cnt[a];
    {
        std::vector<int> fill(rptr.begin(), rptr.end() - 1);
        for (int A = 0; A < nn; ++A) rnodes[fill[node_map[A]]++] = A;
    }
    double v0 = 0.0;
    for (int e = 0; e < full->ne; ++e) v0 += hex8_volume_host(full->X.data(), full->conn0.data() + 8L * e);
    if (!(v0 > 0.0)) return CEOS_ERR_ARG;
    ceos_periodic* p = new ceos_periodic();
    p->full = full;
    p->red = red;
    p->nn = nn;
    p->nn_red = nr;
    p->nnz_red = red->nnz;
    p->v0 = v0;
    p->nblk = std::min((nn + kPerBlock - 1) / kPerBlock, 256);
    hipError_t e = hipSetDevice(full->device);
    if (e == hipSuccess) e = dalloc(p->dmap, nn);
    if (e == hipSuccess) e = dalloc(p->dkptr, kptr.size());
    if (e == hipSuccess) e = dalloc(p->dkidx, kidx.size());
    if (e == hipSuccess) e = dalloc(p->drptr, rptr.size());
    if (e == hipSuccess) e = dalloc(p->drnodes, rnodes.size());
    if (e == hipSuccess) e = dalloc(p->dpart, 9L * p->nblk);
    if (e == hipSuccess) e = hipMemcpy(p->dmap, node_map, sizeof(int) * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->dkptr, kptr.data(), sizeof(int) * kptr.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->dkidx, kidx.data(), sizeof(uint32_t) * kidx.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->drptr, rptr.data(), sizeof(int) * rptr.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->drnodes, rnodes.data(), sizeof(int) * rnodes.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ceos_periodic_destroy(p);
        return CEOS_ERR_HIP;
    }
    *out = p;
    return CEOS_OK;
}

int ceos_periodic_info(const ceos_periodic* p, int* nn_red, long long* contributions, double* volume) {
    if (!p) return CEOS_ERR_ARG;
    if (nn_red) *nn_red = p->nn_red;
    if (contributions) *contributions = 0;
    if (contributions) {
        int last = 0;
        HIP_OK(hipMemcpy(&last, p->dkptr + p->nnz_red, sizeof(int), hipMemcpyDeviceToHost));
        *contributions = last;
    }
    if (volume) *volume = p->v0;
    return CEOS_OK;
}

int ceos_periodic_reduce_device(ceos_periodic* p, const double* dK, const double* dF, double* dKred, double* dRred,
                                int mask_constrained, void* stream) {
    if (!p || (!dK && !dF) || (dK && !dKred) || (dF && !dRred)) return CEOS_ERR_ARG;
    if (mask_constrained && !p->red->sys_ready) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : p->red->stream;
    if (dK)
        hipLaunchKernelGGL(k_periodic_reduce_k, dim3((unsigned)((p->nnz_red + kPerBlock - 1) / kPerBlock)),
                           dim3(kPerBlock), 0, s, p->nnz_red, (const int*)p->dkptr, (const uint32_t*)p->dkidx, dK,
                           dKred);
    if (dF)
        hipLaunchKernelGGL(k_periodic_reduce_r, dim3((unsigned)((p->nn_red + kPerBlock - 1) / kPerBlock)),
                           dim3(kPerBlock), 0, s, p->nn_red, (const int*)p->drptr, (const int*)p->drnodes, dF,
                           mask_constrained ? (const uint8_t*)p->red->ddof_mask : nullptr, dRred);
    HIP_OK(hipGetLastError());
    return CEOS_OK;
}

int ceos_periodic_expand_device(ceos_periodic* p, const double* dw, double scale, double* dU, void* stream) {
    if (!p || !dw || !dU) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : p->red->stream;
    hipLaunchKernelGGL(k_periodic_expand, dim3((unsigned)((p->nn + kPerBlock - 1) / kPerBlock)), dim3(kPerBlock), 0,
                       s, p->nn, (const int*)p->dmap, dw, scale, dU);
    HIP_OK(hipGetLastError());
    return CEOS_OK;
}

int ceos_periodic_homogenize_device(ceos_periodic* p, const double* dF, double* dP, void* stream) {
    if (!p || !dF || !dP) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : p->red->stream;
    hipLaunchKernelGGL(k_periodic_moment_partial, dim3(p->nblk), dim3(kPerBlock), 0, s, p->nn,
                       (const double*)p->full->dX, dF, p->dpart);
    hipLaunchKernelGGL(k_periodic_moment_final, dim3(1), dim3(64), 0, s, p->nblk, (const double*)p->dpart,
                       1.0 / p->v0, dP);
    HIP_OK(hipGetLastError());
    return CEOS_OK;
}

int ceos_spmv_device(ceos_handle* h, const void* dK, const double* dx, double* dy, void* stream) {
    if (!h || !dK || !dx || !dy || !h->sys_ready) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    const SysGeom g = sys_geom(h);
    const unsigned nb = (unsigned)((h->nn + kSysBlock - 1) / kSysBlock);
    if (h->precision == CEOS_K_FP32_STORED)
        hipLaunchKernelGGL(k_sym_spmv<float>, dim3(nb), dim3(kSysBlock), 0, s, g, (const float*)dK, dx, dy);
    else
        hipLaunchKernelGGL(k_sym_spmv<double>, dim3(nb), dim3(kSysBlock), 0, s, g, (const double*)dK, dx, dy);
    HIP_OK(hipGetLastError());
    return note_caller_stream(h, s);
}

int ceos_linear_solve(ceos_handle* h, const double* b, double* x, double rtol, int maxit, int* iters,
                      double* relres, int* flag) {
    if (!h || !b || !x || !h->k_valid) return CEOS_ERR_ARG;
    const long n = 3L * h->nn;
    if (!h->dcg_b) {
        HIP_OK(dalloc(h->dcg_b, n));
        HIP_OK(dalloc(h->dcg_x, n));
    }
    HIP_OK(hipMemcpyAsync(h->dcg_b, b, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    int rc = linear_solve(h, kbuf(h), h->dcg_b, h->dcg_x, rtol, maxit, h->stream, iters, relres, flag);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(x, h->dcg_x, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return CEOS_OK;
}

int ceos_newton_update_device(ceos_handle* h, double* dU, const double* dDX, double scale, void* stream) {
    if (!h || !dU || !dDX) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    const long n = 3L * h->nn;
    hipLaunchKernelGGL(k_newton_update, dim3((unsigned)((n + kSysBlock - 1) / kSysBlock)), dim3(kSysBlock), 0, s, n,
                       scale, dDX, dU, (const double*)h->dX, h->nl ? h->dx : nullptr);
    HIP_OK(hipGetLastError());
    h->f_valid = h->k_valid = false;
    return note_caller_stream(h, s);
}

}  // extern "C"
