// Host side of the batched periodic RVE solver (C ABI in include/ceos_gpu.h).
//
// A batch holds n_rve voxel RVEs of sides 4..16 (even), their phase arrays, the converged state
// of each (macro F and fluctuation) and the per-workgroup scratch of the persistent solver.
// ceos_rve_solve runs every RVE to equilibrium at its new macro F in one launch of
// k_rve_solve (rve_kernel.hpp); queue order is by decreasing voxel count (longest first), so
// the last RVEs to start are the cheapest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "ceos_gpu.h"
#include "rve_kernel.hpp"

using namespace ceos;

#define RVE_HIP_OK(x)                                     \
    do {                                                  \
        if ((x) != hipSuccess) return CEOS_ERR_HIP;       \
    } while (0)

struct ceos_rve_batch {
    int n_rve = 0;
    int model = 0;
    int slots = 0;
    long long max_nodes = 0, total_nodes = 0;
    std::vector<int> side;
    hipStream_t stream = nullptr;
    double newton_rtol = 1e-10, newton_atol = 1e-13, pcg_rtol = 1e-10;
    int newton_maxit = 12, pcg_maxit = 5000, max_cutbacks = 20;
    int inexact = 1;
    int precision = CEOS_K_FP64;  // storage of the reduced tangent
    int cg_group = kCgGroupDefault;  // nodes per thread and operand batch in the CG sweeps
    double tangent_rtol = -1.0;   // < 0: pcg_rtol
    int coarse = 1;               // two-level CG preconditioner (aggregate translations)
    int warm = 1;                 // condensation solves start from the last ones' solutions
    // device
    int *dorder = nullptr, *dside = nullptr, *dinfo = nullptr, *dqueue = nullptr;
    long long *dnode_off = nullptr, *dphase_off = nullptr;
    uint8_t* dphase = nullptr;
    double *dprops = nullptr, *dFbar = nullptr, *dFconv = nullptr, *dw = nullptr, *dP = nullptr, *dA = nullptr;
    double *dM = nullptr, *dV = nullptr;
    float* dC = nullptr;  // [slots][kRveMaxCoarse^2] coarse inverses
    float* dVs = nullptr;  // [total nodes][27] the last condensation's solutions
    int* dvvalid = nullptr;  // [n_rve]
    long long* dprof = nullptr;
    // subset solves: selected ids in queue order and in caller order, compact F / P / A / info
    int *dsel_q = nullptr, *dsel_u = nullptr, *dinfo_s = nullptr;
    double *dF_s = nullptr, *dP_s = nullptr, *dA_s = nullptr;
};

namespace {

template <typename T>
hipError_t rmalloc(T*& p, size_t n) {
    return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
}

template <typename T>
void rfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_all(ceos_rve_batch* b) {
    rfree(b->dorder); rfree(b->dside); rfree(b->dinfo); rfree(b->dqueue);
    rfree(b->dnode_off); rfree(b->dphase_off); rfree(b->dphase);
    rfree(b->dprops); rfree(b->dFbar); rfree(b->dFconv); rfree(b->dw); rfree(b->dP); rfree(b->dA);
    rfree(b->dM); rfree(b->dV); rfree(b->dC); rfree(b->dVs); rfree(b->dvvalid); rfree(b->dprof);
    rfree(b->dsel_q); rfree(b->dsel_u); rfree(b->dinfo_s); rfree(b->dF_s); rfree(b->dP_s); rfree(b->dA_s);
}

__global__ void k_rve_reset(int n_rve, long long nodes3, double* Fconv, double* w, long long* prof, int* vvalid) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2LL * n_rve) prof[t] = 0;
    if (t < n_rve) vvalid[t] = 0;
    if (t < 9LL * n_rve) {
        const int k = (int)(t % 9);
        Fconv[t] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
    }
    if (t < nodes3) w[t] = 0.0;
}

__global__ void k_zero_queue(int* q) { q[0] = 0; }

// macro F of the selected RVEs into their rows of Fbar
__global__ void k_rve_scatter_F(int n_sel, const int* sel, const double* Fs, double* Fbar) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * n_sel) return;
    Fbar[9LL * sel[t / 9] + t % 9] = Fs[t];
}

// results of the selected RVEs into compact arrays (caller order)
__global__ void k_rve_gather(int n_sel, const int* sel, const double* P, const double* A, const int* info,
                             double* Ps, double* As, int* infos) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 81 * n_sel) return;
    const int k = t / 81, j = t % 81, r = sel[k];
    if (As) As[t] = A[81LL * r + j];
    if (j < 9) Ps[9 * k + j] = P[9LL * r + j];
    if (j < 4) infos[4 * k + j] = info[4LL * r + j];
}

// solve = false: the condensation only (the autotuner's probe)
template <int MODEL, typename MT, int G, bool NT>
void launch_nt(const RveParams& p, int grid, bool solve, bool tangent, hipStream_t s) {
    if (solve) hipLaunchKernelGGL((k_rve_solve<MODEL, MT, G, NT>), dim3(grid), dim3(kRveThreads), 0, s, p);
    if (tangent) {
        hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(1), 0, s, p.queue);
        hipLaunchKernelGGL((k_rve_tangent<MODEL, MT, G, NT>), dim3(grid), dim3(kRveThreads), 0, s, p);
    }
}

template <int MODEL, typename MT, int G>
void launch(const RveParams& p, int grid, bool solve, bool tangent, hipStream_t s) {
    if (p.nt_tangent) launch_nt<MODEL, MT, G, true>(p, grid, solve, tangent, s);
    else launch_nt<MODEL, MT, G, false>(p, grid, solve, tangent, s);
}

template <int MODEL, typename MT>
void launch_g(const RveParams& p, int grid, bool solve, bool tangent, int g, hipStream_t s) {
    if (g == 2) launch<MODEL, MT, 2>(p, grid, solve, tangent, s);
    else if (g == 8) launch<MODEL, MT, 8>(p, grid, solve, tangent, s);
    else launch<MODEL, MT, 4>(p, grid, solve, tangent, s);
}

template <int MODEL>
void launch_p(const RveParams& p, int grid, bool solve, bool tangent, bool f32, int g, hipStream_t s) {
    if (f32) launch_g<MODEL, float>(p, grid, solve, tangent, g, s);
    else launch_g<MODEL, double>(p, grid, solve, tangent, g, s);
}

int reset(ceos_rve_batch* b) {
    const long long n = std::max<long long>(9LL * b->n_rve, 3 * b->total_nodes);
    hipLaunchKernelGGL(k_rve_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, b->n_rve,
                       3 * b->total_nodes, b->dFconv, b->dw, b->dprof, b->dvvalid);
    RVE_HIP_OK(hipGetLastError());
    RVE_HIP_OK(hipStreamSynchronize(b->stream));
    return CEOS_OK;
}

constexpr int kRveNtQueue = 128;  // queue length from which the CG reads tangents non-temporally

// one persistent launch over the queue `order` (n_queue RVE ids, largest first); solve = false
// runs the condensation alone at the RVEs' current states; probe = true runs it as the tile probe
// (one single-right-hand-side solve per RVE; states, A and info unchanged)
int run(ceos_rve_batch* b, bool want_A, hipStream_t s, const int* order, int n_queue, bool solve = true,
        int cg_group = 0, bool probe = false) {
    RveParams p;
    p.n_rve = n_queue;
    p.order = order;
    p.side = b->dside;
    p.node_off = b->dnode_off;
    p.phase_off = b->dphase_off;
    p.phase = b->dphase;
    p.props = b->dprops;
    p.Fbar = b->dFbar;
    p.Fconv = b->dFconv;
    p.w = b->dw;
    p.P = b->dP;
    p.A = want_A ? b->dA : nullptr;
    p.info = b->dinfo;
    p.Mbuf = b->dM;
    p.Vbuf = b->dV;
    p.max_nodes = b->max_nodes;
    p.queue = b->dqueue;
    p.newton_rtol = b->newton_rtol;
    p.newton_atol = b->newton_atol;
    p.pcg_rtol = b->pcg_rtol;
    p.tangent_rtol = b->tangent_rtol > 0.0 ? b->tangent_rtol : b->pcg_rtol;
    p.newton_maxit = b->newton_maxit;
    p.pcg_maxit = b->pcg_maxit;
    p.max_cutbacks = b->max_cutbacks;
    p.inexact = b->inexact;
    p.probe = probe ? 1 : 0;
    // A full batch streams far more reduced tangents than the caches hold (256 slots x up to 4 MB
    // in FP32 at 16^3): its CG reads them non-temporally (the plain FE^2 evaluation: 17.5 -> 15.6 s).
    // A short queue (judges, fallbacks: tens of RVEs) keeps ordinary loads, whose tangents the
    // 256 MB last-level cache can hold between CG iterations.
    p.nt_tangent = n_queue >= kRveNtQueue ? 1 : 0;
    p.prof = b->dprof;
    p.Cbuf = b->dC;
    p.coarse = b->coarse;
    p.V = b->dVs;
    p.vvalid = b->dvvalid;
    p.warm = b->warm;
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(1), 0, s, b->dqueue);
    const int grid = std::min(n_queue, b->slots);
    const bool f32 = b->precision == CEOS_K_FP32_STORED;
    const int g = cg_group ? cg_group : b->cg_group;
    switch (b->model) {
        case kNH: launch_p<kNH>(p, grid, solve, want_A, f32, g, s); break;
        case kCNH: launch_p<kCNH>(p, grid, solve, want_A, f32, g, s); break;
        case kSVK: launch_p<kSVK>(p, grid, solve, want_A, f32, g, s); break;
        default: return CEOS_ERR_ARG;
    }
    RVE_HIP_OK(hipGetLastError());
    return CEOS_OK;
}

}  // namespace

extern "C" {

int ceos_rve_batch_create(int n_rve, const int* side, const unsigned char* phase, int model,
                          const double* props, int nprops, ceos_rve_batch** out) {
    if (!out) return CEOS_ERR_ARG;
    *out = nullptr;
    if (n_rve <= 0 || !side || !phase || !props || nprops <= 0 || nprops > 8) return CEOS_ERR_ARG;
    if (model != kNH && model != kCNH && model != kSVK) return CEOS_ERR_ARG;
    for (int i = 0; i < n_rve; ++i)
        if (side[i] < 4 || side[i] > kRveMaxSide || (side[i] & 1)) return CEOS_ERR_ARG;
    ceos_rve_batch* b = new ceos_rve_batch();
    b->n_rve = n_rve;
    b->model = model;
    b->side.assign(side, side + n_rve);
    std::vector<long long> node_off(n_rve), phase_off(n_rve);
    long long tot = 0;
    for (int i = 0; i < n_rve; ++i) {
        const long long nn = (long long)side[i] * side[i] * side[i];
        node_off[i] = tot;
        phase_off[i] = tot;
        tot += nn;
        b->max_nodes = std::max(b->max_nodes, nn);
    }
    b->total_nodes = tot;
    for (long long k = 0; k < tot; ++k)
        if (phase[k] > 1) {
            delete b;
            return CEOS_ERR_ARG;
        }
    std::vector<int> order(n_rve);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return side[a] > side[c]; });
    std::vector<double> pr(16, 0.0);
    for (int ph = 0; ph < 2; ++ph)
        for (int k = 0; k < nprops; ++k) pr[8 * ph + k] = props[nprops * ph + k];

    int dev = 0, cus = 0;
    int rc = CEOS_OK;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0 ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return CEOS_ERR_HIP;
    }
    b->slots = cus;  // ~155 KB of LDS per 256-thread workgroup: one per CU
    const long long slots = std::min<long long>(b->slots, n_rve);
    bool okm =
        rmalloc(b->dorder, n_rve) == hipSuccess && rmalloc(b->dside, n_rve) == hipSuccess &&
        rmalloc(b->dinfo, 4LL * n_rve) == hipSuccess && rmalloc(b->dqueue, 4) == hipSuccess &&
        rmalloc(b->dnode_off, n_rve) == hipSuccess && rmalloc(b->dphase_off, n_rve) == hipSuccess &&
        rmalloc(b->dphase, tot) == hipSuccess && rmalloc(b->dprops, 16) == hipSuccess &&
        rmalloc(b->dFbar, 9LL * n_rve) == hipSuccess && rmalloc(b->dFconv, 9LL * n_rve) == hipSuccess &&
        rmalloc(b->dw, 3 * tot) == hipSuccess && rmalloc(b->dP, 9LL * n_rve) == hipSuccess &&
        rmalloc(b->dA, 81LL * n_rve) == hipSuccess && rmalloc(b->dprof, 2LL * n_rve) == hipSuccess &&
        rmalloc(b->dM, (size_t)(243 * b->max_nodes * slots)) == hipSuccess &&
        rmalloc(b->dV, (size_t)(kRveVecPerNode * b->max_nodes * slots)) == hipSuccess &&
        rmalloc(b->dC, (size_t)kRveMaxCoarse * kRveMaxCoarse * slots) == hipSuccess &&
        rmalloc(b->dVs, (size_t)27 * tot) == hipSuccess && rmalloc(b->dvvalid, n_rve) == hipSuccess &&
        rmalloc(b->dsel_q, n_rve) == hipSuccess && rmalloc(b->dsel_u, n_rve) == hipSuccess &&
        rmalloc(b->dinfo_s, 4LL * n_rve) == hipSuccess && rmalloc(b->dF_s, 9LL * n_rve) == hipSuccess &&
        rmalloc(b->dP_s, 9LL * n_rve) == hipSuccess && rmalloc(b->dA_s, 81LL * n_rve) == hipSuccess;
    if (!okm) rc = CEOS_ERR_HIP;
    if (rc == CEOS_OK) {
        bool okc =
            hipMemcpy(b->dorder, order.data(), sizeof(int) * n_rve, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->dside, side, sizeof(int) * n_rve, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->dnode_off, node_off.data(), sizeof(long long) * n_rve, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->dphase_off, phase_off.data(), sizeof(long long) * n_rve, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->dphase, phase, (size_t)tot, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->dprops, pr.data(), sizeof(double) * 16, hipMemcpyHostToDevice) == hipSuccess;
        if (!okc) rc = CEOS_ERR_HIP;
    }
    if (rc == CEOS_OK) rc = reset(b);
    if (rc != CEOS_OK) {
        free_all(b);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return rc;
    }
    *out = b;
    return CEOS_OK;
}

void ceos_rve_batch_destroy(ceos_rve_batch* b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    free_all(b);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

int ceos_rve_set_solver(ceos_rve_batch* b, double newton_rtol, double newton_atol, int newton_maxit,
                        double pcg_rtol, int pcg_maxit, int max_cutbacks) {
    if (!b || !(newton_rtol > 0.0) || !(newton_atol >= 0.0) || newton_maxit < 1 || newton_maxit > 100 ||
        !(pcg_rtol > 0.0) || pcg_maxit < 1 || max_cutbacks < 0 || max_cutbacks > 60)
        return CEOS_ERR_ARG;
    b->newton_rtol = newton_rtol;
    b->newton_atol = newton_atol;
    b->newton_maxit = newton_maxit;
    b->pcg_rtol = pcg_rtol;
    b->pcg_maxit = pcg_maxit;
    b->max_cutbacks = max_cutbacks;
    return CEOS_OK;
}

int ceos_rve_set_inexact(ceos_rve_batch* b, int inexact) {
    if (!b || (inexact != 0 && inexact != 1)) return CEOS_ERR_ARG;
    b->inexact = inexact;
    return CEOS_OK;
}

int ceos_rve_set_precision(ceos_rve_batch* b, int mode) {
    if (!b || (mode != CEOS_K_FP64 && mode != CEOS_K_FP32_STORED)) return CEOS_ERR_ARG;
    b->precision = mode;
    return CEOS_OK;
}

int ceos_rve_set_tangent_rtol(ceos_rve_batch* b, double rtol) {
    if (!b || !(rtol > 0.0 && rtol < 1.0)) return CEOS_ERR_ARG;
    b->tangent_rtol = rtol;
    return CEOS_OK;
}

int ceos_rve_set_warm_start(ceos_rve_batch* b, int on) {
    if (!b || (on != 0 && on != 1)) return CEOS_ERR_ARG;
    b->warm = on;
    return CEOS_OK;
}

int ceos_rve_set_coarse(ceos_rve_batch* b, int on) {
    if (!b || (on != 0 && on != 1)) return CEOS_ERR_ARG;
    b->coarse = on;
    return CEOS_OK;
}

int ceos_rve_reset(ceos_rve_batch* b) {
    if (!b) return CEOS_ERR_ARG;
    return reset(b);
}

int ceos_rve_solve(ceos_rve_batch* b, const double* Fbar, double* P, double* A, int* info) {
    if (!b || !Fbar || !P) return CEOS_ERR_ARG;
    const int n = b->n_rve;
    RVE_HIP_OK(hipMemcpyAsync(b->dFbar, Fbar, sizeof(double) * 9 * n, hipMemcpyHostToDevice, b->stream));
    int rc = run(b, A != nullptr, b->stream, b->dorder, n);
    if (rc) return rc;
    RVE_HIP_OK(hipMemcpyAsync(P, b->dP, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, b->stream));
    if (A) RVE_HIP_OK(hipMemcpyAsync(A, b->dA, sizeof(double) * 81 * n, hipMemcpyDeviceToHost, b->stream));
    std::vector<int> inf(4 * (size_t)n);
    RVE_HIP_OK(hipMemcpyAsync(inf.data(), b->dinfo, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, b->stream));
    RVE_HIP_OK(hipStreamSynchronize(b->stream));
    if (info) std::memcpy(info, inf.data(), sizeof(int) * 4 * n);
    for (int i = 0; i < n; ++i)
        if (inf[4 * i] != 0) return CEOS_ERR_LOCAL_NR;
    return CEOS_OK;
}

int ceos_rve_solve_device(ceos_rve_batch* b, const double* dFbar, double* dP, double* dA, int* dinfo, void* stream) {
    if (!b || !dFbar || !dP) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    const int n = b->n_rve;
    RVE_HIP_OK(hipMemcpyAsync(b->dFbar, dFbar, sizeof(double) * 9 * n, hipMemcpyDeviceToDevice, s));
    int rc = run(b, dA != nullptr, s, b->dorder, n);
    if (rc) return rc;
    RVE_HIP_OK(hipMemcpyAsync(dP, b->dP, sizeof(double) * 9 * n, hipMemcpyDeviceToDevice, s));
    if (dA) RVE_HIP_OK(hipMemcpyAsync(dA, b->dA, sizeof(double) * 81 * n, hipMemcpyDeviceToDevice, s));
    if (dinfo) RVE_HIP_OK(hipMemcpyAsync(dinfo, b->dinfo, sizeof(int) * 4 * n, hipMemcpyDeviceToDevice, s));
    return CEOS_OK;
}

int ceos_rve_solve_subset(ceos_rve_batch* b, int n_sel, const int* sel, const double* Fsel, double* Psel,
                          double* Asel, int* info_sel) {
    if (!b || n_sel < 0 || n_sel > b->n_rve || (n_sel > 0 && (!sel || !Fsel || !Psel))) return CEOS_ERR_ARG;
    if (n_sel == 0) return CEOS_OK;
    std::vector<char> seen(b->n_rve, 0);
    for (int k = 0; k < n_sel; ++k) {
        if (sel[k] < 0 || sel[k] >= b->n_rve || seen[sel[k]]) return CEOS_ERR_ARG;  // a duplicate would race
        seen[sel[k]] = 1;
    }
    std::vector<int> q(sel, sel + n_sel);
    std::stable_sort(q.begin(), q.end(), [&](int a, int c) { return b->side[a] > b->side[c]; });
    hipStream_t s = b->stream;
    RVE_HIP_OK(hipMemcpyAsync(b->dsel_q, q.data(), sizeof(int) * n_sel, hipMemcpyHostToDevice, s));
    RVE_HIP_OK(hipMemcpyAsync(b->dsel_u, sel, sizeof(int) * n_sel, hipMemcpyHostToDevice, s));
    RVE_HIP_OK(hipMemcpyAsync(b->dF_s, Fsel, sizeof(double) * 9 * n_sel, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_rve_scatter_F, dim3((9 * n_sel + 255) / 256), dim3(256), 0, s, n_sel, b->dsel_u, b->dF_s,
                       b->dFbar);
    int rc = run(b, Asel != nullptr, s, b->dsel_q, n_sel);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rve_gather, dim3((81 * n_sel + 255) / 256), dim3(256), 0, s, n_sel, b->dsel_u, b->dP,
                       Asel ? b->dA : nullptr, b->dinfo, b->dP_s, Asel ? b->dA_s : nullptr, b->dinfo_s);
    RVE_HIP_OK(hipGetLastError());
    std::vector<int> inf(4 * (size_t)n_sel);
    RVE_HIP_OK(hipMemcpyAsync(Psel, b->dP_s, sizeof(double) * 9 * n_sel, hipMemcpyDeviceToHost, s));
    if (Asel) RVE_HIP_OK(hipMemcpyAsync(Asel, b->dA_s, sizeof(double) * 81 * n_sel, hipMemcpyDeviceToHost, s));
    RVE_HIP_OK(hipMemcpyAsync(inf.data(), b->dinfo_s, sizeof(int) * 4 * n_sel, hipMemcpyDeviceToHost, s));
    RVE_HIP_OK(hipStreamSynchronize(s));
    if (info_sel) std::memcpy(info_sel, inf.data(), sizeof(int) * 4 * n_sel);
    for (int k = 0; k < n_sel; ++k)
        if (inf[4 * k] != 0) return CEOS_ERR_LOCAL_NR;
    return CEOS_OK;
}

int ceos_rve_autotune(ceos_rve_batch* b, int max_rves, int* chosen, double* seconds) {
    if (!b || max_rves < 1) return CEOS_ERR_ARG;
    // sample: up to max_rves RVEs, largest first (the CG of the largest sets the batch's time)
    std::vector<int> q(b->n_rve);
    std::iota(q.begin(), q.end(), 0);
    std::stable_sort(q.begin(), q.end(), [&](int a, int c) { return b->side[a] > b->side[c]; });
    const int n = std::min(max_rves, b->n_rve);
    hipStream_t s = b->stream;
    RVE_HIP_OK(hipMemcpyAsync(b->dsel_q, q.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    RVE_HIP_OK(hipEventCreate(&e0));
    RVE_HIP_OK(hipEventCreate(&e1));
    // the probe only needs the time per CG iteration: a loose tolerance keeps it short (iteration
    // counts do not depend on the tile)
    const double keep_rtol = b->tangent_rtol;
    b->tangent_rtol = 1e-4;
    int best = kCgGroupDefault;
    double tbest = 1e300;
    int rc = run(b, true, s, b->dsel_q, n, false, kCgGroupDefault, true);  // warm-up
    for (int k = 0; k < 3 && rc == CEOS_OK; ++k) {
        const int g = kCgGroups[k];
        float ms = 0.0f;
        if (hipEventRecord(e0, s) != hipSuccess) rc = CEOS_ERR_HIP;
        if (rc == CEOS_OK) rc = run(b, true, s, b->dsel_q, n, false, g, true);
        if (rc == CEOS_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = CEOS_ERR_HIP;
        if (rc != CEOS_OK) break;
        if (seconds) seconds[k] = 1e-3 * ms;
        if (ms < tbest) {
            tbest = ms;
            best = g;
        }
    }
    b->tangent_rtol = keep_rtol;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != CEOS_OK) return rc;
    b->cg_group = best;
    if (chosen) *chosen = best;
    return CEOS_OK;
}

int ceos_rve_set_cg_group(ceos_rve_batch* b, int g) {
    if (!b || (g != 2 && g != 4 && g != 8)) return CEOS_ERR_ARG;
    b->cg_group = g;
    return CEOS_OK;
}

int ceos_rve_profile(ceos_rve_batch* b, long long* ticks) {
    if (!b || !ticks) return CEOS_ERR_ARG;
    RVE_HIP_OK(hipStreamSynchronize(b->stream));
    RVE_HIP_OK(hipMemcpy(ticks, b->dprof, sizeof(long long) * 2 * b->n_rve, hipMemcpyDeviceToHost));
    return CEOS_OK;
}

// Converged state of every RVE of the batch: F_conv (the macro F of the last converged solve),
// the fluctuation w and the last homogenized P and dP/dF.  The file starts with a header (magic,
// RVE count, total torus nodes) and the sides, so a state only loads into a batch of the same
// RVEs.  Both calls synchronise the device first: a solve enqueued on any stream has finished.
static const int kRveMagic = 0x43525645;

int ceos_rve_state_save(ceos_rve_batch* b, const char* path) {
    if (!b || !path) return CEOS_ERR_ARG;
    RVE_HIP_OK(hipDeviceSynchronize());
    const long long n = b->n_rve, tot = b->total_nodes;
    std::vector<double> Fc(9 * n), w(3 * tot), P(9 * n), A(81 * n);
    std::vector<int> vv(n);
    std::vector<float> V(27 * tot);
    RVE_HIP_OK(hipMemcpy(Fc.data(), b->dFconv, sizeof(double) * Fc.size(), hipMemcpyDeviceToHost));
    RVE_HIP_OK(hipMemcpy(w.data(), b->dw, sizeof(double) * w.size(), hipMemcpyDeviceToHost));
    RVE_HIP_OK(hipMemcpy(P.data(), b->dP, sizeof(double) * P.size(), hipMemcpyDeviceToHost));
    RVE_HIP_OK(hipMemcpy(A.data(), b->dA, sizeof(double) * A.size(), hipMemcpyDeviceToHost));
    RVE_HIP_OK(hipMemcpy(vv.data(), b->dvvalid, sizeof(int) * vv.size(), hipMemcpyDeviceToHost));
    RVE_HIP_OK(hipMemcpy(V.data(), b->dVs, sizeof(float) * V.size(), hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) return CEOS_ERR_ARG;
    const long long hdr[3] = {kRveMagic, n, tot};
    bool ok = fwrite(hdr, sizeof(hdr), 1, f) == 1 && fwrite(b->side.data(), sizeof(int), n, f) == (size_t)n;
    for (const auto* v : {&Fc, &w, &P, &A})
        ok = ok && fwrite(v->data(), sizeof(double), v->size(), f) == v->size();
    // the warm-start solutions: a restarted batch then continues bitwise
    ok = ok && fwrite(vv.data(), sizeof(int), vv.size(), f) == vv.size() &&
         fwrite(V.data(), sizeof(float), V.size(), f) == V.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? CEOS_OK : CEOS_ERR_ARG;
}

int ceos_rve_state_load(ceos_rve_batch* b, const char* path) {
    if (!b || !path) return CEOS_ERR_ARG;
    RVE_HIP_OK(hipDeviceSynchronize());
    FILE* f = fopen(path, "rb");
    if (!f) return CEOS_ERR_ARG;
    const long long n = b->n_rve, tot = b->total_nodes;
    long long hdr[3];
    std::vector<int> side(n);
    bool ok = fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == kRveMagic && hdr[1] == n && hdr[2] == tot &&
              fread(side.data(), sizeof(int), n, f) == (size_t)n && side == b->side;
    std::vector<double> Fc(9 * n), w(3 * tot), P(9 * n), A(81 * n);
    std::vector<int> vv(n);
    std::vector<float> V(27 * tot);
    for (auto* v : {&Fc, &w, &P, &A})
        ok = ok && fread(v->data(), sizeof(double), v->size(), f) == v->size();
    ok = ok && fread(vv.data(), sizeof(int), vv.size(), f) == vv.size() &&
         fread(V.data(), sizeof(float), V.size(), f) == V.size();
    fclose(f);
    if (!ok) return CEOS_ERR_ARG;
    RVE_HIP_OK(hipMemcpy(b->dFconv, Fc.data(), sizeof(double) * Fc.size(), hipMemcpyHostToDevice));
    RVE_HIP_OK(hipMemcpy(b->dw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    RVE_HIP_OK(hipMemcpy(b->dP, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice));
    RVE_HIP_OK(hipMemcpy(b->dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice));
    RVE_HIP_OK(hipMemcpy(b->dvvalid, vv.data(), sizeof(int) * vv.size(), hipMemcpyHostToDevice));
    RVE_HIP_OK(hipMemcpy(b->dVs, V.data(), sizeof(float) * V.size(), hipMemcpyHostToDevice));
    return CEOS_OK;
}

int ceos_rve_get_fluctuation(ceos_rve_batch* b, int rve, double* w) {
    if (!b || !w || rve < 0 || rve >= b->n_rve) return CEOS_ERR_ARG;
    long long off = 0;
    for (int i = 0; i < rve; ++i) off += (long long)b->side[i] * b->side[i] * b->side[i];
    const long long nn = (long long)b->side[rve] * b->side[rve] * b->side[rve];
    RVE_HIP_OK(hipStreamSynchronize(b->stream));
    RVE_HIP_OK(hipMemcpy(w, b->dw + 3 * off, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost));
    return CEOS_OK;
}

}  // extern "C"
