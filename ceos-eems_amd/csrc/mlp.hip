// MLP surrogate of the homogenized RVE response, trained and evaluated on the GPU.
//
// Network: n_in -> width -> width -> width -> n_out, SiLU after each hidden layer, parameters in
// one FP32 array [W1 b1 W2 b2 W3 b3 W4 b4], W_l row-major [out][in].
//
// Training (ceos_mlp_train_step): BF16 MFMA GEMMs with FP32 accumulation
// (v_mfma_f32_32x32x16_bf16, operands rounded to BF16 when loaded, activations and gradients
// kept in FP32), mean squared error, Adam on the FP32 master parameters.
// Inference (ceos_mlp_forward): FP32 GEMMs, and the input Jacobian dY/dX by reverse mode with
// the n_out unit seeds stacked into one batch.
//
// GEMM convention: C[m][n] = sum_k A(m,k) B(k,n) with A(m,k) = a[m*sam + k*sak] and
// B(k,n) = b[k*sbk + n*sbn]; the epilogue adds the bias / applies SiLU or its derivative.
// MFMA lane layout of the 32x32x16 BF16 instruction (wave64): lane l holds A row l%32, columns
// k = 8(l/32)..8(l/32)+7, B column l%32, rows 8(l/32)..+7; accumulator j of lane l is
// C[8(j/4) + 4(l/32) + j%4][l%32].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ceos_gpu.h"

#define MLP_HIP_OK(x)                               \
    do {                                            \
        if ((x) != hipSuccess) return CEOS_ERR_HIP; \
    } while (0)

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum Epi : int {
    kEpiStore = 0,   // C = acc
    kEpiBiasSilu,    // Z = acc + bias; C = Z, A = silu(Z)
    kEpiBias,        // C = acc + bias
    kEpiDSilu,       // C = acc * silu'(Z[m % zrows][n])
};

struct Gemm {
    int M, N, K;
    const float* a;
    long long sam, sak;
    const float* b;
    long long sbk, sbn;
    float* c;          // [M][N] row-major (ldc = N), or split partials [split][M][N]
    float* act;        // kEpiBiasSilu: silu(Z)
    const float* bias; // [N]
    const float* z;    // kEpiDSilu: pre-activation [zrows][N]
    int zrows;
    int ksplit;        // K range per split (split-K partial sums)
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

template <int EPI>
__device__ __forceinline__ void epilogue(const Gemm& g, float* c, int m, int n, float acc) {
    if (m >= g.M || n >= g.N) return;
    const long long o = (long long)m * g.N + n;
    if constexpr (EPI == kEpiStore) {
        c[o] = acc;
    } else if constexpr (EPI == kEpiBiasSilu) {
        const float zz = acc + g.bias[n];
        c[o] = zz;
        g.act[o] = silu(zz);
    } else if constexpr (EPI == kEpiBias) {
        c[o] = acc + g.bias[n];
    } else {
        c[o] = acc * dsilu(g.z[(long long)(m % g.zrows) * g.N + n]);
    }
}

// BF16 MFMA GEMM: 256 threads = 4 waves, 64x64 tile of C per workgroup (blockIdx.x over M,
// blockIdx.y over N), wave w owns the 32x32 sub-tile (32 (w/2), 32 (w%2)); blockIdx.z = split of
// the K range (partials at c + z M N).
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(Gemm g) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m0 = blockIdx.x * 64 + 32 * (w >> 1), n0 = blockIdx.y * 64 + 32 * (w & 1);
    const int kb = blockIdx.z * g.ksplit, ke = min(g.K, kb + g.ksplit);
    const int mr = m0 + (l & 31), nc = n0 + (l & 31), kl = 8 * (l >> 5);
    f32x16 acc;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += 16) {
        bf16x8 av, bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + kl + j;
            const bool kin = k < ke;
            av[j] = (__bf16)((kin && mr < g.M) ? g.a[(long long)mr * g.sam + (long long)k * g.sak] : 0.0f);
            bv[j] = (__bf16)((kin && nc < g.N) ? g.b[(long long)k * g.sbk + (long long)nc * g.sbn] : 0.0f);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    float* c = g.c + (long long)blockIdx.z * g.M * g.N;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int m = m0 + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
        epilogue<EPI>(g, c, m, nc, acc[j]);
    }
}

// FP32 GEMM (inference): 256 threads, 64x64 tile, 4x4 outputs per thread, 16-deep LDS stages
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32(Gemm g) {
    __shared__ float As[16][64 + 1];
    __shared__ float Bs[16][64 + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + 256 * r;  // 0..1023: (kk, mm) of the 16x64 stage
            const int kk = e >> 6, mm = e & 63;
            const int k = k0 + kk;
            const int m = m0 + mm, n = n0 + mm;
            As[kk][mm] = (k < g.K && m < g.M) ? g.a[(long long)m * g.sam + (long long)k * g.sak] : 0.0f;
            Bs[kk][mm] = (k < g.K && n < g.N) ? g.b[(long long)k * g.sbk + (long long)n * g.sbn] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) epilogue<EPI>(g, g.c, m0 + ty + 16 * i, n0 + tx + 16 * j, acc[i][j]);
}

// sum of split-K partials in split order: out[t] = sum_s part[s][t]
__global__ void k_reduce_splits(const float* part, int splits, long long n, float* out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(long long)k * n + t];
    out[t] = s;
}

// column sums of dZ [rows][cols] (bias gradients), one thread per column, rows in order
__global__ void k_colsum(const float* dz, int rows, int cols, float* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += dz[(long long)r * cols + c];
    out[c] = s;
}

// dY = 2 W (Y - T) / (B n_out), per-block partial sums of W (Y - T)^2 (W = 1 without weights)
__global__ void k_mse_grad(const float* y, const float* t, const float* w, long long n, float scale, float* dy,
                           float* part) {
    __shared__ float red[256];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    float e = 0.0f;
    if (i < n) {
        const float d = y[i] - t[i];
        const float wi = w ? w[i] : 1.0f;
        dy[i] = 2.0f * wi * d * scale;
        e = wi * d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float* p, const float* g, float* m1, float* m2, long long n, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// Jacobian seeds: S[o*B + b][n] = W_last[o][n] * silu'(Z_last[b][n])
__global__ void k_seed(const float* wlast, const float* z, int B, int width, int n_out, float* s) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)n_out * B * width;
    if (t >= tot) return;
    const int n = (int)(t % width);
    const long long r = t / width;
    const int b = (int)(r % B), o = (int)(r / B);
    s[t] = wlast[(long long)o * width + n] * dsilu(z[(long long)b * width + n]);
}

// J[b][o][i] from the stacked [o*B + b][i]
__global__ void k_jac_layout(const float* in, int B, int n_out, int n_in, float* J) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)B * n_out * n_in) return;
    const int i = (int)(t % n_in);
    const long long r = t / n_in;
    const int o = (int)(r % n_out), b = (int)(r / n_out);
    J[t] = in[((long long)o * B + b) * n_in + i];
}

}  // namespace

struct ceos_mlp {
    int n_in = 0, width = 0, n_out = 0, max_batch = 0;
    long long nparams = 0;
    long long off_w[4] = {0, 0, 0, 0}, off_b[4] = {0, 0, 0, 0};
    int din[4] = {0, 0, 0, 0}, dout[4] = {0, 0, 0, 0};
    hipStream_t stream = nullptr;
    long long step = 0;
    bool train_fp32 = false;  // reference mode: the training GEMMs in FP32 (k_gemm_f32)
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    // activations for max_batch (FP32): Z1..Z3, A1..A3, Y, dY, dZ (2 buffers), split partials
    float *Z[3] = {nullptr, nullptr, nullptr}, *A[3] = {nullptr, nullptr, nullptr};
    float *Y = nullptr, *dY = nullptr, *dZa = nullptr, *dZb = nullptr, *part = nullptr, *lossp = nullptr;
    // inference Jacobian: n_out stacked batches
    float *J0 = nullptr, *J1 = nullptr, *Jx = nullptr;
};

namespace {

constexpr int kSplits = 16;

template <typename T>
hipError_t mmalloc(T*& p, size_t n) {
    return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
}
template <typename T>
void mfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_all(ceos_mlp* m) {
    mfree(m->params); mfree(m->grads); mfree(m->m1); mfree(m->m2);
    for (int i = 0; i < 3; ++i) { mfree(m->Z[i]); mfree(m->A[i]); }
    mfree(m->Y); mfree(m->dY); mfree(m->dZa); mfree(m->dZb); mfree(m->part); mfree(m->lossp);
    mfree(m->J0); mfree(m->J1); mfree(m->Jx);
}

Gemm gemm(int M, int N, int K, const float* a, long long sam, long long sak, const float* b, long long sbk,
          long long sbn, float* c) {
    Gemm g;
    std::memset(&g, 0, sizeof(g));
    g.M = M; g.N = N; g.K = K;
    g.a = a; g.sam = sam; g.sak = sak;
    g.b = b; g.sbk = sbk; g.sbn = sbn;
    g.c = c;
    g.zrows = 1;
    g.ksplit = K;
    return g;
}

template <int EPI>
void run_bf16(const Gemm& g, hipStream_t s, int splits = 1) {
    dim3 grid((g.M + 63) / 64, (g.N + 63) / 64, splits);
    hipLaunchKernelGGL(k_gemm_bf16<EPI>, grid, dim3(256), 0, s, g);
}

template <int EPI>
void run_f32(const Gemm& g, hipStream_t s) {
    dim3 grid((g.M + 63) / 64, (g.N + 63) / 64, 1);
    hipLaunchKernelGGL(k_gemm_f32<EPI>, grid, dim3(256), 0, s, g);
}

// W (row-major [out][in]) as the B operand of X W^T: B(k, n) = W[n][k]
Gemm fwd(const ceos_mlp* m, int l, int rows, const float* x, float* z) {
    Gemm g = gemm(rows, m->dout[l], m->din[l], x, m->din[l], 1, m->params + m->off_w[l], 1, m->din[l], z);
    g.bias = m->params + m->off_b[l];
    return g;
}

// training GEMM: BF16 MFMA (split-K partials at c + z M N), or FP32 in the reference mode (one split)
template <int EPI>
void run_train(const ceos_mlp* m, const Gemm& g, hipStream_t s, int splits = 1) {
    if (m->train_fp32) {
        Gemm h = g;
        h.ksplit = g.K;
        run_f32<EPI>(h, s);
    } else {
        run_bf16<EPI>(g, s, splits);
    }
}

// forward for training: keeps Z and A of the hidden layers, Y
void forward_train(ceos_mlp* m, const float* X, int B, hipStream_t s) {
    const float* in = X;
    for (int l = 0; l < 3; ++l) {
        Gemm g = fwd(m, l, B, in, m->Z[l]);
        g.act = m->A[l];
        run_train<kEpiBiasSilu>(m, g, s);
        in = m->A[l];
    }
    run_train<kEpiBias>(m, fwd(m, 3, B, in, m->Y), s);
}

// weight gradient dW_l = dZ^T A_in (split-K over the batch, deterministic reduction) and bias
void weight_grad(ceos_mlp* m, int l, const float* dz, const float* ain, int B, hipStream_t s) {
    const int no = m->dout[l], ni = m->din[l];
    Gemm g = gemm(no, ni, B, dz, 1, no, ain, ni, 1, m->part);
    g.ksplit = (((B + kSplits - 1) / kSplits) + 15) / 16 * 16;
    const int splits = m->train_fp32 ? 1 : (B + g.ksplit - 1) / g.ksplit;
    run_train<kEpiStore>(m, g, s, splits);
    const long long n = (long long)no * ni;
    hipLaunchKernelGGL(k_reduce_splits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->part, splits, n,
                       m->grads + m->off_w[l]);
    hipLaunchKernelGGL(k_colsum, dim3((no + 255) / 256), dim3(256), 0, s, dz, B, no, m->grads + m->off_b[l]);
}

// forward + loss + backward into m->grads; loss partials in m->lossp
int grads(ceos_mlp* m, const float* X, const float* T, const float* W, int B, hipStream_t s) {
    forward_train(m, X, B, s);
    const long long ny = (long long)B * m->n_out;
    const unsigned nb = (unsigned)((ny + 255) / 256);
    hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(256), 0, s, m->Y, T, W, ny, 1.0f / (float)ny, m->dY, m->lossp);
    // layer 4
    weight_grad(m, 3, m->dY, m->A[2], B, s);
    // dZ3 = (dY W4) * silu'(Z3)
    const float* dcur = m->dY;
    float* bufs[2] = {m->dZa, m->dZb};
    for (int l = 3; l >= 1; --l) {
        float* dprev = bufs[l & 1];
        Gemm g = gemm(B, m->din[l], m->dout[l], dcur, m->dout[l], 1, m->params + m->off_w[l], m->din[l], 1, dprev);
        g.z = m->Z[l - 1];
        g.zrows = B;
        run_train<kEpiDSilu>(m, g, s);
        weight_grad(m, l - 1, dprev, l - 1 == 0 ? X : m->A[l - 2], B, s);
        dcur = dprev;
    }
    return hipGetLastError() == hipSuccess ? CEOS_OK : CEOS_ERR_HIP;
}

int read_loss(ceos_mlp* m, int B, float* loss) {
    const long long ny = (long long)B * m->n_out;
    const int nb = (int)((ny + 255) / 256);
    std::vector<float> part(nb);
    MLP_HIP_OK(hipMemcpyAsync(part.data(), m->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_OK(hipStreamSynchronize(m->stream));
    double sum = 0.0;
    for (float v : part) sum += v;
    *loss = (float)(sum / (double)ny);
    return CEOS_OK;
}

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int ceos_mlp_create(int n_in, int width, int n_out, int max_batch, unsigned long long seed, ceos_mlp** out) {
    if (!out) return CEOS_ERR_ARG;
    *out = nullptr;
    if (n_in < 1 || n_in > 64 || width < 16 || width > 1024 || n_out < 1 || n_out > 64 || max_batch < 1 ||
        max_batch > (1 << 20))
        return CEOS_ERR_ARG;
    ceos_mlp* m = new ceos_mlp();
    m->n_in = n_in;
    m->width = width;
    m->n_out = n_out;
    m->max_batch = max_batch;
    const int din[4] = {n_in, width, width, width}, dout[4] = {width, width, width, n_out};
    long long off = 0;
    for (int l = 0; l < 4; ++l) {
        m->din[l] = din[l];
        m->dout[l] = dout[l];
        m->off_w[l] = off;
        off += (long long)din[l] * dout[l];
        m->off_b[l] = off;
        off += dout[l];
    }
    m->nparams = off;
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases from splitmix64(seed)
    std::vector<float> p(off);
    uint64_t st = seed;
    for (int l = 0; l < 4; ++l) {
        const float r = 1.0f / std::sqrt((float)din[l]);
        for (long long i = m->off_w[l]; i < m->off_b[l] + dout[l]; ++i) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[i] = (float)((2.0 * u - 1.0) * r);
        }
    }
    const long long B = max_batch, H = width;
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess &&
              mmalloc(m->params, off) == hipSuccess && mmalloc(m->grads, off) == hipSuccess &&
              mmalloc(m->m1, off) == hipSuccess && mmalloc(m->m2, off) == hipSuccess &&
              mmalloc(m->Y, B * n_out) == hipSuccess && mmalloc(m->dY, B * n_out) == hipSuccess &&
              mmalloc(m->dZa, B * H) == hipSuccess && mmalloc(m->dZb, B * H) == hipSuccess &&
              mmalloc(m->part, (long long)kSplits * H * H) == hipSuccess &&
              mmalloc(m->lossp, (B * n_out + 255) / 256) == hipSuccess &&
              mmalloc(m->J0, (long long)n_out * B * H) == hipSuccess &&
              mmalloc(m->J1, (long long)n_out * B * H) == hipSuccess &&
              mmalloc(m->Jx, (long long)n_out * B * n_in) == hipSuccess;
    for (int i = 0; i < 3 && ok; ++i) ok = mmalloc(m->Z[i], B * H) == hipSuccess && mmalloc(m->A[i], B * H) == hipSuccess;
    ok = ok && hipMemcpy(m->params, p.data(), sizeof(float) * off, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(m->m1, 0, sizeof(float) * off) == hipSuccess && hipMemset(m->m2, 0, sizeof(float) * off) == hipSuccess;
    if (!ok) {
        free_all(m);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        return CEOS_ERR_HIP;
    }
    *out = m;
    return CEOS_OK;
}

void ceos_mlp_destroy(ceos_mlp* m) {
    if (!m) return;
    (void)hipStreamSynchronize(m->stream);
    free_all(m);
    (void)hipStreamDestroy(m->stream);
    delete m;
}

long long ceos_mlp_num_params(const ceos_mlp* m) { return m ? m->nparams : -1; }

int ceos_mlp_set_training_precision(ceos_mlp* m, int fp32) {
    if (!m || (fp32 != 0 && fp32 != 1)) return CEOS_ERR_ARG;
    m->train_fp32 = fp32 == 1;
    return CEOS_OK;
}

int ceos_mlp_shape(const ceos_mlp* m, int* n_in, int* width, int* n_out, int* max_batch) {
    if (!m) return CEOS_ERR_ARG;
    if (n_in) *n_in = m->n_in;
    if (width) *width = m->width;
    if (n_out) *n_out = m->n_out;
    if (max_batch) *max_batch = m->max_batch;
    return CEOS_OK;
}

int ceos_mlp_get_params(ceos_mlp* m, float* p) {
    if (!m || !p) return CEOS_ERR_ARG;
    MLP_HIP_OK(hipStreamSynchronize(m->stream));
    MLP_HIP_OK(hipMemcpy(p, m->params, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
    return CEOS_OK;
}

int ceos_mlp_set_params(ceos_mlp* m, const float* p) {
    if (!m || !p) return CEOS_ERR_ARG;
    MLP_HIP_OK(hipStreamSynchronize(m->stream));
    MLP_HIP_OK(hipMemcpy(m->params, p, sizeof(float) * m->nparams, hipMemcpyHostToDevice));
    MLP_HIP_OK(hipMemset(m->m1, 0, sizeof(float) * m->nparams));
    MLP_HIP_OK(hipMemset(m->m2, 0, sizeof(float) * m->nparams));
    m->step = 0;
    return CEOS_OK;
}

int ceos_mlp_gradients_weighted(ceos_mlp* m, const float* dX, const float* dT, const float* dW, int batch,
                                float* gout, float* loss) {
    if (!m || !dX || !dT || batch < 1 || batch > m->max_batch) return CEOS_ERR_ARG;
    int rc = grads(m, dX, dT, dW, batch, m->stream);
    if (rc) return rc;
    if (gout) MLP_HIP_OK(hipMemcpyAsync(gout, m->grads, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->stream));
    if (loss) return read_loss(m, batch, loss);
    MLP_HIP_OK(hipStreamSynchronize(m->stream));
    return CEOS_OK;
}

int ceos_mlp_gradients(ceos_mlp* m, const float* dX, const float* dT, int batch, float* gout, float* loss) {
    return ceos_mlp_gradients_weighted(m, dX, dT, nullptr, batch, gout, loss);
}

int ceos_mlp_train_step_weighted(ceos_mlp* m, const float* dX, const float* dT, const float* dW, int batch, float lr,
                                 float* loss) {
    if (!m || !dX || !dT || batch < 1 || batch > m->max_batch || !(lr > 0.0f)) return CEOS_ERR_ARG;
    int rc = grads(m, dX, dT, dW, batch, m->stream);
    if (rc) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->nparams + 255) / 256)), dim3(256), 0, m->stream, m->params,
                       m->grads, m->m1, m->m2, m->nparams, lr, b1, b2, c1, c2, 1e-8f);
    MLP_HIP_OK(hipGetLastError());
    if (loss) return read_loss(m, batch, loss);
    return CEOS_OK;
}

int ceos_mlp_train_step(ceos_mlp* m, const float* dX, const float* dT, int batch, float lr, float* loss) {
    return ceos_mlp_train_step_weighted(m, dX, dT, nullptr, batch, lr, loss);
}

int ceos_mlp_forward(ceos_mlp* m, const float* dX, int n, float* dY, float* dJ, void* stream) {
    if (!m || !dX || !dY || n < 1 || n > m->max_batch) return CEOS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : m->stream;
    const float* in = dX;
    for (int l = 0; l < 3; ++l) {
        Gemm g = fwd(m, l, n, in, m->Z[l]);
        g.act = m->A[l];
        run_f32<kEpiBiasSilu>(g, s);
        in = m->A[l];
    }
    run_f32<kEpiBias>(fwd(m, 3, n, in, dY), s);
    if (dJ) {
        const long long nb = (long long)m->n_out * n;
        const long long tot = nb * m->width;
        hipLaunchKernelGGL(k_seed, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m->params + m->off_w[3],
                           m->Z[2], n, m->width, m->n_out, m->J0);
        float* cur = m->J0;
        float* nxt = m->J1;
        for (int l = 2; l >= 1; --l) {
            Gemm g = gemm((int)nb, m->din[l], m->dout[l], cur, m->dout[l], 1, m->params + m->off_w[l], m->din[l], 1, nxt);
            g.z = m->Z[l - 1];
            g.zrows = n;
            run_f32<kEpiDSilu>(g, s);
            float* t = cur;
            cur = nxt;
            nxt = t;
        }
        run_f32<kEpiStore>(gemm((int)nb, m->n_in, m->width, cur, m->width, 1, m->params + m->off_w[0], m->n_in, 1, m->Jx), s);
        const long long nj = nb * m->n_in;
        hipLaunchKernelGGL(k_jac_layout, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, s, m->Jx, n, m->n_out,
                           m->n_in, dJ);
    }
    MLP_HIP_OK(hipGetLastError());
    if (!stream) MLP_HIP_OK(hipStreamSynchronize(s));
    return CEOS_OK;
}

}  // extern "C"
