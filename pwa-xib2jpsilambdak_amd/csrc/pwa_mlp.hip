// Learned start-point proposer (include/pwa_mlp.h): an MLP trained with
// hand-written BF16 MFMA kernels (v_mfma_f32_16x16x16_bf16, FP32
// accumulation), FP32 master weights and Adam.
//
// Kernels:
//  k_gemm<TA,TB,EPI> : C = op(A) op(B), bf16 operands, fp32 accumulation;
//                      64x64 block tile, 32x32 per wave (2x2 MFMA tiles);
//                      epilogues: store fp32 (+bias), or z = acc + bias ->
//                      Z (fp32) and A = gelu(z) (bf16)
//  k_gelu_back       : dZ = dA * gelu'(Z)  (fp32 and bf16 copies)
//  k_loss            : dOut = 2 w (out - y) / B, per-element loss terms
//  k_colsum          : bias gradients
//  k_adam            : Adam update of the flat fp32 parameters + bf16 copy
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "pwa_amp.h"
#include "pwa_mlp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi { EPI_STORE = 0, EPI_GELU = 1 };

// C[M x N] = op(A)[M x K] * op(B)[K x N]; op(A)(m,k) = TA ? A[k*lda+m] : A[m*lda+k],
// op(B)(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  MFMA 16x16x16 bf16 operand layout:
// lane l holds A[l%16][4(l/16)+i] and B[4(l/16)+i][l%16], i = 0..3, and
// receives C[4(l/16)+i][l%16].
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, const float* __restrict__ bias,
                                              float* __restrict__ C, int ldc, uint16_t* __restrict__ Cbf) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int col0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r16 = lane & 15, q = lane >> 4;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a[2], b[2];
        for (int i = 0; i < 2; ++i) {
            const int m = row0 + 16 * i + r16;
            for (int e = 0; e < 4; ++e) {
                const int k = k0 + 4 * q + e;
                uint16_t v = 0;
                if (m < M && k < K) v = TA ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
                a[i][e] = (short)v;
            }
        }
        for (int j = 0; j < 2; ++j) {
            const int n = col0 + 16 * j + r16;
            for (int e = 0; e < 4; ++e) {
                const int k = k0 + 4 * q + e;
                uint16_t v = 0;
                if (n < N && k < K) v = TB ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n];
                b[j][e] = (short)v;
            }
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int n = col0 + 16 * j + r16;
            if (n >= N) continue;
            const float bn = bias ? bias[n] : 0.f;
            for (int e = 0; e < 4; ++e) {
                const int m = row0 + 16 * i + 4 * q + e;
                if (m >= M) continue;
                const float z = acc[i][j][e] + bn;
                C[(int64_t)m * ldc + n] = z;
                if (EPI == EPI_GELU) Cbf[(int64_t)m * ldc + n] = to_bf16(gelu(z));
            }
        }
}

__global__ void k_gelu_back(int64_t n, const float* __restrict__ dA, const float* __restrict__ Z, float* __restrict__ dZ,
                            uint16_t* __restrict__ dZbf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dA[i] * gelu_d(Z[i]);
    dZ[i] = d;
    dZbf[i] = to_bf16(d);
}

__global__ void k_loss(int B, int n_out, const float* __restrict__ out, const float* __restrict__ y,
                       const float* __restrict__ w, float* __restrict__ dOut, uint16_t* __restrict__ dObf,
                       float* __restrict__ terms) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * n_out) return;
    const int j = (int)(i % n_out);
    const float d = out[i] - y[i];
    terms[i] = w[j] * d * d;
    const float g = 2.0f * w[j] * d / (float)B;
    dOut[i] = g;
    dObf[i] = to_bf16(g);
}

// out[j] = sum_r X[r * ld + j] (fixed order)
__global__ void k_colsum(int rows, int cols, const float* __restrict__ X, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += X[(int64_t)r * cols + j];
    out[j] = s;
}

// one block: out[0] = sum of x[0..n) (fixed order per thread, then tree)
__global__ void k_sum(int64_t n, const float* __restrict__ x, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.f;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pbf, float lr, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
    p[i] = w;
    pbf[i] = to_bf16(w);
}

__global__ void k_tobf16(int64_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// batch gather: rows idx[0..B) of the training set
__global__ void k_gather(int B, int n_in, int n_out, const int64_t* __restrict__ idx, const uint16_t* __restrict__ X,
                         const float* __restrict__ Y, uint16_t* __restrict__ xb, float* __restrict__ yb) {
    const int r = blockIdx.x;
    if (r >= B) return;
    const int64_t s = idx[r];
    for (int j = threadIdx.x; j < n_in; j += blockDim.x) xb[(int64_t)r * n_in + j] = X[s * n_in + j];
    for (int j = threadIdx.x; j < n_out; j += blockDim.x) yb[(int64_t)r * n_out + j] = Y[s * n_out + j];
}

inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

} // namespace

struct pwa_mlp {
    int device = 0;
    hipStream_t st = nullptr;
    int n_in = 0, H = 0, L = 0, n_out = 0;
    std::vector<int> fan;           // layer widths: n_in, H, ..., H, n_out
    std::vector<int64_t> woff;      // offset of W_l in the flat parameter vector
    std::vector<int64_t> boff;      // offset of b_l
    int64_t np = 0;
    float *p = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;
    uint16_t* pbf = nullptr;
    int64_t adam_t = 0;
    // activations for up to bcap rows
    int bcap = 0;
    uint16_t* xb = nullptr;         // batch input (bf16)
    float* yb = nullptr;            // batch targets
    std::vector<float*> Z;          // pre-activations of each layer (fp32)
    std::vector<uint16_t*> Abf;     // activations (bf16), hidden layers
    float *dA = nullptr, *dZ = nullptr, *terms = nullptr, *dloss = nullptr;
    uint16_t* dZbf = nullptr;
    int64_t* didx = nullptr;
    // training data
    int64_t n_data = 0;
    uint16_t* X = nullptr;
    float* Y = nullptr;
    float* ow = nullptr;
    std::string err;
};

namespace {

int fail(pwa_mlp* m, int code, const std::string& s) {
    if (m) m->err = s;
    return code;
}

#define MLPCHK(m, x)                                                                                             \
    do {                                                                                                         \
        hipError_t e_ = (x);                                                                                     \
        if (e_ != hipSuccess) return fail(m, PWA_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
    } while (0)

int ensure_batch(pwa_mlp* m, int B) {
    if (B <= m->bcap) return PWA_OK;
    (void)hipStreamSynchronize(m->st);
    auto fr = [](void* q) { if (q) (void)hipFree(q); };
    fr(m->xb); fr(m->yb); fr(m->dA); fr(m->dZ); fr(m->dZbf); fr(m->terms); fr(m->didx);
    for (float* z : m->Z) fr(z);
    for (uint16_t* a : m->Abf) fr(a);
    m->Z.assign(m->L + 1, nullptr);
    m->Abf.assign(m->L, nullptr);
    const int wmax = std::max(std::max(m->H, m->n_out), m->n_in);
    MLPCHK(m, hipMalloc(&m->xb, (size_t)B * m->n_in * sizeof(uint16_t)));
    MLPCHK(m, hipMalloc(&m->yb, (size_t)B * m->n_out * sizeof(float)));
    for (int l = 0; l <= m->L; ++l) MLPCHK(m, hipMalloc(&m->Z[l], (size_t)B * m->fan[l + 1] * sizeof(float)));
    for (int l = 0; l < m->L; ++l) MLPCHK(m, hipMalloc(&m->Abf[l], (size_t)B * m->H * sizeof(uint16_t)));
    MLPCHK(m, hipMalloc(&m->dA, (size_t)B * wmax * sizeof(float)));
    MLPCHK(m, hipMalloc(&m->dZ, (size_t)B * wmax * sizeof(float)));
    MLPCHK(m, hipMalloc(&m->dZbf, (size_t)B * wmax * sizeof(uint16_t)));
    MLPCHK(m, hipMalloc(&m->terms, (size_t)B * m->n_out * sizeof(float)));
    MLPCHK(m, hipMalloc(&m->didx, (size_t)B * sizeof(int64_t)));
    m->bcap = B;
    return PWA_OK;
}

template <bool TA, bool TB, int EPI>
hipError_t gemm(hipStream_t st, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
                const float* bias, float* C, int ldc, uint16_t* Cbf) {
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, bias, C, ldc, Cbf);
    return hipGetLastError();
}

// forward over the B rows in m->xb; output layer pre-activation in m->Z[L]
int forward(pwa_mlp* m, int B) {
    const uint16_t* in = m->xb;
    for (int l = 0; l <= m->L; ++l) {
        const int fi = m->fan[l], fo = m->fan[l + 1];
        const uint16_t* W = m->pbf + m->woff[l];
        const float* b = m->p + m->boff[l];
        if (l < m->L) {
            MLPCHK(m, (gemm<false, false, EPI_GELU>(m->st, B, fo, fi, in, fi, W, fo, b, m->Z[l], fo, m->Abf[l])));
            in = m->Abf[l];
        } else {
            MLPCHK(m, (gemm<false, false, EPI_STORE>(m->st, B, fo, fi, in, fi, W, fo, b, m->Z[l], fo, nullptr)));
        }
    }
    return PWA_OK;
}

// forward + backward on the batch in m->xb / m->yb; gradient into m->g
int loss_grad(pwa_mlp* m, int B) {
    int rc = forward(m, B);
    if (rc) return rc;
    const int no = m->n_out;
    hipLaunchKernelGGL(k_loss, dim3(blocks_for((int64_t)B * no)), dim3(256), 0, m->st, B, no, m->Z[m->L], m->yb, m->ow,
                       m->dZ, m->dZbf, m->terms);
    MLPCHK(m, hipGetLastError());
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->st, (int64_t)B * no, m->terms, m->dloss);
    MLPCHK(m, hipGetLastError());
    // dZ (fp32 + bf16) of layer l is in m->dZ / m->dZbf
    for (int l = m->L; l >= 0; --l) {
        const int fi = m->fan[l], fo = m->fan[l + 1];
        const uint16_t* in = l == 0 ? m->xb : m->Abf[l - 1];
        // dW_l = in^T dZ_l  [fi x fo];  db_l = column sums of dZ_l
        MLPCHK(m, (gemm<true, false, EPI_STORE>(m->st, fi, fo, B, in, fi, m->dZbf, fo, nullptr, m->g + m->woff[l], fo,
                                                 nullptr)));
        hipLaunchKernelGGL(k_colsum, dim3(blocks_for(fo)), dim3(256), 0, m->st, B, fo, m->dZ, m->g + m->boff[l]);
        MLPCHK(m, hipGetLastError());
        if (l == 0) break;
        // dA_{l-1} = dZ_l W_l^T  [B x fi]; dZ_{l-1} = dA * gelu'(Z_{l-1})
        MLPCHK(m, (gemm<false, true, EPI_STORE>(m->st, B, fi, fo, m->dZbf, fo, m->pbf + m->woff[l], fo, nullptr, m->dA,
                                                 fi, nullptr)));
        hipLaunchKernelGGL(k_gelu_back, dim3(blocks_for((int64_t)B * fi)), dim3(256), 0, m->st, (int64_t)B * fi, m->dA,
                           m->Z[l - 1], m->dZ, m->dZbf);
        MLPCHK(m, hipGetLastError());
    }
    return PWA_OK;
}

int load_batch(pwa_mlp* m, const std::vector<int64_t>& idx) {
    const int B = (int)idx.size();
    MLPCHK(m, hipMemcpyAsync(m->didx, idx.data(), B * sizeof(int64_t), hipMemcpyHostToDevice, m->st));
    hipLaunchKernelGGL(k_gather, dim3(B), dim3(256), 0, m->st, B, m->n_in, m->n_out, m->didx, m->X, m->Y, m->xb, m->yb);
    MLPCHK(m, hipGetLastError());
    return PWA_OK;
}

} // namespace

extern "C" {

int pwa_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed, int device,
                   pwa_mlp** out) {
    if (!out || n_in < 1 || n_hidden < 1 || n_layers < 1 || n_out < 1 || n_in > 4096 || n_hidden > 4096 || n_out > 4096)
        return PWA_E_ARG;
    *out = nullptr;
    pwa_mlp* m = new pwa_mlp();
    m->device = device;
    m->n_in = n_in;
    m->H = n_hidden;
    m->L = n_layers;
    m->n_out = n_out;
    m->fan.push_back(n_in);
    for (int l = 0; l < n_layers; ++l) m->fan.push_back(n_hidden);
    m->fan.push_back(n_out);
    for (int l = 0; l <= n_layers; ++l) {
        m->woff.push_back(m->np);
        m->np += (int64_t)m->fan[l] * m->fan[l + 1];
        m->boff.push_back(m->np);
        m->np += m->fan[l + 1];
    }
    // LeCun-normal weights, zero biases
    std::vector<float> w(m->np, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l <= n_layers; ++l) {
        std::normal_distribution<float> nd(0.f, 1.f / std::sqrt((float)m->fan[l]));
        for (int64_t k = 0; k < (int64_t)m->fan[l] * m->fan[l + 1]; ++k) w[m->woff[l] + k] = nd(rng);
    }
    auto bad = [&](hipError_t e) {
        if (e == hipSuccess) return false;
        pwa_mlp_destroy(m);
        return true;
    };
    if (bad(hipSetDevice(device))) return PWA_E_HIP;
    if (bad(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking))) return PWA_E_HIP;
    if (bad(hipMalloc(&m->p, m->np * sizeof(float)))) return PWA_E_NOMEM;
    if (bad(hipMalloc(&m->g, m->np * sizeof(float)))) return PWA_E_NOMEM;
    if (bad(hipMalloc(&m->am, m->np * sizeof(float)))) return PWA_E_NOMEM;
    if (bad(hipMalloc(&m->av, m->np * sizeof(float)))) return PWA_E_NOMEM;
    if (bad(hipMalloc(&m->pbf, m->np * sizeof(uint16_t)))) return PWA_E_NOMEM;
    if (bad(hipMalloc(&m->dloss, sizeof(float)))) return PWA_E_NOMEM;
    if (bad(hipMemsetAsync(m->am, 0, m->np * sizeof(float), m->st))) return PWA_E_HIP;
    if (bad(hipMemsetAsync(m->av, 0, m->np * sizeof(float), m->st))) return PWA_E_HIP;
    *out = m;
    int rc = pwa_mlp_set_params(m, w.data());
    if (rc) {
        *out = nullptr;
        pwa_mlp_destroy(m);
        return rc;
    }
    return PWA_OK;
}

int pwa_mlp_destroy(pwa_mlp* m) {
    if (!m) return PWA_E_ARG;
    (void)hipSetDevice(m->device);
    if (m->st) (void)hipStreamSynchronize(m->st);
    auto fr = [](void* q) { if (q) (void)hipFree(q); };
    fr(m->p); fr(m->g); fr(m->am); fr(m->av); fr(m->pbf); fr(m->dloss);
    fr(m->xb); fr(m->yb); fr(m->dA); fr(m->dZ); fr(m->dZbf); fr(m->terms); fr(m->didx);
    fr(m->X); fr(m->Y); fr(m->ow);
    for (float* z : m->Z) fr(z);
    for (uint16_t* a : m->Abf) fr(a);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return PWA_OK;
}

const char* pwa_mlp_last_error(const pwa_mlp* m) { return m ? m->err.c_str() : "null context"; }

int pwa_mlp_n_params(const pwa_mlp* m, int64_t* n) {
    if (!m || !n) return PWA_E_ARG;
    *n = m->np;
    return PWA_OK;
}

int pwa_mlp_get_params(const pwa_mlp* m, float* w) {
    if (!m || !w) return PWA_E_ARG;
    pwa_mlp* mm = const_cast<pwa_mlp*>(m);
    MLPCHK(mm, hipSetDevice(m->device));
    MLPCHK(mm, hipMemcpyAsync(w, m->p, m->np * sizeof(float), hipMemcpyDeviceToHost, m->st));
    MLPCHK(mm, hipStreamSynchronize(m->st));
    return PWA_OK;
}

int pwa_mlp_set_params(pwa_mlp* m, const float* w) {
    if (!m || !w) return PWA_E_ARG;
    MLPCHK(m, hipSetDevice(m->device));
    MLPCHK(m, hipMemcpyAsync(m->p, w, m->np * sizeof(float), hipMemcpyHostToDevice, m->st));
    hipLaunchKernelGGL(k_tobf16, dim3(blocks_for(m->np)), dim3(256), 0, m->st, m->np, m->p, m->pbf);
    MLPCHK(m, hipGetLastError());
    MLPCHK(m, hipStreamSynchronize(m->st));
    m->adam_t = 0;
    MLPCHK(m, hipMemsetAsync(m->am, 0, m->np * sizeof(float), m->st));
    MLPCHK(m, hipMemsetAsync(m->av, 0, m->np * sizeof(float), m->st));
    MLPCHK(m, hipStreamSynchronize(m->st));
    return PWA_OK;
}

int pwa_mlp_set_data(pwa_mlp* m, int64_t n, const float* x, const float* y, const float* out_weight) {
    if (!m || n < 1 || !x || !y || !out_weight) return PWA_E_ARG;
    MLPCHK(m, hipSetDevice(m->device));
    auto fr = [](void* q) { if (q) (void)hipFree(q); };
    fr(m->X); fr(m->Y); fr(m->ow);
    m->X = nullptr;
    m->Y = nullptr;
    m->ow = nullptr;
    std::vector<uint16_t> xb((size_t)n * m->n_in);
    for (size_t k = 0; k < xb.size(); ++k) xb[k] = to_bf16(x[k]);
    MLPCHK(m, hipMalloc(&m->X, xb.size() * sizeof(uint16_t)));
    MLPCHK(m, hipMalloc(&m->Y, (size_t)n * m->n_out * sizeof(float)));
    MLPCHK(m, hipMalloc(&m->ow, m->n_out * sizeof(float)));
    MLPCHK(m, hipMemcpyAsync(m->X, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, m->st));
    MLPCHK(m, hipMemcpyAsync(m->Y, y, (size_t)n * m->n_out * sizeof(float), hipMemcpyHostToDevice, m->st));
    MLPCHK(m, hipMemcpyAsync(m->ow, out_weight, m->n_out * sizeof(float), hipMemcpyHostToDevice, m->st));
    MLPCHK(m, hipStreamSynchronize(m->st));
    m->n_data = n;
    return PWA_OK;
}

int pwa_mlp_loss_grad(pwa_mlp* m, int64_t first, int32_t batch, float* loss, float* grad) {
    if (!m || batch < 1 || first < 0 || first + batch > m->n_data) return PWA_E_ARG;
    MLPCHK(m, hipSetDevice(m->device));
    int rc = ensure_batch(m, batch);
    if (rc) return rc;
    std::vector<int64_t> idx(batch);
    std::iota(idx.begin(), idx.end(), first);
    if ((rc = load_batch(m, idx))) return rc;
    if ((rc = loss_grad(m, batch))) return rc;
    float l = 0.f;
    MLPCHK(m, hipMemcpyAsync(&l, m->dloss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    if (grad) MLPCHK(m, hipMemcpyAsync(grad, m->g, m->np * sizeof(float), hipMemcpyDeviceToHost, m->st));
    MLPCHK(m, hipStreamSynchronize(m->st));
    if (loss) *loss = l / (float)batch;
    return PWA_OK;
}

int pwa_mlp_train(pwa_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss) {
    if (!m || epochs < 0 || batch < 1 || m->n_data < 1) return PWA_E_ARG;
    MLPCHK(m, hipSetDevice(m->device));
    const int B = (int)std::min<int64_t>(batch, m->n_data);
    int rc = ensure_batch(m, B);
    if (rc) return rc;
    std::mt19937_64 rng(seed);
    std::vector<int64_t> perm(m->n_data), idx(B);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<float> lbuf;
    float last = 0.f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        const int64_t nb = m->n_data / B;
        double acc = 0.0;
        for (int64_t b = 0; b < nb; ++b) {
            std::copy(perm.begin() + b * B, perm.begin() + (b + 1) * B, idx.begin());
            if ((rc = load_batch(m, idx))) return rc;
            if ((rc = loss_grad(m, B))) return rc;
            ++m->adam_t;
            const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->adam_t));
            const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->adam_t));
            hipLaunchKernelGGL(k_adam, dim3(blocks_for(m->np)), dim3(256), 0, m->st, m->np, m->p, m->g, m->am, m->av,
                               m->pbf, lr, c1, c2);
            MLPCHK(m, hipGetLastError());
            if (ep == epochs - 1) {
                float l = 0.f;
                MLPCHK(m, hipMemcpyAsync(&l, m->dloss, sizeof(float), hipMemcpyDeviceToHost, m->st));
                MLPCHK(m, hipStreamSynchronize(m->st));
                acc += l / B;
            }
        }
        if (ep == epochs - 1 && nb > 0) last = (float)(acc / nb);
    }
    MLPCHK(m, hipStreamSynchronize(m->st));
    if (loss) *loss = last;
    return PWA_OK;
}

int pwa_mlp_predict(pwa_mlp* m, int64_t n, const float* x, float* y) {
    if (!m || n < 0 || (n > 0 && (!x || !y))) return PWA_E_ARG;
    MLPCHK(m, hipSetDevice(m->device));
    const int chunk = 4096;
    int rc = ensure_batch(m, (int)std::min<int64_t>(std::max<int64_t>(n, 1), chunk));
    if (rc) return rc;
    std::vector<uint16_t> xb;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int B = (int)std::min<int64_t>(chunk, n - r0);
        xb.resize((size_t)B * m->n_in);
        for (size_t k = 0; k < xb.size(); ++k) xb[k] = to_bf16(x[(size_t)r0 * m->n_in + k]);
        MLPCHK(m, hipMemcpyAsync(m->xb, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, m->st));
        if ((rc = forward(m, B))) return rc;
        MLPCHK(m, hipMemcpyAsync(y + (size_t)r0 * m->n_out, m->Z[m->L], (size_t)B * m->n_out * sizeof(float),
                                 hipMemcpyDeviceToHost, m->st));
        MLPCHK(m, hipStreamSynchronize(m->st));
    }
    return PWA_OK;
}

} // extern "C"
