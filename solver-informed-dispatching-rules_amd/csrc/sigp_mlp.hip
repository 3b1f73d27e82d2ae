// sigp_mlp.hip -- the surrogate network: a 512 -> 256 -> 256 -> 1 ReLU MLP
// trained and run on MI355X.  Weights are kept in FP32 (Adam state too) with
// BF16 copies that feed the matrix cores; every product is a BF16 x BF16
// MFMA (v_mfma_f32_16x16x32_bf16) accumulated in FP32.  Activations that
// feed a later product (inputs, ReLU outputs, back-propagated deltas) are
// rounded to BF16; pre-activations, the output and all gradients stay FP32.
// No counterpart in the reference (the surrogate is new).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sigp.h"
#include "sigp_internal.h"

#define MLP_TRY(x)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            return sigp_internal_fail(SIGP_DEVICE_ERROR, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr int IN = SIGP_MLP_IN, H = SIGP_MLP_HIDDEN;
constexpr int MAXB = 1024;   // rows per device batch

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (the inputs are finite).
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

enum Epi { EPI_STORE = 0, EPI_BIAS_RELU = 1, EPI_RELU_GRAD = 2 };

// C[M x N] = A(M x K) . B(K x N) on BF16 operands (uint16 bits), FP32
// accumulate.  A(i, k) = TA ? A[k * lda + i] : A[i * lda + k];
// B(k, j) = TB ? B[j * ldb + k] : B[k * ldb + j].  One wavefront per 16 x 16
// output tile, K in steps of 32 (M, N multiples of 16, K of 32).  MFMA
// operand layout (16x16x32 bf16): lane l supplies row l % 16 of A and column
// l % 16 of B at k = 8 * (l / 16) .. + 7, and receives rows 4 * (l / 16) .. + 3
// of column l % 16 of the result.
//   EPI_STORE:      out32 = C
//   EPI_BIAS_RELU:  out32 = C + bias[j] (pre-activation), out16 = bf16(max(that, 0))
//   EPI_RELU_GRAD:  out16 = bf16(C * (gate[i][j] > 0))   (gate: pre-activation, fp32)
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(const uint16_t *__restrict__ A, int lda,
                                                   const uint16_t *__restrict__ B, int ldb, int M,
                                                   int N, int K, const float *__restrict__ bias,
                                                   const float *__restrict__ gate,
                                                   float *__restrict__ out32,
                                                   uint16_t *__restrict__ out16) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tiles_n = N / 16;
    if (tile >= (M / 16) * tiles_n) return;
    const int i0 = (tile / tiles_n) * 16, j0 = (tile % tiles_n) * 16;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = k0 + kq + t;
            const uint16_t av = TA ? A[(size_t)k * lda + i0 + r] : A[(size_t)(i0 + r) * lda + k];
            const uint16_t bv = TB ? B[(size_t)(j0 + r) * ldb + k] : B[(size_t)k * ldb + j0 + r];
            a[t] = __builtin_bit_cast(__bf16, av);
            b[t] = __builtin_bit_cast(__bf16, bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int i = i0 + (lane >> 4) * 4 + t, j = j0 + r;
        const size_t at = (size_t)i * N + j;
        float v = acc[t];
        if (EPI == EPI_STORE) {
            out32[at] = v;
        } else if (EPI == EPI_BIAS_RELU) {
            v += bias[j];
            out32[at] = v;
            out16[at] = to_bf16(v > 0.f ? v : 0.f);
        } else {
            out16[at] = to_bf16(gate[at] > 0.f ? v : 0.f);
        }
    }
}

// Output layer and its backward pass for rows b < nb of a batch padded to B
// rows: y = A2 . w3 + b3 (FP32 dot in index order), MSE loss against t,
// dy = 2 (y - t) / nb, delta2 = bf16(dy * w3 * (H2 > 0)); padded rows get
// zero deltas.  One wavefront per row.
__global__ void head_kernel(const uint16_t *__restrict__ A2, const float *__restrict__ H2,
                            const float *__restrict__ w3, const float *__restrict__ b3,
                            const float *__restrict__ t, int nb, int B, float *__restrict__ y,
                            float *__restrict__ dy, float *__restrict__ sqerr,
                            uint16_t *__restrict__ delta2, bool backward) {
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (b >= B) return;
    float part = 0.f;
    for (int j = lane; j < H; j += 64) part += from_bf16(A2[(size_t)b * H + j]) * w3[j];
    for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
    const float yb = part + b3[0];
    const bool live = b < nb;
    float g = 0.f;
    if (lane == 0) y[b] = yb;
    if (backward) {
        if (live) g = 2.f * (yb - t[b]) / (float)nb;
        if (lane == 0) {
            dy[b] = g;
            sqerr[b] = live ? (yb - t[b]) * (yb - t[b]) : 0.f;
        }
        for (int j = lane; j < H; j += 64) {
            const size_t at = (size_t)b * H + j;
            delta2[at] = to_bf16(H2[at] > 0.f ? g * w3[j] : 0.f);
        }
    }
}

// Gradients of the output layer and of the biases: gw3[j] = sum_b A2[b][j] dy[b],
// gb3 = sum_b dy[b], gb2[j] = sum_b delta2[b][j], gb1[j] = sum_b delta1[b][j]
// (each summed in row order by one thread).
__global__ void small_grads_kernel(const uint16_t *__restrict__ A2, const float *__restrict__ dy,
                                   const uint16_t *__restrict__ delta2,
                                   const uint16_t *__restrict__ delta1, int B,
                                   float *__restrict__ gw3, float *__restrict__ gb3,
                                   float *__restrict__ gb2, float *__restrict__ gb1) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < H) {
        float s3 = 0.f, s2 = 0.f, s1 = 0.f;
        for (int b = 0; b < B; ++b) {
            s3 += from_bf16(A2[(size_t)b * H + j]) * dy[b];
            s2 += from_bf16(delta2[(size_t)b * H + j]);
            s1 += from_bf16(delta1[(size_t)b * H + j]);
        }
        gw3[j] = s3;
        gb2[j] = s2;
        gb1[j] = s1;
    }
    if (j == 0) {
        float s = 0.f;
        for (int b = 0; b < B; ++b) s += dy[b];
        gb3[0] = s;
    }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on the FP32
// parameters; refreshes the BF16 copies of W1 and W2.
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float c1, float c2,
                            uint16_t *__restrict__ w1_16, uint16_t *__restrict__ w2_16) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = 0.9f * m[k] + 0.1f * gk;
    const float vk = 0.999f * v[k] + 0.001f * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float pk = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
    p[k] = pk;
    if (k < IN * H) w1_16[k] = to_bf16(pk);
    else if (k < IN * H + H + H * H && k >= IN * H + H) w2_16[k - IN * H - H] = to_bf16(pk);
}

__global__ void gather_kernel(const uint16_t *__restrict__ X, const float *__restrict__ T,
                              const int64_t *__restrict__ idx, int nb, int B,
                              uint16_t *__restrict__ xb, float *__restrict__ tb) {
    const int b = blockIdx.x;
    if (b >= B) return;
    const bool live = b < nb;
    const int64_t src = live ? idx[b] : 0;
    for (int f = threadIdx.x; f < IN; f += blockDim.x)
        xb[(size_t)b * IN + f] = live ? X[(size_t)src * IN + f] : (uint16_t)0;
    if (threadIdx.x == 0) tb[b] = live ? T[src] : 0.f;
}

__global__ void to_bf16_kernel(const float *__restrict__ x, int64_t n, uint16_t *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = to_bf16(x[k]);
}

// Parameter layout (FP32, SIGP_MLP_NPARAMS floats):
//   W1 [512][256] | b1 [256] | W2 [256][256] | b2 [256] | w3 [256] | b3 [1]
constexpr int O_W1 = 0, O_B1 = O_W1 + IN * H, O_W2 = O_B1 + H, O_B2 = O_W2 + H * H, O_W3 = O_B2 + H,
              O_B3 = O_W3 + H, NP = O_B3 + 1;
static_assert(NP == SIGP_MLP_NPARAMS, "parameter count");

}  // namespace

struct sigp_mlp {
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // NP each
    uint16_t *w1 = nullptr, *w2 = nullptr;                          // BF16 copies
    uint16_t *xb = nullptr, *a1 = nullptr, *a2 = nullptr, *d1 = nullptr, *d2 = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *tb = nullptr, *y = nullptr, *dy = nullptr, *se = nullptr;
    int64_t *idx = nullptr;
    uint16_t *data = nullptr;   // resident training inputs (BF16)
    float *targets = nullptr;
    int64_t data_cap = 0;
    long long steps = 0;        // Adam step count
};

static void free_mlp(sigp_mlp *n) {
    void *ptrs[] = {n->p, n->g, n->m, n->v, n->w1, n->w2, n->xb, n->a1, n->a2, n->d1, n->d2,
                    n->h1, n->h2, n->tb, n->y, n->dy, n->se, n->idx, n->data, n->targets};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete n;
}

// Glorot-uniform weights, zero biases, from a splitmix64 stream of `seed`.
static void init_params(uint64_t seed, std::vector<float> &p) {
    p.assign(NP, 0.f);
    uint64_t s = seed;
    auto next = [&]() {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int count, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int k = 0; k < count; ++k) p[off + k] = (float)((2.0 * next() - 1.0) * a);
    };
    fill(O_W1, IN * H, IN, H);
    fill(O_W2, H * H, H, H);
    fill(O_W3, H, H, 1);
}

static sigp_status refresh_bf16(sigp_mlp *n) {
    to_bf16_kernel<<<(IN * H + 255) / 256, 256>>>(n->p + O_W1, IN * H, n->w1);
    to_bf16_kernel<<<(H * H + 255) / 256, 256>>>(n->p + O_W2, H * H, n->w2);
    MLP_TRY(hipGetLastError());
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_create(uint64_t seed, sigp_mlp **out) {
    if (!out) return sigp_internal_fail(SIGP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    sigp_mlp *n = new sigp_mlp();
    auto alloc = [&](void **q, size_t bytes) { return hipMalloc(q, bytes) == hipSuccess; };
    bool ok = alloc((void **)&n->p, NP * 4) && alloc((void **)&n->g, NP * 4) && alloc((void **)&n->m, NP * 4) &&
              alloc((void **)&n->v, NP * 4) && alloc((void **)&n->w1, IN * H * 2) &&
              alloc((void **)&n->w2, H * H * 2) && alloc((void **)&n->xb, (size_t)MAXB * IN * 2) &&
              alloc((void **)&n->a1, (size_t)MAXB * H * 2) && alloc((void **)&n->a2, (size_t)MAXB * H * 2) &&
              alloc((void **)&n->d1, (size_t)MAXB * H * 2) && alloc((void **)&n->d2, (size_t)MAXB * H * 2) &&
              alloc((void **)&n->h1, (size_t)MAXB * H * 4) && alloc((void **)&n->h2, (size_t)MAXB * H * 4) &&
              alloc((void **)&n->tb, MAXB * 4) && alloc((void **)&n->y, MAXB * 4) &&
              alloc((void **)&n->dy, MAXB * 4) && alloc((void **)&n->se, MAXB * 4) &&
              alloc((void **)&n->idx, MAXB * 8);
    if (!ok) { free_mlp(n); return sigp_internal_fail(SIGP_OOM, "MLP allocation failed"); }
    std::vector<float> p;
    init_params(seed, p);
    if (hipMemcpy(n->p, p.data(), NP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, NP * 4) != hipSuccess || hipMemset(n->v, 0, NP * 4) != hipSuccess ||
        refresh_bf16(n) != SIGP_OK) {
        free_mlp(n);
        return sigp_internal_fail(SIGP_DEVICE_ERROR, "MLP initialisation failed");
    }
    *out = n;
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_destroy(sigp_mlp *n) {
    if (n) free_mlp(n);
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_get_params(sigp_mlp *n, float *out) {
    if (!n || !out) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    MLP_TRY(hipMemcpy(out, n->p, NP * 4, hipMemcpyDeviceToHost));
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_set_params(sigp_mlp *n, const float *in) {
    if (!n || !in) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    MLP_TRY(hipMemcpy(n->p, in, NP * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemset(n->m, 0, NP * 4));
    MLP_TRY(hipMemset(n->v, 0, NP * 4));
    n->steps = 0;
    return refresh_bf16(n);
}

// Forward pass (and, with targets, backward pass into n->g) of the batch
// already in n->xb: nb live rows padded to B (a multiple of 16).
static sigp_status run_batch(sigp_mlp *n, int nb, int B, bool backward, hipStream_t stream = nullptr) {
    const int tiles_h = (B / 16) * (H / 16);
    const int grid_h = (tiles_h + 3) / 4;
    gemm_kernel<false, false, EPI_BIAS_RELU><<<grid_h, 256, 0, stream>>>(n->xb, IN, n->w1, H, B, H, IN, n->p + O_B1,
                                                              nullptr, n->h1, n->a1);
    gemm_kernel<false, false, EPI_BIAS_RELU><<<grid_h, 256, 0, stream>>>(n->a1, H, n->w2, H, B, H, H, n->p + O_B2,
                                                              nullptr, n->h2, n->a2);
    head_kernel<<<(B + 3) / 4, 256, 0, stream>>>(n->a2, n->h2, n->p + O_W3, n->p + O_B3, n->tb, nb, B, n->y, n->dy,
                                      n->se, n->d2, backward);
    if (backward) {
        // delta1 = (delta2 . W2^T) * (H1 > 0)
        gemm_kernel<false, true, EPI_RELU_GRAD><<<grid_h, 256, 0, stream>>>(n->d2, H, n->w2, H, B, H, H, nullptr,
                                                                 n->h1, nullptr, n->d1);
        // dW2 = A1^T . delta2, dW1 = X^T . delta1
        gemm_kernel<true, false, EPI_STORE><<<((H / 16) * (H / 16) + 3) / 4, 256, 0, stream>>>(
            n->a1, H, n->d2, H, H, H, B, nullptr, nullptr, n->g + O_W2, nullptr);
        gemm_kernel<true, false, EPI_STORE><<<((IN / 16) * (H / 16) + 3) / 4, 256, 0, stream>>>(
            n->xb, IN, n->d1, H, IN, H, B, nullptr, nullptr, n->g + O_W1, nullptr);
        small_grads_kernel<<<1, 256, 0, stream>>>(n->a2, n->dy, n->d2, n->d1, B, n->g + O_W3, n->g + O_B3,
                                       n->g + O_B2, n->g + O_B1);
    }
    MLP_TRY(hipGetLastError());
    return SIGP_OK;
}

static int padded(int nb) { return (nb + 31) / 32 * 32; }

// Phenotype ranks as network inputs: rank / (n_cand - 1) in FP32, rounded
// to BF16 (what the host computes and uploads), rows [0, nb) of the batch and
// zero rows up to B.
__global__ void ranks_to_bf16_kernel(const int32_t *__restrict__ ranks, int nb, int B, int n_cand,
                                     uint16_t *__restrict__ xb) {
    const float den = (float)(n_cand > 1 ? n_cand - 1 : 1);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < (size_t)B * IN;
         k += (size_t)gridDim.x * blockDim.x)
        xb[k] = k < (size_t)nb * IN ? to_bf16((float)ranks[k] / den) : (uint16_t)0;
}

__global__ void copy_out_kernel(const float *__restrict__ y, int nb, float *__restrict__ out) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < nb) out[k] = y[k];
}

// Scores of `rows` rules from their device-resident phenotype ranks
// (int32 [rows][512]) into y (device), all on `stream`: the same features
// and forward pass as sigp_mlp_predict, so the same values.
sigp_status sigp_internal_mlp_predict_ranks(sigp_mlp *n, const int32_t *ranks, int64_t rows, int n_cand,
                                            float *y, hipStream_t stream) {
    for (int64_t r0 = 0; r0 < rows; r0 += MAXB) {
        const int nb = (int)std::min<int64_t>(MAXB, rows - r0);
        const int B = (nb + 31) / 32 * 32;
        ranks_to_bf16_kernel<<<(B * IN + 255) / 256, 256, 0, stream>>>(ranks + (size_t)r0 * IN, nb, B, n_cand,
                                                                       n->xb);
        sigp_status st = run_batch(n, nb, B, false, stream);
        if (st != SIGP_OK) return st;
        copy_out_kernel<<<(nb + 255) / 256, 256, 0, stream>>>(n->y, nb, y + r0);
    }
    MLP_TRY(hipGetLastError());
    return SIGP_OK;
}

// Uploads host features (converted to BF16) and targets for rows [0, nb).
static sigp_status load_batch(sigp_mlp *n, const float *x, const float *t, int nb) {
    const int B = padded(nb);
    std::vector<uint16_t> xb((size_t)B * IN, 0);
    for (size_t k = 0; k < (size_t)nb * IN; ++k) xb[k] = to_bf16(x[k]);
    MLP_TRY(hipMemcpy(n->xb, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    std::vector<float> tb(B, 0.f);
    if (t) std::memcpy(tb.data(), t, (size_t)nb * 4);
    MLP_TRY(hipMemcpy(n->tb, tb.data(), (size_t)B * 4, hipMemcpyHostToDevice));
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_gradients(sigp_mlp *n, const float *x, const float *t, int32_t nb,
                                          float *grad_out, float *loss_out) {
    if (!n || !x || !t || !grad_out) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    if (nb < 1 || nb > MAXB) return sigp_internal_fail(SIGP_INVALID_ARG, "batch must hold 1..1024 rows");
    sigp_status st = load_batch(n, x, t, nb);
    if (st != SIGP_OK) return st;
    if ((st = run_batch(n, nb, padded(nb), true)) != SIGP_OK) return st;
    MLP_TRY(hipMemcpy(grad_out, n->g, NP * 4, hipMemcpyDeviceToHost));
    if (loss_out) {
        std::vector<float> se(nb);
        MLP_TRY(hipMemcpy(se.data(), n->se, (size_t)nb * 4, hipMemcpyDeviceToHost));
        double s = 0.0;
        for (float e : se) s += e;
        *loss_out = (float)(s / nb);
    }
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_predict(sigp_mlp *n, const float *x, int64_t rows, float *y_out) {
    if (!n || (!x && rows) || (!y_out && rows)) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    for (int64_t r0 = 0; r0 < rows; r0 += MAXB) {
        const int nb = (int)std::min<int64_t>(MAXB, rows - r0);
        sigp_status st = load_batch(n, x + (size_t)r0 * IN, nullptr, nb);
        if (st != SIGP_OK) return st;
        if ((st = run_batch(n, nb, padded(nb), false)) != SIGP_OK) return st;
        MLP_TRY(hipMemcpy(y_out + r0, n->y, (size_t)nb * 4, hipMemcpyDeviceToHost));
    }
    return SIGP_OK;
}

extern "C" sigp_status sigp_mlp_train(sigp_mlp *n, const float *x, const float *t, int64_t rows,
                                      int32_t epochs, int32_t batch, float lr, uint64_t shuffle_seed,
                                      float *epoch_loss) {
    if (!n || !x || !t) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    if (rows < 1 || batch < 1 || batch > MAXB || epochs < 0)
        return sigp_internal_fail(SIGP_INVALID_ARG, "rows >= 1, 1 <= batch <= 1024, epochs >= 0");
    // resident BF16 inputs and FP32 targets
    if (n->data_cap < rows) {
        if (n->data) (void)hipFree(n->data);
        if (n->targets) (void)hipFree(n->targets);
        n->data = nullptr;
        n->targets = nullptr;
        n->data_cap = 0;
        MLP_TRY(hipMalloc((void **)&n->data, (size_t)rows * IN * 2));
        MLP_TRY(hipMalloc((void **)&n->targets, (size_t)rows * 4));
        n->data_cap = rows;
    }
    {
        std::vector<uint16_t> xb((size_t)rows * IN);
        for (size_t k = 0; k < xb.size(); ++k) xb[k] = to_bf16(x[k]);
        MLP_TRY(hipMemcpy(n->data, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
        MLP_TRY(hipMemcpy(n->targets, t, (size_t)rows * 4, hipMemcpyHostToDevice));
    }
    std::vector<int64_t> order(rows);
    for (int64_t k = 0; k < rows; ++k) order[k] = k;
    uint64_t s = shuffle_seed;
    auto next = [&]() {   // splitmix64
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    std::vector<float> se(MAXB);
    for (int e = 0; e < epochs; ++e) {
        for (int64_t k = rows - 1; k > 0; --k) std::swap(order[k], order[next() % (uint64_t)(k + 1)]);
        double loss = 0.0;
        for (int64_t r0 = 0; r0 < rows; r0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, rows - r0);
            const int B = padded(nb);
            MLP_TRY(hipMemcpy(n->idx, order.data() + r0, (size_t)nb * 8, hipMemcpyHostToDevice));
            gather_kernel<<<B, 256>>>(n->data, n->targets, n->idx, nb, B, n->xb, n->tb);
            sigp_status st = run_batch(n, nb, B, true);
            if (st != SIGP_OK) return st;
            ++n->steps;
            const float c1 = 1.f - std::pow(0.9f, (float)n->steps), c2 = 1.f - std::pow(0.999f, (float)n->steps);
            adam_kernel<<<(NP + 255) / 256, 256>>>(n->p, n->g, n->m, n->v, NP, lr, c1, c2, n->w1, n->w2);
            MLP_TRY(hipGetLastError());
            if (epoch_loss) {
                MLP_TRY(hipMemcpy(se.data(), n->se, (size_t)nb * 4, hipMemcpyDeviceToHost));
                for (int b = 0; b < nb; ++b) loss += se[b];
            }
        }
        if (epoch_loss) epoch_loss[e] = (float)(loss / rows);
    }
    MLP_TRY(hipDeviceSynchronize());
    return SIGP_OK;
}
