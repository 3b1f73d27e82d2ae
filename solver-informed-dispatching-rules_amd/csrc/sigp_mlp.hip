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
// accumulate, with both operands stored k-contiguous: A(i, k) = A[i * lda + k]
// and B(k, j) = B[j * ldb + k] (B is held as B^T).  The producers of every
// operand write it in that layout -- transposed BF16 copies of the weights
// (Adam refreshes them) and of the activations and deltas that feed the
// weight gradients (the epilogues write them) -- so each lane fetches its
// MFMA operands as one 16-byte load per matrix per k-step, instead of 8
// two-byte loads strided by a row.  One wavefront per 16 x 16 output tile,
// K in steps of 32 (M, N multiples of 16, K of 32).  MFMA operand layout
// (16x16x32 bf16): lane l supplies row l % 16 of A and column l % 16 of B at
// k = 8 * (l / 16) .. + 7, and receives rows 4 * (l / 16) .. + 3 of column
// l % 16 of the result.  The k-steps accumulate in order, as before the
// layout change, so the results are the same bits.
//   EPI_STORE:      out32 = C
//   EPI_BIAS_RELU:  out32 = C + bias[j] (pre-activation), out16 = bf16(max(that, 0))
//   EPI_RELU_GRAD:  out16 = bf16(C * (gate[i][j] > 0))   (gate: pre-activation, fp32)
// out16t (nullable) receives out16 transposed: out16t[j * ldt + i].
template <int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(const uint16_t *__restrict__ A, int lda,
                                                   const uint16_t *__restrict__ B, int ldb, int M,
                                                   int N, int K, const float *__restrict__ bias,
                                                   const float *__restrict__ gate,
                                                   float *__restrict__ out32,
                                                   uint16_t *__restrict__ out16,
                                                   uint16_t *__restrict__ out16t, int ldt) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tiles_n = N / 16;
    if (tile >= (M / 16) * tiles_n) return;
    const int i0 = (tile / tiles_n) * 16, j0 = (tile % tiles_n) * 16;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    const bf16x8 *ap = (const bf16x8 *)(A + (size_t)(i0 + r) * lda + kq);
    const bf16x8 *bp = (const bf16x8 *)(B + (size_t)(j0 + r) * ldb + kq);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    int k0 = 0;
    for (; k0 + 128 <= K; k0 += 128) {   // four k-steps' loads in flight together
        bf16x8 a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = ap[k0 / 8 + 4 * u];
            b[u] = bp[k0 / 8 + 4 * u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u], b[u], acc, 0, 0, 0);
    }
    for (; k0 < K; k0 += 32) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[k0 / 8], bp[k0 / 8], acc, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int i = i0 + (lane >> 4) * 4 + t, j = j0 + r;
        const size_t at = (size_t)i * N + j;
        float v = acc[t];
        uint16_t h = 0;
        if (EPI == EPI_STORE) {
            out32[at] = v;
        } else if (EPI == EPI_BIAS_RELU) {
            v += bias[j];
            out32[at] = v;
            h = to_bf16(v > 0.f ? v : 0.f);
            out16[at] = h;
        } else {
            h = to_bf16(gate[at] > 0.f ? v : 0.f);
            out16[at] = h;
        }
        if (EPI != EPI_STORE && out16t) out16t[(size_t)j * ldt + i] = h;
    }
}

// dst[c][r] = src[r][c] for a rows x cols BF16 matrix.
__global__ void transpose16_kernel(const uint16_t *__restrict__ src, int rows, int cols,
                                   uint16_t *__restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)rows * cols) return;
    const int rr = (int)(k / cols), c = (int)(k - (int64_t)rr * cols);
    dst[(size_t)c * rows + rr] = src[k];
}

// Output layer and its backward pass for rows b < nb of a batch padded to B
// rows: y = A2 . w3 + b3 (FP32 dot in index order), MSE loss against t,
// dy = 2 (y - t) / nb, delta2 = bf16(dy * w3 * (H2 > 0)); padded rows get
// zero deltas; delta2t = delta2 transposed ([H][B]).  One wavefront per row.
__global__ void head_kernel(const uint16_t *__restrict__ A2, const float *__restrict__ H2,
                            const float *__restrict__ w3, const float *__restrict__ b3,
                            const float *__restrict__ t, int nb, int B, float *__restrict__ y,
                            float *__restrict__ dy, float *__restrict__ sqerr,
                            uint16_t *__restrict__ delta2, uint16_t *__restrict__ delta2t, bool backward) {
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
            const uint16_t h = to_bf16(H2[at] > 0.f ? g * w3[j] : 0.f);
            delta2[at] = h;
            delta2t[(size_t)j * B + b] = h;
        }
    }
}

// Gradients of the output layer and of the biases: gw3[j] = sum_b A2[b][j] dy[b],
// gb3 = sum_b dy[b], gb2[j] = sum_b delta2[b][j], gb1[j] = sum_b delta1[b][j].
// Workgroup = 16 columns j; its 256 threads split the B rows into 16 chunks
// (thread t: column t % 16, chunk t / 16), each summed in row order, and the
// 16 partial sums of a column are then added in chunk order -- a fixed order,
// so the result is deterministic.  (One thread per column summing all B rows
// serially took 71 us per training step, more than the five matrix products
// together.)
__global__ __launch_bounds__(256) void small_grads_kernel(const uint16_t *__restrict__ A2,
                                                          const float *__restrict__ dy,
                                                          const uint16_t *__restrict__ delta2,
                                                          const uint16_t *__restrict__ delta1, int B,
                                                          float *__restrict__ gw3, float *__restrict__ gb3,
                                                          float *__restrict__ gb2, float *__restrict__ gb1) {
    __shared__ float part[4][16][16];
    const int jj = threadIdx.x & 15, c = threadIdx.x >> 4;
    const int j = blockIdx.x * 16 + jj;
    const int rows = B / 16, b0 = c * rows;
    float s3 = 0.f, s2 = 0.f, s1 = 0.f, sd = 0.f;
    for (int b = b0; b < b0 + rows; ++b) {
        const float d = dy[b];
        s3 += from_bf16(A2[(size_t)b * H + j]) * d;
        s2 += from_bf16(delta2[(size_t)b * H + j]);
        s1 += from_bf16(delta1[(size_t)b * H + j]);
        sd += d;
    }
    part[0][c][jj] = s3;
    part[1][c][jj] = s2;
    part[2][c][jj] = s1;
    part[3][c][jj] = sd;
    __syncthreads();
    if (c == 0) {
        float t3 = 0.f, t2 = 0.f, t1 = 0.f, td = 0.f;
        for (int q = 0; q < 16; ++q) {
            t3 += part[0][q][jj];
            t2 += part[1][q][jj];
            t1 += part[2][q][jj];
            td += part[3][q][jj];
        }
        gw3[j] = t3;
        gb2[j] = t2;
        gb1[j] = t1;
        if (blockIdx.x == 0 && jj == 0) gb3[0] = td;
    }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on the FP32
// parameters; refreshes the BF16 copies of W1 and W2 and their transposes.
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float c1, float c2,
                            uint16_t *__restrict__ w1_16, uint16_t *__restrict__ w2_16,
                            uint16_t *__restrict__ w1t_16, uint16_t *__restrict__ w2t_16) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = 0.9f * m[k] + 0.1f * gk;
    const float vk = 0.999f * v[k] + 0.001f * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float pk = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
    p[k] = pk;
    if (k < IN * H) {
        const uint16_t h = to_bf16(pk);
        w1_16[k] = h;
        w1t_16[(size_t)(k % H) * IN + k / H] = h;
    } else if (k < IN * H + H + H * H && k >= IN * H + H) {
        const int q = k - IN * H - H;
        const uint16_t h = to_bf16(pk);
        w2_16[q] = h;
        w2t_16[(size_t)(q % H) * H + q / H] = h;
    }
}

// Training batch rows b < nb from rows idx[b] of the resident set, zero rows
// up to B: xb [B][IN], its transpose xt [IN][B] (the W1 gradient's operand)
// and the targets tb.
__global__ void gather_kernel(const uint16_t *__restrict__ X, const float *__restrict__ T,
                              const int64_t *__restrict__ idx, int nb, int B,
                              uint16_t *__restrict__ xb, uint16_t *__restrict__ xt, float *__restrict__ tb) {
    const int b = blockIdx.x;
    if (b >= B) return;
    const bool live = b < nb;
    const int64_t src = live ? idx[b] : 0;
    for (int f = threadIdx.x; f < IN; f += blockDim.x) {
        const uint16_t h = live ? X[(size_t)src * IN + f] : (uint16_t)0;
        xb[(size_t)b * IN + f] = h;
        xt[(size_t)f * B + b] = h;
    }
    if (threadIdx.x == 0) tb[b] = live ? T[src] : 0.f;
}

// Sum of the batch's squared errors into the epoch's loss, in row order, in
// double (the order and type the host summed in); the rows are staged in
// LDS first so that the serial sum does not wait on one global load per row.
__global__ __launch_bounds__(256) void loss_acc_kernel(const float *__restrict__ se, int nb,
                                                       double *__restrict__ acc) {
    __shared__ float sv[MAXB];
    for (int b = threadIdx.x; b < nb; b += blockDim.x) sv[b] = se[b];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s = *acc;
    for (int b = 0; b < nb; ++b) s += sv[b];
    *acc = s;
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
    uint16_t *w1 = nullptr, *w2 = nullptr;                          // BF16 copies [in][out]
    uint16_t *w1t = nullptr, *w2t = nullptr;                        // and transposed [out][in]
    uint16_t *xb = nullptr, *a1 = nullptr, *a2 = nullptr, *d1 = nullptr, *d2 = nullptr;   // [B][.]
    uint16_t *xt = nullptr, *a1t = nullptr, *d1t = nullptr, *d2t = nullptr;              // [.][B]
    float *h1 = nullptr, *h2 = nullptr, *tb = nullptr, *y = nullptr, *dy = nullptr, *se = nullptr;
    uint16_t *data = nullptr;   // resident training inputs (BF16)
    float *targets = nullptr;
    int64_t *order = nullptr;   // the epoch's shuffled row order
    int64_t data_cap = 0;
    long long steps = 0;        // Adam step count
};

static void free_mlp(sigp_mlp *n) {
    void *ptrs[] = {n->p, n->g, n->m, n->v, n->w1, n->w2, n->w1t, n->w2t, n->xb, n->a1, n->a2, n->d1,
                    n->d2, n->xt, n->a1t, n->d1t, n->d2t, n->h1, n->h2, n->tb, n->y, n->dy, n->se,
                    n->data, n->targets, n->order};
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
    transpose16_kernel<<<(IN * H + 255) / 256, 256>>>(n->w1, IN, H, n->w1t);
    transpose16_kernel<<<(H * H + 255) / 256, 256>>>(n->w2, H, H, n->w2t);
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
              alloc((void **)&n->w2, H * H * 2) && alloc((void **)&n->w1t, IN * H * 2) &&
              alloc((void **)&n->w2t, H * H * 2) && alloc((void **)&n->xt, (size_t)MAXB * IN * 2) &&
              alloc((void **)&n->a1t, (size_t)MAXB * H * 2) && alloc((void **)&n->d1t, (size_t)MAXB * H * 2) &&
              alloc((void **)&n->d2t, (size_t)MAXB * H * 2) && alloc((void **)&n->xb, (size_t)MAXB * IN * 2) &&
              alloc((void **)&n->a1, (size_t)MAXB * H * 2) && alloc((void **)&n->a2, (size_t)MAXB * H * 2) &&
              alloc((void **)&n->d1, (size_t)MAXB * H * 2) && alloc((void **)&n->d2, (size_t)MAXB * H * 2) &&
              alloc((void **)&n->h1, (size_t)MAXB * H * 4) && alloc((void **)&n->h2, (size_t)MAXB * H * 4) &&
              alloc((void **)&n->tb, MAXB * 4) && alloc((void **)&n->y, MAXB * 4) &&
              alloc((void **)&n->dy, MAXB * 4) && alloc((void **)&n->se, MAXB * 4);
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
// already in n->xb (and, for the backward pass, n->xt): nb live rows padded
// to B (a multiple of 32).  Every product takes its operands k-contiguous
// (gemm_kernel): the forward pass reads the transposed weights, the
// backward pass W2 itself and the transposed activations and deltas the
// forward epilogues and the head wrote.
static sigp_status run_batch(sigp_mlp *n, int nb, int B, bool backward, hipStream_t stream = nullptr) {
    const int tiles_h = (B / 16) * (H / 16);
    const int grid_h = (tiles_h + 3) / 4;
    gemm_kernel<EPI_BIAS_RELU><<<grid_h, 256, 0, stream>>>(n->xb, IN, n->w1t, IN, B, H, IN, n->p + O_B1,
                                                           nullptr, n->h1, n->a1, backward ? n->a1t : nullptr, B);
    gemm_kernel<EPI_BIAS_RELU><<<grid_h, 256, 0, stream>>>(n->a1, H, n->w2t, H, B, H, H, n->p + O_B2,
                                                           nullptr, n->h2, n->a2, nullptr, B);
    head_kernel<<<(B + 3) / 4, 256, 0, stream>>>(n->a2, n->h2, n->p + O_W3, n->p + O_B3, n->tb, nb, B, n->y, n->dy,
                                                 n->se, n->d2, n->d2t, backward);
    if (backward) {
        // delta1 = (delta2 . W2^T) * (H1 > 0): B(k, j) = W2[j][k]
        gemm_kernel<EPI_RELU_GRAD><<<grid_h, 256, 0, stream>>>(n->d2, H, n->w2, H, B, H, H, nullptr, n->h1,
                                                               nullptr, n->d1, n->d1t, B);
        // dW2 = A1^T . delta2, dW1 = X^T . delta1 (K = B)
        gemm_kernel<EPI_STORE><<<((H / 16) * (H / 16) + 3) / 4, 256, 0, stream>>>(
            n->a1t, B, n->d2t, B, H, H, B, nullptr, nullptr, n->g + O_W2, nullptr, nullptr, 0);
        gemm_kernel<EPI_STORE><<<((IN / 16) * (H / 16) + 3) / 4, 256, 0, stream>>>(
            n->xt, B, n->d1t, B, IN, H, B, nullptr, nullptr, n->g + O_W1, nullptr, nullptr, 0);
        small_grads_kernel<<<H / 16, 256, 0, stream>>>(n->a2, n->dy, n->d2, n->d1, B, n->g + O_W3,
                                                       n->g + O_B3, n->g + O_B2, n->g + O_B1);
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
    const int B = padded(nb);
    transpose16_kernel<<<(B * IN + 255) / 256, 256>>>(n->xb, B, IN, n->xt);
    if ((st = run_batch(n, nb, B, true)) != SIGP_OK) return st;
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
        if (n->order) (void)hipFree(n->order);
        n->data = nullptr;
        n->targets = nullptr;
        n->order = nullptr;
        n->data_cap = 0;
        MLP_TRY(hipMalloc((void **)&n->data, (size_t)rows * IN * 2));
        MLP_TRY(hipMalloc((void **)&n->targets, (size_t)rows * 4));
        MLP_TRY(hipMalloc((void **)&n->order, (size_t)rows * 8));
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
    // Per epoch the shuffled order goes up once; each step then runs on the
    // device with no host round trip (the batch rows come from the order,
    // the squared errors add up on the device in the host's order and type).
    double *acc = nullptr;
    MLP_TRY(hipMalloc((void **)&acc, sizeof(double) * (size_t)std::max(epochs, 1)));
    if (hipMemset(acc, 0, sizeof(double) * (size_t)std::max(epochs, 1)) != hipSuccess) {
        (void)hipFree(acc);
        return sigp_internal_fail(SIGP_DEVICE_ERROR, "hipMemset failed");
    }
    sigp_status st = SIGP_OK;
    for (int e = 0; e < epochs && st == SIGP_OK; ++e) {
        for (int64_t k = rows - 1; k > 0; --k) std::swap(order[k], order[next() % (uint64_t)(k + 1)]);
        if (hipMemcpy(n->order, order.data(), (size_t)rows * 8, hipMemcpyHostToDevice) != hipSuccess) {
            st = sigp_internal_fail(SIGP_DEVICE_ERROR, "order upload failed");
            break;
        }
        for (int64_t r0 = 0; r0 < rows && st == SIGP_OK; r0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, rows - r0);
            const int B = padded(nb);
            gather_kernel<<<B, 256>>>(n->data, n->targets, n->order + r0, nb, B, n->xb, n->xt, n->tb);
            st = run_batch(n, nb, B, true);
            if (st != SIGP_OK) break;
            ++n->steps;
            const float c1 = 1.f - std::pow(0.9f, (float)n->steps), c2 = 1.f - std::pow(0.999f, (float)n->steps);
            adam_kernel<<<(NP + 255) / 256, 256>>>(n->p, n->g, n->m, n->v, NP, lr, c1, c2, n->w1, n->w2, n->w1t,
                                                   n->w2t);
            if (epoch_loss) loss_acc_kernel<<<1, 256>>>(n->se, nb, acc + e);
            if (hipGetLastError() != hipSuccess) st = sigp_internal_fail(SIGP_DEVICE_ERROR, "training launch failed");
        }
    }
    if (st == SIGP_OK && epoch_loss && epochs > 0) {
        std::vector<double> sums(epochs);
        if (hipMemcpy(sums.data(), acc, sizeof(double) * epochs, hipMemcpyDeviceToHost) != hipSuccess)
            st = sigp_internal_fail(SIGP_DEVICE_ERROR, "loss download failed");
        for (int e = 0; e < epochs && st == SIGP_OK; ++e) epoch_loss[e] = (float)(sums[e] / rows);
    }
    (void)hipFree(acc);
    if (st != SIGP_OK) return st;
    MLP_TRY(hipDeviceSynchronize());
    return SIGP_OK;
}
