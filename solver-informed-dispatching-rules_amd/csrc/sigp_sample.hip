// Stochastic service-time scenarios on the device (engine/sim_host.hpp:43-58
// gen_durations, engine/rng.hpp:135-147 gauss; the Gamma mode has no
// counterpart in the reference).
//
// Row r draws from its own CPython-compatible Mersenne Twister, seeded
// like Mt19937(seed(r)) on the device; customer i (in index order) gets
// max(0, gauss(mu_i, sigma mu_i)) (Box-Muller pair, second value cached) or
// Gamma(k, mu_i / k) by Marsaglia & Tsang's squeeze method (ACM TOMS 26(3),
// 2000) on the same stream; customers with mu_i <= 0 draw nothing.  One
// wavefront per row runs the row's sequential stream in lockstep.
//
// The uniform stream is the reference's bit for bit.  log, sin and cos are
// the correctly rounded ones of sigp_ddmath.h; where glibc's (which the
// reference calls) are not correctly rounded -- about 0.1 % of calls on these
// arguments, by one unit in the last place -- a duration can differ from the
// reference's host sampler by an ulp, or, in the Gamma mode, through an
// acceptance test decided the other way.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/sigp.h"
#include "sigp_ddmath.h"
#include "sigp_internal.h"
#include "sigp_mt.h"

using namespace sigp_mt;

namespace {

struct SampleSpec {
    const double *mu;              // n expected durations
    int n;
    int mode;                      // SIGP_SAMPLE_GAUSS / SIGP_SAMPLE_GAMMA
    double sigma;                  // Gauss: sd = sigma * mu
    double shape, d, c;            // Gamma: shape k >= 1, d = k - 1/3, c = 1 / sqrt(9 d)
    // row r draws from Mt19937(group_seed[r / T] (or seed0) + (offset + *gen) * stride + r % T)
    const long long *group_seed;   // null: seed0 for every group
    long long seed0;
    int T;
    long long stride, offset;
    const int *gen;                // null: 0
};

struct Spare {
    bool has;
    double z;
};

// engine/rng.hpp:135-147
__device__ __forceinline__ double gauss(Rng &g, Spare &sp, double mu, double sigma) {
    double z;
    if (sp.has) {
        z = sp.z;
        sp.has = false;
    } else {
        const double x2pi = g.random() * (2.0 * 3.14159265358979323846);
        const double g2rad = sqrt(-2.0 * sigp_dd::cr_log(1.0 - g.random()));
        double sn, cs;
        sigp_dd::cr_sincos(x2pi, &sn, &cs);
        z = cs * g2rad;
        sp.z = sn * g2rad;
        sp.has = true;
    }
    return mu + z * sigma;
}

// Marsaglia & Tsang, shape >= 1, in the operation order of the host sampler
__device__ __forceinline__ double gamma_draw(Rng &g, Spare &sp, const SampleSpec &S, double scale) {
    for (;;) {
        const double x = gauss(g, sp, 0.0, 1.0);
        double v = 1.0 + S.c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        const double u = g.random();
        if (u < 1.0 - 0.0331 * (x * x) * (x * x)) return scale * S.d * v;
        if (u > 0.0 && sigp_dd::cr_log(u) < 0.5 * x * x + S.d * (1.0 - v + sigp_dd::cr_log(v)))
            return scale * S.d * v;
    }
}

__global__ __launch_bounds__(64) void sample_kernel(SampleSpec S, double *out) {
    __shared__ uint32_t mt[624];
    const int row = (int)blockIdx.x;
    const long long gen = S.gen ? *S.gen : 0;
    const long long base = S.group_seed ? S.group_seed[row / S.T] : S.seed0;
    seed_lds(mt, (uint64_t)(base + (S.offset + gen) * S.stride + row % S.T));
    Rng g;
    g.mt = mt;
    g.idx = 624;
    g.base = -1000;
    g.win = 0;
    Spare sp{false, 0.0};
    double *o = out + (size_t)row * S.n;
    for (int i = 0; i < S.n; ++i) {
        const double e = S.mu[i];
        double v;
        if (e <= 0.0) {
            v = 0.0;
        } else if (S.mode == SIGP_SAMPLE_GAUSS) {
            v = gauss(g, sp, e, S.sigma * e);
            v = v > 0.0 ? v : 0.0;
        } else {
            v = gamma_draw(g, sp, S, e / S.shape);
        }
        if (threadIdx.x == 0) o[i] = v;
    }
}

}  // namespace

static sigp_status check_sampler(const sigp_sampler *smp) {
    if (!smp) return sigp_internal_fail(SIGP_INVALID_ARG, "sampler is NULL");
    if (smp->mode == SIGP_SAMPLE_GAUSS) {
        if (!(smp->sigma >= 0.0)) return sigp_internal_fail(SIGP_INVALID_ARG, "sigma must be >= 0");
    } else if (smp->mode == SIGP_SAMPLE_GAMMA) {
        if (!(smp->shape >= 1.0) || !std::isfinite(smp->shape))
            return sigp_internal_fail(SIGP_INVALID_ARG, "the device Gamma sampler needs a finite shape >= 1");
    } else {
        return sigp_internal_fail(SIGP_INVALID_ARG, "mode must be SIGP_SAMPLE_GAUSS or SIGP_SAMPLE_GAMMA");
    }
    return SIGP_OK;
}

sigp_status sigp_internal_sample(const sigp_instance *inst, const sigp_sampler *smp, const long long *group_seed,
                                 long long seed0, int T, long long stride, long long offset, const int *gen,
                                 int rows, double *out, hipStream_t stream) {
    sigp_status st = check_sampler(smp);
    if (st != SIGP_OK) return st;
    if (rows <= 0) return SIGP_OK;
    SampleSpec S;
    S.mu = sigp_internal_instance_dur(inst);
    S.n = sigp_internal_instance_n(inst);
    S.mode = smp->mode;
    S.sigma = smp->sigma;
    S.shape = smp->shape;
    S.d = smp->shape - 1.0 / 3.0;
    S.c = 1.0 / std::sqrt(9.0 * S.d);
    S.group_seed = group_seed;
    S.seed0 = seed0;
    S.T = T > 0 ? T : rows;
    S.stride = stride;
    S.offset = offset;
    S.gen = gen;
    sample_kernel<<<rows, 64, 0, stream>>>(S, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return sigp_internal_fail(SIGP_DEVICE_ERROR, std::string("sampler: ") + hipGetErrorString(e));
    return SIGP_OK;
}

extern "C" sigp_status sigp_sample_durations_async(sigp_instance *inst, const sigp_sampler *smp, int64_t base_seed,
                                                   int64_t n_rows, double *out_dev, void *stream) {
    if (!inst || !out_dev) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    if (n_rows < 0 || n_rows > 0x7fffffff) return sigp_internal_fail(SIGP_INVALID_ARG, "n_rows out of range");
    if (base_seed < 0) return sigp_internal_fail(SIGP_INVALID_ARG, "base_seed must be >= 0");
    return sigp_internal_sample(inst, smp, nullptr, base_seed, (int)n_rows, 0, 0, nullptr, (int)n_rows, out_dev,
                                (hipStream_t)stream);
}

extern "C" sigp_status sigp_sample_durations(sigp_instance *inst, const sigp_sampler *smp, int64_t base_seed,
                                             int64_t n_rows, double *out) {
    if (!inst || (!out && n_rows > 0)) return sigp_internal_fail(SIGP_INVALID_ARG, "NULL argument");
    if (n_rows <= 0) return check_sampler(smp);
    const size_t bytes = (size_t)n_rows * sigp_internal_instance_n(inst) * sizeof(double);
    double *d = nullptr;
    if (hipMalloc((void **)&d, bytes) != hipSuccess) return sigp_internal_fail(SIGP_OOM, "sampler allocation failed");
    sigp_status st = sigp_sample_durations_async(inst, smp, base_seed, n_rows, d, nullptr);
    hipError_t e = hipSuccess;
    if (st == SIGP_OK) e = hipDeviceSynchronize();
    if (st == SIGP_OK && e == hipSuccess) e = hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (st == SIGP_OK && e != hipSuccess)
        st = sigp_internal_fail(SIGP_DEVICE_ERROR, std::string("sampler: ") + hipGetErrorString(e));
    return st;
}
