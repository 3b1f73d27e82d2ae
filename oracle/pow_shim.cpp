// CPU test infrastructure only: the device's glibc-compatible pow
// (solver-informed-dispatching-rules_amd/csrc/sigp_pow.h, the function the
// rmod operator calls on the GPU) compiled for the host, beside the host's
// own libm pow, so that tests/test_pow_glibc.py can compare the two bit for
// bit on millions of arguments without a GPU.
#include <math.h>
#include <stdint.h>
#include "../solver-informed-dispatching-rules_amd/csrc/sigp_exp_table.h"
#include "../solver-informed-dispatching-rules_amd/csrc/sigp_pow.h"
#include "../solver-informed-dispatching-rules_amd/csrc/sigp_pow_log_table.h"

static const double kLogTab[384] = SIGP_POW_LOG_TAB_INIT;
static const uint64_t kExpTab[256] = SIGP_EXP_TAB_INIT;

extern "C" {
double orc_pow_restated(double x, double y) { return sigp_pow::pow(x, y, kLogTab, kExpTab); }

// mine[k] = the restatement, libm_out[k] = the C library's pow; returns the
// number of k whose bit patterns differ (two NaNs count as equal).
long orc_pow_sweep(const double *x, const double *y, long n, double *mine, double *libm_out) {
    long bad = 0;
    for (long k = 0; k < n; k++) {
        const double a = sigp_pow::pow(x[k], y[k], kLogTab, kExpTab);
        const double b = pow(x[k], y[k]);
        mine[k] = a;
        libm_out[k] = b;
        if (__builtin_bit_cast(uint64_t, a) != __builtin_bit_cast(uint64_t, b) && !(a != a && b != b)) bad++;
    }
    return bad;
}
}
