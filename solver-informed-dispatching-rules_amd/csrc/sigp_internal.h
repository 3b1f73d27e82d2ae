// Interfaces between the library's translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sigp.h"

// Rules already prepared for the kernels and resident on the device (the
// device GP loop writes them between launches): dead leaves are the constant
// 0, live leaves index the compact value store; per rule the read set, the
// fused-scan flags and the launch order.  KL and SL must bound every rule the
// arrays will ever hold (value-store and stack rows per wavefront), and
// has_rmod / has_rare select the kernel variants for all of them.
struct sigp_device_rules {
    const int32_t *ops = nullptr, *args = nullptr, *off = nullptr, *len = nullptr;
    const uint64_t *rd = nullptr, *flags = nullptr;
    const int32_t *order = nullptr;   // rules longest first; null = index order
    // rep[p] != p: rule p repeats rule rep[p] (same tokens, same scenario
    // group) and is not simulated; the caller copies rep[p]'s results.  Null:
    // every rule is simulated.
    const int32_t *rep = nullptr;
    int P = 0, KL = 1, SL = 1;
    bool has_rmod = false, has_rare = false;
};

// Decision situations of the phenotype, resident on the device.
struct sigp_device_situations {
    int n_sit = 0, n_cand = 0;
    const int32_t *cur = nullptr, *rem = nullptr, *cand = nullptr;   // n_sit, n_sit, n_sit x n_cand
    const double *ct = nullptr;                                      // n_sit
    const uint32_t *vis = nullptr;                                   // n_sit x 16 words
};

sigp_status sigp_internal_fail(sigp_status st, const std::string &msg);

// A batch over device-resident rules: rule p reads scenario rows
// (p / group_size) * T + s; sigp_batch_run launches it without allocating
// when given a pairs buffer.  The rule arrays stay owned by the caller.
sigp_status sigp_internal_batch_device(sigp_instance *inst, const sigp_device_rules &rules, int T,
                                       int group_size, const sigp_launch_config &cfg, sigp_batch **out);

// Phenotype ranks of a device batch's rules (int32 [P][n_sit][n_cand]) on
// `stream`; n_sit * n_cand <= 1024.
sigp_status sigp_internal_phenotype_device(const sigp_batch *bt, const sigp_device_situations &sit,
                                           int32_t *out_dev, hipStream_t stream);

int sigp_internal_instance_n(const sigp_instance *inst);
const double *sigp_internal_instance_dur(const sigp_instance *inst);   // device: expected durations

// Scenario rows on `stream` (csrc/sigp_sample.hip): row r of `rows` draws from
// Mt19937(group_seed[r / T] (seed0 if null) + (offset + *gen (0 if null)) * stride + r % T).
sigp_status sigp_internal_sample(const sigp_instance *inst, const sigp_sampler *smp, const long long *group_seed,
                                 long long seed0, int T, long long stride, long long offset, const int *gen,
                                 int rows, double *out, hipStream_t stream);

// Surrogate scores of `rows` rules from device-resident phenotype ranks
// (int32 [rows][SIGP_MLP_IN]) into y (device floats), on `stream`; the same
// values sigp_mlp_predict gives for the host features rank / (n_cand - 1).
sigp_status sigp_internal_mlp_predict_ranks(sigp_mlp *net, const int32_t *ranks, int64_t rows, int n_cand,
                                            float *y, hipStream_t stream);
