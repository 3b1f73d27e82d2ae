// CPU ORACLE -- test infrastructure only: the correctly rounded log / sincos
// of the device scenario sampler (solver-informed-dispatching-rules_amd/csrc/
// sigp_ddmath.h) compiled for the host, so that a host restatement of the
// sampler (sigp/scenarios.py with these functions injected) can check the
// device's rows bit for bit.  The functions themselves are pinned against
// 60-digit decimal values in tests/test_crmath.py.
#include "../solver-informed-dispatching-rules_amd/csrc/sigp_ddmath.h"

extern "C" double orc_cr_log(double x) { return sigp_dd::cr_log(x); }
extern "C" void orc_cr_sincos(double x, double *s, double *c) { sigp_dd::cr_sincos(x, s, c); }
