"""The reference's own GPU gate (verify/verify_cuda.cu), compiled unchanged
against libsigp.so through include/sigp_factory.hpp (oracle/Makefile target
`gate`), run on the device: (a) replays every case of the sim fixture and must
match the recorded rewards bit for bit; (b) compares the reference's CPU
simulator (compiled into the same binary) with this library on the all-61-slot
sum rule, 4 instances x 2 seeds x 64 scenarios, and must be bit-exact."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GATE = os.path.join(ROOT, "oracle", "_ref", "verify_cuda_sigp")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(GATE),
                    reason="oracle/_ref/verify_cuda_sigp not built (needs the reference sources)")
def test_reference_gate_bit_exact():
    env = dict(os.environ, GITC_DATA=os.path.join(ROOT, "tests", "golden", "instances"))
    print("running", GATE, flush=True)
    r = subprocess.run([GATE, os.path.join(ROOT, "tests", "golden", "sim.txt")], env=env,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "part (a) GPU vs recorded fixtures: 240 compared" in r.stdout
    last = r.stdout.strip().splitlines()[-1]
    assert "OK bit-exact" in last and ", 0 nonzero," in last, last
