"""Host build of the kernels' correctly rounded fifth root (diffrax_amd/csrc/erk_common.hpp,
cr_pow_fifth), the PIDController power (1/scaled_error)^0.2 of pid.py:512-515 in every
default-controller step of the small-system kernels.

The header is compiled with g++ into a throwaway library; its host branch differs from the
device one only in the float guess (libm exp2f/log2f instead of v_exp_f32/v_log_f32), whose
error the third-order step and the double-double residual correction remove.  Pins:
- 60-digit decimal x^0.2 (the same inputs as the oracle's pin), exact;
- random float64 bit patterns (subnormals, extremes) against the oracle's independent
  Newton-with-division restatement, bit for bit;
- zero, negative, NaN and infinite inputs (the branch-free special-value selects).
The device branch is pinned by the GPU tests' bit parity with the oracle over every step
of the C1/C2 workloads."""
import ctypes
import math
import shutil
import subprocess
from decimal import Decimal, localcontext
from pathlib import Path

import numpy as np
import pytest

import oracle

CSRC = Path(__file__).resolve().parents[1] / "diffrax_amd" / "csrc"


@pytest.fixture(scope="module")
def fifth(tmp_path_factory):
    if shutil.which("g++") is None or not Path("/opt/rocm/include/hip/hip_runtime.h").exists():
        pytest.skip("g++ or the HIP headers are not available")
    d = tmp_path_factory.mktemp("hostpow")
    src = d / "hostpow.cpp"
    src.write_text('#include "erk_common.hpp"\n'
                   'extern "C" double host_cr_pow_fifth(double x) { return dfx::cr_pow_fifth(x); }\n')
    so = d / "hostpow.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}", str(src),
                    "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.host_cr_pow_fifth.argtypes = [ctypes.c_double]
    lib.host_cr_pow_fifth.restype = ctypes.c_double
    return lib.host_cr_pow_fifth


def test_against_decimal(fifth):
    rng = np.random.default_rng(0)
    xs = np.concatenate([10 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2, 3000),
                         [5e-324, 1e-310, 2.2250738585072014e-308, 1.0, 2.0, 32.0, 1e308,
                          1.7976931348623157e308]])
    with localcontext() as ctx:
        ctx.prec = 60
        for x in xs:
            assert fifth(float(x)) == float(Decimal(float(x)) ** Decimal(0.2)), x


def test_bit_patterns_against_oracle(fifth):
    L = oracle.lib()
    L.orc_pow.argtypes = [ctypes.c_double, ctypes.c_double]
    L.orc_pow.restype = ctypes.c_double
    rng = np.random.default_rng(1)
    bits = rng.integers(1, 0x7FF0000000000000, size=20000, dtype=np.int64)
    for x in bits.view(np.float64):
        a, b = fifth(float(x)), L.orc_pow(float(x), 0.2)
        assert a == b, (x, a, b)


def test_special_inputs(fifth):
    assert fifth(0.0) == 0.0 and math.copysign(1.0, fifth(-0.0)) == 1.0
    assert math.isinf(fifth(math.inf)) and fifth(math.inf) > 0
    assert math.isnan(fifth(-1.0)) and math.isnan(fifth(-math.inf)) and math.isnan(fifth(math.nan))
    assert fifth(32.0) == 2.0 and fifth(1.0) == 1.0 and fifth(2.0 ** -1070) > 0
