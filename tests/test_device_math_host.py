"""Host build of the kernels' correctly rounded fifth root (diffrax_amd/csrc/erk_common.hpp,
cr_pow_fifth), the PIDController power (1/scaled_error)^0.2 of pid.py:512-515 in every
default-controller step of the small-system kernels, and of their error norm with the
division by sqrt(d) written out (rms_norm_rc).

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
def hostlib(tmp_path_factory):
    if shutil.which("g++") is None or not Path("/opt/rocm/include/hip/hip_runtime.h").exists():
        pytest.skip("g++ or the HIP headers are not available")
    d = tmp_path_factory.mktemp("hostpow")
    src = d / "hostpow.cpp"
    src.write_text('#include "erk_common.hpp"\n'
                   'extern "C" double host_cr_pow_fifth(double x) { return dfx::cr_pow_fifth(x); }\n'
                   'template <int D> static double rn(const double* e) {\n'
                   '  double v[D]; for (int j = 0; j < D; ++j) v[j] = e[j];\n'
                   '  const double c = sqrt((double)D);\n'
                   '  return dfx::rms_norm_rc<D>(v, c, dfx::refined_rcp(c)); }\n'
                   'extern "C" double host_rms_norm_rc(const double* e, int d) {\n'
                   '  switch (d) { case 2: return rn<2>(e); case 3: return rn<3>(e);\n'
                   '    case 4: return rn<4>(e); case 10: return rn<10>(e); case 16: return rn<16>(e); }\n'
                   '  return -1.0; }\n')
    so = d / "hostpow.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}", str(src),
                    "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.host_cr_pow_fifth.argtypes = [ctypes.c_double]
    lib.host_cr_pow_fifth.restype = ctypes.c_double
    lib.host_rms_norm_rc.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.host_rms_norm_rc.restype = ctypes.c_double
    return lib


@pytest.fixture(scope="module")
def fifth(hostlib):
    return hostlib.host_cr_pow_fifth


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


def test_rms_norm_division_written_out(hostlib):
    """rms_norm_rc equals the oracle's sqrt(sum e^2) / sqrt(d) (two correctly rounded
    operations) bit for bit over the whole reachable range of the norm: sums of squares
    from the smallest subnormal (norm 2^-537) to overflow, sign and magnitude mixes, zero,
    and inf / NaN components."""
    rng = np.random.default_rng(2)
    f = hostlib.host_rms_norm_rc
    for d in (2, 3, 4, 10, 16):
        cases = [np.zeros(d), np.full(d, 5e-324), np.full(d, 1e-160), np.full(d, 1e154),
                 np.full(d, 1.3e154)]
        for _ in range(4000):
            e = rng.normal(size=d) * 10.0 ** rng.uniform(-330, 160)
            e[rng.random(d) < 0.3] = 0.0
            cases.append(e)
        for _ in range(500):   # sums of squares in the subnormal range
            cases.append(rng.uniform(-1, 1, size=d) * 2.0 ** rng.uniform(-560, -537))
        for e in cases:
            e = np.ascontiguousarray(e, dtype=np.float64)
            got = f(e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), d)
            s = 0.0
            for v in e.tolist():
                s = s + v * v
            want = math.sqrt(s) / math.sqrt(d)    # the kernel's summation order
            assert got == want or (math.isnan(got) and math.isnan(want)), (d, e, got, want)
        inf = np.full(d, 0.5)
        inf[0] = math.inf
        assert f(inf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), d) == math.inf
        nan = np.full(d, 0.5)
        nan[-1] = math.nan
        assert math.isnan(f(nan.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), d))
