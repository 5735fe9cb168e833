"""Accuracy of the path-sum kernels' FP64 cos / sin of 2 pi x and exp (fp64math.h),
built for the host with g++, against long-double libm (80-bit on x86-64) on dense random
samples, plus the exact special values and the flush-to-zero / overflow ends of exp."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("fp64math") / "libfp64math_host.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    os.path.join(HERE, "fp64math_host.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.fp64_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return L


def _eval(L, which, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(x.size * (2 if which == 2 else 1))
    L.fp64_eval(which, x.ctypes.data, y.ctypes.data, x.size)
    return y


TWO_PI = 2 * np.longdouble("3.14159265358979323846264338327950288")   # 2 pi in long double


def test_cos_sin_2pi(lib):
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-40, 40, 400_000), rng.uniform(-1e-6, 1e-6, 50_000),
                        np.arange(-8, 8.001, 0.125)])   # quarter points: exact zeros / ones
    xl = x.astype(np.longdouble)
    ref_c, ref_s = np.cos(TWO_PI * xl), np.sin(TWO_PI * xl)
    assert np.abs(_eval(lib, 0, x) - ref_c).max() < 4e-16
    assert np.abs(_eval(lib, 1, x) - ref_s).max() < 4e-16
    sc = _eval(lib, 2, x).reshape(-1, 2)
    assert np.abs(sc[:, 0] - ref_s).max() < 4e-16
    assert np.abs(sc[:, 1] - ref_c).max() < 4e-16
    q = np.array([0.0, 0.25, 0.5, 0.75, 1.0, -0.25, 3.0])
    np.testing.assert_array_equal(_eval(lib, 0, q), [1, 0, -1, 0, 1, 0, 1])


def test_exp(lib):
    rng = np.random.default_rng(12)
    x = np.concatenate([rng.uniform(-700, 0, 400_000), rng.uniform(-2, 2, 100_000),
                        rng.uniform(0, 700, 50_000)])
    ref = np.exp(x.astype(np.longdouble))
    rel = np.abs((_eval(lib, 3, x) - ref) / ref)
    assert rel.max() < 2e-16
    assert _eval(lib, 3, np.array([0.0]))[0] == 1.0
    np.testing.assert_array_equal(_eval(lib, 3, np.array([-800.0, -2000.0, 800.0])),
                                  [0.0, 0.0, np.inf])
    # huge negative arguments (c1 = -2 / lambda with lambda near 0): 0, as exp gives, not a
    # wrapped exponent; NaN stays NaN
    big = np.array([-1100.0, -5e8, -1.6e9, -3e9, -1e12, -1e300, -np.inf])
    np.testing.assert_array_equal(_eval(lib, 3, big), np.zeros(big.size))
    assert np.isnan(_eval(lib, 3, np.array([np.nan]))[0])
