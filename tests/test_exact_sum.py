"""Exact, order-independent sums of the adjoint (eos_amd/csrc/exact_sum.h),
compiled for the host and checked against exact rational arithmetic: every
value becomes sign * trunc(|c| * 2^64) split into 32-bit limbs, and sums of
limbs are exact integer sums whatever the order."""
import ctypes as C
import math
import os
import random
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "exact_sum.h"
extern "C" int es_limbs(float c, long long* out) { return eos::to_limbs(c, out[0], out[1], out[2]) ? 1 : 0; }
extern "C" void es_sum(const float* v, int n, unsigned long long* limbs) {
    for (int i = 0; i < n; ++i) {
        long long l[3];
        if (!eos::to_limbs(v[i], l[0], l[1], l[2])) continue;
        for (int k = 0; k < 3; ++k) limbs[k] += (unsigned long long)l[k];   // wraps like the GPU's integer atomics
    }
}
extern "C" double es_value(const unsigned long long* limbs) { return eos::limb_value(limbs, 1, 0); }
'''


@pytest.fixture(scope="module")
def es(tmp_path_factory):
    d = tmp_path_factory.mktemp("es")
    (d / "es.cpp").write_text(SRC)
    so = d / "libes.so"
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-I", os.path.join(ROOT, "eos_amd", "csrc"),
                           str(d / "es.cpp"), "-o", str(so)])
    L = C.CDLL(str(so))
    L.es_limbs.argtypes = [C.c_float, C.c_void_p]
    L.es_sum.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.es_value.argtypes = [C.c_void_p]
    L.es_value.restype = C.c_double
    return L


def _exact(c: float) -> int:
    """sign(c) * trunc(|c| * 2^64) in exact arithmetic."""
    f = abs(Fraction(float(np.float32(c)))) * 2 ** 64
    return int(f) * (-1 if c < 0 else 1)


def _values(n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        e = rng.randint(-60, 30)
        out.append(np.float32(rng.choice((-1.0, 1.0)) * rng.random() * 2.0 ** e))
    out += [np.float32(0.0), np.float32(-0.0), np.float32(1e-45), np.float32(2.0 ** 30 * 1.999),
            np.float32(-(2.0 ** -41)), np.float32(2.0 ** -64), np.float32(3.0 * 2.0 ** -66)]
    return np.array(out, np.float32)


def test_limbs_are_the_truncated_fixed_point_value(es):
    out = np.zeros(3, np.int64)
    for c in _values(5000, 1):
        assert es.es_limbs(float(c), out.ctypes.data) == 1
        l0, l1, l2 = (int(x) for x in out)
        assert 0 <= abs(l0) < 2 ** 32 and 0 <= abs(l1) < 2 ** 32 and abs(l2) < 2 ** 31
        assert l2 * 2 ** 64 + l1 * 2 ** 32 + l0 == _exact(c), float(c)


def test_out_of_range_values_are_refused(es):
    out = np.zeros(3, np.int64)
    for c in (2.0 ** 31, -(2.0 ** 33), math.inf, -math.inf, math.nan):
        assert es.es_limbs(c, out.ctypes.data) == 0
    assert es.es_limbs(float(np.nextafter(np.float32(2.0 ** 31), np.float32(0))), out.ctypes.data) == 1


def test_sums_are_exact_and_order_independent(es):
    v = _values(100000, 2)
    limbs = np.zeros(3, np.uint64)
    es.es_sum(v.ctypes.data, len(v), limbs.ctypes.data)
    want = sum(_exact(c) for c in v)
    # the limbs, read as signed and carried, hold exactly the integer sum
    l0, l1, l2 = (int(x) - (1 << 64 if int(x) >= 1 << 63 else 0) for x in limbs)
    assert l2 * 2 ** 64 + l1 * 2 ** 32 + l0 == want
    rng = np.random.default_rng(3)
    for _ in range(3):
        p = v[rng.permutation(len(v))]
        other = np.zeros(3, np.uint64)
        es.es_sum(p.ctypes.data, len(p), other.ctypes.data)
        assert np.array_equal(other, limbs)
    # and the double it stands for is the exact sum to FP64 rounding
    got = es.es_value(limbs.ctypes.data)
    exact = Fraction(want, 2 ** 64)
    assert abs(Fraction(got) - exact) <= abs(exact) * 2 ** -52 + Fraction(1, 2 ** 64)
    assert abs(got - math.fsum(float(c) for c in v)) <= len(v) * 2.0 ** -64 + abs(got) * 2 ** -50
