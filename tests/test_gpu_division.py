"""The scorer forms the five EM quotients f_b r_b / s of a read from one correctly rounded
reciprocal of s and an FMA correction per quotient.  Bit-for-bit agreement with ordinary division
over 2^31 random and adversarial operand pairs (EM terms, frequency updates, the whole exponent
range, quotients near 1, significands of all ones / all zeros) on the device."""
import ctypes as C
import sys

import pytest

from breseq_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_shared_reciprocal_division_is_correctly_rounded():
    print("[gpu] shared-reciprocal division self-check", flush=True)
    sys.stdout.flush()
    n = 1 << 31
    bad, fast = C.c_uint64(), C.c_uint64()
    st = L.Status()
    L.check(L.lib().bsq_selftest_division(0, n, 12345, C.byref(bad), C.byref(fast), C.byref(st)), st)
    assert bad.value == 0, bad.value
    assert fast.value > n // 2, fast.value
