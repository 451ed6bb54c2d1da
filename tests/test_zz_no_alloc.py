"""The allocation-free trims entry point (bsq_sequence_trims_static) against bsq_sequence_trims, bit
for bit, through the C ABI: two calls in a row, the second smaller than the first, so a stale tail
of the first call's scratch contents would show; and the size limit reported as an error."""
import ctypes as C

import numpy as np
import pytest

from breseq_amd import _lib as L

pytestmark = pytest.mark.gpu

MAX = 8 << 20


def _trims(fn, seq):
    b = seq.encode()
    out = np.zeros(2 * len(b), dtype=np.uint8)
    st = L.Status()
    rc = fn(b, len(b), out.ctypes.data_as(L.c_u8p), C.byref(st))
    assert rc == 0, st.message
    return out


def _sequence(rng, n):
    s = rng.choice(list("ACGT"), n)
    # tandem repeats of every unit size up to 18 and beyond the 255 cap, at random places
    for unit in list(range(1, 19)) + [3, 7, 1]:
        at = int(rng.integers(0, n - 600))
        rep = rng.choice(list("ACGT"), unit)
        k = int(rng.integers(2, 600 // unit + 2))
        block = np.tile(rep, k)[: n - at]
        s[at:at + len(block)] = block
    s[rng.random(n) < 1e-4] = "N"
    return "".join(s)


def test_static_trims_equal_allocating_entry_point():
    lib = L.lib()
    fn = lib.bsq_sequence_trims_static
    fn.restype = C.c_int
    fn.argtypes = [C.c_char_p, C.c_int64, L.c_u8p, C.POINTER(L.Status)]
    rng = np.random.default_rng(9)
    first = _sequence(rng, 6_000_000)
    second = _sequence(rng, 250_000)
    for seq in (first, second):
        print("[gpu] static trims, %d bases" % len(seq), flush=True)
        want = _trims(lib.bsq_sequence_trims, seq)
        got = _trims(fn, seq)
        assert np.array_equal(got, want)
    # beyond the scratch capacity: a clear error, nothing launched
    st = L.Status()
    out = np.zeros(2, dtype=np.uint8)
    rc = fn(b"A", MAX + 1, out.ctypes.data_as(L.c_u8p), C.byref(st))
    assert rc == 5 and b"capacity" in st.message
