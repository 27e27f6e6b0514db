"""fmh_mul_device_noalloc (the allocation-free variant of fmh_mul_device,
include/fmh.h) against fmh_mul_device through the C-ABI: two calls in a row
on one stream, the second with a smaller input, then the largest product it
takes and a square; every result equals the existing entry point's and the
oracle's, word for word.  (Its refusal of longer products is checked in the
CPU emulation, tests/emu/emu_check.py, which never hands a GPU an undersized
buffer.)"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(fm, s, a, b, square=False):
    """(noalloc result, fmh_mul_device result) of a * b (b ignored for a square)."""
    na, nb = a.size, (a.size if square else b.size)
    with fm.DeviceBuffer.from_host(a, s) as da, fm.DeviceBuffer.from_host(a if square else b, s) as db, \
            fm.DeviceBuffer(na + nb) as d1, fm.DeviceBuffer(na + nb) as d2:
        second = da if square else db
        fm.mul_device_noalloc(da, na, second, nb, d1, s)
        s.synchronize()
        got = d1.download()
        fm.mul_device(da, na, second, nb, d2, s)
        s.synchronize()
        return got, d2.download()


def test_noalloc_matches_device_entry_point(fm, orc):
    lib = fm.lib()
    assert lib.fmh_init(0) == 0
    with fm.Stream() as s:
        # two calls in a row, the second smaller (the array keeps the first's contents)
        for n1, n2 in [(300000, 200001), (70000, 1000)]:
            a, b = orc.splitmix(n1 + 11, n1), orc.splitmix(n2 + 13, n2)
            got, want = _both(fm, s, a, b)
            assert np.array_equal(got, want), (n1, n2)
            assert np.array_equal(got, orc.mul(a, b)), (n1, n2)
        # the largest product it takes (2^20 coefficients), an unbalanced one
        # the other entry point would split, and a square
        for n1, n2, sq in [(1 << 19, (1 << 19) + 1, False), ((1 << 20) - 40, 41, False), (1 << 19, 0, True)]:
            a = orc.splitmix(n1 + 17, n1)
            b = None if sq else orc.splitmix(n2 + 19, n2)
            got, want = _both(fm, s, a, b, square=sq)
            assert np.array_equal(got, want), (n1, n2, sq)
            assert np.array_equal(got, orc.sqr(a) if sq else orc.mul(a, b)), (n1, n2, sq)
