"""The pair term's square root (the hardware square root corrected to correct rounding,
without sqrtf's rescaling of inputs below 2^-96) equals sqrtf bit for bit on every
float from 2^-96 to 2^24 and on 0: the distances that pick table bins in the step
kernels are the reference's bits. Runs the library's self-test over those patterns."""
import ctypes as C

import pytest

from marsmd import _capi as api

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lo,hi", [(0x00000000, 0x00000001),   # +0
                                   (0x0F800000, 0x4B800000)])  # [2^-96, 2^24)
def test_pair_sqrt_is_correctly_rounded(cuda, lo, hi):
    bad = C.c_uint64(123)
    api.call("marsmd_selftest_pair_sqrt", lo, hi, C.byref(bad), None)
    assert bad.value == 0, f"{bad.value} patterns in [{lo:#010x}, {hi:#010x}) differ from sqrtf"
