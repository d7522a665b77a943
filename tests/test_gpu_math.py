"""GPU: the transport kernels' fast reciprocal equals IEEE division on every
binary32 input, so replacing 1.0f / x by it keeps GPU trajectories identical
to the oracle's (which divides)."""
from __future__ import annotations

import ctypes as C
import os

import pytest

from nmc_amd.build import PROBE_LIB

pytestmark = pytest.mark.gpu


def test_rcp_rn_equals_division_for_all_inputs(gpu):
    assert os.path.exists(PROBE_LIB), "libnmc_probe.so is not built"
    L = C.CDLL(PROBE_LIB)
    L.nmcp_rcp_check.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]
    bad, first = C.c_ulonglong(), C.c_uint()
    assert L.nmcp_rcp_check(C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} inputs differ, first 0x{first.value:08x}"
