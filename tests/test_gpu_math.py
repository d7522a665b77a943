"""GPU: the transport kernels' short reciprocal, division and square root give
the IEEE results bit for bit, so using them keeps GPU trajectories identical
to the oracle's (which divides and takes square roots the IEEE way).

- reciprocal: every binary32 input;
- division: every operand pair of log_pos, every Henyey-Greenstein divisor of
  256 anisotropies, 2^32 pairs shaped like the rotation's, 2^32 general pairs
  with exponents in [-60, 60];
- square root: every input in [2^-100, 2^126], and +0;
- the voxel walk's one-path expm1 against the branched one: every input
  in [-inf, -0];
- the walks' table log and table azimuth sine and cosine against the
  oracle's: every uniform draw."""
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


@pytest.mark.parametrize("domain,name", [(0, "log_pos operands"), (1, "Henyey-Greenstein divisors"),
                                         (2, "rotation operands"), (3, "general pairs")])
def test_div_rn_equals_division(gpu, domain, name):
    L = C.CDLL(PROBE_LIB)
    L.nmcp_div_check.argtypes = [C.c_int] + [C.POINTER(C.c_ulonglong)] * 3
    cases, bad, first = C.c_ulonglong(), C.c_ulonglong(), C.c_ulonglong()
    assert L.nmcp_div_check(domain, C.byref(cases), C.byref(bad), C.byref(first)) == 0
    assert cases.value >= 1 << 23
    assert bad.value == 0, f"{name}: {bad.value} of {cases.value} differ, first case {first.value}"


def test_sqrt_rn_equals_sqrtf_on_its_domain(gpu):
    L = C.CDLL(PROBE_LIB)
    L.nmcp_sqrt_check.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]
    bad, first = C.c_ulonglong(), C.c_uint()
    assert L.nmcp_sqrt_check(C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} inputs differ, first 0x{first.value:08x}"


def test_expm1_neg_equals_expm1_for_all_nonpositive_inputs(gpu):
    L = C.CDLL(PROBE_LIB)
    L.nmcp_expm1_check.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]
    bad, first = C.c_ulonglong(), C.c_uint()
    assert L.nmcp_expm1_check(C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} inputs differ, first 0x{first.value:08x}"


def test_log_tab_equals_oracle_on_every_uniform_draw(gpu):
    """The voxel walk's step-length log (table in LDS) gives the oracle's bits
    for every uniform draw k 2^-24, k = 1..2^24."""
    import numpy as np

    from oracle import oracle as O

    L = C.CDLL(PROBE_LIB)
    L.nmcp_log_tab.argtypes = [C.c_void_p]
    got = np.empty(1 << 24, np.float32)
    assert L.nmcp_log_tab(got.ctypes.data) == 0
    u = (np.arange(1, 2 ** 24 + 1, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    ref = O.logf_tab(u)
    bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, f"{bad.size} draws differ, first k = {bad[0] + 1}"


def test_sincos_tab_equals_oracle_on_every_draw(gpu):
    """The walks' azimuth sine and cosine (table in LDS) give the oracle's bits
    for every random word's top 24 bits."""
    import numpy as np

    from oracle import oracle as O

    L = C.CDLL(PROBE_LIB)
    L.nmcp_sincos_tab.argtypes = [C.c_void_p, C.c_void_p]
    s = np.empty(1 << 24, np.float32)
    c = np.empty(1 << 24, np.float32)
    assert L.nmcp_sincos_tab(s.ctypes.data, c.ctypes.data) == 0
    k = np.arange(0, 1 << 24, dtype=np.uint32)
    rs, rc = O.sincos2pi_bits(k << np.uint32(8) | np.uint32(0x5A))
    for got, ref, name in ((s, rs, "sin"), (c, rc, "cos")):
        bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, f"{name}: {bad.size} words differ, first k = {bad[0]}"
