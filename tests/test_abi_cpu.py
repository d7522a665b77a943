"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/nmc.h, its structs match the reference layout, and the
validation layer turns bad configurations into typed errors."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd._lib import EXPORTED, NmcError, lib
from nmc_amd.metal_mc import validate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions() -> set[str]:
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(nmc_\w+)\s*\(", src))
    return names


def test_every_declared_symbol_is_exported():
    L = lib()
    declared = header_functions()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    assert set(EXPORTED) <= declared


def test_struct_layout_matches_reference():
    assert C.sizeof(P.RunParams) == 172 and C.sizeof(P.TissueParams) == 608
    off = {n: getattr(P.RunParams, n).offset for n, _ in P.RunParams._fields_}
    # offsets of definitions.h:186-209 with 32-byte myname and 4-byte enums
    assert off["mcflag"] == 32 and off["xfocus"] == 52 and off["ux0"] == 64
    assert off["radius"] == 76 and off["zsurf"] == 92 and off["n1"] == 96
    assert off["na"] == 104 and off["det_state"] == 116 and off["ph_quant"] == 120
    assert off["dx"] == 124 and off["Nx"] == 136 and off["Nt"] == 148
    assert off["xs"] == 152 and off["time_min"] == 164 and off["mckernelflag"] == 168
    assert [getattr(P.TissueParams, n).offset for n in ("muav", "musv", "gv", "gf", "CC")] == \
        [0, 76, 152, 228, 532]
    assert C.sizeof(P.Totals) == 23 * 8


def test_ext_config_defaults_match_library():
    e = P.ExtConfig()
    lib().nmc_ext_config_init(C.byref(e))
    d = P.ExtConfig.default()
    for name, _ in P.ExtConfig._fields_:
        if name == "reserved":
            continue
        assert getattr(e, name) == pytest.approx(getattr(d, name)), name
    assert e.threshold == pytest.approx(1e-4) and e.chance == pytest.approx(0.1)


def test_version_string():
    assert b"gfx950" in lib().nmc_version()


def _expect(code, fn):
    with pytest.raises(NmcError) as ei:
        fn()
    assert ei.value.code == code, str(ei.value)
    return str(ei.value)


def test_valid_configs_pass():
    for w in (W.skin_vessel(n=20), W.mcsub_semi_infinite(), W.tiny_voxel_case("two_term")):
        validate(w.sim, w.rp, w.tp, w.v, w.ext)


def test_validation_errors():
    w = W.tiny_voxel_case("mismatch")
    v = w.v.copy()
    v[5] = 19
    assert "tissue id 19" in _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, w.tp, v, w.ext))
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, w.tp, None, w.ext))
    tp = W._tp({1: (0.01, 1.0, 1.0), 2: (2.0, -60.0, 0.8), 3: (8.0, 30.0, 0.5)})
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, tp, w.v, w.ext))
    tp = W._tp({1: (0.01, 1.0, 1.0), 2: (float("nan"), 60.0, 0.8), 3: (8.0, 30.0, 0.5)})
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, tp, w.v, w.ext))
    tp = W._tp({1: (0.01, 1.0, 1.0), 2: (2.0, 60.0, 1.5), 3: (8.0, 30.0, 0.5)})
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, tp, w.v, w.ext))
    for field, bad in (("boundaryflag", 3), ("mcflag", 7), ("det_state", 2), ("dx", 0.0),
                       ("Nx", 0), ("n1", -1.0)):
        rp = P.RunParams.from_buffer_copy(w.rp)
        setattr(rp, field, bad)
        _expect(P.ERR_ARG, lambda: validate(w.sim, rp, w.tp, w.v, w.ext))
    rp = P.RunParams.from_buffer_copy(w.rp)
    rp.launchflag, rp.ux0, rp.uy0, rp.uz0 = 1, 1.0, 1.0, 0.0
    _expect(P.ERR_ARG, lambda: validate(w.sim, rp, w.tp, w.v, w.ext))
    ext = P.ExtConfig.default(version=99)
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, w.tp, w.v, ext))
    ext = P.ExtConfig.default(chance=0.0)
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, w.tp, w.v, ext))
    ext = P.ExtConfig.default(max_events=0)
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, w.tp, w.v, ext))
    _expect(P.ERR_UNSUPPORTED, lambda: validate(5, w.rp, w.tp, w.v, w.ext))


def test_mcsub_validation_errors():
    w = W.mcsub_semi_infinite()
    for kw in (dict(nr=0), dict(na=65), dict(dr=0.0)):
        ext = P.ExtConfig.default(**kw)
        _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, w.tp, None, ext))
    tp = W._tp({1: (0.0, 0.0, 0.9)})
    _expect(P.ERR_ARG, lambda: validate(w.sim, w.rp, tp, None, w.ext))


def test_mcsub_source_depth_rule():
    """The mcsub medium's surface is z = 0 for every source: a beam (mcflag 0,
    1) must enter there, a point source (mcflag 2) sits at zs >= 0 inside."""
    w = W.mcsub_semi_infinite()
    for mcflag in (0, 1):
        rp = P.RunParams.from_buffer_copy(w.rp)
        rp.mcflag, rp.zs = mcflag, 0.05
        assert "zs" in _expect(P.ERR_ARG, lambda: validate(w.sim, rp, w.tp, None, w.ext))
        rp.zs = 0.0
        validate(w.sim, rp, w.tp, None, w.ext)
    rp = P.RunParams.from_buffer_copy(w.rp)
    rp.mcflag, rp.zs = 2, 0.05
    validate(w.sim, rp, w.tp, None, w.ext)
    for bad in (-0.01, float("inf"), float("nan")):
        rp.zs = bad
        _expect(P.ERR_ARG, lambda: validate(w.sim, rp, w.tp, None, w.ext))


def test_create_without_device_fails_cleanly():
    h = C.c_void_p()
    if os.environ.get("HIP_VISIBLE_DEVICES") is None:
        st = lib().nmc_create(1, 0, C.byref(h))
        if st != 0:  # no GPU in this process: a typed error, not a crash
            assert st == P.ERR_DEVICE and lib().nmc_last_error()
        else:
            lib().nmc_destroy(h)
    st = lib().nmc_create(99, 0, C.byref(h))
    assert st == P.ERR_UNSUPPORTED


def test_workload_geometry():
    w = W.skin_vessel()
    n = 200
    v = w.v.reshape(n, n, n)        # (iz, ix, iy) = linear iz*Ny*Nx + ix*Ny + iy
    assert v.dtype == np.uint8 and w.v.size == n ** 3
    assert (v[:2] == 1).all()                       # air: z < 0.01
    assert (v[2][v[2] != 4] == 2).all()             # epidermis row iz=2 outside the vessel
    zc = (np.arange(n) + 0.5) / n
    yc = (np.arange(n) + 0.5 - n / 2) / n
    vessel = (yc[None, :] ** 2 + (zc[:, None] - 0.06) ** 2) <= 0.05 ** 2   # (iz, iy)
    vessel[zc < 0.01] = False
    # the vessel runs along x: same cross-section for every ix
    assert np.array_equal(v[:, 17, :] == 4, vessel)
    assert np.array_equal(v[:, 150, :] == 4, vessel)
    assert (v[zc >= 0.11] == 3).all()


def test_noalloc_entry_point_rejects_bad_arguments_before_device_work():
    """nmc_mcxyz_run_noalloc checks its arguments on the host first."""
    w = W.tiny_voxel_case("mismatch")
    L = lib()
    null = C.c_void_p(0)
    st = L.nmc_mcxyz_run_noalloc(P.SIM_MCSUB, C.byref(w.rp), C.byref(w.tp), C.byref(w.ext),
                                 C.POINTER(P.Tuning)(), null, 0, 10, null, null, null, null, null)
    assert st == P.ERR_UNSUPPORTED
    st = L.nmc_mcxyz_run_noalloc(w.sim, C.byref(w.rp), C.byref(w.tp), C.byref(w.ext),
                                 C.POINTER(P.Tuning)(), null, 0, 10, null, null, null, null, null)
    assert st == P.ERR_ARG and b"required" in L.nmc_last_error()
    rp = P.RunParams.from_buffer_copy(w.rp)
    rp.Nx = 0
    st = L.nmc_mcxyz_run_noalloc(w.sim, C.byref(rp), C.byref(w.tp), C.byref(w.ext),
                                 C.POINTER(P.Tuning)(), null, 0, 10, null, null, null, null, null)
    assert st == P.ERR_ARG
