"""The allocation-free entry point nmc_mcxyz_run_noalloc against the handle
path (nmc_prepare + nmc_set_seed + nmc_run) on the same photons: integer
tallies bit-exact, float tallies to FP64 summation order, on consecutive
calls that share the library's static scratch array (the second call on a
smaller grid), plus its rejections."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd._lib import lib
from nmc_amd.metal_mc import simulate

from parity_util import assert_same_tallies

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceTallies:
    """Zeroed device arrays for one mcxyz tally set (torch as the allocator)."""

    def __init__(self, w, fluence=True):
        import torch
        rp = w.rp
        dev = torch.device("cuda:0")
        self.v = torch.from_numpy(np.ascontiguousarray(w.v, np.uint8)).to(dev)
        self.fluence = (torch.zeros(rp.Nz * rp.Nx * rp.Ny, dtype=torch.float64, device=dev)
                        if fluence else None)
        self.ryx = torch.zeros(rp.Nx * rp.Ny, dtype=torch.float64, device=dev)
        self.det2d = torch.zeros(rp.Nx * rp.Ny, dtype=torch.int32, device=dev)
        self.totals = torch.zeros(C.sizeof(P.Totals) // 8, dtype=torch.float64, device=dev)

    def host(self, w) -> P.TallyArrays:
        import torch
        torch.cuda.synchronize()
        t = P.TallyArrays(w.rp, w.ext, w.sim, fluence=self.fluence is not None)
        if self.fluence is not None:
            t.fluence[:] = self.fluence.cpu().numpy()
        t.ryx[:] = self.ryx.cpu().numpy()
        t.det2d[:] = self.det2d.cpu().numpy().view(np.uint32)
        raw = self.totals.cpu().numpy().tobytes()
        C.memmove(C.byref(t.res.totals), raw, C.sizeof(P.Totals))
        return t


def noalloc(w, d: DeviceTallies, n: int, first: int = 0, tuning: P.Tuning | None = None,
            v=None) -> int:
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    vt = d.v if v is None else v
    st = lib().nmc_mcxyz_run_noalloc(
        w.sim, C.byref(w.rp), C.byref(w.tp), C.byref(w.ext),
        C.byref(tuning) if tuning is not None else C.POINTER(P.Tuning)(),
        C.c_void_p(vt.data_ptr()), first, n,
        C.c_void_p(d.fluence.data_ptr() if d.fluence is not None else 0),
        C.c_void_p(d.ryx.data_ptr()), C.c_void_p(d.det2d.data_ptr()),
        C.c_void_p(d.totals.data_ptr()), C.c_void_p(stream))
    torch.cuda.synchronize()
    return st


def test_noalloc_matches_handle_path(gpu):
    big = W.skin_vessel()
    small = W.tiny_voxel_case("two_term", n=9)
    n_big, n_small = 20_000, 3_000
    _say("[noalloc] call 1: skin-vessel 200^3, 20000 photons")
    d_big = DeviceTallies(big)
    assert noalloc(big, d_big, n_big) == 0, lib().nmc_last_error()
    got = d_big.host(big)
    ref = simulate(big, n_big)
    assert_same_tallies(got, ref, "noalloc skin-vessel")

    _say("[noalloc] call 2: two-term 9^3, 3000 photons")
    d_small = DeviceTallies(small)
    assert noalloc(small, d_small, n_small) == 0, lib().nmc_last_error()
    assert_same_tallies(d_small.host(small), simulate(small, n_small), "noalloc two-term 9^3")

    _say("[noalloc] call 3: skin-vessel photons [20000, 25000) added to call 1")
    assert noalloc(big, d_big, 5_000, first=n_big) == 0, lib().nmc_last_error()
    assert_same_tallies(d_big.host(big), simulate(big, n_big + 5_000), "noalloc accumulated")


def test_noalloc_fp32_tile_split_and_no_fluence(gpu):
    w = W.tiny_voxel_case("mismatch")
    n = 6_000
    _say("[noalloc] fp32 tile split into launches; no-fluence variant")
    d = DeviceTallies(w)
    t = P.Tuning.make(mode=1, split=3)
    assert noalloc(w, d, n, tuning=t) == 0, lib().nmc_last_error()
    got, ref = d.host(w), simulate(w, n)
    for k in P.TOTAL_INT_FIELDS:
        assert got.totals[k] == ref.totals[k], k
    assert np.array_equal(got.det2d, ref.det2d)
    np.testing.assert_allclose(got.fluence, ref.fluence, rtol=2e-5, atol=1e-9)
    d2 = DeviceTallies(w, fluence=False)
    assert noalloc(w, d2, n) == 0, lib().nmc_last_error()
    assert_same_tallies(d2.host(w), simulate(w, n, fluence=False), "noalloc no fluence")


def test_noalloc_rejections(gpu):
    import torch
    w = W.skin_vessel()
    d = DeviceTallies(w)
    _say("[noalloc] rejections")
    # an explicit tile whose block partials exceed the scratch array (10 000
    # FP64 voxels fit the LDS; x one workgroup per CU is over 2 Mi doubles)
    t = P.Tuning.make(tile=(50, 50, 0, 100, 100, 1))
    assert noalloc(w, d, 1000, tuning=t) == P.ERR_ARG
    assert b"scratch" in lib().nmc_last_error()
    # a voxel map with a tissue id the table does not have (checked on the device)
    bad = d.v.clone()
    bad[12345] = 19
    assert noalloc(w, d, 1000, v=bad) == P.ERR_ARG
    assert b"tissue id 19" in lib().nmc_last_error()
    # nothing was transported by the rejected calls
    assert float(d.totals.abs().sum()) == 0.0 and float(d.fluence.abs().sum()) == 0.0
    assert torch.count_nonzero(d.det2d) == 0
