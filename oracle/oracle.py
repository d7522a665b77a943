"""ctypes loader for the CPU oracle (oracle/libnmc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
benchmark's CPU-baseline leg, never by the nmc_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from nmc_amd import params as P
from nmc_amd.workloads import Workload

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnmc_oracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f32p = C.POINTER(C.c_float)
        L.nmco_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        for name in ("nmco_logf", "nmco_expf", "nmco_expm1f"):
            getattr(L, name).argtypes = [C.c_float]
            getattr(L, name).restype = C.c_float
        for name in ("nmco_logf_n", "nmco_expf_n", "nmco_expm1f_n"):
            getattr(L, name).argtypes = [f32p, f32p, C.c_int64]
        L.nmco_sincos2pi_n.argtypes = [f32p, f32p, f32p, C.c_int64]
        L.nmco_fresnel.argtypes = [C.c_float, C.c_float, C.c_float, f32p]
        L.nmco_fresnel.restype = C.c_float
        L.nmco_hg_cos.argtypes = [C.c_float, C.c_float]
        L.nmco_hg_cos.restype = C.c_float
        L.nmco_tt_cos.argtypes = [C.c_float] * 6
        L.nmco_tt_cos.restype = C.c_float
        L.nmco_spin.argtypes = [C.c_float, C.c_float, f32p, f32p, f32p]
        L.nmco_mcxyz.argtypes = [C.POINTER(P.RunParams), C.POINTER(P.TissueParams),
                                 C.POINTER(C.c_uint8), C.POINTER(P.ExtConfig), C.c_uint64,
                                 C.c_uint64, C.c_int, C.POINTER(P.Results)]
        L.nmco_mcsub.argtypes = [C.POINTER(P.RunParams), C.POINTER(P.TissueParams),
                                 C.POINTER(P.ExtConfig), C.c_uint64, C.c_uint64, C.c_int,
                                 C.POINTER(P.Results)]
        L.nmco_last_seconds.restype = C.c_double
        L.nmco_max_threads.restype = C.c_int
        _lib = L
    return _lib


def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().nmco_philox4x32_10(c, k, o)
    return list(o)


def _vec(fn, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    fp = C.POINTER(C.c_float)
    getattr(lib(), fn)(x.ctypes.data_as(fp), y.ctypes.data_as(fp), x.size)
    return y


def logf(x): return _vec("nmco_logf_n", x)
def expf(x): return _vec("nmco_expf_n", x)
def expm1f(x): return _vec("nmco_expm1f_n", x)


def sincos2pi(u):
    u = np.ascontiguousarray(u, np.float32)
    s = np.empty_like(u)
    c = np.empty_like(u)
    fp = C.POINTER(C.c_float)
    lib().nmco_sincos2pi_n(u.ctypes.data_as(fp), s.ctypes.data_as(fp), c.ctypes.data_as(fp), u.size)
    return s, c


def fresnel(n1, n2, ca1):
    ca2 = C.c_float()
    r = lib().nmco_fresnel(n1, n2, ca1, C.byref(ca2))
    return r, ca2.value


def run(w: Workload, n_photons: int | None = None, first: int = 0, threads: int = 0,
        fluence: bool = True) -> P.TallyArrays:
    """Transport photons [first, first+n) of a workload; returns host tallies."""
    n = w.n_photons if n_photons is None else n_photons
    t = P.TallyArrays(w.rp, w.ext, w.sim, fluence=fluence)
    L = lib()
    if w.sim == P.SIM_MCSUB:
        st = L.nmco_mcsub(C.byref(w.rp), C.byref(w.tp), C.byref(w.ext), first, n, threads,
                          C.byref(t.res))
    else:
        v = np.ascontiguousarray(w.v, np.uint8)
        st = L.nmco_mcxyz(C.byref(w.rp), C.byref(w.tp), v.ctypes.data_as(C.POINTER(C.c_uint8)),
                          C.byref(w.ext), first, n, threads, C.byref(t.res))
    if st != 0:
        raise RuntimeError(f"oracle returned status {st}")
    t.seconds = L.nmco_last_seconds()
    return t


def max_threads() -> int:
    return lib().nmco_max_threads()
