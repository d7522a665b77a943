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
    """Builds the oracle library (always recompiles)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


def _bind(L):
    f32p = C.POINTER(C.c_float)
    L.nmco_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
    for name in ("nmco_logf", "nmco_logf_tab", "nmco_expf", "nmco_expm1f"):
        getattr(L, name).argtypes = [C.c_float]
        getattr(L, name).restype = C.c_float
    for name in ("nmco_logf_n", "nmco_logf_tab_n", "nmco_expf_n", "nmco_expm1f_n"):
        getattr(L, name).argtypes = [f32p, f32p, C.c_int64]
    L.nmco_sincos2pi_n.argtypes = [f32p, f32p, f32p, C.c_int64]
    L.nmco_sincos2pi_bits_n.argtypes = [C.POINTER(C.c_uint32), f32p, f32p, C.c_int64]
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
    L.nmco_mcxyz_sens.argtypes = [C.POINTER(P.RunParams), C.POINTER(P.TissueParams),
                                  C.POINTER(C.c_uint8), C.POINTER(P.ExtConfig), C.c_uint64,
                                  C.c_uint64, C.c_int, C.POINTER(P.Results),
                                  C.POINTER(C.c_int32), C.c_int, C.c_void_p]
    L.nmco_mcsub.argtypes = [C.POINTER(P.RunParams), C.POINTER(P.TissueParams),
                             C.POINTER(P.ExtConfig), C.c_uint64, C.c_uint64, C.c_int,
                             C.POINTER(P.Results)]
    L.nmco_xyz_source.argtypes = [C.POINTER(P.RunParams), C.POINTER(P.TissueParams),
                                  C.POINTER(C.c_uint8), C.POINTER(P.ExtConfig), C.c_uint64,
                                  C.c_uint64, f32p]
    L.nmco_sub_source.argtypes = [C.POINTER(P.RunParams), C.POINTER(P.TissueParams),
                                  C.POINTER(P.ExtConfig), C.c_uint64, C.c_uint64, f32p]
    L.nmco_layered.argtypes = [C.POINTER(P.LayeredConfig), C.POINTER(P.Stack), C.c_int,
                               C.c_uint64, C.c_int, C.c_void_p, C.c_void_p,
                               C.POINTER(P.Totals), C.POINTER(C.c_uint64), C.c_uint64,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p]
    L.nmco_layered_backward.argtypes = [C.POINTER(P.LayeredConfig), C.POINTER(P.Stack), C.c_int,
                                        C.c_uint64, C.c_uint64] + [C.c_void_p] * 8
    L.nmco_last_seconds.restype = C.c_double
    L.nmco_max_threads.restype = C.c_int
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


MUTANTS = {1: "flat beam radius R*xi instead of R*sqrt(xi)",
           2: "remaining step keeps its physical length across a tissue/layer change",
           3: "mcsub drops its weight at the pre-hop position",
           4: "mcsub bins counted from 1",
           5: "absorbed sensitivity to mus without the within-segment score term",
           6: "voxel sensitivity to mua without the own-tissue W s e^-x term",
           7: "voxel sensitivity to mus without the own-tissue W s h(x) term"}
_mutants: dict = {}


def mutant(k: int, workdir: str):
    """A copy of the oracle built with planted mistake k (MUTANTS) into
    workdir, for tests that show a pin fails on it."""
    if k not in _mutants:
        out = os.path.join(workdir, f"libnmc_oracle_mutant{k}.so")
        subprocess.run(["make", "-s", "-B", "-C", HERE, "mutant", f"K={k}", f"OUT={out}"],
                       check=True)
        _mutants[k] = _bind(C.CDLL(out))
    return _mutants[k]


def xyz_source(w: Workload, n: int, first: int = 0, L=None) -> np.ndarray:
    """Launch states of an mcxyz workload's photons: (n, 8) x, y, z, ux, uy, uz,
    W, specular loss."""
    L = L or lib()
    out = np.zeros((n, 8), np.float32)
    v = np.ascontiguousarray(w.v, np.uint8)
    st = L.nmco_xyz_source(C.byref(w.rp), C.byref(w.tp), v.ctypes.data_as(C.POINTER(C.c_uint8)),
                           C.byref(w.ext), first, n, out.ctypes.data_as(C.POINTER(C.c_float)))
    if st != 0:
        raise RuntimeError(f"oracle returned status {st}")
    return out


def sub_source(w: Workload, n: int, first: int = 0, L=None) -> np.ndarray:
    """Launch states of an mcsub workload's photons: (n, 7) x, y, z, ux, uy, uz,
    specular reflectance."""
    L = L or lib()
    out = np.zeros((n, 7), np.float32)
    st = L.nmco_sub_source(C.byref(w.rp), C.byref(w.tp), C.byref(w.ext), first, n,
                           out.ctypes.data_as(C.POINTER(C.c_float)))
    if st != 0:
        raise RuntimeError(f"oracle returned status {st}")
    return out


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
def logf_tab(x): return _vec("nmco_logf_tab_n", x)
def expf(x): return _vec("nmco_expf_n", x)
def expm1f(x): return _vec("nmco_expm1f_n", x)


def sincos2pi(u):
    u = np.ascontiguousarray(u, np.float32)
    s = np.empty_like(u)
    c = np.empty_like(u)
    fp = C.POINTER(C.c_float)
    lib().nmco_sincos2pi_n(u.ctypes.data_as(fp), s.ctypes.data_as(fp), c.ctypes.data_as(fp), u.size)
    return s, c


def sincos2pi_bits(x):
    x = np.ascontiguousarray(x, np.uint32)
    s = np.empty(x.shape, np.float32)
    c = np.empty(x.shape, np.float32)
    fp = C.POINTER(C.c_float)
    lib().nmco_sincos2pi_bits_n(x.ctypes.data_as(C.POINTER(C.c_uint32)), s.ctypes.data_as(fp),
                                c.ctypes.data_as(fp), x.size)
    return s, c


def fresnel(n1, n2, ca1):
    ca2 = C.c_float()
    r = lib().nmco_fresnel(n1, n2, ca1, C.byref(ca2))
    return r, ca2.value


def run(w: Workload, n_photons: int | None = None, first: int = 0, threads: int = 0,
        fluence: bool = True, L=None) -> P.TallyArrays:
    """Transport photons [first, first+n) of a workload; returns host tallies
    (L: another build of the oracle, e.g. a mutant)."""
    n = w.n_photons if n_photons is None else n_photons
    t = P.TallyArrays(w.rp, w.ext, w.sim, fluence=fluence)
    L = L or lib()
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


def run_sens(w: Workload, ids, n_photons: int | None = None, first: int = 0, threads: int = 0,
             L=None):
    """mcxyz run with per-voxel sensitivities to mua and mus of tissues `ids`:
    (tallies, d_mua [len(ids), Nz, Nx, Ny], d_mus [same])."""
    n = w.n_photons if n_photons is None else n_photons
    t = P.TallyArrays(w.rp, w.ext, w.sim, fluence=True)
    L = L or lib()
    ids_a = np.ascontiguousarray(ids, np.int32)
    shape = (w.rp.Nz, w.rp.Nx, w.rp.Ny)
    sens = np.zeros((2, len(ids_a)) + shape)
    v = np.ascontiguousarray(w.v, np.uint8)
    st = L.nmco_mcxyz_sens(C.byref(w.rp), C.byref(w.tp), v.ctypes.data_as(C.POINTER(C.c_uint8)),
                           C.byref(w.ext), first, n, threads, C.byref(t.res),
                           ids_a.ctypes.data_as(C.POINTER(C.c_int32)), len(ids_a),
                           sens.ctypes.data)
    if st != 0:
        raise RuntimeError(f"oracle returned status {st}")
    t.seconds = L.nmco_last_seconds()
    return t, sens[0], sens[1]


def max_threads() -> int:
    return lib().nmco_max_threads()


class LayeredOut:
    """Host results of a layered oracle run."""


def run_layered(cfg: P.LayeredConfig, stacks, first: int = 0, threads: int = 0,
                records: bool = True, L=None, sens: bool = False) -> LayeredOut:
    """Oracle run of an array of Stack structs; records (if asked) in photon
    order; sens: the absorbed-weight sensitivities o.sens_mua, o.sens_mus
    [n_stacks, MAXL (absorbing layer), MAXL (parameter layer)]."""
    L = L or lib()
    ns = len(stacks)
    arr = (P.Stack * ns)(*stacks)
    o = LayeredOut()
    o.stacks = arr
    o.cfg = cfg
    o.rd = np.zeros((ns, cfg.nr + 1))
    o.absorbed = np.zeros((ns, P.MAX_LAYERS))
    tots = (P.Totals * ns)()
    nrec = C.c_uint64(0)
    cap = int(sum(s.n_photons for s in stacks)) if records else 0
    o.cap = cap
    if records:
        o.rec_item = np.zeros(cap, np.uint32)
        o.rec_bin = np.zeros(cap, np.uint32)
        o.rec_w = np.zeros(cap, np.float32)
        o.rec_L = np.zeros((P.MAX_LAYERS, cap), np.float32)
        o.rec_K = np.zeros((P.MAX_LAYERS, cap), np.uint32)
        ptrs = [o.rec_item.ctypes.data, o.rec_bin.ctypes.data, o.rec_w.ctypes.data,
                o.rec_L.ctypes.data, o.rec_K.ctypes.data]
    else:
        ptrs = [None] * 5
    sens_arr = np.zeros((ns, 2, P.MAX_LAYERS, P.MAX_LAYERS)) if sens else None
    st = L.nmco_layered(C.byref(cfg), arr, ns, first, threads, o.rd.ctypes.data,
                        o.absorbed.ctypes.data, tots, C.byref(nrec), cap, *ptrs,
                        sens_arr.ctypes.data if sens else None)
    if sens:
        o.sens_mua, o.sens_mus = sens_arr[:, 0], sens_arr[:, 1]
    if st != 0:
        raise RuntimeError(f"oracle returned status {st}")
    o.totals = [t.as_dict() for t in tots]
    o.n_records = nrec.value
    o.seconds = L.nmco_last_seconds()
    return o


def layered_backward(o: LayeredOut, grad_rd: np.ndarray):
    """The perturbation-MC gradient over the oracle's records."""
    L = lib()
    ns = len(o.stacks)
    g = np.ascontiguousarray(grad_rd, np.float64).reshape(ns, o.cfg.nr + 1)
    gmua = np.zeros((ns, P.MAX_LAYERS))
    gmus = np.zeros((ns, P.MAX_LAYERS))
    L.nmco_layered_backward(C.byref(o.cfg), o.stacks, ns, o.n_records, o.cap,
                            o.rec_item.ctypes.data, o.rec_bin.ctypes.data, o.rec_w.ctypes.data,
                            o.rec_L.ctypes.data, o.rec_K.ctypes.data, g.ctypes.data,
                            gmua.ctypes.data, gmus.ctypes.data)
    return gmua, gmus
