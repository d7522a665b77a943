"""Layered media runner (forward transport + perturbation-MC backward).

`LayeredMC` drives an NMC_SIM_LAYERED handle of the C ABI: a set of layer
stacks (each with its own optical properties, Philox stream and photon
budget) is transported by one persistent launch; the per-photon path records
kept in HBM then give dLoss/dmua and dLoss/dmus per layer for any loss of the
Rd(r) profiles (nmc_layered_backward).  There is no counterpart in the
reference (NMC/NMCCode has a single global n1/n2 and no gradients).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import params as P
from ._lib import check, lib


class LayeredMC:
    def __init__(self, cfg: P.LayeredConfig, stacks, device: int = 0,
                 tuning: P.Tuning | None = None):
        self._h = C.c_void_p()
        check(lib().nmc_create(P.SIM_LAYERED, int(device), C.byref(self._h)))
        if tuning is not None:
            check(lib().nmc_set_tuning(self._h, C.byref(tuning)))
        self.cfg = cfg
        # bumped by every change of properties and every run: a backward pass
        # is valid only for the records of the run it follows
        self.epoch = 0
        self._stacks = (P.Stack * len(stacks))(*stacks)
        check(lib().nmc_layered_prepare(self._h, C.byref(cfg), self._stacks, len(stacks)))

    @property
    def n_stacks(self) -> int:
        return len(self._stacks)

    def stacks(self):
        return list(self._stacks)

    def set_properties(self, mua: np.ndarray, mus: np.ndarray) -> None:
        """New mua/mus per (stack, layer) for the same stack shapes."""
        mua = np.asarray(mua, np.float64)
        mus = np.asarray(mus, np.float64)
        for j, s in enumerate(self._stacks):
            for l in range(s.n_layers):
                s.layer[l].mua = float(mua[j, l])
                s.layer[l].mus = float(mus[j, l])
        self.epoch += 1
        check(lib().nmc_layered_update(self._h, self._stacks, len(self._stacks)))

    def set_seed(self, seed: int, first_photon: int = 0) -> None:
        check(lib().nmc_set_seed(self._h, int(seed), int(first_photon)))

    def run(self) -> None:
        self.epoch += 1
        check(lib().nmc_layered_run(self._h))

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(lib().nmc_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def results(self):
        """(rd raw weight [n, nr+1], absorbed [n, MAX_LAYERS], totals list, n_records)."""
        n = self.n_stacks
        rd = np.zeros((n, self.cfg.nr + 1))
        ab = np.zeros((n, P.MAX_LAYERS))
        tots = (P.Totals * n)()
        nrec = C.c_uint64()
        check(lib().nmc_layered_results(self._h, rd.ctypes.data, ab.ctypes.data, tots, C.byref(nrec)))
        return rd, ab, [t.as_dict() for t in tots], nrec.value

    def records(self) -> dict:
        """The last run's path records (unspecified order): stack, bin, w,
        L [n, rec_nl], K [n, rec_nl]."""
        nrec = self.results()[3]
        n_out, nl = C.c_uint64(), C.c_int()
        check(lib().nmc_layered_records(self._h, None, 0, C.byref(n_out), C.byref(nl)))
        rs = 3 + 2 * nl.value
        buf = np.zeros((max(nrec, 1), rs), np.uint32)
        check(lib().nmc_layered_records(self._h, buf.ctypes.data, nrec, C.byref(n_out),
                                        C.byref(nl)))
        buf = buf[:n_out.value]
        k = nl.value
        return {"stack": buf[:, 0], "bin": buf[:, 1], "w": buf[:, 2].view(np.float32),
                "L": buf[:, 3:3 + k].view(np.float32), "K": buf[:, 3 + k:3 + 2 * k]}

    def absorbed_grad(self):
        """Sensitivities of the raw weight absorbed per layer (config
        absorbed_grad = 1): (d/dmua, d/dmus) [n_stacks, MAX_LAYERS (absorbing
        layer l), MAX_LAYERS (parameter layer t)]."""
        n = self.n_stacks
        da = np.zeros((n, P.MAX_LAYERS, P.MAX_LAYERS))
        ds = np.zeros((n, P.MAX_LAYERS, P.MAX_LAYERS))
        check(lib().nmc_layered_absorbed_grad(self._h, da.ctypes.data, ds.ctypes.data))
        return da, ds

    def get_tuning(self) -> P.Tuning:
        """Effective launch values; `mode` 1 = Rd(r) in LDS, 2 = global replicas."""
        t = P.Tuning()
        check(lib().nmc_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, tuning: P.Tuning) -> None:
        """Refill batch and queue chunk for later runs (the grid is fixed at
        prepare; pass the whole tuning to the constructor to change it)."""
        check(lib().nmc_set_tuning(self._h, C.byref(tuning)))

    def rd_per_photon(self) -> np.ndarray:
        """Rd(r) per launched photon, [n_stacks, nr+1] (the backward's variable)."""
        rd = self.results()[0]
        budgets = np.array([s.n_photons for s in self._stacks], np.float64)
        return rd / np.maximum(budgets, 1.0)[:, None]

    def backward(self, grad_rd: np.ndarray):
        """grad_rd = dLoss/d(rd_per_photon) -> (dLoss/dmua, dLoss/dmus) [n, MAX_LAYERS]."""
        g = np.ascontiguousarray(grad_rd, np.float64).reshape(self.n_stacks, self.cfg.nr + 1)
        gmua = np.zeros((self.n_stacks, P.MAX_LAYERS))
        gmus = np.zeros((self.n_stacks, P.MAX_LAYERS))
        check(lib().nmc_layered_backward(self._h, g.ctypes.data, gmua.ctypes.data, gmus.ctypes.data))
        return gmua, gmus

    def close(self) -> None:
        if self._h:
            lib().nmc_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
