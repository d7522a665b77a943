"""Python host interface mirroring the reference's MetalMC object.

The reference driver (NMC/NMCCode/main.m:151-191) creates a MetalMC for one
simulation kind, hands it the parameter structs and voxel map, runs batches
and collects the tallies.  `MetalMC` below keeps those method names
(snake_cased) and their meaning, on top of the C ABI in include/nmc.h:

    initWithDevice::     -> MetalMC(kernel_type, device)      MetalMC.h:25
    prepareData:::       -> prepare_data(rp, tp, v, ext)      MetalMC.h:26
    setNumRunsSeed::     -> set_num_runs_seed(seed, first)    MetalMC.h:30
    sendComputeCommand   -> send_compute_command(n)           MetalMC.h:27
    getComputeResults:.. -> get_compute_results()             MetalMC.h:28

Errors raise `NmcError` (the reference logs and continues or exits).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import params as P
from ._lib import NmcError, check, lib


class MetalMC:
    def __init__(self, kernel_type: int, device: int = 0):
        self._h = C.c_void_p()
        check(lib().nmc_create(int(kernel_type), int(device), C.byref(self._h)))
        self.kernel_type = int(kernel_type)
        self.rp = self.tp = self.ext = None
        self._sens_ids = np.zeros(0, np.int32)

    # -- configuration -------------------------------------------------
    def prepare_data(self, run_params: P.RunParams, tiss_params: P.TissueParams,
                     v: np.ndarray | None = None, ext: P.ExtConfig | None = None) -> None:
        self.rp, self.tp = run_params, tiss_params
        self.ext = ext if ext is not None else P.ExtConfig.default()
        self._sens_ids = np.zeros(0, np.int32)   # a new preparation selects none
        vp = C.POINTER(C.c_uint8)()
        if v is not None:
            self._v = np.ascontiguousarray(v, np.uint8)
            vp = self._v.ctypes.data_as(C.POINTER(C.c_uint8))
        check(lib().nmc_prepare(self._h, C.byref(run_params), C.byref(tiss_params), vp,
                                C.byref(self.ext)))

    def set_num_runs_seed(self, seed: int, first_photon: int = 0) -> None:
        check(lib().nmc_set_seed(self._h, int(seed), int(first_photon)))

    # -- execution ----------------------------------------------------
    def send_compute_command(self, n_photons: int) -> None:
        """Transports the next n_photons photons and waits for completion."""
        check(lib().nmc_run(self._h, int(n_photons)))

    def send_compute_command_async(self, n_photons: int) -> None:
        check(lib().nmc_run_async(self._h, int(n_photons)))

    def synchronize(self) -> None:
        check(lib().nmc_synchronize(self._h))

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(lib().nmc_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def last_transport_ms(self) -> float:
        """The last transport-kernel launch alone (mcxyz, eager runs)."""
        ms = C.c_float()
        check(lib().nmc_last_transport_ms(self._h, C.byref(ms)))
        return ms.value

    def last_launch_count(self) -> int:
        """Transport-kernel launches of the last run."""
        n = C.c_uint32()
        check(lib().nmc_last_launch_count(self._h, C.byref(n)))
        return n.value

    # -- results --------------------------------------------------------
    def get_compute_results(self, fluence: bool = True) -> P.TallyArrays:
        if self.rp is None:
            raise NmcError(P.ERR_STATE, "prepare_data has not been called")
        no_fl = bool(self.ext.flags & P.FLAG_NO_FLUENCE)
        t = P.TallyArrays(self.rp, self.ext, self.kernel_type, fluence=fluence and not no_fl)
        check(lib().nmc_get_results(self._h, C.byref(t.res)))
        return t

    def add_compute_results(self, t: P.TallyArrays) -> None:
        """getComputeResults' semantics: adds the accumulated tallies into the
        arrays and totals of t (e.g. summing several handles)."""
        check(lib().nmc_add_results(self._h, C.byref(t.res)))

    def reset_tallies(self) -> None:
        check(lib().nmc_reset_tallies(self._h))

    # -- per-tissue fluence sensitivities (mcxyz) ----------------------------
    def set_sensitivity(self, tissue_ids) -> None:
        """Later runs also accumulate d(absorbed weight per voxel)/d(mua, mus)
        of these tissue ids (at most NMC_SENS_MAX; empty: off)."""
        ids = np.ascontiguousarray(list(tissue_ids), np.int32)
        self._sens_ids = ids
        check(lib().nmc_mcxyz_sens(self._h, ids.ctypes.data_as(C.POINTER(C.c_int32)), len(ids)))

    def sensitivity_results(self):
        """(d_mua, d_mus), each [n_tissues, Nz, Nx, Ny]: derivatives of the raw
        fluence tally (get_compute_results().fluence) since the selection or
        the last reset_tallies."""
        shape = (len(self._sens_ids), self.rp.Nz, self.rp.Nx, self.rp.Ny)
        da = np.zeros(shape)
        ds = np.zeros(shape)
        check(lib().nmc_mcxyz_sens_results(self._h, da.ctypes.data, ds.ctypes.data))
        return da, ds

    # -- launch tuning / graphs ---------------------------------------------
    def set_tuning(self, tuning: P.Tuning) -> None:
        check(lib().nmc_set_tuning(self._h, C.byref(tuning)))

    def get_tuning(self) -> P.Tuning:
        t = P.Tuning()
        check(lib().nmc_get_tuning(self._h, C.byref(t)))
        return t

    def set_graph_mode(self, enable: bool) -> None:
        check(lib().nmc_set_graph_mode(self._h, int(bool(enable))))

    # -- batched mcsub (many media, one launch) -------------------------------
    def prepare_batch(self, run_params: P.RunParams, items, ext: P.ExtConfig | None = None) -> None:
        """items: sequence of WorkItem (or dicts with its fields)."""
        self.rp = run_params
        self.ext = ext if ext is not None else P.ExtConfig.default()
        arr = (P.WorkItem * len(items))()
        for i, it in enumerate(items):
            if isinstance(it, dict):
                for k, v in it.items():
                    setattr(arr[i], k, v)
            else:
                C.memmove(C.byref(arr[i]), C.byref(it), C.sizeof(P.WorkItem))
        self._items = arr
        check(lib().nmc_prepare_batch(self._h, C.byref(run_params), C.byref(self.ext), arr, len(items)))

    def run_batch(self) -> None:
        check(lib().nmc_run_batch(self._h))

    def run_batch_async(self) -> None:
        check(lib().nmc_run_batch_async(self._h))

    def get_batch_results(self):
        """(tally[n_items, per_item] float64, list of totals dicts)."""
        n = len(self._items)
        per = (self.ext.nr + 1) * self.ext.na
        if not self.ext.flags & P.FLAG_NO_FLUENCE:
            per += (self.ext.nr + 1) * (self.ext.nz + 1)
        tally = np.zeros((n, per), np.float64)
        tots = (P.Totals * n)()
        check(lib().nmc_get_batch_results(self._h, tally.ctypes.data_as(C.POINTER(C.c_double)), tots))
        return tally, [t.as_dict() for t in tots]

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


def validate(sim: int, rp: P.RunParams, tp: P.TissueParams, v: np.ndarray | None = None,
             ext: P.ExtConfig | None = None) -> None:
    """Runs the library's configuration checks (no device needed)."""
    vp = C.POINTER(C.c_uint8)()
    if v is not None:
        v = np.ascontiguousarray(v, np.uint8)
        vp = v.ctypes.data_as(C.POINTER(C.c_uint8))
    check(lib().nmc_validate(int(sim), C.byref(rp), C.byref(tp), vp,
                             C.byref(ext) if ext is not None else C.POINTER(P.ExtConfig)()))


def simulate(w, n_photons: int | None = None, seed: int | None = None, first_photon: int = 0,
             device: int = 0, fluence: bool = True) -> P.TallyArrays:
    """One-shot run of a workload; returns the host tallies (with .kernel_ms)."""
    ext = P.ExtConfig.default()
    C.memmove(C.byref(ext), C.byref(w.ext), C.sizeof(ext))
    if not fluence:
        ext.flags |= P.FLAG_NO_FLUENCE
    with MetalMC(w.sim, device) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, ext)
        mc.set_num_runs_seed(ext.seed if seed is None else seed, first_photon)
        mc.send_compute_command(w.n_photons if n_photons is None else n_photons)
        t = mc.get_compute_results(fluence=fluence)
        t.kernel_ms = mc.last_kernel_ms()
    return t
