"""Launch autotuner for the transport kernels.

For an mcxyz workload it
  1. runs a short pilot on the GPU with NMC_FLAG_COUNT_VISITS, giving the number
     of voxel visits (= fluence flushes) per voxel;
  2. searches, with a 3-D summed-area table, the boxes of several aspect ratios
     that fit the LDS tile budget and cover the most visits;
  3. times short trial launches over (deposit scheme: FP64 tile, FP32 tile,
     FP32 tile + binned records x tile box x block size x refill batch)
     and keeps the fastest.
Results are cached in a JSON file keyed by a hash of the workload, so a
repeated configuration is tuned once.  For a batched mcsub launch
(`autotune_mcsub_batch`) the search is over block size x tally placement (LDS
or global replicas) x refill batch, each trial timed on the batch itself.
Layered handles fix their grid at prepare and use the library's defaults.

    from nmc_amd.autotune import autotune_mcxyz, autotune_mcsub_batch
    tuning = autotune_mcxyz(workload)           # -> params.Tuning
    tuning = autotune_mcsub_batch(rp, items, ext)
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import params as P
from .metal_mc import MetalMC

CACHE = os.environ.get("NMC_TUNING_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "nmc_amd", "tuning.json"))


def workload_key(w, extra: str = "") -> str:
    h = hashlib.sha1()
    h.update(bytes(w.rp))
    h.update(bytes(w.tp))
    if w.v is not None:
        h.update(np.ascontiguousarray(w.v).tobytes())
    h.update(bytes(w.ext))
    h.update(extra.encode())
    return h.hexdigest()


def _load_cache() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, value: dict) -> None:
    d = _load_cache()
    d[key] = value
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1)
    os.replace(tmp, CACHE)


def visit_histogram(w, n_pilot: int = 200_000, device: int = 0, block: int = 512) -> np.ndarray:
    """Voxel-visit counts [Nz, Nx, Ny] from a GPU pilot run (launched with its
    own block size, so a profile tells its dispatches from the transport's)."""
    ext = P.ExtConfig.from_buffer_copy(w.ext)
    ext.flags = (ext.flags | P.FLAG_COUNT_VISITS) & ~P.FLAG_NO_FLUENCE
    with MetalMC(w.sim, device) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, ext)
        mc.set_tuning(P.Tuning.make(block=block, tile=(0, 0, 0, -1, 0, 0)))
        mc.set_num_runs_seed(ext.seed, 1 << 41)
        mc.send_compute_command(n_pilot)
        t = mc.get_compute_results()
    return t.fluence.reshape(w.rp.Nz, w.rp.Nx, w.rp.Ny)


def best_boxes(hist: np.ndarray, budget: int, n_best: int = 3):
    """Boxes (x0, y0, z0, nx, ny, nz) with nx*ny*nz <= budget covering the most
    of `hist` (axes z, x, y), over a family of aspect ratios."""
    nz_, nx_, ny_ = hist.shape
    S = np.zeros((nz_ + 1, nx_ + 1, ny_ + 1))
    S[1:, 1:, 1:] = hist.astype(np.float64).cumsum(0).cumsum(1).cumsum(2)
    total = max(float(S[-1, -1, -1]), 1.0)
    cands = []
    for nz in sorted({max(1, int(round(budget ** (1 / 3) * f))) for f in (0.25, 0.4, 0.55, 0.7, 0.9, 1.2, 1.6)}):
        nz = min(nz, nz_)
        for ratio in (1.0, 1.3, 1.6, 0.75, 0.6):
            area = budget // nz
            nx = int(min(nx_, max(1, round((area / ratio) ** 0.5))))
            ny = int(min(ny_, max(1, area // nx)))
            if nx * ny * nz > budget:
                continue
            box = (S[nz:, nx:, ny:] - S[:-nz, nx:, ny:] - S[nz:, :-nx, ny:] - S[nz:, nx:, :-ny]
                   + S[:-nz, :-nx, ny:] + S[:-nz, nx:, :-ny] + S[nz:, :-nx, :-ny] - S[:-nz, :-nx, :-ny])
            k = int(np.argmax(box))
            z0, x0, y0 = np.unravel_index(k, box.shape)
            cands.append((float(box.flat[k]) / total, (int(x0), int(y0), int(z0), nx, ny, nz)))
    cands.sort(key=lambda c: -c[0])
    out, seen = [], set()
    for cov, b in cands:
        if b not in seen:
            seen.add(b)
            out.append((cov, b))
        if len(out) == n_best:
            break
    return out


def box_coverage(hist: np.ndarray, tile) -> float:
    """Fraction of the visits in `hist` [Nz, Nx, Ny] inside the tile box
    (x0, y0, z0, nx, ny, nz)."""
    x0, y0, z0, nx, ny, nz = (int(t) for t in tile)
    if nx <= 0 or ny <= 0 or nz <= 0:
        return 0.0
    total = float(hist.sum())
    inside = float(hist[z0:z0 + nz, x0:x0 + nx, y0:y0 + ny].sum())
    return inside / total if total > 0 else 0.0


def _time(mc: MetalMC, n: int, seed: int) -> float:
    mc.set_num_runs_seed(seed, 0)
    mc.send_compute_command(n)
    return mc.last_kernel_ms()


def autotune_mcxyz(w, n_trial: int = 20_000_000, device: int = 0, use_cache: bool = True,
                   verbose: bool = False) -> P.Tuning:
    key = workload_key(w, "mcxyz-v3")
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return P.Tuning.make(**hit["tuning"])
    hist = visit_histogram(w, device=device)
    with MetalMC(w.sim, device) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        trials = []
        # FP64 tile, FP32 tile (twice the voxels), FP32 tile + binned records
        for mode in (0, 1, 2):
            mc.set_tuning(P.Tuning.make(mode=mode))
            default = mc.get_tuning()
            budget = default.tile[3] * default.tile[4] * default.tile[5]
            boxes = [b for _, b in best_boxes(hist, budget)] if budget > 0 else []
            trials += [P.Tuning.make(block=1024, refill_min=r, tile=b, mode=mode)
                       for b in boxes + [tuple(default.tile)] for r in (8, 16, 32)]
            trials += [P.Tuning.make(block=512, refill_min=16, tile=b, mode=mode)
                       for b in boxes[:1]]
        best, best_ms, log = None, float("inf"), []
        _time(mc, n_trial // 4, w.ext.seed)   # warm-up
        for t in trials:
            try:
                mc.set_tuning(t)
            except Exception as e:   # e.g. a box that does not fit this block size
                log.append({"tuning": t.as_dict(), "error": str(e)})
                continue
            ms = _time(mc, n_trial, w.ext.seed)
            log.append({"tuning": mc.get_tuning().as_dict(), "ms": ms})
            if verbose:
                print(f"[autotune] {mc.get_tuning().as_dict()} {ms:.2f} ms", flush=True)
            if ms < best_ms:
                best, best_ms = mc.get_tuning(), ms
    if best is None:
        raise RuntimeError(f"no mcxyz launch shape fits this workload: {log}")
    res = {"tuning": best.as_dict(), "ms": best_ms, "n_trial": n_trial, "trials": log,
           "time": time.time()}
    res["tuning"].pop("blocks_per_cu", None)
    if use_cache:
        _store_cache(key, res)
    return P.Tuning.make(**res["tuning"])


def _item_bytes(items) -> bytes:
    arr = (P.WorkItem * len(items))()
    for i, it in enumerate(items):
        if isinstance(it, dict):
            for k, v in it.items():
                setattr(arr[i], k, v)
        else:
            arr[i] = it
    return bytes(arr)


def autotune_mcsub_batch(rp: P.RunParams, items, ext: P.ExtConfig | None = None, device: int = 0,
                         use_cache: bool = True, verbose: bool = False) -> P.Tuning:
    """Fastest (block size, tally placement, refill batch) for one batched
    mcsub launch of `items`; every trial transports the same photons (seed
    reset before each), so only the launch shape differs."""
    ext = ext if ext is not None else P.ExtConfig.default()
    h = hashlib.sha1()
    h.update(bytes(rp))
    h.update(bytes(ext))
    h.update(_item_bytes(items))
    h.update(b"mcsub-batch-v1")
    key = h.hexdigest()
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return P.Tuning.make(**hit["tuning"])
    trials = [P.Tuning.make(block=b, mode=m, refill_min=r)
              for b in (1024, 512, 256) for m in (0, 2) for r in (1, 4)]
    best, best_ms, log = None, float("inf"), []
    with MetalMC(P.SIM_MCSUB, device) as mc:
        mc.prepare_batch(rp, items, ext)
        mc.set_num_runs_seed(ext.seed, 0)
        mc.run_batch()                               # warm-up
        for t in trials:
            try:
                mc.set_tuning(t)
            except Exception as e:   # e.g. tallies that do not fit the block's LDS share
                log.append({"tuning": t.as_dict(), "error": str(e)})
                continue
            mc.set_num_runs_seed(ext.seed, 0)
            mc.run_batch()
            ms = mc.last_kernel_ms()
            eff = mc.get_tuning()
            log.append({"tuning": eff.as_dict(), "ms": ms})
            if verbose:
                print(f"[autotune] {eff.as_dict()} {ms:.2f} ms", flush=True)
            if ms < best_ms:
                best, best_ms = eff, ms
    if best is None:
        raise RuntimeError(f"no mcsub launch shape fits this batch: {log}")
    res = {"tuning": best.as_dict(), "ms": best_ms, "trials": log, "time": time.time()}
    res["tuning"].pop("blocks_per_cu", None)
    if use_cache:
        _store_cache(key, res)
    return P.Tuning.make(**res["tuning"])
