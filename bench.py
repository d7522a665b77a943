#!/usr/bin/env python3
"""Benchmark: photons/s of the mcxyz transport kernel on the 200^3 skin-vessel
workload (1e8 photons per step), on one MI355X.

A step is one nmc_run over 1e8 photons with the voxel map and tallies already
resident in HBM.  The kernel time of every step is taken from HIP events
recorded by the library on the stream the kernel runs on.  Prints one JSON
line (see the keys below); the CPU baseline is the oracle run on this host's
cores over a bounded sample of the same workload.

The timed steps use the FP32 LDS fluence tile on the box the visit-count
pilot finds, with the voxel visits outside it written as binned deposit
records and summed per bucket in FP64 (tuning.mode 2; sums within a
workgroup's tile in FP32, everything else in FP64; the reference adds every
deposit into an FP32 global array).  The records of one step are summed by
the next step's workgroups once their own photons are done (the fused
reduction), those of the last step by the synchronize that closes the timed
region; the `tile` leg also times the records with their own reduction
kernel after each step.  `--tile fp32`
times the FP32 tile with FP64 atomics outside it, `--tile fp64` the all-FP64
tile; the `tile` leg reports the two modes not timed.  Every timed step (and
the warm-up) launches mcxyz_kernel with 1024-lane blocks; the other launches
of this run (visit-count pilots, Rd parity run) use 512-lane blocks, so a
kernel-trace profile of this command reports the timed kernel under its own
name.  The roofline divides by the transport kernel's own duration (HIP
events around that kernel on its stream); `value` counts the whole run,
reductions included.

    python bench.py                      # the measurement
    python bench.py --calibrate-copy 256 # a 256 MiB copy, to scale PMC byte counters
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s on 1 MI355X (mcxyz 200^3) at <=0.5% Rd err; grad cost vs forward"

# Lane-op cost of one event of the mcxyz loop (fixed table: a same-voxel piece
# 44, a voxel-crossing piece 97, an interaction 145; transcendental = 4).
OPS_SAME, OPS_CROSS, OPS_INT = 44, 97, 145
# gfx950: 256 CUs x 64 FP32 lanes x 2.4 GHz, non-packed lane-ops/s (T).
VALU_PEAK_T = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0
# HBM bytes per photon of the timed kernel on this workload, by tuning.mode
# (0 FP64 tile, 1 FP32 tile, 2 FP32 tile + binned records), from the
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes over one 1e8-photon launch,
# scaled by the copy calibration (FETCH x2.0, WRITE x1.0): FP64
# (18 955 068 KiB x 2 + 98 872 849 KiB) x 1024 B / 1e8, FP32 (18 945 480 KiB
# x 2 + 92 328 830 KiB) x 1024 B / 1e8, records (50 714 041 KiB x 2 +
# 62 456 345 KiB) x 1024 B / 1e8 for a launch that also reduces the previous
# launch's records (the timed steps after the first), each on the visit-pilot
# box (profiles/mcxyz_skinvessel_200_pmc_traffic.md).
TRAFFIC_BYTES_PER_PHOTON = {0: 1400.7, 1: 1333.5, 2: 1678.2}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def probe_lib():
    from nmc_amd.build import PROBE_LIB
    if not os.path.exists(PROBE_LIB):
        raise RuntimeError(f"{PROBE_LIB} is not built (run __graft_entry__.build())")
    L = C.CDLL(PROBE_LIB)
    L.nmcp_atomic_rate.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_float),
                                   C.POINTER(C.c_double)]
    L.nmcp_copy.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_float)]
    return L


def atomic_ceiling(slots: int) -> dict:
    """FP64 atomic adds/s to scattered addresses over a fluence-sized array."""
    L = probe_lib()
    ms, n = C.c_float(), C.c_double()
    rc = L.nmcp_atomic_rate(slots, 64, C.byref(ms), C.byref(n))
    if rc != 0:
        raise RuntimeError(f"nmcp_atomic_rate failed ({rc})")
    return {"per_s": n.value / (ms.value * 1e-3), "slots": slots, "atomics": n.value,
            "ms": ms.value}


def calibrate_copy(mib: int) -> int:
    L = probe_lib()
    ms = C.c_float()
    nbytes = mib << 20
    rc = L.nmcp_copy(nbytes, 3, C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"nmcp_copy failed ({rc})")
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": 3,
                      "ms": ms.value, "GB_per_s": 2 * nbytes / (ms.value * 1e-3) / 1e9}),
          flush=True)
    return 0


def cpu_baseline(w, target_s: float) -> dict:
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    # the oracle's per-thread tallies cost a fixed set-up and merge time, so a
    # small probe underestimates the rate: size the sample from the probe and
    # enlarge it once if it ran short
    n = 200_000
    t = O.run(w, n, threads=cores, fluence=True)
    for _ in range(2):
        if t.seconds >= 0.5 * target_s:
            break
        n = int(min(2e8, n * target_s / max(t.seconds, 1e-3)))
        log(f"[bench] cpu baseline: oracle on {cores} threads, {n} photons")
        t = O.run(w, n, threads=cores, fluence=True)
    n1 = max(1000, int(n / cores / 4))
    log(f"[bench] cpu baseline: oracle on 1 thread, {n1} photons")
    t1 = O.run(w, n1, threads=1, fluence=True)
    return {"value": n / t.seconds, "unit": "photons/s", "cores": cores, "kind": "port",
            "sample": f"photons [0,{n}) of the same workload in {t.seconds:.1f} s "
                      "(oracle, OpenMP, FP64 tallies)",
            "single_thread": {"value": n1 / t1.seconds, "photons": n1},
            "_tallies": t, "_n": n}


TILE_MODES = {"rec": 2, "fp32": 1, "fp64": 0}
MODE_NAMES = {v: k for k, v in TILE_MODES.items()}
# fluence-variant template argument of the timed kernel per tuning.mode
MODE_FMODE = {0: 1, 1: 1, 2: 3}


def kernel_name(mode: int, block: int = 1024) -> str:
    elem = "float" if mode else "double"
    return f"mcxyz_kernel<false, {block}, {elem}, {MODE_FMODE[mode]}>"


def pilot_tuning(mc, hist, mode: int, rec_fuse: int = 0):
    """Tuning with tile mode `mode` (0 FP64, 1 FP32, 2 FP32 + records) and the
    box of that mode's LDS budget holding the most pilot visits."""
    from nmc_amd import params as P
    from nmc_amd.autotune import best_boxes

    mc.set_tuning(P.Tuning.make(mode=mode))
    budget = int(np.prod(mc.get_tuning().tile[3:6]))
    box = best_boxes(hist, budget, 1)[0][1]
    return P.Tuning.make(mode=mode, tile=box, rec_fuse=rec_fuse)


def tile_leg(mc, hist, n: int, mode: int, steps: int = 2, rec_fuse: int = 0) -> dict:
    """The same workload with another tile mode (0 FP64 tile, 1 FP32 tile,
    2 FP32 tile + records; rec_fuse -1: records reduced by their own kernel
    after each launch), box from the visit pilot.  Blocking runs: each step
    ends with its own reductions."""
    from nmc_amd.autotune import box_coverage

    t = pilot_tuning(mc, hist, mode, rec_fuse)
    mc.set_tuning(t)
    box = tuple(t.tile)
    log(f"[bench] {MODE_NAMES[mode]} tile {box} rec_fuse {rec_fuse}: {steps} x {n} photons")
    mc.reset_tallies()
    mc.set_num_runs_seed(1, 0)
    ms, kms, wall = [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        mc.send_compute_command(n)
        wall.append(time.perf_counter() - t0)
        ms.append(mc.last_kernel_ms())
        kms.append(mc.last_transport_ms())
    avg = sum(ms) / len(ms)
    out = {"kernel": kernel_name(mode), "tile": list(box),
           "lds_tile_fraction": box_coverage(hist, box), "run_ms": avg,
           "kernel_ms": sum(kms) / len(kms), "photons_per_s": n / (avg * 1e-3),
           "fluence_parity": ("integer tallies bit-exact; fluence within 2e-5 relative of the "
                              "oracle (tests/test_gpu_parity.py)") if mode else
                             "every tally within 1e-9 relative of the oracle"}
    if mode == 2:
        out["wall_ms"] = 1e3 * sum(wall) / len(wall)
        out["photons_per_s"] = n / (out["wall_ms"] * 1e-3)
        out["reduction_ms"] = avg - out["kernel_ms"]
    return out

def sweep_leg(budget_scale: float, reps: int = 3) -> dict:
    """128-wavelength sweep: one batched launch vs 128 plain launches (each
    with its own configure, run and read-back), same photons.  The pair is
    measured `reps` times, alternating, and the median ratio reported (the
    GPU's clock state drifts by a few per cent over a run)."""
    from nmc_amd import params as P
    from nmc_amd import workloads as W
    from nmc_amd.metal_mc import MetalMC

    w, items, _ = W.spectral_sweep(budget_scale=budget_scale)
    total = sum(i["n_photons"] for i in items)
    log(f"[bench] sweep: 128 wavelengths, {total} photons, {reps} x (batched, plain)")
    pairs = []
    same = True
    for _ in range(reps):
        with MetalMC(P.SIM_MCSUB) as mc:
            mc.prepare_batch(w.rp, items, w.ext)
            mc.synchronize()
            t0 = time.perf_counter()
            mc.run_batch()
            tally, tots = mc.get_batch_results()
            t_batch = time.perf_counter() - t0
            k_batch = mc.last_kernel_ms()
        t0 = time.perf_counter()
        k_plain = 0.0
        with MetalMC(P.SIM_MCSUB) as mc:
            for j, it in enumerate(items):
                w.rp.n1, w.rp.n2 = it["n1"], it["n2"]
                w.tp.muav[1], w.tp.musv[1], w.tp.gv[1] = it["mua"], it["mus"], it["g"]
                w.ext.stream_id = it["stream_id"]
                mc.prepare_data(w.rp, w.tp, None, w.ext)
                mc.send_compute_command(it["n_photons"])
                res = mc.get_compute_results()
                k_plain += mc.last_kernel_ms()
                same &= all(res.totals[k] == tots[j][k] for k in P.TOTAL_INT_FIELDS)
                same &= bool(np.allclose(res.rd_ra, tally[j, :w.ext.nr + 1], rtol=1e-9,
                                         atol=1e-12))
        t_plain = time.perf_counter() - t0
        pairs.append((t_plain / t_batch, t_batch * 1e3, k_batch, t_plain * 1e3, k_plain))
    pairs.sort()
    sp, t_batch_ms, k_batch, t_plain_ms, k_plain = pairs[len(pairs) // 2]
    return {"photons": total, "batched_ms": t_batch_ms, "batched_kernel_ms": k_batch,
            "plain_128_ms": t_plain_ms, "plain_kernel_ms_sum": k_plain,
            "speedup": sp, "speedups": [p[0] for p in pairs], "reps": reps,
            "tallies_identical": bool(same), "batched_photons_per_s": total / (t_batch_ms * 1e-3)}


def sweep_cpu_baseline(budget_scale: float, target_s: float) -> dict:
    """The oracle over the 128-wavelength sweep's items on this host's cores,
    each item's budget scaled so the whole sample takes about target_s."""
    from oracle import oracle as O
    from nmc_amd import workloads as W

    cores = len(os.sched_getaffinity(0))
    w, items, _ = W.spectral_sweep(budget_scale=budget_scale)

    def run(scale):
        photons, events, secs = 0, 0, 0.0
        for it in items:
            w.rp.n1, w.rp.n2 = it["n1"], it["n2"]
            w.tp.muav[1], w.tp.musv[1], w.tp.gv[1] = it["mua"], it["mus"], it["g"]
            w.ext.stream_id = it["stream_id"]
            m = max(1, int(it["n_photons"] * scale))
            t = O.run(w, m, threads=cores, fluence=False)
            photons += m
            events += t.totals["substeps_same"] + t.totals["substeps_cross"]
            secs += t.seconds
        return photons, events, secs

    scale = 1e-3
    p, e, sec = run(scale)
    scale *= min(50.0, target_s / max(sec, 1e-3))
    log(f"[bench] sweep cpu baseline: oracle on {cores} threads, item budgets x {scale:.2e}")
    p, e, sec = run(scale)
    return {"value": p / sec, "unit": "photons/s", "events_per_s": e / sec, "cores": cores,
            "kind": "port", "sample": f"all 128 items at {scale:.2e} of their budgets: {p} photons, "
                                      f"{e} events in {sec:.1f} s (oracle, OpenMP, FP64 tallies)"}


def gradient_leg(n_photons: int) -> dict:
    """5-layer differentiable case: (forward with path records + backward)
    against the plain forward, kernel times on the library stream."""
    import numpy as np

    from nmc_amd import workloads as W
    from nmc_amd.layered import LayeredMC

    cfg, stacks = W.multilayer_5(n_photons=n_photons)
    log(f"[bench] gradient: 5-layer stack, {n_photons} photons")
    cfg.records = 0
    with LayeredMC(cfg, stacks) as mc:
        mc.run()
        mc.run()
        fwd = mc.last_kernel_ms()
    cfg.records = 1
    with LayeredMC(cfg, stacks) as mc:
        mc.run()
        mc.run()
        fwd_rec = mc.last_kernel_ms()
        g = np.ones((1, cfg.nr + 1))
        mc.backward(g)
        bwd = mc.last_kernel_ms()
        nrec = mc.results()[3]
    return {"photons": n_photons, "forward_ms": fwd, "forward_with_records_ms": fwd_rec,
            "backward_ms": bwd, "records": nrec, "grad_over_forward": (fwd_rec + bwd) / fwd,
            "forward_photons_per_s": n_photons / fwd * 1e3,
            "grad_photons_per_s": n_photons / (fwd_rec + bwd) * 1e3}


def voxel_sens_leg(n_photons: int) -> dict:
    """Per-tissue fluence sensitivities on the skin-vessel grid: a run with
    d(absorbed weight per voxel)/d(mua, mus) of all 4 tissues against the
    same photons' plain run (FP64 tile, the variant whose deposits the
    sensitivity kernel shares), transport kernel times on the library
    stream."""
    from nmc_amd import params as P
    from nmc_amd import workloads as W
    from nmc_amd.metal_mc import MetalMC

    w = W.skin_vessel(n_photons)
    log(f"[bench] voxel sensitivities: skin-vessel, 4 tissues, {n_photons} photons")
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=0))
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(n_photons)
        mc.send_compute_command(n_photons)
        plain = mc.last_transport_ms()
        mc.set_sensitivity([1, 2, 3, 4])
        mc.send_compute_command(n_photons)
        mc.send_compute_command(n_photons)
        sens = mc.last_transport_ms()
    return {"photons": n_photons, "tissues": 4, "plain_fp64_tile_ms": plain, "sens_ms": sens,
            "sens_over_plain": sens / plain, "sens_photons_per_s": n_photons / sens * 1e3}


def gradient_cpu_baseline(target_s: float) -> dict:
    """The oracle's forward with path records plus its pMC backward on the
    5-layer stack, on this host's cores, over a sample of about target_s."""
    from oracle import oracle as O
    from nmc_amd import workloads as W

    cores = len(os.sched_getaffinity(0))

    def run(n):
        cfg, stacks = W.multilayer_5(n_photons=n)
        t0 = time.perf_counter()
        o = O.run_layered(cfg, stacks, threads=cores)
        O.layered_backward(o, np.ones((1, cfg.nr + 1)))
        return time.perf_counter() - t0

    n = 20_000
    sec = run(n)
    n = int(min(5e7, n * target_s / max(sec, 1e-3)))
    log(f"[bench] gradient cpu baseline: oracle on {cores} threads, {n} photons")
    sec = run(n)
    return {"value": n / sec, "unit": "photons/s", "cores": cores, "kind": "port",
            "sample": f"photons [0,{n}) of the 5-layer stack, forward with records and pMC "
                      f"backward in {sec:.1f} s (oracle, OpenMP)"}


def fit_leg(steps: int) -> dict:
    """Batched inverse fit: 256 three-layer instances, `steps` Adam steps of
    one forward (with path records) and one backward launch each, against the
    same step done as 256 per-instance launches; per-layer median relative
    errors at the start, after the Adam steps and after 8 Levenberg-Marquardt
    steps on the same posterior, and the Cramer-Rao bound of each parameter
    for the noise the loss weights by (the target's 1 % noise and the Monte
    Carlo variance of target and model runs, measured from 32 target batches)
    and the start-value prior: what an efficient fit of this data reaches."""
    from nmc_amd.inverse_fit import fit

    log(f"[bench] inverse fit: 256 instances x {steps} Adam steps")
    r = fit(256, steps, plain_compare=True)
    return r


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8, help="photons per step")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", choices=tuple(TILE_MODES), default="rec",
                    help="deposit scheme of the timed steps: rec (FP32 LDS tile + binned "
                         "records), fp32 (FP32 tile + FP64 atomics), fp64 (FP64 tile + atomics)")
    ap.add_argument("--legs", default="tile,sweep,grad,sens,fit",
                    help="extra measurements: tile (the other deposit schemes), sweep "
                         "(batched vs plain), grad (pMC cost), sens (per-tissue voxel "
                         "sensitivities), fit (batched inverse fit)")
    ap.add_argument("--fit-steps", type=int, default=50)
    ap.add_argument("--sweep-scale", type=float, default=1.0, help="scale of the sweep budgets")
    ap.add_argument("--grad-photons", type=float, default=1e7)
    ap.add_argument("--autotune", action="store_true",
                    help="run the launch autotuner first (its trial launches share the "
                         "timed kernel's name in a profile)")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB",
                    help="only run a plain copy of MIB MiB (scales PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding the "
              "photon-id range over several GPUs is not run here.", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)

    from nmc_amd import params as P
    from nmc_amd import workloads as W
    from nmc_amd.autotune import box_coverage, visit_histogram
    from nmc_amd.metal_mc import MetalMC

    n = int(args.photons)
    w = W.skin_vessel(n_photons=n)
    tuning = None
    if args.autotune:
        from nmc_amd.autotune import autotune_mcxyz
        t0 = time.perf_counter()
        tuning = autotune_mcxyz(w)
        log(f"[bench] autotune {tuning.as_dict()} in {time.perf_counter() - t0:.1f} s")
    log("[bench] atomic ceiling probe")
    ceiling = atomic_ceiling(w.rp.Nx * w.rp.Ny * w.rp.Nz)
    log("[bench] visit-count pilot (512-lane blocks)")
    hist = visit_histogram(w)
    mode = TILE_MODES[args.tile]
    mc = MetalMC(w.sim, 0)
    mc.prepare_data(w.rp, w.tp, w.v, w.ext)
    if tuning is None:
        tuning = pilot_tuning(mc, hist, mode)
    mc.set_tuning(tuning)
    eff = mc.get_tuning()
    coverage = box_coverage(hist, eff.tile)
    mc.set_num_runs_seed(1, 0)
    log(f"[bench] warmup {args.warmup} x {n} photons")
    for _ in range(args.warmup):
        mc.send_compute_command(n)
    mc.reset_tallies()
    mc.set_num_runs_seed(1, 0)
    run_ms, kern_ms = [], []
    log(f"[bench] timed {args.steps} x {n} photons")
    mc.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mc.send_compute_command_async(n)
        run_ms.append(mc.last_kernel_ms())      # waits for this step's end event
        kern_ms.append(mc.last_transport_ms())  # the transport kernel alone
    mc.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    res = mc.get_compute_results(fluence=False)
    tot = res.totals
    value = n * args.steps / wall
    # the tally download a caller does after a step (fluence 200^3 FP64,
    # R(x,y), detection counts, totals), timed apart from the steps above
    rb = []
    for _ in range(3):
        r0 = time.perf_counter()
        mc.get_compute_results()
        rb.append(time.perf_counter() - r0)
    readback_s = min(rb)
    nv, n2 = w.rp.Nx * w.rp.Ny * w.rp.Nz, w.rp.Nx * w.rp.Ny
    readback_bytes = nv * 8 + n2 * 12
    ksec = sum(kern_ms) * 1e-3
    avg_ms = sum(kern_ms) / len(kern_ms)
    avg_run_ms = sum(run_ms) / len(run_ms)

    ops = OPS_SAME * tot["substeps_same"] + OPS_CROSS * tot["substeps_cross"] \
        + OPS_INT * tot["interactions"]
    achieved = ops / ksec / 1e12
    # voxel-visit flushes: one per voxel left plus one per photon ending in the grid
    visits = tot["substeps_cross"] - tot["surface_hits"] + tot["photons"] \
        - tot["escapes_bottom"] - tot["escapes_side"] - tot["escapes_top"]
    out_of_tile = visits * (1.0 - coverage)
    records = out_of_tile if eff.mode == 2 else 0.0
    global_atomics = (out_of_tile - records) + tot["detections"]
    rd = tot["rd"] / tot["photons"]
    per_step = tot["photons"] / args.steps

    roof = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_T,
            "unit": "Tlane-op/s", "frac": achieved / VALU_PEAK_T, "traffic": None,
            # template arguments: two-term phase function, block, tile element,
            # fluence variant (1 deposits, 3 deposits with records), as in a
            # kernel trace
            "kernel": kernel_name(eff.mode, eff.block),
            "kernel_ms": avg_ms, "run_ms": avg_run_ms,
            "ops_per_photon": ops / tot["photons"],
            "events_per_photon": {"same": tot["substeps_same"] / tot["photons"],
                                  "cross": tot["substeps_cross"] / tot["photons"],
                                  "interactions": tot["interactions"] / tot["photons"]},
            "atomic": {"bound": "fp64 atomic adds, scattered", "ceiling_per_s": ceiling["per_s"],
                       "visits_per_photon": visits / tot["photons"],
                       "lds_tile_fraction": coverage,
                       "global_atomics_per_photon": global_atomics / tot["photons"],
                       "global_atomics_per_s": global_atomics / ksec,
                       "frac": global_atomics / ksec / ceiling["per_s"]},
            "tuning": eff.as_dict()}
    if eff.mode == 2:
        # the out-of-tile visits leave as 8-byte records, written by the
        # transport kernel and read once by the bucket reduction (pilot-sized
        # slices: the atomic fallback takes none here).  Fused (rec_fuse 1):
        # each launch's workgroups reduce the previous launch's records once
        # their own photons are done, so kernel_ms holds that work; the last
        # step's records are reduced in the synchronize that ends the timed
        # region.
        roof["records"] = {"per_photon": records / tot["photons"],
                           "bytes_per_photon": 8.0 * records / tot["photons"],
                           "reduction": ("fused into the next launch's drained tail"
                                         if eff.rec_fuse == 1 else "own kernel after each launch")}
        if eff.rec_fuse != 1:
            roof["records"]["reduction_ms"] = avg_run_ms - avg_ms
            roof["records"]["reduction_GB_per_s"] = (8.0 * records / args.steps
                                                     / ((avg_run_ms - avg_ms) * 1e-3) / 1e9)
    else:
        roof["atomic"]["photons_per_s_at_ceiling"] = ceiling["per_s"] * tot["photons"] / global_atomics
    if TRAFFIC_BYTES_PER_PHOTON.get(eff.mode) is not None:
        roof["traffic"] = TRAFFIC_BYTES_PER_PHOTON[eff.mode] * per_step
        roof["hbm_GB_per_s"] = roof["traffic"] / (avg_ms * 1e-3) / 1e9
        roof["hbm_frac"] = roof["hbm_GB_per_s"] / HBM_PEAK_GBS
    out = {
        "metric": METRIC, "value": value, "unit": "photons/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mcxyz skin-vessel 200^3 uint8 (0.005 cm voxels; air/epidermis/"
                   "dermis/blood cylinder), flat collimated beam r=0.06 cm, n=1, 1e8 photons/step",
                   "photons_per_step": n, "autotuned": args.autotune,
                   "lds_tile": {2: "FP32 sums within a workgroup, FP64 across workgroups; "
                                   "visits outside the tile as binned records (FP32 weight per "
                                   "visit) summed per bucket in FP64 (the reference adds every "
                                   "deposit into an FP32 global array)",
                                1: "FP32 sums within a workgroup, FP64 across workgroups and "
                                   "outside the tile (the reference adds every deposit into an "
                                   "FP32 global array)",
                                0: "FP64"}[eff.mode],
                   "tile_box": list(eff.tile)},
        "roofline": roof,
        "readback": {"ms": readback_s * 1e3, "bytes": readback_bytes,
                     "GB_per_s": readback_bytes / readback_s / 1e9,
                     "photons_per_s_with_readback_every_step":
                         n / (wall / args.steps + readback_s)},
        "rd": rd,
        "energy_balance": (tot["rd"] + tot["rd_outside"] + tot["absorbed"] + tot["tt"]
                           + tot["lateral"] + tot["top"] + tot["roulette_loss"] + tot["capped"]
                           - tot["roulette_gain"] + tot["specular"]) / tot["photons"],
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(w, args.cpu_seconds)
        ref = cb.pop("_tallies")
        m = cb.pop("_n")
        # the same photon ids on the GPU (512-lane blocks: a separate kernel
        # name in a profile; the launch shape never changes results)
        mc.set_tuning(P.Tuning.make(block=512, tile=tuple(eff.tile), mode=eff.mode))
        mc.reset_tallies()
        mc.set_num_runs_seed(1, 0)
        mc.send_compute_command(m)
        g = mc.get_compute_results(fluence=False).totals
        r = ref.totals
        out["cpu_baseline"] = cb
        out["rd_parity"] = {"photons": m, "gpu_rd": g["rd"] / m, "oracle_rd": r["rd"] / m,
                            "rel_diff": abs(g["rd"] - r["rd"]) / r["rd"],
                            "int_tallies_equal": all(g[k] == r[k] for k in P.TOTAL_INT_FIELDS)}
    legs = [s for s in args.legs.split(",") if s]
    if "tile" in legs:
        if eff.mode == 2 and eff.rec_fuse == 1:
            # the same records with their own reduction kernel after each launch
            out["rec_unfused_tile"] = tile_leg(mc, hist, n, 2, rec_fuse=-1)
        for m in sorted(TILE_MODES.values(), reverse=True):
            if m != eff.mode:
                out[f"{MODE_NAMES[m]}_tile"] = tile_leg(mc, hist, n, m)
    mc.close()
    if "sweep" in legs:
        out["batched_sweep"] = sweep_leg(args.sweep_scale)
        if not args.no_cpu_baseline:
            out["batched_sweep"]["cpu_baseline"] = sweep_cpu_baseline(args.sweep_scale,
                                                                      args.cpu_seconds / 2)
    if "grad" in legs:
        out["gradient"] = gradient_leg(int(args.grad_photons))
        if not args.no_cpu_baseline:
            out["gradient"]["cpu_baseline"] = gradient_cpu_baseline(args.cpu_seconds / 2)
    if "sens" in legs:
        out["voxel_sensitivities"] = voxel_sens_leg(int(args.grad_photons))
    if "fit" in legs:
        out["inverse_fit"] = fit_leg(args.fit_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
