"""Shared comparison helpers for the parity tests (HIP library vs oracle)."""
from __future__ import annotations

import numpy as np

from nmc_amd import params as P

INT_FIELDS = P.TOTAL_INT_FIELDS
FLOAT_FIELDS = P.TOTAL_FLOAT_FIELDS

# Same photons, same float trajectories: only the order of the FP64 sums
# differs (atomics), so the float tallies agree to ~1e-15 relative.  The
# tolerance leaves room for that and nothing else.
RTOL = 1e-9


def assert_same_tallies(gpu: P.TallyArrays, ref: P.TallyArrays, label: str = "") -> None:
    g, r = gpu.totals, ref.totals
    for k in INT_FIELDS:
        assert g[k] == r[k], f"{label}: integer tally {k}: gpu {g[k]} != oracle {r[k]}"
    for k in FLOAT_FIELDS:
        assert np.isclose(g[k], r[k], rtol=RTOL, atol=1e-12 * max(1, r["photons"])), \
            f"{label}: {k}: gpu {g[k]!r} oracle {r[k]!r}"
    if ref.det2d is not None:
        bad = np.flatnonzero(gpu.det2d != ref.det2d)
        assert bad.size == 0, f"{label}: det2d differs at {bad[:10]}"
    for name in ("fluence", "ryx", "rd_ra", "a_rz"):
        a, b = getattr(gpu, name), getattr(ref, name)
        if a is None or b is None:
            continue
        scale = max(1e-300, float(np.abs(b).max()))
        err = np.abs(a - b)
        ok = err <= RTOL * np.abs(b) + 1e-13 * scale
        assert ok.all(), (f"{label}: {name} differs at {np.flatnonzero(~ok)[:10]}, "
                          f"max err {err.max():.3e} scale {scale:.3e}")
        # exact zero pattern: a bin is hit on one side iff on the other
        assert np.array_equal(a == 0, b == 0), f"{label}: {name} zero pattern differs"


def energy_balance(t: dict) -> float:
    """Launched weight accounted for by the tallies, per photon (== 1)."""
    out = (t["specular"] + t["rd"] + t["rd_outside"] + t["absorbed"] + t["absorbed_outside"]
           + t["tt"] + t["lateral"] + t["top"] + t["roulette_loss"] + t["capped"]
           - t["roulette_gain"])
    return out / t["photons"]
