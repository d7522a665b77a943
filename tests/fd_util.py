"""Central finite differences with common random numbers (CRN) against the
perturbation-MC gradient of a layered stack, on Rd(r) bin groups.

The 20 perturbed copies of a stack (mua and mus of each layer, +h and -h)
share the base stack's Philox stream, so each transports the same photons
with the same random numbers; the oracle and the GPU both take all of them
in one run.  The comparison is made per batch of photons (K disjoint photon
ranges), so the noise floor of the difference pMC - FD is measured, not
assumed.

The bar, per component (Rd group g, parameter, layer l), with m the pMC
mean over the batches, d = mean(FD) - m and s the standard error of d:
  - zero: no detected photon reached layer l, so pMC and FD are exactly 0
    in every batch (deep layers of an optically thick stack);
  - 1 %: where the noise floor resolves 1 % (4.5 s <= 0.01 |m|): |d| <= 0.01 |m|;
  - 4.5 s: elsewhere |d| <= 4.5 s (the noise floor at this N is stated).
"""
from __future__ import annotations

import numpy as np

from nmc_amd import params as P

# Rd(r) bin groups (200 bins of 0.005 cm + overflow): 0-0.05, 0.05-0.15,
# 0.15-0.3, 0.3-0.6 cm, and the total
GROUPS = [(0, 10), (10, 30), (30, 60), (60, 120), (0, 201)]
# relative FD steps.  mua: the two-point central difference (mua leaves the
# paths unchanged, so its h^2 error is far below 1 % at h = 1 %).  mus: the
# five-point central stencil (f(-2h) - 8 f(-h) + 8 f(h) - f(2h)) / 12 h, whose
# error is O(h^4): a two-point difference at h = 5 % carries an h^2 error of
# up to ~1 % here, and a smaller h would raise the CRN noise (paths
# decorrelate as mus moves every step length).
H_MUA, H_MUS = 0.01, 0.05
PER_LAYER = 6       # mua +h, -h; mus +h, -h, +2h, -2h


def perturbed(stack, nl: int) -> list:
    """The stack's PER_LAYER perturbed copies per layer, in fd_groups' order."""
    out = []
    for l in range(nl):
        for name, f in (("mua", 1 + H_MUA), ("mua", 1 - H_MUA), ("mus", 1 + H_MUS),
                        ("mus", 1 - H_MUS), ("mus", 1 + 2 * H_MUS), ("mus", 1 - 2 * H_MUS)):
            s = P.Stack.from_buffer_copy(stack)
            setattr(s.layer[l], name, getattr(stack.layer[l], name) * f)
            out.append(s)
    return out


def fd_groups(rd_pert: np.ndarray, stack, nl: int):
    """FD of the grouped Rd per photon: (fd_mua [G, nl], fd_mus [G, nl]) from
    the perturbed runs' Rd per photon rd_pert [PER_LAYER nl, nr + 1]."""
    grp = lambda r: np.array([r[a:b].sum() for a, b in GROUPS])
    fa = np.zeros((len(GROUPS), nl))
    fs = np.zeros((len(GROUPS), nl))
    for l in range(nl):
        r = [grp(rd_pert[PER_LAYER * l + i]) for i in range(PER_LAYER)]
        ha = H_MUA * stack.layer[l].mua
        hs = H_MUS * stack.layer[l].mus
        fa[:, l] = (r[0] - r[1]) / (2 * ha)
        fs[:, l] = (8 * (r[2] - r[3]) - (r[4] - r[5])) / (12 * hs)
    return fa, fs


def fd_values(vals: np.ndarray, stack, nl: int):
    """FD of per-group values (e.g. the absorbed weight of each layer) from
    the perturbed runs' values vals [PER_LAYER nl, G]: (fd_mua [G, nl],
    fd_mus [G, nl])."""
    G = vals.shape[1]
    fa = np.zeros((G, nl))
    fs = np.zeros((G, nl))
    for l in range(nl):
        r = [vals[PER_LAYER * l + i] for i in range(PER_LAYER)]
        ha = H_MUA * stack.layer[l].mua
        hs = H_MUS * stack.layer[l].mus
        fa[:, l] = (r[0] - r[1]) / (2 * ha)
        fs[:, l] = (8 * (r[2] - r[3]) - (r[4] - r[5])) / (12 * hs)
    return fa, fs


def grad_rd_groups(nr: int) -> np.ndarray:
    """One grad_rd row per group: the indicator of its bins."""
    g = np.zeros((len(GROUPS), nr + 1))
    for i, (a, b) in enumerate(GROUPS):
        g[i, a:b] = 1.0
    return g


def judge(pmc: np.ndarray, fd: np.ndarray, label: str) -> dict:
    """pmc, fd: [K, G, nl] per-batch estimates.  Applies the bar; returns
    counts and the table of rows (group, layer, m, d/m, s/m, bar)."""
    K = pmc.shape[0]
    m = pmc.mean(0)
    d = fd.mean(0) - m
    s = (pmc - fd).std(0, ddof=1) / np.sqrt(K)
    rows, counts, bad = [], {"zero": 0, "1%": 0, "4.5s": 0}, []
    for g in range(m.shape[0]):
        for l in range(m.shape[1]):
            if not pmc[:, g, l].any() and not fd[:, g, l].any():
                counts["zero"] += 1
                rows.append((g, l, 0.0, 0.0, 0.0, "zero"))
                continue
            if 4.5 * s[g, l] <= 0.01 * abs(m[g, l]):
                bar, ok = "1%", abs(d[g, l]) <= 0.01 * abs(m[g, l])
            else:
                bar, ok = "4.5s", abs(d[g, l]) <= 4.5 * s[g, l]
            counts[bar] += 1
            rel = lambda x: x / abs(m[g, l]) if m[g, l] != 0 else float("inf")
            rows.append((g, l, m[g, l], rel(d[g, l]), rel(s[g, l]), bar))
            if not ok:
                bad.append((label, g, l, m[g, l], d[g, l], s[g, l], bar))
    assert not bad, bad
    return {"counts": counts, "rows": rows}


def table(label: str, res: dict) -> str:
    lines = [f"{label}: {res['counts']}",
             "  group layer        pMC    (FD-pMC)/|pMC|  noise/|pMC|  bar"]
    for g, l, m, r, s, bar in res["rows"]:
        lines.append(f"  {g:5d} {l:5d} {m:11.4e} {r:14.2e} {s:12.2e}  {bar}")
    return "\n".join(lines)
