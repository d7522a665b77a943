"""CPU checks of the autotuner's box search (no GPU)."""
import numpy as np

from nmc_amd.autotune import best_boxes


def test_box_search_finds_the_blob():
    z, x, y = np.meshgrid(np.arange(40), np.arange(60), np.arange(60), indexing="ij")
    hist = np.exp(-((x - 20) ** 2 + (y - 41) ** 2) / 50.0 - (z - 3) ** 2 / 20.0)
    boxes = best_boxes(hist, 2000, n_best=3)
    assert len(boxes) == 3
    cov, (x0, y0, z0, nx, ny, nz) = boxes[0]
    assert nx * ny * nz <= 2000
    # the reported coverage is the box sum
    s = hist[z0:z0 + nz, x0:x0 + nx, y0:y0 + ny].sum() / hist.sum()
    assert abs(s - cov) < 1e-9
    # centred on the blob and no shifted copy of the box does better
    assert x0 <= 20 < x0 + nx and y0 <= 41 < y0 + ny and z0 <= 3 < z0 + nz
    for dz, dx, dy in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (0, -1, 0), (0, 0, -1)):
        zz, xx, yy = z0 + dz, x0 + dx, y0 + dy
        if min(zz, xx, yy) < 0 or zz + nz > 40 or xx + nx > 60 or yy + ny > 60:
            continue
        assert hist[zz:zz + nz, xx:xx + nx, yy:yy + ny].sum() / hist.sum() <= cov + 1e-12


def test_batch_cache_key_covers_every_item_field():
    """The batched-mcsub tuner keys its cache on the items' bytes: dicts and
    WorkItem structs of the same media agree, and any field change differs."""
    from nmc_amd import params as P
    from nmc_amd import workloads as W
    from nmc_amd.autotune import _item_bytes

    _, items, _ = W.spectral_sweep(n_wavelengths=4, budget_scale=0.001)
    structs = []
    for it in items:
        s = P.WorkItem()
        for k, v in it.items():
            setattr(s, k, v)
        structs.append(s)
    assert _item_bytes(items) == _item_bytes(structs)
    changed = [dict(it) for it in items]
    changed[2]["n_photons"] += 1
    assert _item_bytes(changed) != _item_bytes(items)
