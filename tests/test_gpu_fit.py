"""GPU: the batched inverse fit reduces its loss and moves the parameters of
the top layer toward the hidden values (small, fast version)."""
from __future__ import annotations

import numpy as np
import pytest

from nmc_amd.inverse_fit import fit

pytestmark = pytest.mark.gpu


def test_small_inverse_fit(gpu):
    print("[fit] 8 instances x 20 steps", flush=True)
    r = fit(n_inst=8, steps=20, budget_scale=0.02, lr=0.05, plain_compare=False)
    assert r["loss_last"] < 0.5 * r["loss_first"]
    assert r["median_rel_err_mus_per_layer"][0] < r["median_start_err_mus_per_layer"][0]
    assert np.isfinite(r["wall_s"])
