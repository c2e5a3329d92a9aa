"""The learned warm start on the device.

The operator is trained here by the device kernels (BF16 MFMA, FP32 Adam) on
steady fields of the device step: phonomc_amd/data/heated_film_fields.npz holds
160 heated-film cases (source centre, widths, power and reservoir temperatures
drawn by warmstart.sample_cases, seed 11), each the mean over 4000 steps after
14 000 steps from the linear reservoir field at 5e5 carriers.  It was written by
`python -m phonomc_amd.warmstart --cases 80 --first 0|80` on the device.  The
first 144 cases train, the last 16 are held out.

What the network gains over the method without it, on held-out cases:
- its field is much closer to the steady field than the reservoir field, the
  best guess without it (test_operator_generalizes);
- Monte Carlo started from its field reaches 0.5 K RMS of a long plain run's
  late field at least 3x sooner than plain Monte Carlo from the background
  temperature, at the same carrier count and averaging window, and ends as
  close to that field as the plain run does (test_warm_start_reaches_steady_faster).
"""
import os

import numpy as np
import pytest

from phonomc_amd import warmstart as ws

pytestmark = pytest.mark.gpu
DATA = os.path.join(os.path.dirname(ws.__file__), "data", "heated_film_fields.npz")
NTRAIN = 144


@pytest.fixture(scope="module")
def trained():
    params, x, y = ws.load_dataset(DATA)
    model = ws.WarmStart()
    losses = model.fit(x[:NTRAIN], y[:NTRAIN])
    return model, params, x, y, losses


def test_operator_generalizes(trained):
    model, params, x, y, losses = trained
    assert losses[-1] < 0.05 * losses[0]
    pred = model.predict(x[NTRAIN:])
    rms = np.sqrt(np.mean((pred - y[NTRAIN:]) ** 2, axis=1))
    without = np.sqrt(np.mean(y[NTRAIN:] ** 2, axis=1))  # reservoir field: zero predicted rise
    print(f"held-out RMS error: operator {rms.mean():.3f} K (max {rms.max():.3f}), "
          f"reservoir field {without.mean():.3f} K (max {without.max():.3f}); "
          f"training loss {losses[0]:.3f} -> {losses[-1]:.4f} K^2")
    assert rms.mean() < 0.1 * without.mean()
    assert rms.mean() < 0.5


def test_warm_start_reaches_steady_faster(trained):
    model, params, x, y, losses = trained
    held = ws.cases_from_params(params[NTRAIN:])
    k = int(np.argmax(y[NTRAIN:].max(axis=1)))  # the hottest held-out case
    out = ws.evaluate(model, held[k], particles=1000000, log=print)
    assert out["warm_windows"] is not None and out["plain_windows"] is not None
    print(f"  speedup {out['speedup_steps']:.1f}x in steps, {out['speedup_wall']:.1f}x in wall time")
    assert out["speedup_steps"] >= 3.0
    assert out["speedup_wall"] >= 3.0
    # Converges to the plain field: no farther from T_ref than plain Monte Carlo
    # of the same length (the margin is a few times the noise of a 12 000-step mean).
    assert out["warm_late_rms_k"] <= out["plain_late_rms_k"] + 0.05
