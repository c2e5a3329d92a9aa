"""Initial targeted modal over-occupation (mode_excitation, the reference's
MonteCarloSolver::apply_initial_mode_excitation, src/solver/Initialization.cpp:146-178).

- Closed form: from the same Philox population, every carrier whose frequency
  omega / 2 pi lies in the band starts at multiplier times the occupation it has
  without the excitation, every other carrier unchanged (the oracle's
  initialization, which the device's is compared with carrier by carrier).
- The host's [input] summary keys equal the ones the reference's DomainOutput
  writes for the same TOML, and invalid settings fail with its messages.
"""
import os
import re
import tempfile

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from oracle import ref_ctypes as R
from phonomc_amd import runtime as rt

EXCITE = {"mode_excitation.enabled": "true", "mode_excitation.frequency_min": "4.0",
          "mode_excitation.frequency_max": "9.0", "mode_excitation.occupation_multiplier": "1.75"}


def population(**over):
    p = rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 20000, **over}),
                   mode_sampling=1)
    sim = O.Sim(p)
    sim.initialize()
    return p, sim.carriers()


def test_excitation_multiplies_the_band():
    p, plain = population()
    pe, excited = population(**EXCITE)
    assert pe.views()[3].excite == 1  # pmc_init
    assert np.array_equal(plain["mode"], excited["mode"]) and np.array_equal(plain["uid"], excited["uid"])
    _, m, *_ = p.views()
    f = rt.arr(m.omega, m.n_active)[plain["mode"]] / (2 * np.pi)
    band = (f >= 4.0) & (f <= 9.0)
    assert 0 < band.sum() < len(band)
    assert np.array_equal(excited["occupation"][band], plain["occupation"][band] * 1.75)
    assert np.array_equal(excited["occupation"][~band], plain["occupation"][~band])


def test_multiplier_of_one_changes_nothing():
    p, plain = population()
    pe, same = population(**{**EXCITE, "mode_excitation.occupation_multiplier": "1.0"})
    assert pe.views()[3].excite == 0
    assert np.array_equal(plain["occupation"], same["occupation"])


@pytest.mark.skipif(not R.available(), reason="reference build (oracle/_ref) not present")
def test_summary_keys_match_reference():
    text = rt.config_text("c1_cross_plane.toml", **EXCITE)
    d = tempfile.mkdtemp()
    path = os.path.join(d, "run.toml")
    with open(path, "w") as fh:
        fh.write(text)
    R.domain_summary(path, os.path.join(d, "ref"))
    ref = open(os.path.join(d, "ref", "summary.txt")).read()
    p = rt.Problem(text)
    ours_dir = os.path.join(d, "ours")
    os.makedirs(ours_dir)
    p.write_summary(ours_dir)
    ours = open(os.path.join(ours_dir, "summary.txt")).read()
    pick = lambda t: re.findall(r"^mode_excitation_\w+ = .*$", t, re.M)
    assert pick(ours) == pick(ref) and len(pick(ref)) == 4


@pytest.mark.parametrize("bad,message", [
    ({"mode_excitation.occupation_multiplier": "0"}, "mode_excitation.occupation_multiplier must be finite and positive."),
    ({"mode_excitation.frequency_min": "9", "mode_excitation.frequency_max": "4"},
     "mode_excitation.frequency_max must be no smaller than frequency_min."),
])
def test_invalid_excitation_rejected_with_reference_messages(bad, message):
    with pytest.raises(RuntimeError, match=message):
        rt.Problem(rt.config_text("c1_cross_plane.toml", **{**EXCITE, **bad}))
