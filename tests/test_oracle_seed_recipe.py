"""Pins of the oracle's seed recipes (reference seeds.py:11-103) by the
reference's own initial states with their noise switched off
(tests/golden/make_seed_golden.py, run on the CPU where the reference is
mounted; only its output is committed).  With seed_noise = seed_phase_noise
= 0 the torch-RNG noise factors are exactly 1, so the Gaussian and
Thomas-Fermi envelopes, the imprinted windings and the triangular lattice
(angle and shift from numpy's default_rng(rng), tanh cores, summed phases,
winding sign(Omega), the fallback to the tf seed below |Omega| = 0.3, straight
lines along z in 3D) are what the reference returns, to 1e-12."""
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_seed_golden import CASES, case_cfg
from tests.helpers import cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "seed_reference.npz")


def check_seed(psi, i):
    ref = np.load(GOLD)[f"psi{i}"]
    assert psi.shape == ref.shape
    assert np.abs(psi - ref).max() <= 1e-12 * np.abs(ref).max(), np.abs(psi - ref).max()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_seed_recipe_matches_reference(i):
    c = cfg(tf_cache=False, **case_cfg(CASES[i]))
    check_seed(O.make_seed(O.make_grid(c), c), i)
