"""Pins of the oracle's multi-seed selection (reference solvers.py:238-263) by
the reference's own selections with the seeds' noise switched off
(tests/golden/make_multiseed_golden.py, run on the CPU where the reference is
mounted; only its output is committed): the seed cycle and rng offsets, every
seed's converged energy (to 1e-9), the winner where the energies differ by
more than that (the lattice seed of a rotating condensate), the spread and
the selected E, mu (to the 1e-8 residual's first order) and L_z."""
import json
import os

import pytest

import oracle as O
from tests.golden.make_multiseed_golden import CASES, case_cfg
from tests.helpers import cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "multiseed_reference.json")


def check_selection(obs, i):
    with open(GOLD) as f:
        g = json.load(f)["cases"][i]
    mine = [float(e) for e in obs["seed_energies"].split(";")]
    ref = [float(e) for e in g["seed_energies"].split(";")]
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-9, (mine, ref)    # the strings carry 10 digits
    if max(ref) - min(ref) > 1e-8 * max(abs(e) for e in ref):
        assert obs["seed_used"] == g["seed_used"]
    assert abs(obs["seed_energy_spread"] - g["seed_energy_spread"]) <= 1e-8 * max(1.0, abs(g["E"]))
    # runs stopped at res_tol 1e-8 along different round-off trajectories:
    # E agrees to second order in the residual, mu to first order
    assert abs(obs["E"] - g["E"]) <= 1e-9 * abs(g["E"]), (obs["E"], g["E"])
    assert abs(obs["mu"] - g["mu"]) <= 1e-8 * abs(g["mu"]), (obs["mu"], g["mu"])
    assert abs(obs["Lz"] - g["Lz"]) <= 1e-7 * max(1.0, abs(g["Lz"]))


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_multiseed_matches_reference(i):
    c = cfg(tf_cache=False, **case_cfg(CASES[i]))
    _, _, obs = O.ground_state_multiseed(c, O.make_grid(c))
    check_selection(obs, i)
