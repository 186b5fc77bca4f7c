"""Pins of the oracle's Thomas-Fermi envelopes (reference thomasfermi.py:65-156)
by envelopes the reference computed itself (tests/golden/make_tf_golden.py, run
on the CPU where the reference is mounted; only its output is committed):
the regime chosen (kind, or no reference), mu, R90, R99 and the density for a
repulsive parabola, a trapped cubic-quintic, the self-bound flat top, and the
nonlocal compressed fixed point with the 1/r (2D and 3D), q2d and -ln r
kernels.  The fixed point's mixing, bisection and stopping rule are what these
vectors fix; its L1 tolerance of 1e-9 bounds how closely two runs agree."""
import math
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_tf_golden import CASES, case_cfg
from tests.helpers import cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tf_reference.npz")


def check_against_reference(tf, i, dv, dx, density):
    """tf: a result of extract_tf (density as a numpy array via density(tf))"""
    g = np.load(GOLD)
    if bool(g[f"none{i}"]):
        assert tf is None
        return
    assert tf is not None
    assert tf["kind"] == str(g[f"kind{i}"])
    assert bool(tf["converged"]) == bool(g[f"converged{i}"])
    assert math.isclose(tf["mu"], float(g[f"mu{i}"]), rel_tol=1e-9, abs_tol=1e-12), (tf["mu"], float(g[f"mu{i}"]))
    assert np.abs(density(tf) - g[f"n{i}"]).sum() * dv < 1e-8
    for k in ("R90", "R99"):
        assert abs(tf[k] - float(g[f"{k}{i}"])) <= 1e-9 + 2.0 * dx, (k, tf[k], float(g[f"{k}{i}"]))


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_tf_matches_reference(i):
    c = cfg(tf_cache=False, **case_cfg(CASES[i]))
    G = O.make_grid(c)
    check_against_reference(O.extract_tf(c, G), i, O.dV(G), float(G["dx"]), lambda tf: tf["n"])
