"""The energy's gradient against the reference's own torch autograd of its
energy_components (tests/golden/make_gradient_golden.py, run on the CPU where
the reference is mounted): on the six physics sets' seeded states the
oracle's adjoint 2 H psi dv (torch's convention for the gradient of a real
function of a complex field: dE/dRe psi + i dE/dIm psi) equals it to 1e-12
relative -- the kinetic, trap, contact, three-body, rotation and long-range
(convolution adjoint) terms together."""
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_operator_golden import CASES, state
from tests.helpers import cfg
from tests.test_oracle_operators_reference import close_vec

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gradient_reference.npz")


def reference_gradient(i):
    return np.load(GOLD)[f"grad{i}"]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_adjoint_matches_reference_autograd(i):
    c = cfg(**CASES[i])
    G = O.make_grid(c)
    p = state(CASES[i], seed=700 + i)
    close_vec(2.0 * O.apply_H(p, G, c) * O.dV(G), reference_gradient(i))
