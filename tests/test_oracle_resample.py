"""Pinning the oracle's resample_state (CPU): the reference's own identity
tests (tests/test_audit.py:139-143, 230-234), exact reproduction of
multilinear functions, zeros outside the old span, unit norm."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import cfg


def _unit(psi, G):
    return psi / O.norm_of(psi, G)


@pytest.mark.parametrize("N", [64, 65])
def test_identity_even_and_odd_grids(N):
    G = O.make_grid(cfg(dimension="2D", Ngrid=N))
    psi = _unit(np.exp(-((G["X"] - 1.0) ** 2 + G["Y"] ** 2)) + 0j, G)
    assert np.abs(O.resample_state(psi, G, G) - psi).max() < 1e-12


@pytest.mark.parametrize("dim,n0,L0,n1,L1", [("2D", 40, 6.0, 64, 4.0), ("2D", 48, 6.0, 30, 5.0),
                                              ("3D", 16, 4.0, 20, 3.0)])
def test_multilinear_functions_are_reproduced(dim, n0, L0, n1, L1):
    """Inside the old span, interpolation of (1 + x)(2 - i y)[(3 + z)] is exact;
    a slip in an index, a weight or an axis order breaks it."""
    G0 = O.make_grid(cfg(dimension=dim, Ngrid=n0, L=L0))
    G1 = O.make_grid(cfg(dimension=dim, Ngrid=n1, L=L1))

    def f(G):
        v = (1.0 + G["coords"][0]) * (2.0 - 1j * G["coords"][1])
        if G["ndim"] == 3:
            v = v * (3.0 + 0.5 * G["coords"][2])
        return v

    got = O.resample_state(f(G0), G0, G1)
    want = _unit(f(G1), G1)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_zero_outside_the_old_span():
    G0 = O.make_grid(cfg(dimension="2D", Ngrid=32, L=3.0))
    G1 = O.make_grid(cfg(dimension="2D", Ngrid=64, L=6.0))
    got = O.resample_state(np.ones(G0["R2"].shape, complex), G0, G1)
    x0, x1 = -16 * G0["dx"], 15 * G0["dx"]
    X, Y = G1["coords"]
    outside = (X < x0) | (X > x1) | (Y < x0) | (Y > x1)
    inside = ~outside
    # the first node outside sees the edge node with a weight of rounding size
    assert np.abs(got[outside]).max() < 1e-14 and np.ptp(got[inside]) < 1e-14
    assert abs(O.norm_of(got, G1) - 1.0) < 1e-14


def test_degenerate_grid_gives_none():
    G0 = O.make_grid(cfg(dimension="2D", Ngrid=1, L=3.0))
    G1 = O.make_grid(cfg(dimension="2D", Ngrid=8, L=3.0))
    assert O.resample_state(np.ones((1, 1), complex), G0, G1) is None
