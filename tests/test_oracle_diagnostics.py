"""Pinning the oracle's diagnostics (CPU) by closed forms: the lowest-Landau-
level orbitals and states orthogonal to them, states with phase windings
planted at known points, and density profiles whose 1% edge is known."""
import math

import numpy as np
import pytest

import oracle as O
from tests.helpers import cfg, lll_state, zeros_state


def grid2(N=96, L=8.0):
    return O.make_grid(cfg(dimension="2D", Ngrid=N, L=L))


def unit(psi, G):
    return psi / O.norm_of(psi, G)


def test_lll_orbitals_have_unit_weight_on_their_own_side():
    """vortex.py:61-84: z^m e^{-r^2/2} lies in the LLL of Omega > 0 and is
    orthogonal to it for Omega < 0 (theta -> -theta), and conversely."""
    G = grid2()
    for m in (0, 1, 3, 7):
        assert O.lll_weight(lll_state(G, m), G, Omega=1.0) == pytest.approx(1.0, abs=1e-12)
        assert O.lll_weight(lll_state(G, m, conj=True), G, Omega=-0.5) == pytest.approx(1.0, abs=1e-12)
    for m in (1, 3):
        assert O.lll_weight(lll_state(G, m), G, Omega=-1.0) < 1e-12
        assert O.lll_weight(lll_state(G, m, conj=True), G, Omega=0.7) < 1e-12


def test_lll_weight_of_orthogonal_mixture():
    """(1 - r^2) e^{-r^2/2} is orthogonal to every z^m e^{-r^2/2} (its m = 0
    overlap is int (1 - r^2) e^{-r^2} r dr = 0): a normalised equal mixture with
    z e^{-r^2/2} has weight 1/2; a slip of r^2 for r, or of the Gaussian's
    width, moves it."""
    G = grid2()
    rad = unit((1.0 - G["R2"]) * np.exp(-0.5 * G["R2"]) + 0j, G)
    assert O.lll_weight(rad, G, Omega=1.0) < 1e-12
    mix = rad + unit(lll_state(G, 1), G)
    assert O.lll_weight(mix, G, Omega=1.0) == pytest.approx(0.5, abs=1e-12)


def test_lll_cap_follows_the_box():
    """Mcap = min(max(Mmax, estimate), max(Mmax, int((0.8 L)^2))): with Mmax = 4
    the orbital m = 8 is inside the sum for L = 3.6 (cap 8) and outside it for
    L = 3.5 (cap 7)."""
    for L, inside in ((3.6, True), (3.5, False)):
        G = O.make_grid(cfg(dimension="2D", Ngrid=64, L=L))
        w = O.lll_weight(lll_state(G, 8), G, Mmax=4, Omega=1.0)
        assert (w > 1.0 - 1e-12) if inside else (w < 1e-2), (L, w)


def test_lll_weight_is_nan_outside_2d():
    G = O.make_grid(cfg(dimension="3D", Ngrid=16, L=4.0))
    assert math.isnan(O.lll_weight(np.ones(G["R2"].shape, complex), G))


@pytest.mark.parametrize("plus,minus", [((), ()), ((1.1 + 0.4j,), ()), ((1.1 + 0.4j, -0.3 + 1.35j), (-0.9 - 0.7j,)),
                                        (tuple(1.5 * np.exp(1j * (2 * math.pi * k / 6 + 0.1)) for k in range(6)), ())])
def test_vortex_count_of_planted_zeros(plus, minus):
    """vortex.py:10-48: each simple zero of z - a winds +1 counter-clockwise in
    (x, y), each zero of conj(z - b) winds -1; a sign slip in the angle or the
    plaquette order swaps Nplus and Nminus."""
    G = grid2()
    d = O.vortex_count_2d(zeros_state(G, plus, minus), G)
    assert d == dict(Nv=len(plus) + len(minus), Nplus=len(plus), Nminus=len(minus), Nnet=len(plus) - len(minus))


def test_vortex_outside_the_droplet_is_not_counted():
    """A zero out in the tail (density far below 15% of the smoothed maximum and
    beyond the 99% mass radius) is not counted."""
    G = grid2()
    d = O.vortex_count_2d(zeros_state(G, (6.2 + 0.3j,), (), sigma=1.5), G)
    assert d["Nv"] == 0


def test_droplet_radius_closed_forms():
    """fields.py:40-67: a Gaussian exp(-r^2 / 2 s^2) falls to 1% of its peak
    at r = s sqrt(2 ln 100), and a flat disc of radius a ends in the bin
    holding a; both about their centre of mass (placed off the origin)."""
    G = O.make_grid(cfg(dimension="2D", Ngrid=256, L=8.0))
    X, Y = G["coords"]
    bw = float(G["L"]) / int(min(400, max(20, G["L"] / (2 * G["dx"]))))
    s, x0 = 1.3, 0.75
    rho = np.exp(-((X - x0) ** 2 + Y ** 2) / (2 * s * s))
    assert abs(O.droplet_radius(rho, G) - s * math.sqrt(2 * math.log(100.0))) <= bw
    a = 2.6
    disc = (((X + 0.4) ** 2 + (Y - 0.3) ** 2) <= a * a).astype(float)
    r1 = O.droplet_radius(disc, G)
    assert abs(r1 - a) <= bw
    assert math.isnan(O.droplet_radius(np.zeros_like(rho), G))
