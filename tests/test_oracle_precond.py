"""Pins of the oracle's Sobolev preconditioner (reference solvers.py:28-45).

* Vectors the reference computed itself (tests/golden/make_precond_golden.py,
  run where the reference is mounted): sigma and z for nine cases, covering
  the six physics sets of its FD audit, the floor, a self-bound cubic-quintic
  state and a 3D state above the floor.  They fix the heuristic parts: the
  local stiffness V + 2 beta2 rho + 3 beta3 rho^2 (+ phi), the bulk mask
  rho > 0.05 max rho and the floor.
* A closed form: for any field v, applying the preconditioner to
  g = IFFT((k^2/2 + sigma) FFT(v)) returns proj_tangent(v), which fixes the
  k^2/2 + sigma denominator and the projection.
* A density with a hand-computed lower median.
"""
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_precond_golden import CASES, inputs
from tests.helpers import cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "precond_reference.npz")


def case_cfg(kw):
    c = cfg(**kw)
    return c


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_preconditioner_matches_reference_vectors(i):
    kw = CASES[i]
    gold = np.load(GOLD)
    c = case_cfg(kw)
    G = O.make_grid(c)
    ndim = 3 if kw["dimension"] == "3D" else 2
    psi, grad = inputs(ndim, int(kw["Ngrid"]), float(kw["L"]), seed=100 + i)
    phi = None
    if G["kernel"] != "none" and abs(c["G_C"]) > 1e-15:
        phi = O.long_range_phi(np.abs(psi) ** 2, G, c["G_C"])
    z, sigma = O.preconditioner(grad, psi, G, c, phi)
    s_ref = float(gold[f"sigma{i}"])
    assert abs(sigma - s_ref) <= 1e-13 * max(1.0, abs(s_ref)), (sigma, s_ref)
    zr = gold[f"z{i}"]
    assert np.abs(z - zr).max() <= 1e-12 * np.abs(zr).max()


def test_planted_stiffness_mistake_changes_the_shift():
    """beta2 rho in place of 2 beta2 rho moves sigma off the reference's value."""
    i = 0
    kw = CASES[i]
    gold = np.load(GOLD)
    c = case_cfg(kw)
    G = O.make_grid(c)
    psi, _ = inputs(2, int(kw["Ngrid"]), float(kw["L"]), seed=100 + i)
    rho = np.abs(psi) ** 2
    local = G["V"] + c["beta2"] * rho + 3.0 * c["beta3"] * rho ** 2
    wrong = max(O.lower_median(local[rho > 0.05 * rho.max()]), 0.5)
    assert abs(wrong - float(gold["sigma0"])) > 1e-3 * float(gold["sigma0"])


@pytest.mark.parametrize("ndim,N", [(2, 32), (3, 12)])
def test_preconditioner_inverts_its_operator(ndim, N):
    c = cfg(dimension="3D" if ndim == 3 else "2D", beta2=40.0, beta3=3.0, Ngrid=N, L=4.0)
    G = O.make_grid(c)
    rng = np.random.default_rng(7)
    shape = (N,) * ndim
    psi = np.exp(-0.5 * G["R2"]) * (1.0 + 0.2 * rng.standard_normal(shape))
    psi = (psi / O.norm_of(psi, G)).astype(np.complex128)
    v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    sigma = O.precond_shift(psi, G, c)
    g = O.ifftn((0.5 * G["K2"] + sigma) * O.fftn(v))
    z, s2 = O.preconditioner(g, psi, G, c)
    assert s2 == sigma
    want = O.proj_tangent(v, psi, O.dV(G))
    assert np.abs(z - want).max() <= 1e-12 * np.abs(want).max()
    # the projection leaves z orthogonal to psi (complex inner product)
    assert abs((psi.conj() * z).sum() * O.dV(G)) <= 1e-12 * np.abs(z).max()


def test_shift_is_the_lower_median_of_the_bulk():
    """rho takes the values 1 (10 points), 2 (10), 3 (11) and 0.1 elsewhere,
    which is below 0.05 * max rho = 0.15 and so outside the bulk.  With no
    trap, beta2 = 1, beta3 = 0 the local values are 2 rho; the 31 bulk values
    sorted have the lower median (index 15) at rho = 2, so sigma = 4 (to the
    rounding of |sqrt(2)|^2)."""
    c = cfg(dimension="2D", s=0, trap_coeff=0.0, beta2=1.0, beta3=0.0, Ngrid=16, L=4.0)
    G = O.make_grid(c)
    rho = np.full((16, 16), 0.1)
    flat = rho.reshape(-1)
    flat[0:10] = 1.0
    flat[40:50] = 2.0
    flat[100:111] = 3.0
    psi = np.sqrt(rho).astype(np.complex128)
    assert abs(O.precond_shift(psi, G, c) - 4.0) <= 4e-15
    # with beta3 = 1 the value at rho = 2 is 2*2 + 3*4 = 16
    assert abs(O.precond_shift(psi, G, dict(c, beta3=1.0)) - 16.0) <= 1.6e-14
    # a floor above the median wins
    assert O.precond_shift(psi, G, dict(c, precond_shift_min=5.0)) == 5.0
