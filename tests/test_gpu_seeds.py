"""Seeds on the device (bec_seed, bec_tf_nonlocal) against the oracle's
restatement of seeds.py:11-103 with the same counter-based noise.

Envelopes that need no fixed point (Gaussian, the inverted TF profiles, the
flat top) agree to 1e-12; envelopes from the nonlocal fixed point agree to
the fixed point's own tolerance (its iterates stop at an L1 change of 1e-9).
"""
import math

import numpy as np
import pytest

import oracle as O
from tests.helpers import cfg

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


LOCAL = [
    dict(seed="gaussian"),
    dict(seed="gaussian", seed_winding=2, rng=5),
    dict(seed="tf", beta2=100.0, rng=3),                                    # repulsive parabola
    dict(seed="tf", beta2=30.0, beta3=5.0, Omega=0.2, rng=1),               # local trapped repulsive
    dict(seed="tf", trap_coeff=0.0, beta2=-60.0, beta3=1000.0, L=8.0),      # flat top
    dict(seed="tf", beta2=0.0, beta3=0.0),                                  # no TF density: the Gaussian
    dict(seed="triangular", Omega=0.9, beta2=200.0, L=10.0, Ngrid=128, rng=2),
    dict(seed="triangular", Omega=-0.7, beta2=150.0, L=10.0, Ngrid=96, rng=4),
    dict(seed="tf", dimension="3D", Ngrid=32, L=5.0, beta2=50.0, rng=9),
    dict(seed="triangular", dimension="3D", Ngrid=32, L=6.0, beta2=200.0, Omega=0.8, rng=1),
]


@pytest.mark.parametrize("kw", LOCAL)
def test_local_seeds_match_oracle(kw):
    import becgpp_amd as B
    c = cfg(tf_cache=False, **kw)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    print(f"[seeds] {kw}", flush=True)
    got = B.make_seed(Gd, c).cpu().numpy()
    ref = O.make_seed(Go, c)
    assert _rel(got, ref) < 1e-12, _rel(got, ref)
    assert abs(np.sum(np.abs(got) ** 2) * O.dV(Go) - 1.0) < 1e-13


@pytest.mark.parametrize("kw", [
    dict(seed="tf", beta2=30.0, G_C=3.0, kernel="newton"),
    dict(seed="tf", dimension="quasi2D", beta2=30.0, G_C=2.0, kernel="q2d", l_z=0.5),
    dict(seed="triangular", beta2=100.0, G_C=5.0, kernel="newton", Omega=0.9, L=8.0, Ngrid=96),
])
def test_nonlocal_seeds_match_oracle(kw):
    import becgpp_amd as B
    c = cfg(tf_cache=False, **kw)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    got = B.make_seed(Gd, c).cpu().numpy()
    ref = O.make_seed(Go, c)
    assert _rel(got, ref) < 1e-8, _rel(got, ref)


def test_nonlocal_fixed_point_reports_its_iterations():
    """bec_tf_nonlocal's state: the iterations, the final change and mu equal
    the oracle's loop (thomasfermi.py:95-113) on the same grid."""
    import becgpp_amd as B
    c = cfg(beta2=30.0, G_C=3.0, kernel="newton", tf_cache=False)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    ref = O.extract_tf(c, Go)
    got = B.extract_tf(c, Gd)
    assert got["converged"] and ref["converged"]
    assert math.isclose(got["mu"], ref["mu"], rel_tol=1e-10)
    assert got["final_dn"] < 1e-9 and abs(got["final_dn"] - ref["final_dn"]) < 1e-11
    # a capped run stops unconverged after exactly that many iterations
    capped = B.thomasfermi._compressed(c, Gd, B.thomasfermi.tf_effective_potential(Gd, c), 3, 0.4)
    assert capped is not None and not capped["converged"] and capped["final_dn"] > 1e-9


def test_seed_vortices_wind_with_the_rotation():
    """The triangular seed's lattice: every vortex winds once in the sense of
    Omega (the reference's seeds.py:53), counted on the device."""
    import becgpp_amd as B
    for om in (0.9, -0.9):
        c = cfg(seed="triangular", Omega=om, beta2=200.0, L=10.0, Ngrid=128, seed_noise=0.0, tf_cache=False)
        G = B.make_grid(c)
        psi = B.make_seed(G, c)
        d = B.vortex_diagnostic(psi, G, om)
        assert d["Nv"] > 5
        assert (d["Nplus"], d["Nminus"]) == ((d["Nv"], 0) if om > 0 else (0, d["Nv"])), d


@pytest.mark.parametrize("i", range(11))
def test_device_seed_recipe_matches_reference(i):
    """bec_seed with the noise switched off against the reference's own
    initial states (tests/golden/seed_reference.npz): envelopes, windings and
    the triangular lattice to 1e-12."""
    import becgpp_amd as B
    from tests.golden.make_seed_golden import CASES, case_cfg
    from tests.test_oracle_seed_recipe import check_seed
    assert len(CASES) == 11
    c = cfg(tf_cache=False, **case_cfg(CASES[i]))
    print(f"[seeds] reference recipe case {i}: {CASES[i]}", flush=True)
    check_seed(B.make_seed(B.make_grid(c), c).cpu().numpy(), i)
