"""Thomas-Fermi envelopes on the GPU (the chemical potential's bisection in
bec_tf_mu, the nonlocal fixed point with the HIP convolution) against the
oracle's restatement of thomasfermi.py:33-156."""
import math

import numpy as np
import pytest

import oracle as O
from tests.helpers import cfg

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", [
    dict(beta2=100.0),                                            # repulsive parabola
    dict(beta2=30.0, beta3=5.0, Omega=0.4),                       # local trapped repulsive
    dict(trap_coeff=0.0, beta2=-60.0, beta3=1000.0, L=8.0),       # flat-top droplet (no bisection)
    dict(beta2=30.0, G_C=3.0, kernel="newton"),                   # nonlocal compressed, 1/r
    dict(dimension="quasi2D", beta2=30.0, G_C=2.0, kernel="q2d", l_z=0.5),
    dict(dimension="3D", Ngrid=24, L=5, beta2=20.0, G_C=3.0, kernel="newton"),
])
def test_extract_tf_matches_oracle(kw):
    import becgpp_amd as B
    c = cfg(tf_cache=False, **kw)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    ref = O.extract_tf(c, Go)
    got = B.extract_tf(c, Gd)
    assert (ref is None) == (got is None)
    if ref is None:
        return
    assert got["kind"] == ref["kind"]
    assert math.isclose(got["mu"], ref["mu"], rel_tol=1e-9, abs_tol=1e-12), (got["mu"], ref["mu"])
    n_got = got["n"].cpu().numpy()
    dv = O.dV(Go)
    assert np.abs(n_got - ref["n"]).sum() * dv < 1e-8
    for k in ("R90", "R99"):
        assert abs(got[k] - ref[k]) <= 1e-9 + 2.0 * float(Go["dx"]), (k, got[k], ref[k])


def test_tf_bisection_matches_host_bisection():
    """The device bisection takes the reference's steps: the same widening of
    hi and the same 90 halvings as the oracle's bisect_mu_for_norm."""
    import becgpp_amd as B
    from becgpp_amd.thomasfermi import _bisect_mu_for_norm
    c = cfg(beta2=50.0, beta3=2.0)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    Veff = O.tf_effective_potential(Go, c)
    for lo, hi in ((float(Veff.min()), float(Veff.max()) + 53.0), (float(Veff.min()), float(Veff.min()) + 1e-3)):
        ref = O.bisect_mu_for_norm(lambda m: O.invert_local(m, Veff, 50.0, 2.0), Go, lo, hi)
        got = _bisect_mu_for_norm(torch.as_tensor(Veff).cuda(), 50.0, 2.0, Gd, lo, hi)
        assert math.isclose(got, ref, rel_tol=1e-13), (got, ref)


@pytest.mark.parametrize("i", range(8))
def test_extract_tf_matches_reference_envelopes(i):
    """The device envelopes against those the reference computed itself
    (tests/golden/tf_reference.npz): regime, mu, radii and density."""
    import becgpp_amd as B
    from tests.golden.make_tf_golden import CASES, case_cfg
    from tests.test_oracle_tf import check_against_reference
    assert len(CASES) == 8
    c = cfg(tf_cache=False, **case_cfg(CASES[i]))
    Gd = B.make_grid(c)
    print(f"[tf] reference envelope case {i}: {CASES[i]}", flush=True)
    check_against_reference(B.extract_tf(c, Gd), i, float(Gd["dx"]) ** int(Gd["ndim"]), float(Gd["dx"]),
                            lambda tf: tf["n"].cpu().numpy())
