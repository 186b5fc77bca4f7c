"""Pins of the oracle's diagnostic functions by the bundles and resampled
states the reference computed itself (tests/golden/make_diag_golden.py, run
on the CPU where the reference is mounted; only its output is committed).

Eight states (Gaussian envelopes with planted vortices of either sign and
seeded noise; contact, newton, log and q2d kernels, rotation, a self-bound
cubic-quintic, a quartic trap at Omega = 1 and two 3D grids): every key of
the reference's diagnostics (diagnostics.py:18-47) -- the observables with
their virial terms, rrms, peak, 3D radii and oblateness (operators.py:55-103),
the mass radii and 1% edge (fields.py), the vortex counts on the plane or the
3D midplane and the LLL weight (vortex.py), and the Thomas-Fermi fields --
from the oracle's functions composed as the reference composes them; and
three resampled states (grid.py:140-161)."""
import math
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_diag_golden import CASES, KEYS, RESAMPLE, state
from tests.helpers import cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "diag_reference.npz")
INT_KEYS = ("Nv", "Nplus", "Nminus", "Nnet")
RADII = ("R50", "R90", "R99", "R_1pct", "points_per_R90")


def case_state(i, G):
    kw, zeros = CASES[i]
    psi = state(kw, zeros, seed=500 + i)
    return psi / math.sqrt(float((np.abs(psi) ** 2).sum()) * O.dV(G))


def oracle_bundle(psi, G, c):
    """diagnostics.py:18-47 from the oracle's functions"""
    d = dict(O.observables(psi, G, c))
    rho = np.abs(psi) ** 2
    d["R50"], d["R90"], d["R99"] = O.mass_quantiles(rho, G, (0.50, 0.90, 0.99))
    d["R_1pct"] = O.droplet_radius(rho, G)
    Om = float(c.get("Omega", 0.0))
    if int(G["ndim"]) == 2:
        d.update(O.vortex_count_2d(psi, G, Om))
    else:                                              # the z = 0 midplane (vortex.py:51-59)
        G2 = O.make_grid(dict(c, dimension="2D", kernel="none", G_C=0.0))
        d.update(O.vortex_count_2d(psi[:, :, int(G["N"]) // 2], G2, Om))
    lll_ok = (int(G["ndim"]) == 2 and abs(float(G["s"]) - 2.0) < 1e-12
              and abs(float(G.get("trap_coeff", 0.5)) - 0.5) < 1e-12 and abs(Om) > 0.0)
    d["w_LLL"] = O.lll_weight(psi, G, 60, Om) if lll_ok else float("nan")
    d["lll_param"] = abs(float(c["beta2"])) * d["peak"] / (1.0 + abs(Om)) if lll_ok else float("nan")
    tf = O.extract_tf(c, G)
    nan = float("nan")
    d["tf_kind"] = tf["kind"] if tf else ""
    for k in ("R90", "R99", "mu", "rho0"):
        d["tf_" + k] = float(tf.get(k, nan)) if tf else nan
    d["points_per_R90"] = d["R90"] / float(G["dx"]) if math.isfinite(d["R90"]) else nan
    return d


def check_bundle(d, i, dx, grid_step_radii=False):
    """grid_step_radii: radii picked by a cumulative sum or a bin may be one
    grid step apart where sums summed in another order tie at rounding size
    (the device path); the oracle's agree to rounding"""
    g = np.load(GOLD)
    assert str(d["tf_kind"]) == str(g[f"tf_kind{i}"])
    for k in KEYS:
        ref, got = float(g[f"{k}{i}"]), float(d[k])
        if math.isnan(ref):
            assert math.isnan(got), (k, got)
        elif k in INT_KEYS:
            assert got == ref, (k, got, ref)
        elif k in RADII or k in ("tf_R90", "tf_R99"):
            step = (1.0 if k == "points_per_R90" else dx) if grid_step_radii else 0.0
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)) + step, (k, got, ref)
        elif k.startswith("tf_"):
            assert abs(got - ref) <= 1e-9 * max(1.0, abs(ref)), (k, got, ref)
        else:
            assert abs(got - ref) <= 1e-11 * max(1.0, abs(ref)), (k, got, ref)


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_bundle_matches_reference(i):
    c = cfg(tf_cache=False, **CASES[i][0])
    G = O.make_grid(c)
    check_bundle(oracle_bundle(case_state(i, G), G, c), i, float(G["dx"]))


@pytest.mark.parametrize("j", range(len(RESAMPLE)))
def test_oracle_resample_matches_reference(j):
    i, L2, N2 = RESAMPLE[j]
    c = cfg(**CASES[i][0])
    G = O.make_grid(c)
    G2 = O.make_grid(dict(c, L=L2, Ngrid=N2))
    got = O.resample_state(case_state(i, G), G, G2)
    ref = np.load(GOLD)[f"resampled{j}"]
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
