"""The oracle's operators against the reference's own applications
(tests/golden/make_operator_golden.py, run on the CPU where the reference is
mounted; only its output is committed): H psi, every energy component with
phi, the residual triple, L_z psi and the kernel convolution for the six
physics sets of the reference's FD audit, to 1e-12 relative (vectors: the
largest difference against the largest entry)."""
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_operator_golden import CASES, COMPONENTS, state
from tests.helpers import cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "operator_reference.npz")
TOL = 1e-12


def close_vec(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    assert np.abs(a - b).max() <= tol * np.abs(b).max(), np.abs(a - b).max() / np.abs(b).max()


def close(a, b, scale, tol=TOL):
    assert abs(float(a) - float(b)) <= tol * max(abs(float(b)), scale), (a, b)


def check_operators(i, H, comp, phi, res, lzp, conv):
    g = np.load(GOLD)
    escale = max(abs(float(g[f"{k}{i}"])) for k in COMPONENTS)       # components share one scale
    for k in COMPONENTS:
        close(comp[k], g[f"{k}{i}"], escale)
    if f"conv{i}" in g.files:
        close_vec(conv, g[f"conv{i}"])
        close_vec(phi, g[f"phi{i}"])
    close_vec(H, g[f"H{i}"])
    ra, mu, rr = g[f"res{i}"]
    close(res[0], ra, ra)
    close(res[1], mu, mu)
    close(res[2], rr, rr)
    close_vec(lzp, g[f"Lz{i}psi"])


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_operators_match_reference(i):
    c = cfg(**CASES[i])
    G = O.make_grid(c)
    p = state(CASES[i], seed=700 + i)
    comp = O.energy_components(p, G, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    conv = O.kernel_convolution(np.abs(p) ** 2, G) if G["kernel"] != "none" else None
    check_operators(i, O.apply_H(p, G, c), comp, comp["phi"], O.compute_residual(p, G, c), O.Lz_op(p, G), conv)
