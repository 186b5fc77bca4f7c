"""Closed-form pins of the oracle's observables (reference operators.py:55-103).

On a Gaussian psi ~ exp(-r^2/(2 w^2)) (isotropic, or with widths a in-plane
and c along z in 3D), well inside the box, every moment and energy term has a
closed form: <r^2> = d w^2/2, Ekin = d/(4 w^2), Etrap = tc <r^2> (s = 2),
Econtact = beta2/2 (2 pi w^2)^(-d/2), Ethree = beta3/3 3^(-d/2) (pi w^2)^(-d),
R_perp_rms = a, R_z_rms = c/sqrt(2), oblateness = a/c, and the virial
combination 2 Ekin - s Etrap + d Econtact + 2 d Ethree (+ T_lr: -G_C/2 for the
log kernel at unit mass, Egrav for 1/r).  A wrong coefficient or a swapped
radius fails one of them.
"""
import math

import numpy as np
import pytest

import oracle as O
from tests.helpers import cfg


def gauss(G, a, c=None):
    C = G["coords"]
    q = (C[0] ** 2 + C[1] ** 2) / (2 * a * a)
    if G["ndim"] == 3:
        q = q + C[2] ** 2 / (2 * (c if c is not None else a) ** 2)
    p = np.exp(-q).astype(np.complex128)
    return p / O.norm_of(p, G)


@pytest.mark.parametrize("dim,N,L", [("2D", 64, 8.0), ("3D", 48, 8.0)])
def test_gaussian_observables_closed_forms(dim, N, L):
    w, b2, b3, tc = 1.1, 7.0, 3.0, 0.5
    c = cfg(dimension=dim, Ngrid=N, L=L, beta2=b2, beta3=b3, trap_coeff=tc, s=2)
    G = O.make_grid(c)
    d = G["ndim"]
    ob = O.observables(gauss(G, w), G, c)
    Ekin = d / (4 * w * w)
    Etrap = tc * d * w * w / 2
    Ec = 0.5 * b2 * (2 * math.pi * w * w) ** (-d / 2)
    Et = (b3 / 3.0) * 3.0 ** (-d / 2) * (math.pi * w * w) ** (-d)
    exp = dict(Ekin=Ekin, Etrap=Etrap, Econtact=Ec, Ethree=Et, E=Ekin + Etrap + Ec + Et,
               mu=Ekin + Etrap + 2 * Ec + 3 * Et, rrms=math.sqrt(d * w * w / 2),
               peak=(math.pi * w * w) ** (-d / 2),
               virial=2 * Ekin - 2 * Etrap + d * Ec + 2 * d * Et)
    for k, v in exp.items():
        assert abs(ob[k] - v) <= 1e-11 * max(1.0, abs(v)), (k, ob[k], v)
    assert ob["virial_lr_term"] == 0.0 and math.isnan(ob["E_bind"])
    vscale = Ekin + Etrap + Ec + Et
    assert abs(ob["virial_rel"] - abs(exp["virial"]) / max(1.0, vscale)) <= 1e-11


def test_anisotropic_radii_and_oblateness():
    a, cz = 1.3, 0.8
    c = cfg(dimension="3D", Ngrid=48, L=8.0, beta2=0.0)
    G = O.make_grid(c)
    ob = O.observables(gauss(G, a, cz), G, c)
    assert abs(ob["R_perp_rms"] - a) <= 1e-12
    assert abs(ob["R_z_rms"] - cz / math.sqrt(2.0)) <= 1e-12
    assert abs(ob["oblateness"] - a / cz) <= 1e-11
    assert abs(ob["rrms"] - math.sqrt(a * a + cz * cz / 2)) <= 1e-12


def test_2d_radii_are_nan_and_binding_energy_at_critical_rotation():
    c = cfg(dimension="2D", Ngrid=48, L=6.0, beta2=5.0, Omega=1.0, s=2)
    G = O.make_grid(c)
    ob = O.observables(gauss(G, 1.0), G, c)
    assert math.isnan(ob["R_perp_rms"]) and math.isnan(ob["R_z_rms"]) and math.isnan(ob["oblateness"])
    assert abs(ob["E_bind"] - (ob["E"] - 1.0)) <= 1e-15
    ob2 = O.observables(gauss(G, 1.0), G, dict(c, Omega=0.9))
    assert math.isnan(ob2["E_bind"])


@pytest.mark.parametrize("kernel", ["log", "newton"])
def test_long_range_virial_terms(kernel):
    gc = 2.5
    c = cfg(dimension="2D", Ngrid=48, L=6.0, beta2=5.0, G_C=gc, kernel=kernel)
    G = O.make_grid(c)
    ob = O.observables(gauss(G, 1.0), G, c)
    want = -0.5 * gc if kernel == "log" else ob["Egrav"]
    assert abs(ob["virial_lr_term"] - want) <= 1e-12 * max(1.0, abs(want))
    base = 2 * ob["Ekin"] - 2 * ob["Etrap"] + 2 * ob["Econtact"] + 4 * ob["Ethree"]
    assert abs(ob["virial"] - (base + want)) <= 1e-12 * max(1.0, abs(ob["virial"]))
