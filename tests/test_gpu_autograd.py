"""Differentiation on the GPU: the energy's autograd backward (2 H psi dV)
against central finite differences and the oracle's H psi, and the
Hellmann-Feynman derivatives of converged ground-state energies against
central differences of the solver's energies."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import PHYSICS, cfg, random_state

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", PHYSICS)
def test_energy_backward_is_hamiltonian(kw):
    import becgpp_amd as B
    c = cfg(**kw)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    psi, rng = random_state(Go, 2)
    psi = psi / O.norm_of(psi, Go)
    args = (c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    p = torch.as_tensor(psi).cuda().requires_grad_(True)
    E = B.energy_t(p, Gd, *args)
    E.backward()
    g = p.grad.detach().cpu().numpy()
    # the backward equals the oracle's 2 H psi dV
    ref = 2.0 * O.dV(Go) * O.apply_H(psi, Go, c)
    assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()
    # and the directional derivative matches central differences of E
    v = (rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape)) * np.exp(-0.5 * Go["R2"])
    eps = 1e-5

    def En(x):
        return B.energy_components(torch.as_tensor(x).cuda(), Gd, *args)["E"].item()

    fd = (En(psi + eps * v) - En(psi - eps * v)) / (2 * eps)
    an = float(np.real(np.vdot(g, v)))
    assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)


def _converged_E(B, c, G, psi0):
    _, _, obs = B.ground_state(c, G, psi0, verbose=False)
    return obs


@pytest.mark.parametrize("param,kw,key", [
    ("beta2", dict(beta2=40.0), "dE_dbeta2"),
    ("beta3", dict(beta2=20.0, beta3=5.0), "dE_dbeta3"),
    ("G_C", dict(beta2=30.0, G_C=3.0, kernel="newton"), "dE_dG_C"),
    ("Omega", dict(beta2=30.0, Omega=0.6), "dE_dOmega"),
])
def test_hellmann_feynman(param, kw, key):
    """dE/dp of the converged ground state equals the expectation the
    sensitivities() report (Hellmann-Feynman), to the accuracy of a central
    difference of two converged energies."""
    import becgpp_amd as B
    c = cfg(res_tol=1e-9, maxit=20000, **kw)
    G = B.make_grid(c)
    psi0 = torch.as_tensor(O.noisy_gaussian(O.make_grid(c), 1e-3, seed=4)).cuda()
    psi, _, obs = B.ground_state(c, G, psi0, verbose=False)
    hf = B.sensitivities(psi, G, c)[key]
    h = 1e-3 * max(1.0, abs(c[param]))
    Ep = _converged_E(B, dict(c, **{param: c[param] + h}), G, psi)["E"]
    Em = _converged_E(B, dict(c, **{param: c[param] - h}), G, psi)["E"]
    fd = (Ep - Em) / (2 * h)
    assert abs(fd - hf) <= 1e-5 * max(abs(hf), 1e-3), (key, fd, hf)


@pytest.mark.parametrize("kw", [dict(beta2=40.0, beta3=3.0, Omega=0.4, G_C=2.0, kernel="newton"),
                                dict(dimension="3D", Ngrid=24, L=5, beta2=20.0, G_C=3.0, kernel="newton")])
def test_energy_parameter_gradients(kw):
    """Autograd derivatives of E with respect to beta2, beta3, Omega and G_C at
    fixed psi equal central differences of E in each parameter."""
    import becgpp_amd as B
    c = cfg(**kw)
    Go, Gd = O.make_grid(c), B.make_grid(c)
    psi, _ = random_state(Go, 9)
    psi = psi / O.norm_of(psi, Go)
    p = torch.as_tensor(psi).cuda()
    names = ("beta2", "beta3", "Omega", "G_C")
    par = [torch.tensor(float(c[k]), dtype=torch.float64, requires_grad=True) for k in names]
    E = B.energy_t(p, Gd, *par)
    E.backward()
    for i, k in enumerate(names):
        h = 1e-4 * max(1.0, abs(float(c[k])))
        vals = [float(c[n]) for n in names]
        up, dn = list(vals), list(vals)
        up[i] += h
        dn[i] -= h
        fd = (B.energy_components(p, Gd, *up)["E"].item() - B.energy_components(p, Gd, *dn)["E"].item()) / (2 * h)
        assert abs(par[i].grad.item() - fd) <= 1e-7 * max(1.0, abs(fd)), (k, par[i].grad.item(), fd)
