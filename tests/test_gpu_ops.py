"""Operator parity on the GPU: each libbecgpp_hip operator against the oracle
on the same seeded input, to 1e-12 relative (max-norm, relative to the
largest reference value)."""
import math

import numpy as np
import pytest

import oracle as O
from tests.helpers import PHYSICS, cfg, random_state, relerr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(kw):
    import becgpp_amd as B
    c = cfg(**kw)
    Go = O.make_grid(c)
    Gd = B.make_grid(c)
    return c, Go, Gd


def _dev(a):
    return torch.as_tensor(a).to("cuda")


@pytest.mark.parametrize("kw", [dict(Ngrid=48), dict(Ngrid=64, L=8), dict(Ngrid=96), dict(Ngrid=128, L=9),
                                dict(Ngrid=384, L=12), dict(dimension="3D", Ngrid=24, L=5),
                                dict(dimension="3D", Ngrid=32, L=6),
                                dict(Ngrid=80, L=8), dict(Ngrid=160, L=10), dict(Ngrid=120, L=9),
                                dict(dimension="3D", Ngrid=40, L=6)])
def test_lz_op(kw):
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 3)
    ref = O.Lz_op(psi, Go)
    got = B.Lz_op(_dev(psi), Gd).cpu().numpy()
    assert relerr(got, ref) < TOL


@pytest.mark.parametrize("kw", PHYSICS)
def test_apply_H(kw):
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 1)
    ref = O.apply_H(psi, Go, c)
    got = B.apply_H(_dev(psi), Gd, c).cpu().numpy()
    assert relerr(got, ref) < TOL


# line lengths whose transforms use every stage kind: N = 384 (radices 8, 6, 8
# through LDS on 64 threads per line), conv M = 384 from N = 192, N = 160
# (8, 5, 4), and the register-resident 2^a lengths elsewhere in this file
@pytest.mark.parametrize("kw", [dict(dimension="2D", Ngrid=384, L=12, beta2=200, Omega=0.9),
                                dict(dimension="2D", Ngrid=192, L=10, beta2=30, G_C=3, kernel="newton",
                                     Omega=0.3),
                                dict(dimension="2D", Ngrid=160, L=10, beta2=30, G_C=2, kernel="log")])
def test_apply_H_long_lines(kw):
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 5)
    ref = O.apply_H(psi, Go, c)
    got = B.apply_H(_dev(psi), Gd, c).cpu().numpy()
    assert relerr(got, ref) < TOL


# grids whose transforms need the radices 7, 11 and 13 (the sizes auto_grid
# picks at 224, 352 and 416, scaled down), including odd N; the conv lines are
# M = 2N.  H, the energy components and the preconditioner against the oracle.
@pytest.mark.parametrize("kw", [dict(dimension="2D", Ngrid=224, L=10, beta2=100, Omega=0.5),
                                dict(dimension="2D", Ngrid=88, L=8, beta2=30, G_C=3, kernel="newton", Omega=0.3),
                                dict(dimension="2D", Ngrid=104, L=8, beta2=30, G_C=2, kernel="log"),
                                dict(dimension="2D", Ngrid=63, L=6, beta2=30, beta3=5),
                                dict(dimension="2D", Ngrid=65, L=6, beta2=30, G_C=2, kernel="newton"),
                                dict(dimension="quasi2D", Ngrid=77, L=6, beta2=30, G_C=2, kernel="q2d", l_z=0.5),
                                dict(dimension="3D", Ngrid=28, L=5, beta2=20, beta3=2, G_C=3, kernel="newton",
                                     Omega=0.5)])
def test_odd_radix_grids(kw):
    import becgpp_amd as B
    from becgpp_amd.operators import preconditioner
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 4)
    psi = psi / O.norm_of(psi, Go)
    ref = O.apply_H(psi, Go, c)
    got = B.apply_H(_dev(psi), Gd, c).cpu().numpy()
    assert relerr(got, ref) < TOL
    e_ref = O.energy_components(psi, Go, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    e_got = B.energy_components(_dev(psi), Gd, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    scale = max(abs(e_ref[k]) for k in ("Ekin", "Etrap", "Econtact", "Ethree", "Egrav", "Lz"))
    for k in ("E", "Ekin", "Etrap", "Econtact", "Ethree", "Erot", "Egrav", "Lz"):
        assert abs(e_got[k].item() - e_ref[k]) <= TOL * scale, (k, e_got[k].item(), e_ref[k])
    pre_ref, sig_ref = O.preconditioner(ref, psi, Go, c)
    pre_got, sig_got = preconditioner(_dev(ref), _dev(psi), Gd, c)
    assert abs(sig_got - sig_ref) <= 1e-14 * abs(sig_ref)
    assert relerr(pre_got.cpu().numpy(), pre_ref) < TOL


@pytest.mark.parametrize("kw", PHYSICS)
def test_energy_components(kw):
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 1)
    ref = O.energy_components(psi, Go, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    got = B.energy_components(_dev(psi), Gd, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    scale = max(abs(ref[k]) for k in ("Ekin", "Etrap", "Econtact", "Ethree", "Egrav", "Lz"))
    for k in ("E", "Ekin", "Etrap", "Econtact", "Ethree", "Erot", "Egrav", "Lz"):
        assert abs(got[k].item() - ref[k]) <= TOL * scale, (k, got[k].item(), ref[k])
    assert relerr(got["phi"].cpu().numpy(), ref["phi"]) < TOL if np.abs(ref["phi"]).max() > 0 else True


@pytest.mark.parametrize("kw", PHYSICS[1:5] + [dict(dimension="2D", G_C=2.0, kernel="newton", Ngrid=96, L=8),
                                               dict(dimension="3D", G_C=2.0, kernel="newton", Ngrid=32, L=6),
                                               dict(dimension="2D", G_C=2.0, kernel="log", Ngrid=80, L=8),
                                               dict(dimension="3D", G_C=2.0, kernel="newton", Ngrid=20, L=5)])
def test_kernel_convolution(kw):
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    rng = np.random.default_rng(2)
    a = rng.random(Go["R2"].shape)
    ref = O.kernel_convolution(a, Go)
    got = B.kernel_convolution(_dev(a), Gd).cpu().numpy()
    assert relerr(got, ref) < TOL


def test_q2d_virial_convolution():
    import becgpp_amd as B
    c, Go, Gd = _pair(PHYSICS[3])
    rng = np.random.default_rng(5)
    a = rng.random(Go["R2"].shape)
    ref = O.kernel_convolution(a, Go, "W")
    got = B.kernel_convolution(_dev(a), Gd, "W").cpu().numpy()
    assert relerr(got, ref) < TOL


@pytest.mark.parametrize("kw", PHYSICS)
def test_residual(kw):
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 1)
    psi = psi / O.norm_of(psi, Go)
    ra, mu, rr = O.compute_residual(psi, Go, c)
    ga, gmu, grr = B.compute_residual(_dev(psi), Gd, c)
    assert abs(gmu - mu) <= TOL * max(1.0, abs(mu))
    assert abs(ga - ra) <= 1e-11 * ra


@pytest.mark.parametrize("kw", PHYSICS)
def test_preconditioner(kw):
    from becgpp_amd.operators import preconditioner
    c, Go, Gd = _pair(kw)
    psi, rng = random_state(Go, 1)
    psi = psi / O.norm_of(psi, Go)
    g = O.apply_H(psi, Go, c)
    zr, sr = O.preconditioner(g, psi, Go, c)
    zg, sg = preconditioner(_dev(g), _dev(psi), Gd, c)
    if Gd["kernel"] == "none":
        assert sg == sr                        # identical inputs: the same element is selected
    else:
        assert abs(sg - sr) <= 1e-14 * abs(sr)  # phi enters the local potential at round-off
    assert relerr(zg.cpu().numpy(), zr) < TOL


def test_shift_is_exact_lower_median():
    """The device selection returns exactly the lower median of the masked
    local potential, on a density with many bulk points (candidate path) and
    on a flat-top density (all bulk values equal)."""
    from becgpp_amd.operators import preconditioner
    for kw in (dict(Ngrid=128, L=8, beta2=200.0), dict(Ngrid=64, L=8, beta2=100.0, s=0)):
        c, Go, Gd = _pair(kw)
        psi, _ = random_state(Go, 11)
        if kw.get("s") == 0:
            psi = np.where(Go["R2"] < 9.0, 1.0 + 0j, 0.0 + 0j)
        psi = psi / O.norm_of(psi, Go)
        _, sr = O.preconditioner(psi, psi, Go, c)
        _, sg = preconditioner(_dev(psi), _dev(psi), Gd, c)
        assert sg == sr


@pytest.mark.parametrize("kw", PHYSICS)
def test_mu_equals_expectation_of_H(kw):
    """operators.observables mu vs <psi,H psi> (reference tests/test_audit.py:99-106)."""
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, _ = random_state(Go, 1)
    psi = psi / O.norm_of(psi, Go)
    p = _dev(psi)
    Hp = B.apply_H(p, Gd, c)
    mu_H = ((p.conj() * Hp).real.sum() * B.dV(Gd)).item()
    assert abs(B.observables(p, Gd, c)["mu"] - mu_H) < 1e-10 * max(1.0, abs(mu_H))


@pytest.mark.parametrize("kw", PHYSICS)
def test_energy_gradient_matches_hamiltonian(kw):
    """Central FD of E vs 2 Re<H psi, v> dv (reference tests/test_audit.py:68-85)."""
    import becgpp_amd as B
    c, Go, Gd = _pair(kw)
    psi, rng = random_state(Go, 1)
    env = np.exp(-0.5 * Go["R2"])
    v = env * rng.standard_normal(env.shape) + 1j * env * rng.standard_normal(env.shape)

    def E(p):
        return B.energy_components(_dev(p), Gd, c["beta2"], c["beta3"], c["Omega"], c["G_C"])["E"].item()
    eps = 1e-5
    fd = (E(psi + eps * v) - E(psi - eps * v)) / (2 * eps)
    an = 2.0 * ((B.apply_H(_dev(psi), Gd, c).conj() * _dev(v)).real.sum() * B.dV(Gd)).item()
    assert abs(fd - an) / abs(an) < 1e-7


def test_norm_of():
    import becgpp_amd as B
    c, Go, Gd = _pair(dict(Ngrid=64, L=8))
    psi, _ = random_state(Go, 4)
    assert abs(B.norm_of(_dev(psi), Gd).item() - O.norm_of(psi, Go)) < 1e-14 * O.norm_of(psi, Go)
    _ = math


def test_kernel_none_grid_ignores_G_C():
    """A grid whose kernel resolves to "none" never applies a long-range term,
    even when a grid of the same geometry loaded a spectrum into the shared
    plan before (the reference decides from G["kernel"], operators.py:21-22)."""
    import becgpp_amd as B
    c_lr = cfg(G_C=3.0, kernel="newton", beta2=30.0)
    c_none = cfg(G_C=0.0, kernel="newton", beta2=30.0)
    G_lr, G_none = B.make_grid(c_lr), B.make_grid(c_none)
    assert G_lr["kernel"] == "newton" and G_none["kernel"] == "none"
    Go = O.make_grid(c_none)
    psi, _ = random_state(Go, 6)
    B.energy_components(_dev(psi), G_lr, 30.0, 0.0, 0.0, 3.0)
    e_gc = B.energy_components(_dev(psi), G_none, 30.0, 0.0, 0.0, 3.0)
    e_0 = B.energy_components(_dev(psi), G_none, 30.0, 0.0, 0.0, 0.0)
    eo = O.energy_components(psi, Go, 30.0, 0.0, 0.0, 3.0)
    assert e_gc["E"].item() == e_0["E"].item()
    assert abs(e_gc["E"].item() - eo["E"]) <= TOL * abs(eo["E"])
    h = B.apply_H(_dev(psi), G_none, dict(c_none, G_C=3.0)).cpu().numpy()
    assert relerr(h, O.apply_H(psi, Go, dict(c_none, G_C=3.0))) < TOL


@pytest.mark.parametrize("kw", [dict(dimension="2D", Ngrid=96, L=8, beta2=30, G_C=3, kernel="newton", Omega=0.3),
                                dict(dimension="3D", Ngrid=32, L=6, beta2=20, G_C=3, kernel="newton")])
def test_autotuned_schedule_keeps_parity(kw):
    """The autotuner may pick any stage order and lines per workgroup: the
    operators stay within the parity bound under the schedule it keeps."""
    import becgpp_amd as B
    from becgpp_amd.interactions import ensure_kernel
    c, Go, Gd = _pair(kw)
    pl = ensure_kernel(Gd, "K")
    pl.schedule = None
    sched = pl.autotune(2)
    assert sched["step_passes_ms"] > 0 and sched["step_passes_ms"] <= sched["default_ms"] * 1.0000001
    assert "stages [" in sched["schedule"]
    psi, _ = random_state(Go, 8)
    ref = O.apply_H(psi, Go, c)
    got = B.apply_H(_dev(psi), Gd, c).cpu().numpy()
    assert relerr(got, ref) < TOL
    e_ref = O.energy_components(psi, Go, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    e_got = B.energy_components(_dev(psi), Gd, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    assert abs(e_got["E"].item() - e_ref["E"]) <= TOL * max(1.0, abs(e_ref["E"]))


def test_long_range_refused_beyond_padded_limit():
    """A grid whose padded length 2N exceeds 1024 solves without long range
    (the N = 1024 ladder) but refuses a kernel spectrum with a ValueError."""
    import becgpp_amd as B
    c = cfg(dimension="2D", Ngrid=576, L=8.0, beta2=30.0, G_C=2.0, kernel="newton")
    G = B.make_grid(c)
    psi = torch.ones((576, 576), dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="padded length"):
        B.apply_H(psi, G, c)
    c0 = cfg(dimension="2D", Ngrid=576, L=8.0, beta2=30.0)
    ref = O.apply_H(np.ones((576, 576), complex), O.make_grid(c0), c0)
    assert relerr(B.apply_H(psi, B.make_grid(c0), c0).cpu().numpy(), ref) < TOL
