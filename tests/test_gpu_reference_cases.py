"""The reference's own test cases for the path (tests/test_audit.py and
tests/test_kernels.py of becGPP), run through this build on the GPU."""
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _cfg(B, **kw):
    """tests/test_audit.py:35-38."""
    return B.default_cfg(**dict(dict(mode="single", L=6, Ngrid=48, save_figs=False, save_ckpt=False), **kw))


def _small(B, **kw):
    """tests/test_kernels.py:21-26."""
    base = dict(mode="single", dimension="2D", s=2, Omega=0.0, beta2=50, beta3=0, L=8, Ngrid=64, seed="tf",
                maxit=6000, res_tol=1e-6, save_figs=False, save_ckpt=False)
    base.update(kw)
    return B.default_cfg(**base)


@pytest.mark.parametrize("m", [1, -1, 2])
def test_vortex_count_of_imprinted_vortex(m):
    """tests/test_audit.py:122-129."""
    import becgpp_amd as B
    c = _cfg(B, Ngrid=64, L=6)
    G = B.make_grid(c)
    psi = (torch.exp(-0.5 * G["R2"] / 4.0) * torch.tanh(torch.sqrt(G["R2"]) / 0.5)).to(torch.complex128)
    psi = psi * torch.exp(1j * m * torch.atan2(G["Y"] + 1e-9, G["X"] + 1e-9))
    assert B.vortex_diagnostic(psi, G, 0.5)["Nnet"] == m


def test_lll_weight_of_lll_state():
    """tests/test_audit.py:132-136."""
    import becgpp_amd as B
    G = B.make_grid(_cfg(B, Ngrid=96, L=8))
    z = (G["X"] + 1j * G["Y"]).to(torch.complex128)
    assert abs(B.lll_weight(z ** 3 * torch.exp(-0.5 * G["R2"]), G) - 1.0) < 1e-8


def test_lll_weight_negative_rotation():
    """tests/test_audit.py:221-227: through the diagnostic bundle."""
    import becgpp_amd as B
    c = _cfg(B, Omega=-0.9, beta2=0.0, Ngrid=96, L=8)
    G = B.make_grid(c)
    zb = (G["X"] - 1j * G["Y"]).to(torch.complex128)
    psi = zb ** 3 * torch.exp(-0.5 * G["R2"])
    psi = psi / torch.sqrt((psi.abs() ** 2).sum() * B.dV(G))
    assert abs(B.diagnostics(psi, G, c)["w_LLL"] - 1.0) < 1e-6


def test_triangular_seed_follows_rotation_sign():
    """tests/test_audit.py:214-218."""
    import becgpp_amd as B
    from becgpp_amd.seeds import triangular_seed
    c = _cfg(B, Omega=-0.9, beta2=200, L=10, Ngrid=96, seed="triangular")
    G = B.make_grid(c)
    Lz = B.energy_components(triangular_seed(G, c), G, 200.0, 0.0, -0.9, 0.0)["Lz"].item()
    assert Lz < 0.0, Lz


def test_tf_trapped_pure_quintic():
    """tests/test_audit.py:247-249."""
    import becgpp_amd as B
    c = _cfg(B, beta2=0.0, beta3=50.0)
    assert B.extract_tf(c, B.make_grid(c)) is not None


def test_log_kernel_virial():
    """tests/test_kernels.py:30-36: for K = -ln r the virial term is -G_C/2."""
    import becgpp_amd as B
    cfg = _small(B, G_C=5.0, kernel="log")
    psi, G, obs = B.ground_state(cfg, verbose=False)
    d = B.diagnostics(psi, G, cfg)
    assert abs(d["virial_lr_term"] + 2.5) < 1e-6
    assert d["virial_rel"] < 5e-3, d["virial_rel"]


@pytest.mark.parametrize("kw", [dict(kernel="newton"), dict(dimension="quasi2D", kernel="q2d", l_z=0.5)])
def test_long_range_virial(kw):
    """tests/test_kernels.py:39-51: newton 2D and q2d ground states satisfy
    the virial identity."""
    import becgpp_amd as B
    cfg = _small(B, G_C=5.0, **kw)
    psi, G, obs = B.ground_state(cfg, verbose=False)
    assert B.diagnostics(psi, G, cfg)["virial_rel"] < 5e-3


def test_quasi2d_auto_kernel():
    """tests/test_kernels.py:66-72: kernel 'auto' resolves to q2d in quasi2D
    with l_z > 0, newton with l_z = 0, log in 2D."""
    import becgpp_amd as B
    assert B.make_grid(B.default_cfg(dimension="quasi2D", G_C=1.0, l_z=0.3, Ngrid=32))["kernel"] == "q2d"
    assert B.make_grid(B.default_cfg(dimension="quasi2D", G_C=1.0, l_z=0.0, Ngrid=32))["kernel"] == "newton"
    assert B.make_grid(B.default_cfg(dimension="2D", G_C=1.0, Ngrid=32))["kernel"] == "log"


def test_single_ground_state_converges():
    """tests/test_smoke.py:24-36: grid -> operators -> solver -> diagnostics ->
    Thomas-Fermi on a small grid."""
    import math
    import becgpp_amd as B
    cfg = B.default_cfg(mode="single", dimension="2D", s=2, Omega=0.0, beta2=100, beta3=0, G_C=0, kernel="none",
                        L=10, Ngrid=96, seed="tf", maxit=4000, res_tol=1e-4, save_figs=False, save_ckpt=False)
    G = B.make_grid(cfg)
    psi, G, obs = B.ground_state(cfg, G=G, verbose=False)
    assert obs["converged"] and obs["resid_rel"] < 2e-4
    d = B.diagnostics(psi, G, cfg)
    assert d["tf_kind"] == "repulsive_parabola"
    assert math.isfinite(d["E"]) and d["virial_rel"] < 1e-2


def test_flattop_tf_exact():
    """tests/test_smoke.py:39-45: the self-bound flat top rho0 = -3 b2 / (4 b3)."""
    import becgpp_amd as B
    cfg = B.default_cfg(dimension="2D", s=2, Omega=1.0, beta2=-1.0, beta3=1.0, G_C=0, kernel="none", L=20,
                        Ngrid=192)
    tf = B.extract_tf(cfg, B.make_grid(cfg))
    assert tf is not None and tf["kind"] == "local_flattop"
    assert abs(tf["rho0"] - 0.75) < 1e-9
