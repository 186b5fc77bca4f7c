"""Solver parity on the GPU: the device loop against the oracle's loop from the
same psi0 (converged E, mu, L_z to 1e-8 relative, density to 1e-6 in L2), the
reference's exact limits and its stop-rule regressions."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import cfg

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _both(c, psi0):
    import becgpp_amd as B
    Go = O.make_grid(c)
    Gd = B.make_grid(c)
    po, _, oo = O.ground_state(c, Go, psi0)
    pd, _, od = B.ground_state(c, Gd, torch.as_tensor(psi0).cuda(), verbose=False)
    return Go, po, oo, pd.cpu().numpy(), od


def _dens_l2(a, b, G):
    ra, rb = np.abs(a) ** 2, np.abs(b) ** 2
    return float(np.sqrt(((ra - rb) ** 2).sum()) / np.sqrt((rb ** 2).sum()))


@pytest.mark.parametrize("kw", [
    dict(beta2=30.0, res_tol=1e-8),
    dict(beta2=30.0, beta3=5.0, Omega=0.4, res_tol=1e-8),
    dict(beta2=30.0, G_C=3.0, kernel="newton", res_tol=1e-8),
    dict(beta2=30.0, G_C=3.0, kernel="log", res_tol=1e-8, cg_beta="pr"),
    dict(dimension="quasi2D", beta2=30.0, G_C=3.0, kernel="q2d", l_z=0.5, res_tol=1e-8),
    dict(dimension="3D", beta2=20.0, G_C=3.0, kernel="newton", Ngrid=24, L=5, res_tol=1e-8),
])
def test_converged_state_matches_oracle(kw):
    c = cfg(**kw)
    Go = O.make_grid(c)
    psi0 = O.noisy_gaussian(Go, 1e-3, seed=7)
    Go, po, oo, pd, od = _both(c, psi0)
    for k in ("E", "mu", "Lz"):
        assert abs(od[k] - oo[k]) <= 1e-8 * max(1.0, abs(oo[k])), (k, od[k], oo[k])
    assert _dens_l2(pd, po, Go) < 1e-6
    # res_tol=1e-8 is at the FP64 residual floor: either side may stop by stall just above it
    assert od["resid_rel"] < 1e-7 and oo["resid_rel"] < 1e-7


def test_first_iterations_track_oracle():
    """Before round-off separates them, the device and oracle trajectories agree."""
    c = cfg(beta2=50.0, maxit=15, res_tol=1e-14, record_trace=True)
    Go = O.make_grid(c)
    psi0 = O.noisy_gaussian(Go, 1e-3, seed=3)
    Eo = []
    O.ground_state(c, Go, psi0, callback=lambda it, E, r, t: Eo.append((it, E, r)))
    import becgpp_amd as B
    _, _, od = B.ground_state(c, B.make_grid(c), torch.as_tensor(psi0).cuda(), verbose=False)
    tr = od["trace"][1:]
    assert len(tr) == len(Eo)
    for (t, it, E, r), (ito, Eor, ro) in zip(tr, Eo):
        assert it == ito + 1
        assert abs(E - Eor) <= 1e-11 * abs(Eor)
        assert abs(r - ro) <= 1e-6 * ro


@pytest.mark.parametrize("kw,exact", [
    (dict(dimension="2D"), 1.0),
    (dict(dimension="3D", Ngrid=32), 1.5),
    (dict(dimension="2D", trap_coeff=0.125), 0.5),
    (dict(dimension="2D", Omega=0.6), 1.0),
])
def test_noninteracting_ground_state_energy(kw, exact):
    """Exact harmonic-oscillator energies (reference tests/test_audit.py:109-119)."""
    import becgpp_amd as B
    L = 9.0 if kw.get("trap_coeff") else 6.0
    kw = dict(dict(Ngrid=48), **kw)
    c = B.default_cfg(mode="single", L=L, beta2=0.0, res_tol=1e-7, maxit=600, **kw)
    _, _, obs = B.ground_state(c, verbose=False)
    assert abs(obs["E"] - exact) < 1e-7, obs["E"]


def test_solver_stops_when_line_search_stalls():
    """reference tests/test_audit.py:165-172, with one Armijo trial per loop
    step: each trial's energy is summed over the grid, as in the reference,
    and the search stalls once the energy cannot resolve the decrease."""
    import becgpp_amd as B
    c = B.default_cfg(mode="single", L=6, Ngrid=48, beta2=0.0, res_tol=1e-12, energy_tol=1e-9,
                      conv_window=200, maxit=1000, line_search="trial")
    _, _, obs = B.ground_state(c, verbose=False)
    assert abs(obs["E"] - 1.0) < 1e-10
    assert obs["stop_reason"] in ("line_search_stalled", "energy_converged"), obs["stop_reason"]
    assert obs["iters"] < 400, obs["iters"]


def test_polynomial_search_at_an_unreachable_tolerance_stops():
    """The same case with the polynomial line search: a trial's energy comes
    from a few operations on carried sums, whose rounding is far below that
    of a sum over the grid, so the search keeps resolving the decrease where
    the reference's stalls and the solve may reach res_tol 1e-12; it must
    still stop early, with the exact energy."""
    import becgpp_amd as B
    c = B.default_cfg(mode="single", L=6, Ngrid=48, beta2=0.0, res_tol=1e-12, energy_tol=1e-9,
                      conv_window=200, maxit=1000)
    _, _, obs = B.ground_state(c, verbose=False)
    assert abs(obs["E"] - 1.0) < 1e-10
    assert obs["stop_reason"] in ("converged", "line_search_stalled", "energy_converged"), obs["stop_reason"]
    assert obs["iters"] < 400, obs["iters"]


def test_iteration_count():
    """reference tests/test_audit.py:252-255."""
    import becgpp_amd as B
    c = B.default_cfg(mode="single", L=6, Ngrid=48, beta2=50.0, res_tol=1e-14, maxit=5)
    _, _, obs = B.ground_state(c, verbose=False)
    assert obs["iters"] == 5


@pytest.mark.parametrize("kw", [dict(beta2=30.0, Omega=0.4), dict(beta2=30.0, G_C=3.0, kernel="newton"),
                                dict(dimension="3D", beta2=20.0, G_C=3.0, kernel="newton", Ngrid=24, L=5)])
def test_device_and_host_loops_agree(kw):
    """The loop's scalar logic on the device (default) and on the host: the
    same trajectory to round-off over the first iterations, and the same
    converged state."""
    import becgpp_amd as B
    c = cfg(res_tol=1e-8, record_trace=True, **kw)
    Gd = B.make_grid(c)
    psi0 = torch.as_tensor(O.noisy_gaussian(O.make_grid(c), 1e-3, seed=9)).cuda()
    _, _, od = B.ground_state(c, Gd, psi0, verbose=False)
    _, _, oh = B.ground_state(dict(c, host_loop=True), Gd, psi0, verbose=False)
    for (_, i1, E1, _), (_, i2, E2, _) in list(zip(od["trace"], oh["trace"]))[:10]:
        assert i1 == i2 and abs(E1 - E2) <= 1e-12 * abs(E2)
    # two converged states at res_tol 1e-8: E agrees to second order, mu and
    # L_z to first order in the residual
    assert abs(od["E"] - oh["E"]) <= 1e-9 * max(1.0, abs(oh["E"])), (od["E"], oh["E"])
    for k in ("mu", "Lz"):
        assert abs(od[k] - oh[k]) <= 1e-8 * max(1.0, abs(oh[k])), (k, od[k], oh[k])


@pytest.mark.parametrize("kw", [dict(beta2=30.0, Omega=0.4, Ngrid=56, res_tol=1e-8),
                                dict(beta2=30.0, G_C=3.0, kernel="newton", Ngrid=63, res_tol=1e-8),
                                dict(dimension="3D", beta2=20.0, G_C=3.0, kernel="newton", Ngrid=28, L=5,
                                     res_tol=1e-8)])
def test_converged_state_on_odd_radix_grids(kw):
    """Grids whose transforms take the radices 7 and 11 / odd N (56 = 8 7,
    63 = 7 9 with conv lines 126, 28 = 4 7 with 56): the converged state
    matches the oracle; the FP32 warm phase, not built for these radices, is
    declined and the solve runs in FP64 throughout."""
    import becgpp_amd as B
    c = cfg(**kw)
    Go = O.make_grid(c)
    psi0 = O.noisy_gaussian(Go, 1e-3, seed=7)
    po, _, oo = O.ground_state(c, Go, psi0)
    pd, _, od = B.ground_state(dict(c, warm_fp32=True), B.make_grid(c), torch.as_tensor(psi0).cuda(), verbose=False)
    assert od.get("warm_iters", 0) == 0
    for k in ("E", "mu", "Lz"):
        assert abs(od[k] - oo[k]) <= 1e-8 * max(1.0, abs(oo[k])), (k, od[k], oo[k])
    assert _dens_l2(pd.cpu().numpy(), po, Go) < 1e-6


@pytest.mark.parametrize("kw", [dict(beta2=100.0), dict(beta2=30.0, beta3=5.0, Omega=0.4),
                                dict(trap_coeff=0.0, beta2=-60.0, beta3=1000.0, L=8.0)])
def test_fp32_warm_phase_then_fp64_polish(kw):
    """warm_fp32: the loop first runs on complex64 fields to res 1e-4, then the
    FP64 loop polishes; the converged state matches the oracle's FP64 solve
    (E, mu, Lz to 1e-8, density L2 to 1e-6)."""
    import becgpp_amd as B
    c = cfg(res_tol=1e-8, maxit=20000, linesearch_max=20, **kw)
    Go = O.make_grid(c)
    p0 = O.noisy_gaussian(Go, 1e-3, seed=21)
    po, _, oo = O.ground_state(c, Go, p0)
    pd, _, od = B.ground_state(dict(c, warm_fp32=True), B.make_grid(c), torch.as_tensor(p0).cuda(), verbose=False)
    assert od.get("warm_iters", 0) > 0
    for k in ("E", "mu", "Lz"):
        assert abs(od[k] - oo[k]) <= 1e-8 * max(1.0, abs(oo[k])), (k, od[k], oo[k])
    ro, rd = np.abs(po) ** 2, np.abs(pd.cpu().numpy()) ** 2
    assert np.sqrt(((rd - ro) ** 2).sum() / (ro ** 2).sum()) < 1e-6
    assert od["resid_rel"] < 2e-8


def test_fp32_warm_phase_batched_and_long_range():
    import becgpp_amd as B
    # a grid with the Newton kernel; the three batched instances have G_C = 0
    base = cfg(res_tol=1e-8, maxit=20000, linesearch_max=20, kernel="newton", G_C=3.0)
    cs = [dict(base, beta2=b2, Omega=om, G_C=0.0, warm_fp32=True)
          for b2, om in ((80.0, 0.0), (40.0, 0.3), (20.0, 0.5))]
    Go = O.make_grid(base)
    p0s = [O.noisy_gaussian(Go, 1e-3, seed=60 + i) for i in range(3)]
    G = B.make_grid(base)
    runs = B.ground_state_batch(cs, G=G, psi0s=[torch.as_tensor(p).cuda() for p in p0s])
    for c, p0, (_, _, od) in zip(cs, p0s, runs):
        assert od.get("warm_iters", 0) > 0
        _, _, oo = O.ground_state(c, O.make_grid(c), p0)
        assert abs(od["E"] - oo["E"]) <= 1e-8 * max(1.0, abs(oo["E"]))
    # a long-range instance: the warm phase keeps its convolutions FP64
    lr = dict(cs[0], G_C=3.0)
    _, _, ol = B.ground_state(lr, G, torch.as_tensor(p0s[0]).cuda(), verbose=False)
    assert ol.get("warm_iters", 0) > 0 and ol["Egrav"] != 0.0
    _, _, oo = O.ground_state(lr, O.make_grid(lr), p0s[0])
    for k in ("E", "mu", "Lz"):
        assert abs(ol[k] - oo[k]) <= 1e-8 * max(1.0, abs(oo[k])), (k, ol[k], oo[k])
    # the trial-per-step search has no FP32 phase with long range
    _, _, ot = B.ground_state(dict(lr, line_search="trial"), G, torch.as_tensor(p0s[0]).cuda(), verbose=False)
    assert "warm_iters" not in ot
