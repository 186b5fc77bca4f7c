"""Shared test inputs: configurations and seeded states built identically for
the oracle (numpy) and the device path."""
import numpy as np

import oracle as O

# the six physics sets of the reference's FD-gradient audit (tests/test_audit.py:55-62)
PHYSICS = [
    dict(dimension="2D", beta2=30, beta3=5, Omega=0.4),
    dict(dimension="2D", beta2=30, G_C=3, kernel="newton", Omega=0.3),
    dict(dimension="2D", beta2=30, G_C=3, kernel="log"),
    dict(dimension="quasi2D", beta2=30, G_C=3, kernel="q2d", l_z=0.5),
    dict(dimension="3D", beta2=20, beta3=2, G_C=3, kernel="newton", Omega=0.5, Ngrid=24, L=5),
    dict(dimension="2D", s=4, trap_coeff=0.25, beta2=10),
]

BASE = dict(mode="single", s=2, trap_coeff=0.5, Omega=0.0, beta2=100.0, beta3=0.0, G_C=0.0,
            kernel="auto", l_z=0.0, L=6.0, Ngrid=48, pad=2, maxit=60000, res_tol=1e-4,
            energy_tol=1e-9, conv_window=2000, step=1.0, step_max=3.0, linesearch_max=8,
            linesearch_stall=10, precond_shift_min=0.5, cg_beta="pr_precond", cg_restart=30,
            check=200, seed="tf", seed_winding=0, seed_noise=1e-2, seed_phase_noise=5e-2, rng=0,
            nseeds=1, record_trace=False, tf_cache=True)


def cfg(**kw):
    c = dict(BASE)
    c.update(kw)
    return c


def random_state(G, seed=1, amp=0.3):
    """Gaussian envelope with seeded complex noise (the recipe of the
    reference's tests/test_audit.py:46-52, numpy RNG)."""
    rng = np.random.default_rng(seed)
    env = np.exp(-0.5 * G["R2"])
    sh = G["R2"].shape
    re = env * (1.0 + amp * rng.standard_normal(sh))
    im = amp * env * rng.standard_normal(sh)
    return (re + 1j * im).astype(np.complex128), rng


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def oracle_grid(c):
    return O.make_grid(c)


def zeros_state(G, plus=(), minus=(), sigma=2.5, shift=(0.0, 0.0)):
    """prod (z - a) over plus, prod conj(z - b) over minus, times a Gaussian of
    width sigma centred at shift (z = x + i y in the first two axes; a 3D grid
    gets the factor exp(-z^2/2)): phase windings +1 at each a, -1 at each b."""
    X, Y = G["coords"][0], G["coords"][1]
    z = X + 1j * Y
    psi = np.exp(-((X - shift[0]) ** 2 + (Y - shift[1]) ** 2) / (2.0 * sigma ** 2)).astype(np.complex128)
    for a in plus:
        psi = psi * (z - a)
    for b in minus:
        psi = psi * np.conj(z - b)
    if G["ndim"] == 3:
        psi = psi * np.exp(-0.5 * G["coords"][2] ** 2)
    return psi


def lll_state(G, m, conj=False):
    """z^m exp(-r^2/2) (conj(z)^m with conj): the lowest-Landau-level orbitals."""
    X, Y = G["coords"][0], G["coords"][1]
    z = X - 1j * Y if conj else X + 1j * Y
    return (z ** m * np.exp(-0.5 * G["R2"])).astype(np.complex128)
