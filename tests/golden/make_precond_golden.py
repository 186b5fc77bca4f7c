"""Write reference vectors for the Sobolev preconditioner (solvers.py:28-45).

The preconditioner's local stiffness V + 2 beta2 rho + 3 beta3 rho^2 (+ phi),
its bulk mask rho > 0.05 max rho and its floor precond_shift_min are choices
of the reference that no identity or closed form fixes, so they are pinned
by vectors the reference itself computes.  This script runs in a container
where the reference is mounted (it imports its package on the CPU, with its
output directory redirected to a temporary folder); only the vectors are
committed.  Run:  python tests/golden/make_precond_golden.py [/root/reference]

Inputs are regenerated by the tests from the same recipe (`inputs` below,
numpy default_rng): a Gaussian envelope of width `width` times
(1 + 0.4 xi1) e^{i 0.3 xi2} for psi, and an independent complex field for the
gradient.  Output: tests/golden/precond_reference.npz with sigma and z for
each case (z stored as complex128; 2D grids 24^2, 3D 16^3).
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "precond_reference.npz")

# the six physics sets of the reference's FD audit (tests/test_audit.py:55-62)
# on small grids, plus a floor case (sigma below precond_shift_min) and a
# quintic-dominated one, and a 3D case whose shift lies above its floor
CASES = [
    dict(dimension="2D", beta2=30.0, beta3=5.0, Omega=0.4, L=5.0, Ngrid=24),
    dict(dimension="2D", beta2=30.0, G_C=3.0, kernel="newton", Omega=0.3, L=5.0, Ngrid=24),
    dict(dimension="2D", beta2=30.0, G_C=3.0, kernel="log", L=5.0, Ngrid=24),
    dict(dimension="quasi2D", beta2=30.0, G_C=3.0, kernel="q2d", l_z=0.5, L=5.0, Ngrid=24),
    dict(dimension="3D", beta2=20.0, beta3=2.0, G_C=3.0, kernel="newton", Omega=0.5, L=5.0, Ngrid=16),
    dict(dimension="2D", s=4, trap_coeff=0.25, beta2=10.0, L=5.0, Ngrid=24),
    dict(dimension="2D", beta2=0.05, s=2, trap_coeff=0.01, precond_shift_min=0.5, L=5.0, Ngrid=24),
    dict(dimension="2D", beta2=-20.0, beta3=400.0, s=2, trap_coeff=0.0, precond_shift_min=0.1, L=5.0, Ngrid=24),
    dict(dimension="3D", beta2=400.0, beta3=60.0, G_C=1.0, kernel="newton", Omega=0.2, L=5.0, Ngrid=16),
]


def inputs(ndim, N, L, seed, width=1.2):
    """psi and grad on the grid x_j = (j - N/2) dx, dx = 2L/N, indexing ij."""
    dx = 2.0 * L / N
    x = (np.arange(N) - N // 2) * dx
    C = np.meshgrid(*([x] * ndim), indexing="ij")
    r2 = sum(c * c for c in C)
    rng = np.random.default_rng(seed)
    shape = (N,) * ndim
    env = np.exp(-0.5 * r2 / (width * width))
    psi = env * (1.0 + 0.4 * rng.standard_normal(shape)) * np.exp(0.3j * rng.standard_normal(shape))
    grad = env * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return psi.astype(np.complex128), grad.astype(np.complex128)


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import torch
    import becgpp
    from becgpp.grid import make_grid
    from becgpp.interactions import long_range_phi
    from becgpp.solvers import _preconditioner

    out = {}
    for i, kw in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", **kw)
        G = make_grid(cfg)
        ndim = 3 if kw["dimension"] == "3D" else 2
        psi_np, grad_np = inputs(ndim, int(kw["Ngrid"]), float(kw["L"]), seed=100 + i)
        psi = torch.as_tensor(psi_np)
        grad = torch.as_tensor(grad_np)
        rho = psi.abs() ** 2
        phi = None
        if G["kernel"] != "none" and abs(cfg["G_C"]) > 1e-15:
            phi = long_range_phi(rho, G, cfg["G_C"])
        z, sigma = _preconditioner(grad, psi, G, cfg, phi)
        out[f"z{i}"] = z.numpy().astype(np.complex128)
        out[f"sigma{i}"] = np.array(float(sigma))
        print(f"case {i}: {kw} sigma={sigma!r}")
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
