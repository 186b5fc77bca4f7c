"""Write the reference's operator applications on seeded states.

apply_H, energy_components, compute_residual, Lz_op and kernel_convolution
(operators.py:10-52, grid.py:112-116, interactions.py:154-171) evaluated by
the reference itself on the CPU (its package imported from the reference
root, output directory redirected to a temporary folder) for the six physics
sets of its FD audit (tests/test_audit.py:55-62) on small grids, from states
the tests regenerate (tests/helpers.random_state's recipe: a Gaussian envelope
with seeded complex noise, numpy default_rng).  Output:
tests/golden/operator_reference.npz.
Run:  python tests/golden/make_operator_golden.py [/root/reference]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "operator_reference.npz")

CASES = [
    dict(dimension="2D", beta2=30.0, beta3=5.0, Omega=0.4, L=6.0, Ngrid=32),
    dict(dimension="2D", beta2=30.0, G_C=3.0, kernel="newton", Omega=0.3, L=6.0, Ngrid=32),
    dict(dimension="2D", beta2=30.0, G_C=3.0, kernel="log", L=6.0, Ngrid=32),
    dict(dimension="quasi2D", beta2=30.0, G_C=3.0, kernel="q2d", l_z=0.5, L=6.0, Ngrid=32),
    dict(dimension="3D", beta2=20.0, beta3=2.0, G_C=3.0, kernel="newton", Omega=0.5, Ngrid=24, L=5.0),
    dict(dimension="2D", s=4, trap_coeff=0.25, beta2=10.0, L=6.0, Ngrid=32),
]
COMPONENTS = ("E", "Ekin", "Etrap", "Econtact", "Ethree", "Erot", "Egrav", "Lz")


def state(kw, seed, amp=0.3):
    """exp(-r^2/2) (1 + amp xi1) + i amp exp(-r^2/2) xi2 on the case's grid
    (x_j = (j - N/2) dx, dx = 2L/N, indexing ij), numpy default_rng(seed)."""
    ndim = 3 if kw["dimension"] == "3D" else 2
    N, L = int(kw["Ngrid"]), float(kw["L"])
    dx = 2.0 * L / N
    x = (np.arange(N) - N // 2) * dx
    C = np.meshgrid(*([x] * ndim), indexing="ij")
    r2 = sum(c * c for c in C)
    rng = np.random.default_rng(seed)
    env = np.exp(-0.5 * r2)
    re = env * (1.0 + amp * rng.standard_normal(r2.shape))
    im = amp * env * rng.standard_normal(r2.shape)
    return (re + 1j * im).astype(np.complex128)


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import torch
    import becgpp
    from becgpp.grid import Lz_op, make_grid
    from becgpp.interactions import kernel_convolution
    from becgpp.operators import apply_H, compute_residual, energy_components

    out = {}
    for i, kw in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", **kw)
        G = make_grid(cfg)
        p = torch.as_tensor(state(kw, seed=700 + i))
        comp = energy_components(p, G, cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"])
        for k in COMPONENTS:
            out[f"{k}{i}"] = np.array(float(comp[k]))
        out[f"phi{i}"] = comp["phi"].numpy().astype(np.float64) if torch.is_tensor(comp["phi"]) else np.zeros(1)
        out[f"H{i}"] = apply_H(p, G, cfg).numpy().astype(np.complex128)
        ra, mu, rr = compute_residual(p, G, cfg)
        out[f"res{i}"] = np.array([float(ra), float(mu), float(rr)])
        out[f"Lz{i}psi"] = Lz_op(p, G).numpy().astype(np.complex128)
        if G["kernel"] != "none":
            out[f"conv{i}"] = kernel_convolution(p.abs() ** 2, G).numpy().astype(np.float64)
        print(f"case {i}: {kw}: E={float(comp['E'])!r}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
