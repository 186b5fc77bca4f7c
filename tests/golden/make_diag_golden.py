"""Write the reference's diagnostic bundle and resampled states.

``diagnostics`` (diagnostics.py:18-47) collects observables (operators.py:
55-103: energies, mu, L_z, rrms, peak, the virial with its long-range term,
the 3D radii and oblateness, E_bind), the mass radii and 1% edge
(fields.py), the vortex counts and LLL weight (vortex.py) and the TF fields.
This script runs it on the CPU (the reference's package imported from its
root, output directory redirected to a temporary folder) for states the tests
regenerate from the recipe below -- Gaussian envelopes with planted zeros
(vortices of either sign) and seeded noise -- and runs ``resample_state``
(grid.py:140-161) between grid pairs.  Output: tests/golden/diag_reference.npz
(every numeric key of each bundle, and the resampled states).
Run:  python tests/golden/make_diag_golden.py [/root/reference]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "diag_reference.npz")

# (configuration, planted zeros (x, y, sign)) on small grids
CASES = [
    (dict(dimension="2D", beta2=50.0, Omega=0.6, L=6.0, Ngrid=64), [(0.8, 0.3, 1), (-0.5, -0.7, 1), (0.1, 1.1, -1)]),
    (dict(dimension="2D", beta2=30.0, Omega=0.3, G_C=2.0, kernel="newton", L=6.0, Ngrid=48), [(0.6, -0.4, 1)]),
    (dict(dimension="2D", beta2=30.0, G_C=2.0, kernel="log", L=6.0, Ngrid=48), []),
    (dict(dimension="quasi2D", beta2=30.0, G_C=3.0, kernel="q2d", l_z=0.5, Omega=0.4, L=6.0, Ngrid=48),
     [(-0.7, 0.2, 1), (0.5, 0.6, 1)]),
    (dict(dimension="2D", beta2=-30.0, beta3=400.0, trap_coeff=0.0, L=5.0, Ngrid=48), []),
    (dict(dimension="2D", s=4, trap_coeff=0.25, beta2=20.0, beta3=3.0, Omega=1.0, L=6.0, Ngrid=48), [(0.4, 0.4, 1)]),
    (dict(dimension="3D", beta2=20.0, beta3=2.0, G_C=3.0, kernel="newton", L=5.0, Ngrid=24), []),
    (dict(dimension="3D", beta2=40.0, Omega=0.5, L=5.0, Ngrid=24, trap_coeff=0.5), []),
]
# resampling: (case whose state is resampled, new L, new N)
RESAMPLE = [(0, 8.0, 48), (2, 4.0, 64), (6, 6.0, 20)]
KEYS = ("E", "mu", "Lz", "rrms", "peak", "E_bind", "virial", "virial_rel", "virial_lr_term", "Ekin", "Etrap",
        "Econtact", "Ethree", "Erot", "Egrav", "R_perp_rms", "R_z_rms", "oblateness", "R50", "R90", "R99", "R_1pct",
        "Nv", "Nplus", "Nminus", "Nnet", "w_LLL", "lll_param", "tf_R90", "tf_R99", "tf_mu", "tf_rho0",
        "points_per_R90")


def state(kw, zeros, seed):
    """exp(-r^2/2 w^2) prod (z - z_k) or conj(z - z_k) per planted zero,
    times (1 + 0.05 xi1) e^{i 0.05 xi2} (numpy default_rng(seed)); w = 1.3,
    zeros in the x-y plane (3D: every z slice), indexing ij."""
    ndim = 3 if kw["dimension"] == "3D" else 2
    N, L = int(kw["Ngrid"]), float(kw["L"])
    dx = 2.0 * L / N
    x = (np.arange(N) - N // 2) * dx
    C = np.meshgrid(*([x] * ndim), indexing="ij")
    r2 = sum(c * c for c in C)
    z = C[0] + 1j * C[1]
    psi = np.exp(-0.5 * r2 / 1.69).astype(np.complex128)
    for (a, b, sgn) in zeros:
        f = z - (a + 1j * b)
        psi = psi * (f if sgn > 0 else np.conj(f))
    rng = np.random.default_rng(seed)
    psi = psi * (1.0 + 0.05 * rng.standard_normal(psi.shape)) * np.exp(0.05j * rng.standard_normal(psi.shape))
    return psi


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import torch
    import becgpp
    from becgpp.diagnostics import diagnostics
    from becgpp.grid import make_grid, resample_state

    out = {}
    grids, psis = [], []
    for i, (kw, zeros) in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", tf_cache=False, **kw)
        G = make_grid(cfg)
        psi = torch.as_tensor(state(kw, zeros, seed=500 + i))
        psi = psi / torch.sqrt((psi.abs() ** 2).sum() * float(G["dx"]) ** int(G["ndim"]))
        d = diagnostics(psi, G, cfg)
        for k in KEYS:
            out[f"{k}{i}"] = np.array(float(d[k]))
        out[f"tf_kind{i}"] = np.array(str(d["tf_kind"]))
        grids.append((cfg, G))
        psis.append(psi)
        print(f"case {i}: E={d['E']!r} Nv={d['Nv']} Nplus={d['Nplus']} Nminus={d['Nminus']} w_LLL={d['w_LLL']!r} "
              f"virial_rel={d['virial_rel']!r} tf={d['tf_kind']}")
    for j, (i, L2, N2) in enumerate(RESAMPLE):
        cfg, G = grids[i]
        G2 = make_grid(dict(cfg, L=L2, Ngrid=N2))
        r = resample_state(psis[i], G, G2)
        out[f"resampled{j}"] = r.numpy().astype(np.complex128)
        print(f"resample {j}: case {i} to L={L2} N={N2}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
