"""Write the reference's initial states with their noise switched off.

The seeds (seeds.py:11-103) multiply a deterministic recipe -- a Gaussian or
Thomas-Fermi envelope, an imprinted winding, the triangular vortex lattice
(its angle and shift drawn with numpy's default_rng(rng), its tanh cores and
summed phases) -- by noise from the torch device RNG, which cannot be
reproduced.  With seed_noise = seed_phase_noise = 0 the noise factors are
exactly 1 and the recipe itself is what the reference returns.  This script
runs its make_seed on the CPU (its package imported from the reference root,
output directory redirected to a temporary folder) and writes
tests/golden/seed_reference.npz.
Run:  python tests/golden/make_seed_golden.py [/root/reference]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "seed_reference.npz")

NOISE_OFF = dict(seed_noise=0.0, seed_phase_noise=0.0)
CASES = [
    dict(seed="gaussian", dimension="2D", L=6.0, Ngrid=48),
    dict(seed="gaussian", dimension="2D", seed_winding=2, L=6.0, Ngrid=48),
    dict(seed="gaussian", dimension="3D", seed_winding=-1, L=5.0, Ngrid=24),
    dict(seed="tf", dimension="2D", beta2=100.0, L=6.0, Ngrid=48),
    dict(seed="tf", dimension="2D", beta2=30.0, G_C=3.0, kernel="newton", seed_winding=1, L=6.0, Ngrid=48),
    dict(seed="tf", dimension="2D", beta2=-60.0, beta3=1000.0, trap_coeff=0.0, L=8.0, Ngrid=48),
    dict(seed="triangular", dimension="2D", beta2=200.0, Omega=0.9, rng=0, L=8.0, Ngrid=64),
    dict(seed="triangular", dimension="2D", beta2=100.0, Omega=0.5, rng=7, L=8.0, Ngrid=64),
    dict(seed="triangular", dimension="2D", beta2=150.0, Omega=-0.7, rng=3, L=8.0, Ngrid=64),
    dict(seed="triangular", dimension="2D", beta2=100.0, Omega=0.2, rng=1, L=6.0, Ngrid=48),   # below 0.3: tf seed
    dict(seed="triangular", dimension="3D", beta2=200.0, Omega=0.6, rng=5, L=6.0, Ngrid=32),
]


def case_cfg(kw):
    c = dict(NOISE_OFF)
    c.update(kw)
    return c


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import becgpp
    from becgpp.grid import make_grid
    from becgpp.seeds import make_seed

    out = {}
    for i, kw in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", tf_cache=False, **case_cfg(kw))
        G = make_grid(cfg)
        psi = make_seed(G, cfg)
        out[f"psi{i}"] = psi.numpy().astype(np.complex128)
        print(f"case {i}: {kw}: |psi|max={float(psi.abs().max())!r}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
