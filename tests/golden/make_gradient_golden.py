"""Write the reference's own autograd gradient of the energy.

The reference's energy_components (operators.py:10-29) is built from torch
operations, so torch autograd differentiates it; for the real energy of a
complex field torch returns dE/dRe psi + i dE/dIm psi.  This script evaluates
that gradient on the CPU (the reference's package imported from its root,
output directory redirected to a temporary folder) for the six physics sets
of tests/golden/make_operator_golden.py, on the same seeded states, and writes
tests/golden/gradient_reference.npz.
Run:  python tests/golden/make_gradient_golden.py [/root/reference]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_operator_golden import CASES, state  # noqa: E402

OUT = os.path.join(HERE, "gradient_reference.npz")


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import torch
    import becgpp
    from becgpp.grid import make_grid
    from becgpp.operators import energy_components

    out = {}
    for i, kw in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", **kw)
        G = make_grid(cfg)
        p = torch.as_tensor(state(kw, seed=700 + i)).clone().requires_grad_(True)
        E = energy_components(p, G, cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"])["E"]
        E.backward()
        out[f"grad{i}"] = p.grad.numpy().astype(np.complex128)
        print(f"case {i}: {kw}: |grad|max={float(p.grad.abs().max())!r}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
