"""Write reference solver trajectories (solvers.py:48-235) for small cases.

The loop's scalar logic -- Barzilai-Borwein trial step, Armijo halvings,
Polak-Ribiere variants and restarts, failed-search handling, the stall,
maxit and energy-window stops and the best-residual return -- is fixed by the
reference's own execution, which only running it can show.  This script runs
the reference's ``ground_state`` on the CPU (its package imported from the
reference root, its output directory redirected to a temporary folder) with
``record_trace`` on, from initial states the tests regenerate from the recipe
below, and writes tests/golden/trajectory_reference.json: per case the
iteration count, stop reason, final E, mu, L_z, residuals, shift and the
trace (iteration, E, relative residual) of every accepted step; and
tests/golden/trajectory_density.npz: the returned state's density.
Run:  python tests/golden/make_trajectory_golden.py [/root/reference]
"""
import json
import math
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "trajectory_reference.json")
OUT_RHO = os.path.join(HERE, "trajectory_density.npz")     # the returned states' densities

# small grids; each case exercises one part of the loop
CASES = [
    dict(name="contact", dimension="2D", beta2=30.0, L=6.0, Ngrid=32, res_tol=1e-8),
    dict(name="rotation", dimension="2D", beta2=50.0, Omega=0.5, L=6.0, Ngrid=32, res_tol=1e-7),
    dict(name="newton_pr_restart7", dimension="2D", beta2=20.0, G_C=3.0, kernel="newton", L=6.0, Ngrid=32,
         res_tol=1e-7, cg_beta="pr", cg_restart=7),
    dict(name="log", dimension="2D", beta2=10.0, G_C=2.0, kernel="log", L=6.0, Ngrid=32, res_tol=1e-7),
    dict(name="q2d", dimension="quasi2D", beta2=30.0, G_C=3.0, kernel="q2d", l_z=0.5, L=6.0, Ngrid=32,
         res_tol=1e-8),
    dict(name="newton_3d", dimension="3D", beta2=20.0, beta3=2.0, G_C=3.0, kernel="newton", L=5.0, Ngrid=24,
         res_tol=1e-7),
    dict(name="cubic_quintic", dimension="2D", beta2=-20.0, beta3=400.0, trap_coeff=0.0, precond_shift_min=0.1,
         L=5.0, Ngrid=32, res_tol=1e-7),
    dict(name="quartic_no_cg", dimension="2D", s=4, trap_coeff=0.25, beta2=10.0, L=6.0, Ngrid=32, res_tol=1e-5,
         cg_beta="none"),
    dict(name="stall", dimension="2D", beta2=30.0, L=6.0, Ngrid=32, res_tol=1e-14, linesearch_max=1,
         linesearch_stall=3, step=3.0),
    dict(name="maxit", dimension="2D", beta2=30.0, Omega=0.3, L=6.0, Ngrid=32, res_tol=1e-14, maxit=15),
    dict(name="energy_window", dimension="2D", beta2=30.0, L=6.0, Ngrid=32, res_tol=1e-14, conv_window=5,
         energy_tol=1e-7),
]


def inputs(kw, seed):
    """psi0 on the case's grid (x_j = (j - N/2) dx, dx = 2L/N, indexing ij): a
    Gaussian envelope of width 1.2 times (1 + 0.2 xi1) e^{i 0.3 xi2}, numpy
    default_rng(seed)."""
    ndim = 3 if kw["dimension"] == "3D" else 2
    N, L = int(kw["Ngrid"]), float(kw["L"])
    dx = 2.0 * L / N
    x = (np.arange(N) - N // 2) * dx
    C = np.meshgrid(*([x] * ndim), indexing="ij")
    r2 = sum(c * c for c in C)
    rng = np.random.default_rng(seed)
    shape = (N,) * ndim
    env = np.exp(-0.5 * r2 / 1.44)
    psi = env * (1.0 + 0.2 * rng.standard_normal(shape)) * np.exp(0.3j * rng.standard_normal(shape))
    return psi.astype(np.complex128)


def case_overrides(kw):
    return {k: v for k, v in kw.items() if k != "name"}


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import torch
    import becgpp
    from becgpp.grid import make_grid
    from becgpp.solvers import ground_state

    out = []
    rho = {}
    for i, kw in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", record_trace=True, **case_overrides(kw))
        G = make_grid(cfg)
        psi0 = torch.as_tensor(inputs(kw, seed=300 + i))
        psi, _, obs = ground_state(cfg, G=G, psi0=psi0, verbose=False)
        rho[f"rho{i}"] = (psi.abs() ** 2).numpy().astype(np.float64)

        def num(v):
            v = float(v)
            return v if math.isfinite(v) else None

        out.append(dict(name=kw["name"], case=case_overrides(kw), seed=300 + i, iters=int(obs["iters"]),
                        stop_reason=str(obs["stop_reason"]), converged=bool(obs["converged"]),
                        E=num(obs["E"]), mu=num(obs["mu"]), Lz=num(obs["Lz"]), resid_rel=num(obs["resid_rel"]),
                        min_resid=num(obs["min_resid"]), preconditioner_shift=num(obs["preconditioner_shift"]),
                        trace=[[int(t[1]), num(t[2]), num(t[3])] for t in obs["trace"]]))
        print(f"{kw['name']}: iters={obs['iters']} stop={obs['stop_reason']} E={float(obs['E'])!r}")
    with open(OUT, "w") as f:
        json.dump(dict(_source="becGPP 1.2.1 becgpp/solvers.py ground_state (record_trace), CPU", cases=out), f,
                  indent=1)
    np.savez_compressed(OUT_RHO, **rho)
    print(f"wrote {OUT} and {OUT_RHO}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
