"""Write the reference's sweep rows (modes.py:295-387) with the seeds' noise off.

A sweep keeps, per point, the lower-energy candidate of the previous point's
state (continuation, resampled when the grid changes) and the point's fresh
seeds (multi-seed), and records the diagnostic bundle.  With seed_noise =
seed_phase_noise = 0 every candidate starts from a deterministic state, so
the branch chosen per point, its energy and its diagnostics can be compared.
This script runs mode_sweep on the CPU (the reference's package imported from
its root, output directory redirected to a temporary folder, no checkpoints
or figures) for a G_C sweep on a fixed grid and an autobox sweep whose grid
changes with the point, and writes tests/golden/sweep_reference.json.
Run:  python tests/golden/make_sweep_golden.py [/root/reference]
"""
import json
import math
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "sweep_reference.json")

COMMON = dict(mode="sweep", seed_noise=0.0, seed_phase_noise=0.0, save_ckpt=False, save_figs=False,
              tf_cache=False, res_tol=1e-8)
SWEEPS = [
    dict(dimension="2D", kernel="newton", beta2=40.0, Omega=0.0, nseeds=2, L=6.0, Ngrid=32,
         sweep_param="G_C", sweep_values=[1.0, 2.5, 4.0], sweep_continuation=True),
    dict(dimension="2D", kernel="newton", beta2=0.0, Omega=1.0, nseeds=2, G_C=16.0,
         sweep_param="G_C", sweep_values=[16.0, 24.0], sweep_continuation=True, sweep_autobox=True),
]
ROW_KEYS = ("E", "mu", "Lz", "Nv", "R90", "branch", "candidate_E", "converged", "stop_reason", "L", "N", "iters")


def sweep_cfg(kw):
    c = dict(COMMON)
    c.update(kw)
    return c


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import matplotlib
    matplotlib.use("Agg")
    import becgpp
    from becgpp.modes import mode_sweep

    out = []
    for kw in SWEEPS:
        cfg = becgpp.default_cfg(**sweep_cfg(kw))
        rows = mode_sweep(cfg)
        rec = []
        for r in rows:
            d = {}
            for k in ROW_KEYS + (kw["sweep_param"],):
                v = r[k]
                if hasattr(v, "item"):
                    v = v.item()
                if isinstance(v, float) and not math.isfinite(v):
                    v = None
                d[k] = v
            rec.append(d)
        out.append(dict(sweep=kw, rows=rec))
    with open(OUT, "w") as f:
        json.dump(dict(_source="becGPP 1.2.1 becgpp/modes.py mode_sweep, CPU, seeds' noise off", sweeps=out), f,
                  indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
