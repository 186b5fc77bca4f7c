"""Write the reference's multi-seed selections with the seeds' noise off.

ground_state_multiseed (solvers.py:238-263) cycles the seed kinds
(triangular, tf, gaussian when |Omega| >= 0.3, else tf, gaussian) with rng +
j, solves each and keeps the lowest energy, reporting seed_used,
seed_energies and seed_energy_spread.  With seed_noise = seed_phase_noise = 0
its seeds are the deterministic recipes, so the selection can be compared.
This script runs it on the CPU (the reference's package imported from its
root, output directory redirected to a temporary folder) and writes
tests/golden/multiseed_reference.json.
Run:  python tests/golden/make_multiseed_golden.py [/root/reference]
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "multiseed_reference.json")

NOISE_OFF = dict(seed_noise=0.0, seed_phase_noise=0.0)
CASES = [
    dict(dimension="2D", beta2=60.0, Omega=0.6, nseeds=3, rng=2, L=6.0, Ngrid=48, res_tol=1e-8),
    dict(dimension="2D", beta2=40.0, G_C=2.0, kernel="newton", nseeds=2, rng=0, L=6.0, Ngrid=48, res_tol=1e-8),
    dict(dimension="2D", beta2=-30.0, beta3=400.0, trap_coeff=0.0, nseeds=2, rng=4, L=5.0, Ngrid=48,
         res_tol=1e-8, precond_shift_min=0.1),
]


def case_cfg(kw):
    c = dict(NOISE_OFF)
    c.update(kw)
    return c


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import becgpp
    from becgpp.solvers import ground_state_multiseed

    out = []
    for kw in CASES:
        cfg = becgpp.default_cfg(mode="single", tf_cache=False, **case_cfg(kw))
        _, _, obs = ground_state_multiseed(cfg, verbose=False)
        rec = {k: obs[k] for k in ("E", "mu", "Lz", "seed_used", "seed_energies", "seed_energy_spread",
                                   "converged", "stop_reason")}
        rec = {k: (float(v) if isinstance(v, float) or hasattr(v, "item") and not isinstance(v, (str, bool))
                   else v) for k, v in rec.items()}
        rec["case"] = kw
        out.append(rec)
        print(kw, rec["seed_used"], rec["seed_energies"])
    with open(OUT, "w") as f:
        json.dump(dict(_source="becGPP 1.2.1 becgpp/solvers.py ground_state_multiseed, CPU, seeds' noise off",
                       cases=out), f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
