"""Extract the golden values this build's tests pin against from the
reference's committed result files (run in a container where the reference is
mounted; the output JSON is committed, the reference is not needed after).

Sources (under the reference root):
  benchmarks/results/verification/validation.csv   analytic-gate relative errors
  benchmarks/results/itp_timing/timing.csv          converged energies at res 1e-8
  benchmarks/results/verification/pcg_variants.csv  converged energies at res 1e-6
  benchmarks/results/verification/kernels_log_q2d.csv  log/q2d/thin-disc energies
  benchmarks/results/verification/sweep_omega_Omega.csv  rotation sweep: E, L_z, radii, vortex counts, w_LLL
  benchmarks/results/verification/sweep_gc2d_G_C.csv    G_C sweep (newton, Omega = 1, autobox): the same columns
  data/sample_output/scan/scan_singles.csv               single runs of the 1.0 scan: label, configuration, E, R90, Nv
  benchmarks/results/verification/cq_ladder.csv          self-bound cubic-quintic ladder at N = 1024: E, mu, peak
"""
import csv
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_results.json")


def rows(rel):
    with open(os.path.join(REF, rel), newline="") as f:
        return list(csv.DictReader(f))


def main():
    out = {}
    out["validation"] = [dict(test=r["test"], case=r["case"], quantity=r["quantity"],
                              rel_error=float(r["rel_error"]))
                         for r in rows("benchmarks/results/verification/validation.csv")]
    out["itp_timing"] = [dict(case=r["case"], method=r["method"], N=int(r["N"]), E_final=float(r["E_final"]),
                              E_ref=float(r["E_ref"]), iters=int(r["iters_or_steps"]), stop=r["stop"],
                              resid=float(r["resid"]) if r["resid"] not in ("", "nan") else None)
                         for r in rows("benchmarks/results/itp_timing/timing.csv")]
    out["pcg_variants"] = [dict(case=r["case"], cg_beta=r["cg_beta"], E=float(r["E"]), iters=int(r["iters"]),
                                resid=float(r["resid"]), stop=r["stop"])
                           for r in rows("benchmarks/results/verification/pcg_variants.csv")]
    out["kernels_log_q2d"] = [{k: v for k, v in r.items()}
                              for r in rows("benchmarks/results/verification/kernels_log_q2d.csv")]
    keep = ("Omega", "E", "Lz", "Nv", "Nplus", "Nminus", "Nnet", "w_LLL", "lll_param", "R50", "R90", "R99",
            "R_1pct", "peak", "resid_rel", "iters", "branch", "tf_R90", "tf_R99", "tf_mu", "points_per_R90",
            "mem_estimate_gb")
    out["sweep_omega"] = [{k: r[k] for k in keep} for r in rows("benchmarks/results/verification/sweep_omega_Omega.csv")]
    out["sweep_gc2d"] = [{k: r[k] for k in keep[1:] + ("G_C", "L", "N")}
                         for r in rows("benchmarks/results/verification/sweep_gc2d_G_C.csv")]
    out["scan_singles"] = [{k: r[k] for k in ("label", "dimension", "kernel", "N", "L", "beta2", "beta3", "G_C",
                                               "Omega", "s", "seed", "nseeds", "E", "R90", "Nv", "resid_rel")}
                           for r in rows("data/sample_output/scan/scan_singles.csv")]
    out["cq_ladder"] = [{k: float(r[k]) for k in ("beta2", "beta3", "E", "mu", "peak", "R90", "virial_rel", "resid")}
                        for r in rows("benchmarks/results/verification/cq_ladder.csv")]
    out["_source"] = "becGPP 1.2.1 benchmarks/results (reference repository)"
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
