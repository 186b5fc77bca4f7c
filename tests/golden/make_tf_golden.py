"""Write reference Thomas-Fermi envelopes (thomasfermi.py:65-156).

The nonlocal compressed envelope is a relaxed fixed point (40% mixing, a
90-step bisection of mu per iteration, stop at an L1 change of 1e-9) whose
result no closed form gives; the published sweep rows pin it at a few
points only.  This script runs the reference's ``extract_tf`` on the CPU (its
package imported from the reference root, its output directory redirected to
a temporary folder) for the cases of tests/test_gpu_tf.py and writes
tests/golden/tf_reference.npz: per case the kind, mu, R90, R99, the
convergence flag and the density n.
Run:  python tests/golden/make_tf_golden.py [/root/reference]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "tf_reference.npz")

# the configurations of the GPU test (on the test helpers' L = 6, N = 48
# default grid unless given), and one without a TF reference
CASES = [
    dict(beta2=100.0),                                            # repulsive parabola
    dict(beta2=30.0, beta3=5.0, Omega=0.4),                       # local trapped repulsive
    dict(trap_coeff=0.0, beta2=-60.0, beta3=1000.0, L=8.0),       # flat-top droplet
    dict(beta2=30.0, G_C=3.0, kernel="newton"),                   # nonlocal compressed, 1/r
    dict(dimension="quasi2D", beta2=30.0, G_C=2.0, kernel="q2d", l_z=0.5),
    dict(dimension="3D", Ngrid=24, L=5, beta2=20.0, G_C=3.0, kernel="newton"),
    dict(beta2=30.0, G_C=2.0, kernel="log"),                      # nonlocal compressed, -ln r
    dict(beta2=-5.0, G_C=2.0, kernel="newton"),                   # attraction only: no TF reference
]
GRID = dict(L=6.0, Ngrid=48)


def case_cfg(kw):
    c = dict(GRID)
    c.update(kw)
    return c


def main(ref_root):
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_ref_out_")
    sys.path.insert(0, ref_root)
    import becgpp
    from becgpp.grid import make_grid
    from becgpp.thomasfermi import extract_tf

    out = {}
    for i, kw in enumerate(CASES):
        cfg = becgpp.default_cfg(mode="single", tf_cache=False, **case_cfg(kw))
        G = make_grid(cfg)
        tf = extract_tf(cfg, G)
        out[f"none{i}"] = np.array(tf is None)
        if tf is None:
            print(f"case {i}: {kw}: no TF reference")
            continue
        out[f"kind{i}"] = np.array(tf["kind"])
        for k in ("mu", "R90", "R99"):
            out[f"{k}{i}"] = np.array(float(tf[k]))
        out[f"converged{i}"] = np.array(bool(tf["converged"]))
        out[f"n{i}"] = tf["n"].numpy().astype(np.float64)
        print(f"case {i}: {kw}: {tf['kind']} mu={float(tf['mu'])!r} converged={tf['converged']}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
