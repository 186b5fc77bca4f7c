"""The sweep mode (modes.py:295-387: continuation from the previous point,
resampled when the autobox grid changes, against the point's fresh seeds,
lowest energy kept) against sweeps the reference ran itself with the seeds'
noise switched off (tests/golden/make_sweep_golden.py): per point the grid
(autobox L and N), the branch kept wherever its candidates' energies differ
by more than 1e-8, E to 1e-9, mu to 1e-6, L_z, the vortex count and R90."""
import json
import os

import pytest

from tests.golden.make_sweep_golden import SWEEPS, sweep_cfg

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sweep_reference.json")


def candidates(s):
    return {kv.split("=")[0]: float(kv.split("=")[1]) for kv in s.split(";")}


@pytest.mark.parametrize("k", range(len(SWEEPS)))
def test_sweep_rows_match_reference(k):
    import becgpp_amd as B
    with open(GOLD) as f:
        ref = json.load(f)["sweeps"][k]
    kw = SWEEPS[k]
    print(f"[sweep] reference sweep {k}: {kw['sweep_param']} over {kw['sweep_values']}", flush=True)
    rows = B.run(B.default_cfg(**sweep_cfg(kw)))
    assert len(rows) == len(ref["rows"])
    p = kw["sweep_param"]
    for r, g in zip(rows, ref["rows"]):
        assert float(r[p]) == float(g[p])
        assert float(r["L"]) == float(g["L"]) and int(r["N"]) == int(g["N"])
        cg = candidates(g["candidate_E"])
        assert set(candidates(r["candidate_E"])) == set(cg)
        if len(cg) > 1 and max(cg.values()) - min(cg.values()) > 1e-8 * max(abs(v) for v in cg.values()):
            assert r["branch"] == g["branch"], (r["branch"], g["branch"])
        assert abs(r["E"] - g["E"]) <= 1e-9 * abs(g["E"]), (r["E"], g["E"])
        assert abs(r["mu"] - g["mu"]) <= 1e-6 * abs(g["mu"]), (r["mu"], g["mu"])
        assert abs(r["Lz"] - g["Lz"]) <= 1e-6 * max(1.0, abs(g["Lz"]))
        assert int(r["Nv"]) == int(g["Nv"])
        assert abs(r["R90"] - g["R90"]) <= 1e-9 + 2.0 * (2.0 * float(g["L"]) / int(g["N"]))
