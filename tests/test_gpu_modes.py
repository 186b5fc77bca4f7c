"""run(cfg) on the GPU (becgpp/modes.py): the validation gate passes, single
and sweep return what the solver returns for the same seeds, and a sweep
without continuation (batched) equals the point-by-point solves."""
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _cfg(B, **kw):
    # res_tol 1e-7: a solve to 1e-8 sits at the FP64 residual floor of these
    # small grids and may stop as stalled just above it
    return B.default_cfg(**dict(dict(L=6.0, Ngrid=48, res_tol=1e-7, maxit=20000), **kw))


def test_validate_mode_passes(tmp_path, monkeypatch):
    """The reference's tests/test_validation.py: the gate passes, its error
    magnitudes are within bounds, and validation.csv is written."""
    import becgpp_amd as B
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    g = B.run(B.default_cfg(mode="validate", validate_N=256))
    assert g["passed"], g
    assert g["ml"] < 1e-12 and g["mtf"] < 1e-12
    for k in ("mg", "mp", "mlg", "mlp", "mq", "m3"):
        assert g[k] < 5e-3, k
    assert all(g["kernel_orders"][k] > 1.7 for k in ("2D", "2D_log", "3D"))
    head = (tmp_path / "validation.csv").read_text().splitlines()[0].split(",")
    assert {"test", "case", "quantity", "rel_error"} <= set(head)


def test_single_mode_equals_multiseed():
    import becgpp_amd as B
    c = _cfg(B, mode="single", beta2=40.0, nseeds=2)
    d = B.run(c)
    _, _, obs = B.ground_state_multiseed(c)
    assert abs(d["E"] - obs["E"]) <= 1e-10 * abs(obs["E"])
    assert d["rid"] == B.run_id(c) and d["converged"]


def test_sweep_without_continuation_is_batched_seeds():
    import becgpp_amd as B
    c = _cfg(B, mode="sweep", sweep_param="beta2", sweep_values=[20.0, 60.0], nseeds=2,
             sweep_continuation=False)
    rows = B.run(c)
    assert [r["beta2"] for r in rows] == [20.0, 60.0]
    for r in rows:
        cp = dict(c, beta2=r["beta2"])
        _, _, oo = O.ground_state_multiseed(dict(cp, mode="single"))
        assert abs(r["E"] - oo["E"]) <= 1e-8 * abs(oo["E"]), (r["E"], oo["E"])


def test_sweep_with_continuation_keeps_lowest_candidate(tmp_path, monkeypatch):
    """modes.py:295-387: each point keeps its lowest-energy candidate; the
    point's checkpoint and the sweep CSV are written as the sweep goes."""
    import becgpp_amd as B
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    c = _cfg(B, mode="sweep", sweep_param="Omega", sweep_values=[0.0, 0.2], nseeds=2, beta2=30.0)
    rows = B.run(c)
    assert rows[1]["branch"] in ("continuation", "seed:tf", "seed:gaussian")
    cand = [float(x.split("=")[1]) for x in rows[1]["candidate_E"].split(";")]   # printed to 10 digits
    assert abs(rows[1]["E"] - min(cand)) <= 1e-9 * abs(min(cand))
    lines = (tmp_path / "sweep_Omega.csv").read_text().splitlines()
    assert len(lines) == 3 and "Nv" in lines[0].split(",") and "psi" not in lines[0].split(",")
    for r in rows:
        assert (tmp_path / "ckpt" / f"{r['rid']}.pt").is_file()
        psi = B.records.load_checkpoint_psi(r["rid"])
        assert torch.equal(psi, r["psi"].detach().cpu())


def test_single_mode_writes_records_and_resumes(tmp_path, monkeypatch):
    """modes.py:216-247: the checkpoint of a run is keyed by its run id; a second
    run of the same configuration resumes from it and converges at once; the
    summary CSV and the per-run JSON are written."""
    import becgpp_amd as B
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    c = _cfg(B, mode="single", beta2=70.0, nseeds=1, seed="gaussian")
    d1 = B.run(c)
    rid = B.run_id(c)
    assert (tmp_path / "ckpt" / f"{rid}.pt").is_file()
    assert (tmp_path / "single_summary.csv").is_file() and (tmp_path / f"{rid}.json").is_file()
    d2 = B.run(c)
    assert d2["iters"] < d1["iters"] and d2["converged"]
    assert abs(d2["E"] - d1["E"]) <= 1e-10 * abs(d1["E"])
    # the diagnostic bundle is part of the record (diagnostics.py:18-47)
    header = (tmp_path / "single_summary.csv").read_text().splitlines()[0].split(",")
    for k in ("Nv", "Nplus", "Nminus", "Nnet", "w_LLL", "lll_param", "R50", "R90", "R99", "R_1pct", "tf_kind",
              "tf_R90", "tf_mu", "mem_estimate_gb", "points_per_R90"):
        assert k in d1 and k in header, k
    assert d1["Nv"] == 0 and d1["w_LLL"] != d1["w_LLL"] and d1["tf_kind"] == "repulsive_parabola"


def test_tf_only_mode_writes_the_envelope(tmp_path, monkeypatch):
    """modes.py:264-291: the TF envelope of the configuration (the oracle's
    extract_tf), appended to tf_only.csv; None without one."""
    import becgpp_amd as B
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    c = _cfg(B, mode="tf_only", beta2=100.0, tf_cache=False)
    tf = B.run(c)
    ref = O.extract_tf(c, O.make_grid(c))
    assert tf["kind"] == ref["kind"] and abs(tf["mu"] - ref["mu"]) <= 1e-9 * abs(ref["mu"])
    assert (tmp_path / "tf_only.csv").read_text().splitlines()[0].split(",") == ["R90", "R99", "kind", "mu",
                                                                                 "rho0", "rid"]
    # attractive contact with long range: no envelope (thomasfermi.py:86-87)
    assert B.run(_cfg(B, mode="tf_only", beta2=-1.0, beta3=0.0, G_C=2.0, kernel="newton", tf_cache=False)) is None


def test_convergence_mode_rows(tmp_path, monkeypatch):
    """modes.py:392-436: one row per distinct (L, N, pad) variant with the
    diagnostic bundle; each row equals a direct solve of that variant."""
    import becgpp_amd as B
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    c = _cfg(B, mode="convergence", beta2=50.0, Ngrid=64, L=6.0, conv_Ngrids=[48, 64], conv_box_factors=[1.0, 1.25],
             conv_pads=[2, 3], res_tol=1e-7)
    rows = B.run(c)
    keys = [(r["axis"], r["L"], r["N"], r["pad"]) for r in rows]
    # boxes keep dx in multiples of 32 points, at least 96; pad 2 repeats the
    # grid variant N = 64 and is skipped
    assert keys == [("grid", 6.0, 48, 2), ("grid", 6.0, 64, 2), ("box", 6.0, 96, 2), ("box", 7.5, 96, 2),
                    ("pad", 6.0, 64, 3)]
    assert all("Nv" in r and "R_1pct" in r and r["converged"] for r in rows)
    _, _, obs = B.ground_state(dict(c, L=7.5, Ngrid=96), verbose=False)
    assert abs(rows[3]["E"] - obs["E"]) <= 1e-9 * abs(obs["E"])
    assert (tmp_path / "convergence.csv").is_file()


def test_sweep_over_a_solver_key_solves_each_point_with_its_own_value():
    """Without continuation the sweep batches its points' seeds; a swept solver
    setting (here maxit, and res_tol) must still apply per point, as the
    reference's point-by-point solves apply it."""
    import becgpp_amd as B
    c = _cfg(B, mode="sweep", sweep_param="maxit", sweep_values=[5, 40], nseeds=1, seed="gaussian",
             sweep_continuation=False, beta2=40.0, res_tol=1e-12)
    rows = B.run(c)
    assert [r["iters"] for r in rows] == [5, 40]
    for r in rows:
        _, _, o = B.ground_state(dict(c, mode="single", maxit=int(r["maxit"])), verbose=False)
        assert abs(r["E"] - o["E"]) <= 1e-12 * abs(o["E"])
    c2 = _cfg(B, mode="sweep", sweep_param="res_tol", sweep_values=[1e-2, 1e-7], nseeds=1, seed="gaussian",
              sweep_continuation=False, beta2=40.0)
    r2 = B.run(c2)
    assert r2[0]["iters"] < r2[1]["iters"] and r2[1]["resid_rel"] < 1e-7


def test_continuation_candidate_batched_with_the_seeds_equals_separate_solves():
    """With continuation each point's continuation candidate and its fresh
    seeds form one batch: every candidate's energy equals its own solve."""
    import becgpp_amd as B
    from becgpp_amd.seeds import make_seed
    c = _cfg(B, mode="sweep", sweep_param="beta2", sweep_values=[30.0, 45.0], nseeds=2, res_tol=1e-9)
    rows = B.run(c)
    prev = rows[0]["psi"]
    cp = dict(c, mode="single", beta2=45.0, tag=f"{c.get('tag', 'becgpp')}_beta245")
    G = B.make_grid(cp)
    _, _, oc = B.ground_state(cp, G=G, psi0=prev, verbose=False)
    cand = dict(x.split("=") for x in rows[1]["candidate_E"].split(";"))
    assert abs(float(cand["continuation"]) - oc["E"]) <= 1e-9 * abs(oc["E"])
    seeds = [dict(cp, rng=int(cp["rng"]) + j, seed=("tf", "gaussian")[j]) for j in range(2)]
    Es = [B.ground_state(s, G=G, psi0=make_seed(G, s), verbose=False)[2]["E"] for s in seeds]
    key = next(k for k in cand if k.startswith("seed:"))
    assert abs(float(cand[key]) - min(Es)) <= 1e-9 * abs(min(Es))
