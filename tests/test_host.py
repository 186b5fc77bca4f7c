"""Host-side logic (CPU, no GPU calls): configuration API, geometry and kernel
resolution, and the C-ABI library's exported symbols."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_default_cfg_rejects_unknown_and_maps_aliases():
    """reference tests/test_audit.py:146-151."""
    from becgpp_amd import default_cfg
    assert default_cfg(res=1e-7)["res_tol"] == 1e-7
    assert default_cfg(N=64)["Ngrid"] == 64
    assert default_cfg(etol=1e-6)["energy_tol"] == 1e-6
    assert default_cfg(dim="3D")["dimension"] == "3D"
    assert default_cfg(precond_shift=0.1)["precond_shift_min"] == 0.1
    with pytest.raises(ValueError):
        default_cfg(omega=0.9)
    with pytest.raises(ValueError):
        default_cfg(res=1e-7, res_tol=1e-6)


def test_defaults_are_safe():
    """reference tests/test_kernels.py:80-83."""
    from becgpp_amd.config import CFG_DEFAULTS
    assert CFG_DEFAULTS["Omega"] == 0.0 and CFG_DEFAULTS["G_C"] == 0.0 and CFG_DEFAULTS["Ngrid"] <= 128
    assert CFG_DEFAULTS["cg_beta"] == "pr_precond"


def test_default_cfg_returns_fresh_copies():
    from becgpp_amd import default_cfg
    a = default_cfg()
    a["beta2"] = -1.0
    assert default_cfg()["beta2"] == 100.0


def test_geometry_and_kernel_resolution():
    """reference tests/test_audit.py:237-239 and tests/test_kernels.py:66-73."""
    from becgpp_amd import default_cfg, geometry, resolve_kernel
    with pytest.raises(ValueError):
        geometry(dict(dimension="1D"))
    assert geometry(dict(dimension="quasi2d")) == (2, True)
    assert resolve_kernel(default_cfg(dimension="quasi2D", G_C=1.0, l_z=0.3))[0] == "q2d"
    assert resolve_kernel(default_cfg(dimension="quasi2D", G_C=1.0, l_z=0.0))[0] == "newton"
    assert resolve_kernel(default_cfg(dimension="2D", G_C=1.0))[0] == "log"
    assert resolve_kernel(default_cfg(dimension="3D", G_C=1.0))[0] == "newton"
    assert resolve_kernel(default_cfg(dimension="3D", G_C=0.0, kernel="newton"))[0] == "none"
    with pytest.raises(ValueError):
        resolve_kernel(default_cfg(dimension="3D", G_C=1.0, kernel="log"))
    with pytest.raises(ValueError):
        resolve_kernel(default_cfg(dimension="2D", G_C=1.0, kernel="q2d", l_z=0.0))
    with pytest.raises(ValueError):
        resolve_kernel(default_cfg(dimension="2D", G_C=1.0, kernel="yukawa"))


def test_auto_grid_repulsive_long_range():
    """reference tests/test_audit.py:242-244."""
    from becgpp_amd import auto_grid, default_cfg
    _, _, R = auto_grid(default_cfg(G_C=-1.0, beta2=10.0))
    assert R > 0


def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "becgpp_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(bec_[A-Za-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    from becgpp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libbecgpp_hip.so is not built (python -m becgpp_amd._build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "becgpp_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.replace("the oracle", ""), f


def test_run_dispatch_errors():
    """modes.py:473-478: an unknown mode is a ValueError, every reference mode
    is dispatched; sweep arguments are validated (modes.py:312-323)."""
    import becgpp_amd as B
    with pytest.raises(ValueError):
        B.run(B.default_cfg(mode="single") | {"mode": "nope"})
    assert sorted(B.DISPATCH) == ["convergence", "single", "smoke", "sweep", "tf_only", "validate"]
    with pytest.raises(ValueError):
        B.modes.mode_sweep(B.default_cfg(mode="sweep", sweep_param="seed"))
    with pytest.raises(ValueError):
        B.modes.mode_sweep(B.default_cfg(mode="sweep", sweep_values=[1.0, True]))


def test_run_id_is_the_reference_formula():
    """io.py:22-26: tag + md5 of the code version and the state-defining keys;
    excluded keys (maxit, mode, tag...) leave the hash unchanged."""
    import becgpp_amd as B
    c = B.default_cfg()
    a = B.run_id(c)
    assert a.startswith("becgpp_") and len(a) == len("becgpp_") + 10
    assert B.run_id(dict(c, maxit=5, mode="sweep")) == a
    assert B.run_id(dict(c, beta2=1.0)) != a
    assert B.run_id(dict(c, tag="x")) == "x_" + a.split("_")[1]
    # the reference's tests/test_audit.py:175-176
    assert B.run_id(dict(c, seed_winding=0)) != B.run_id(dict(c, seed_winding=2))


def test_smoke_ok_reflects_convergence(monkeypatch):
    """The reference's tests/test_audit.py:207-211 (modes.py:440-459): smoke is
    ok only when the gate passed and its run converged."""
    import becgpp_amd as B
    monkeypatch.setattr(B.modes, "mode_validate", lambda cfg: dict(passed=True))
    monkeypatch.setattr(B.modes, "mode_single", lambda cfg: dict(converged=False, E=0.0))
    monkeypatch.setattr(B.modes, "mode_tf_only", lambda cfg: None)
    assert B.modes.mode_smoke(B.default_cfg())["ok"] is False
    monkeypatch.setattr(B.modes, "mode_single", lambda cfg: dict(converged=True, E=0.0))
    assert B.modes.mode_smoke(B.default_cfg())["ok"] is True


def test_build_extension_keys():
    """Keys of this build (warm_fp32, warm_res_tol, autotune, host_loop) are
    accepted when given, absent otherwise, and leave the run id unchanged; the
    solver configuration carries the warm-phase tolerance only when asked."""
    import becgpp_amd as B
    from becgpp_amd.config import CFG_DEFAULTS, CFG_EXTENSIONS
    from becgpp_amd.solvers import _solver_cfg
    assert not set(CFG_EXTENSIONS) & set(CFG_DEFAULTS)
    c = B.default_cfg()
    assert "warm_fp32" not in c and "autotune" not in c
    w = B.default_cfg(warm_fp32=True, warm_res_tol=3e-4, autotune=True, host_loop=False)
    assert B.run_id(w) == B.run_id(c)
    assert _solver_cfg(c, False).warm_res_tol == 0.0
    assert _solver_cfg(w, False).warm_res_tol == 3e-4
    assert _solver_cfg(B.default_cfg(warm_fp32=True), False).warm_res_tol == 1e-4
    with pytest.raises(ValueError):
        B.default_cfg(warm_fp64=True)


def test_run_records_round_trip(tmp_path, monkeypatch):
    """records.py: checkpoint {psi, cfg, obs} under $GPP_OUTDIR/ckpt readable
    with weights_only=True; CSV header = sorted union of keys; per-run JSON."""
    import csv
    import json
    import torch
    from becgpp_amd import records
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    psi = torch.randn(8, 8, dtype=torch.complex128)
    cfg = dict(beta2=1.0, sweep_values=[1.0, 2.0], kernel="none")
    obs = dict(E=1.5, converged=True, stop_reason="converged", Elist=torch.tensor([1.0, 2.0]))
    records.save_checkpoint("becgpp_abc", psi, cfg, obs)
    assert os.path.isfile(tmp_path / "ckpt" / "becgpp_abc.pt")
    ck = torch.load(tmp_path / "ckpt" / "becgpp_abc.pt", weights_only=True)
    assert torch.equal(ck["psi"], psi) and ck["cfg"]["sweep_values"] == [1.0, 2.0] and ck["obs"]["E"] == 1.5
    assert torch.equal(records.load_checkpoint_psi("becgpp_abc"), psi)
    assert records.load_checkpoint_psi("becgpp_missing") is None
    records.write_csv(str(tmp_path / "s.csv"), [dict(b=1, a=2), dict(c=3)])
    rows = list(csv.DictReader(open(tmp_path / "s.csv")))
    assert list(rows[0].keys()) == ["a", "b", "c"] and rows[1]["c"] == "3"
    records.write_run_json("becgpp_abc", cfg, obs, "cuda")
    j = json.load(open(tmp_path / "becgpp_abc.json"))
    assert j["cfg"]["beta2"] == 1.0 and j["diagnostics"]["Elist"] == [1.0, 2.0] and "code_version" in j


def test_units_maps():
    """units.py:27-91 and the reference's tests/test_smoke.py:48-52: the trap
    length of 87Rb at omega = 2 pi 100 Hz is 1.078 um (a textbook value), the
    3D contact map is 4 pi N a_s / a_ho, the pancake map sqrt(8 pi) N a_s / l_z,
    self-gravity G m^2 N / (a_ho hbar omega); outputs scale back by hbar omega
    and a_ho; unknown regimes and pancakes without omega_z are ValueErrors."""
    import math
    import becgpp_amd as B
    m, w = 86.909180527 * 1.66053906660e-27, 2.0 * math.pi * 100.0
    sc = B.physical_scales(m, w)
    assert abs(sc["a_ho_m"] - 1.0785e-6) < 5e-10
    a_s, N = 5.3e-9, 1e5
    out = B.sim_params_from_physical("bec3d", m, w, N, a_s=a_s, verbose=False)
    assert out["dimension"] == "3D" and math.isclose(out["beta2"], 4 * math.pi * N * a_s / sc["a_ho_m"], rel_tol=1e-14)
    wz = 2.0 * math.pi * 1000.0
    pc = B.sim_params_from_physical("bec2d_pancake", m, w, N, a_s=a_s, omega_z=wz, verbose=False)
    lz = math.sqrt(1.054571817e-34 / (m * wz))
    assert pc["dimension"] == "quasi2D" and math.isclose(pc["beta2"], math.sqrt(8 * math.pi) * N * a_s / lz,
                                                         rel_tol=1e-14)
    g = B.sim_params_from_physical("selfgrav3d", m=1e-26, omega_perp=100.0, N=1e5, a_s=0.0, verbose=False)
    assert g["dimension"] == "3D" and g["kernel"] == "newton" and g["G_C"] > 0.0 and g["beta2"] == 0.0
    a = math.sqrt(1.054571817e-34 / (1e-26 * 100.0))
    assert math.isclose(g["G_C"], 6.67430e-11 * 1e-52 * 1e5 / (a * 1.054571817e-34 * 100.0), rel_tol=1e-14)
    gp = B.sim_params_from_physical("selfgrav_pancake", 1e-26, 100.0, 1e5, omega_z=1e4, verbose=False)
    assert gp["kernel"] == "q2d" and math.isclose(gp["l_z"], math.sqrt(1.054571817e-34 / (1e-26 * 1e4)) / a,
                                                  rel_tol=1e-14)
    o = B.outputs_to_physical(sc, E=2.0, mu=3.0, R=4.0)
    assert math.isclose(o["E_J"], 2.0 * 1.054571817e-34 * w) and math.isclose(o["R_m"], 4.0 * sc["a_ho_m"])
    with pytest.raises(ValueError):
        B.sim_params_from_physical("bec1d", m, w, N, verbose=False)
    with pytest.raises(ValueError):
        B.sim_params_from_physical("bec2d_pancake", m, w, N, a_s=a_s, verbose=False)


def test_paths_configure(tmp_path, monkeypatch):
    """paths.py: GPP_OUTDIR, then configure(base) at run time; the records go
    where BASE points."""
    import becgpp_amd as B
    monkeypatch.setattr(B.paths, "_configured", None)
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path / "env"))
    assert B.paths.BASE == str(tmp_path / "env") and B.records.out_dir() == str(tmp_path / "env")
    B.paths.configure(tmp_path / "cfg")
    assert B.paths.BASE == str(tmp_path / "cfg") and B.paths.CKPT_DIR == str(tmp_path / "cfg" / "ckpt")
    assert B.records.out_dir() == str(tmp_path / "cfg") and (tmp_path / "cfg" / "ckpt").is_dir()


def test_reference_exports(tmp_path, monkeypatch):
    """Every name of the reference's becgpp.__all__ (becgpp/__init__.py) that
    belongs to the path is exported; zip_results archives the output
    directory (modes.py:462-466)."""
    import zipfile
    import becgpp_amd as B
    for name in ("CODE_VERSION", "DEV", "CFG_DEFAULTS", "default_cfg", "paths", "make_grid", "geometry",
                 "resolve_kernel", "auto_grid", "resample_state", "long_range_phi", "kernel_convolution",
                 "q2d_kernel", "clear_kernel_cache", "energy_components", "apply_H", "compute_residual",
                 "observables", "mass_quantiles", "mass_radius", "radial_average", "droplet_radius", "extract_tf",
                 "make_seed", "ground_state", "ground_state_multiseed", "vortex_diagnostic", "lll_weight",
                 "diagnostics", "physical_scales", "sim_params_from_physical", "outputs_to_physical", "run",
                 "DISPATCH", "zip_results", "mode_validate", "mode_smoke", "mode_single", "mode_tf_only",
                 "mode_sweep", "mode_convergence"):
        assert hasattr(B, name), name
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path / "out"))
    B.records.write_csv(os.path.join(B.records.out_dir(), "x.csv"), [dict(a=1)])
    arch = B.zip_results("arch")
    assert zipfile.ZipFile(arch).namelist().count("x.csv") == 1


def test_reference_audit_host_cases():
    """The reference's host-only cases for the path: an empty sweep returns no
    rows and an unknown sweep key is refused (tests/test_audit.py:197-204); the
    defaults are safe: no rotation, no long-range term, a small grid
    (tests/test_kernels.py:75-78); G_C carries one factor of N, like
    beta2 = g N (tests/test_kernels.py:81-85)."""
    import becgpp_amd as B
    assert B.modes.mode_sweep(B.default_cfg(mode="sweep", sweep_values=[])) == []
    with pytest.raises(ValueError):
        B.modes.mode_sweep(B.default_cfg(mode="sweep", sweep_param="omega", sweep_values=[0.0], Ngrid=32))
    assert B.CFG_DEFAULTS["Omega"] == 0.0 and B.CFG_DEFAULTS["G_C"] == 0.0 and B.CFG_DEFAULTS["Ngrid"] <= 128
    a = B.sim_params_from_physical("selfgrav3d", m=1e-26, omega_perp=100.0, N=1e4, verbose=False)
    b = B.sim_params_from_physical("selfgrav3d", m=1e-26, omega_perp=100.0, N=2e4, verbose=False)
    assert abs(b["G_C"] / a["G_C"] - 2.0) < 1e-12


def test_uneven_schedule_serialises_groups_that_share_a_plan():
    """Groups on one geometry (same N, L, pad, trap) share a native plan and
    must not run concurrently; distinct geometries may."""
    from becgpp_amd.solvers import uneven_schedule

    def g(N, kernel="q2d", l_z=0.5):
        return dict(ndim=2, N=N, L=16.0, pad=2, s=2.0, trap_coeff=0.5, kernel=kernel, l_z=l_z)

    groups, conc = uneven_schedule([g(128), g(192), g(256), g(128), g(192)])
    assert conc and sorted(map(sorted, groups)) == [[0, 3], [1, 4], [2]]
    groups, conc = uneven_schedule([g(128), g(128, kernel="none", l_z=0.0), g(256)])
    assert not conc and len(groups) == 3
    groups, conc = uneven_schedule([g(128), g(128, l_z=0.25)])
    assert not conc and len(groups) == 2
    groups, conc = uneven_schedule([g(128), g(128)])
    assert not conc and groups == [[0, 1]]


def test_uneven_calls_take_one_batch_of_every_grid():
    """ground_state_uneven's native calls: each grid's instances split by
    solver setting and batch cap, call r advancing the r-th batch of every
    grid; every instance solved exactly once; at most eight grids a call."""
    from becgpp_amd.solvers import GRIDS_MAX, uneven_calls
    # grid 0: 5 instances, two settings; grid 1: 3 instances (cap 2); grid 2: 1
    groups = [[0, 1, 2, 3, 4], [5, 6, 7], [8]]
    keys = ["a", "b", "a", "a", "b", "a", "a", "a", "c"]
    calls = uneven_calls(groups, keys, [64, 2, 64])
    assert calls == [[[0, 2, 3], [5, 6], [8]], [[1, 4], [7]]]
    # a batch never mixes settings or grids, and the cap holds
    gid = {k: g for g, ks in enumerate(groups) for k in ks}
    for call in calls:
        assert len({gid[kk[0]] for kk in call}) == len(call)
        for kk in call:
            assert len({keys[k] for k in kk}) == 1 and len({gid[k] for k in kk}) == 1
    assert sorted(k for call in calls for kk in call for k in kk) == list(range(9))
    # more grids than one call takes
    many = [[i] for i in range(GRIDS_MAX + 3)]
    calls = uneven_calls(many, ["a"] * len(many), [64] * len(many))
    assert [len(c) for c in calls] == [GRIDS_MAX, 3]
    assert uneven_calls([], [], []) == []


def test_batches_split_by_solver_settings():
    """A batched loop shares its solver settings: configurations that differ
    in any solver key get different batch keys, the physics does not."""
    import becgpp_amd as B
    from becgpp_amd.solvers import SOLVER_KEYS, _solver_key
    c = B.default_cfg(mode="single")
    assert _solver_key(c) == _solver_key(dict(c, beta2=7.0, Omega=0.3, G_C=2.0, beta3=1.0))
    for k, v in (("res_tol", 1e-9), ("maxit", 17), ("linesearch_max", 3), ("cg_beta", "pr"),
                 ("precond_shift_min", 0.25), ("warm_fp32", True), ("line_search", "trial"), ("step", 0.5)):
        assert k in SOLVER_KEYS
        assert _solver_key(dict(c, **{k: v})) != _solver_key(c), k


def test_checkpoint_with_numpy_scalars_resumes(tmp_path, monkeypatch):
    """The reference saves obs holding numpy float64 scalars: its checkpoints
    load with the weights-only loader."""
    import numpy as np
    import torch
    monkeypatch.setenv("GPP_OUTDIR", str(tmp_path))
    from becgpp_amd import records
    psi = torch.arange(4, dtype=torch.float64).to(torch.complex128)
    torch.save(dict(psi=psi, cfg={"beta2": 1.0}, obs={"R_1pct": np.float64(1.25), "E": 1.0}),
               records.ckpt_path("rid123"))
    got = records.load_checkpoint_psi("rid123")
    assert got is not None and torch.equal(got, psi)


def test_library_noise_equals_the_oracle_generator():
    """bec_noise (host code of the library, no device work) draws the numbers
    the oracle's counter_normals draws: the same integers, so the normals agree
    to the last-bit rounding of log and cos in the two maths libraries."""
    import ctypes
    import numpy as np
    import oracle as O
    from becgpp_amd import _lib
    lib = _lib.load()
    for rng, stream, first in ((0, 0, 0), (3, 1, 12345), (2**33 + 5, 0, 7)):
        out = (ctypes.c_double * 64)()
        assert lib.bec_noise(rng, stream, first, 64, out) == 0
        ref = O.counter_normals(rng, stream, 64, first=first)
        assert np.abs(np.array(out) - ref).max() <= 4e-16 * max(1.0, np.abs(ref).max())


def test_solver_entry_points_reject_bad_arguments_before_device_work():
    """bec_ground_state_queue and bec_ground_state_uneven check their
    arguments on the host and return status 1 with a message (a ValueError
    through _lib.check) before touching a device: null arguments, a grid
    count outside [1, 8], batch sizes outside [1, 64], the host-driven loop
    and two grids of one call sharing a plan.  The plan handles here are
    never dereferenced."""
    import ctypes
    from becgpp_amd import _lib
    lib = _lib.load()

    def err(status):
        assert status == 1, status
        return lib.bec_last_error().decode()

    out = (_lib.SolveOut * 4)()
    ph = (_lib.Phys * 4)()
    sc = _lib.SolverCfg()
    assert "null" in err(lib.bec_ground_state_queue(None, 3, 10, ph, ctypes.byref(sc), None, out))
    fake = [ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)]

    def uneven(ng, plans, nbs, host_loop=0):
        cfg = _lib.SolverCfg()
        cfg.host_loop = host_loop
        cps = (ctypes.POINTER(_lib.SolverCfg) * max(ng, 1))(*([ctypes.pointer(cfg)] * max(ng, 1)))
        psi = (ctypes.c_void_p * max(ng, 1))(*([0x3000] * max(ng, 1)))
        hs = (ctypes.c_void_p * max(ng, 1))(*[p.value for p in plans])
        return lib.bec_ground_state_uneven(ng, hs, (ctypes.c_int * max(ng, 1))(*nbs), ph, cps, psi, out)

    assert "number of grids" in err(uneven(0, fake[:1], [1]))
    assert "number of grids" in err(uneven(9, fake[:1] * 9, [1] * 9))
    assert "batch size" in err(uneven(2, fake, [1, 65]))
    assert "device-driven" in err(uneven(1, fake[:1], [1], host_loop=1))
    assert "plan of its own" in err(uneven(2, [fake[0], fake[0]], [1, 1]))
    with pytest.raises(ValueError):
        _lib.check(uneven(2, [fake[1], fake[1]], [2, 2]))
