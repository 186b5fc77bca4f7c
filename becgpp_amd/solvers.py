"""Ground-state solver (becgpp/solvers.py:48-263).

``ground_state`` hands the whole Riemannian preconditioned CG loop to
``bec_ground_state`` in libbecgpp_hip: fields stay in HBM, scalar decisions
(Polak-Ribiere beta, descent fallbacks, Barzilai-Borwein step, Armijo test,
stop rules) follow the reference in the same order.  The final residual and
observables are recomputed from the returned state, as in solvers.py:227-229.
"""
import ctypes

from . import _lib
from .grid import _ptr, as_field, make_grid, plan_of
from .interactions import ensure_kernel
from .operators import compute_residual, observables
from .seeds import make_seed


def _solver_cfg(cfg, verbose):
    cg_beta = str(cfg.get("cg_beta", "pr_precond")).lower()
    if cg_beta not in _lib.CG_BETA:
        raise ValueError(f"cg_beta must be 'pr', 'pr_precond' or 'none', got {cg_beta!r}")
    return _lib.SolverCfg(
        int(cfg["maxit"]), float(cfg["res_tol"]), float(cfg.get("energy_tol", 1e-9)),
        int(cfg.get("conv_window", 2000)), float(cfg.get("step", 1.0)), float(cfg.get("step_max", 3.0)),
        int(cfg.get("linesearch_max", 8)), int(cfg.get("linesearch_stall", 10)),
        float(cfg.get("precond_shift_min", 0.5)), _lib.CG_BETA[cg_beta], int(cfg.get("cg_restart", 30)),
        1 if verbose else 0, int(cfg.get("check", 200)))


def solve_native(cfg, G, psi0, verbose=False, record=False):
    """Run the device loop on psi0 (copied).  Returns (psi, raw result, trace)."""
    gc = float(cfg["G_C"])
    lr = G["kernel"] != "none" and abs(gc) > 1e-15
    pl = ensure_kernel(G, "K") if lr else plan_of(G)
    psi = as_field(psi0, G).clone()
    sc = _solver_cfg(cfg, verbose)
    ph = _lib.phys(cfg["beta2"], cfg["beta3"], cfg["Omega"], gc)
    res = _lib.SolveOut()
    cap = int(cfg["maxit"]) + 2 if record else 0
    tr = (ctypes.c_double * (4 * cap))() if cap else None
    _lib.check(pl.lib.bec_ground_state(pl.h, ctypes.byref(ph), ctypes.byref(sc), _ptr(psi),
                                       ctypes.byref(res), tr, cap))
    trace = None
    if record:
        trace = [(tr[4 * i], int(tr[4 * i + 1]), tr[4 * i + 2], tr[4 * i + 3]) for i in range(res.n_trace)]
    return psi, res, trace


def ground_state(cfg, G=None, psi0=None, verbose=True):
    """Riemannian preconditioned CG on the unit sphere (solvers.py:48-235)."""
    if G is None:
        G = make_grid(cfg)
    psi_init = psi0 if psi0 is not None else make_seed(G, cfg)
    record = bool(cfg.get("record_trace", False))
    psi, res, trace = solve_native(cfg, G, psi_init, verbose=verbose, record=record)
    resid_abs, _mu, resid_rel = compute_residual(psi, G, cfg)
    converged = bool(res.converged or (resid_rel < float(cfg["res_tol"])))
    obs = observables(psi, G, cfg)
    obs.update(iters=res.iters, walltime=res.walltime, resid_abs=resid_abs, resid_rel=resid_rel,
               min_resid=res.best_resid, converged=converged,
               stop_reason=_lib.STOP_REASONS[res.stop_reason], preconditioner_shift=res.sigma)
    if record:
        obs["trace"] = trace
    return psi.detach(), G, obs


def ground_state_multiseed(cfg, G=None, verbose=False):
    """nseeds solves with the reference's seed cycle; lowest energy wins
    (solvers.py:238-263)."""
    n = int(cfg.get("nseeds", 1))
    if G is None:
        G = make_grid(cfg)
    rotating = abs(float(cfg.get("Omega", 0.0))) >= 0.3
    seed_cycle = ("triangular", "tf", "gaussian") if rotating else ("tf", "gaussian")
    best = None
    Es = []
    for j in range(max(1, n)):
        c = dict(cfg)
        c["rng"] = int(cfg.get("rng", 0)) + j
        c["seed"] = seed_cycle[j % len(seed_cycle)]
        psi, G, obs = ground_state(c, G=G, verbose=verbose)
        Es.append(obs["E"])
        if best is None or obs["E"] < best[0]:
            best = (obs["E"], psi, obs, c)
    _, psi, obs, cbest = best
    obs["seed_energy_spread"] = (max(Es) - min(Es)) if len(Es) > 1 else 0.0
    obs["seed_used"] = cbest["seed"]
    obs["seed_energies"] = ";".join(f"{e:.10g}" for e in Es)
    return psi, G, obs

