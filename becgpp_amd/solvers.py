"""Ground-state solver (becgpp/solvers.py:48-263).

``ground_state`` hands the whole Riemannian preconditioned CG loop to
``bec_ground_state`` in libbecgpp_hip: fields stay in HBM, scalar decisions
(Polak-Ribiere beta, descent fallbacks, Barzilai-Borwein step, Armijo test,
stop rules) follow the reference in the same order.  The final residual and
observables are recomputed from the returned state, as in solvers.py:227-229.
"""
import ctypes

import torch

from . import _lib
from .grid import _ptr, as_field, make_grid, plan_of
from .interactions import ensure_kernel
from .operators import compute_residual, observables, phys_for
from .seeds import make_seed

BATCH_MAX = 64
GRIDS_MAX = 8           # grids advanced together by one native call


def batchable(G):
    """Whether the grid supports batched solves (bec_plan_max_batch): each
    instance's lines must fill whole workgroups of its line passes, which
    excludes small or awkward N (2D N = 8, 20, 50, ...)."""
    pl = plan_of(G)
    nb = ctypes.c_int()
    _lib.check(pl.lib.bec_plan_max_batch(pl.h, ctypes.byref(nb)))
    return nb.value > 1


def _solver_cfg(cfg, verbose):
    cg_beta = str(cfg.get("cg_beta", "pr_precond")).lower()
    if cg_beta not in _lib.CG_BETA:
        raise ValueError(f"cg_beta must be 'pr', 'pr_precond' or 'none', got {cg_beta!r}")
    return _lib.SolverCfg(
        int(cfg["maxit"]), float(cfg["res_tol"]), float(cfg.get("energy_tol", 1e-9)),
        int(cfg.get("conv_window", 2000)), float(cfg.get("step", 1.0)), float(cfg.get("step_max", 3.0)),
        int(cfg.get("linesearch_max", 8)), int(cfg.get("linesearch_stall", 10)),
        float(cfg.get("precond_shift_min", 0.5)), _lib.CG_BETA[cg_beta], int(cfg.get("cg_restart", 30)),
        1 if verbose else 0, int(cfg.get("check", 200)), 1 if cfg.get("host_loop", False) else 0,
        float(cfg.get("warm_res_tol", 1e-4)) if cfg.get("warm_fp32", False) else 0.0,
        _line_poly(cfg))


def _line_poly(cfg):
    mode = str(cfg.get("line_search", "poly")).lower()
    if mode not in ("poly", "trial"):
        raise ValueError(f"line_search must be 'poly' or 'trial', got {mode!r}")
    return 1 if mode == "poly" else 0


def solve_native(cfg, G, psi0, verbose=False, record=False):
    """Run the device loop on psi0 (copied).  Returns (psi, raw result, trace)."""
    gc = float(cfg["G_C"])
    lr = G["kernel"] != "none" and abs(gc) > 1e-15
    pl = ensure_kernel(G, "K") if lr else plan_of(G)
    if cfg.get("autotune", False):
        pl.autotune(1)
    psi = as_field(psi0, G).clone()
    sc = _solver_cfg(cfg, verbose)
    ph = phys_for(G, cfg["beta2"], cfg["beta3"], cfg["Omega"], gc)
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
    obs = _finish(psi, res, G, cfg)
    if record:
        obs["trace"] = trace
    return psi.detach(), G, obs


def solve_native_batch(cfgs, G, psi0s, verbose=False):
    """Run up to 64 solves on one grid together (bec_ground_state_batch).

    cfgs share the grid, kernel and solver knobs; beta2, beta3, Omega and G_C
    may differ per instance.  Returns (list of psi, list of raw results)."""
    nb = len(cfgs)
    if nb != len(psi0s):
        raise ValueError("one initial state per configuration")
    if not 1 <= nb <= BATCH_MAX:
        raise ValueError(f"batch size must be in [1, {BATCH_MAX}]")
    lr_any = G["kernel"] != "none" and any(abs(float(c["G_C"])) > 1e-15 for c in cfgs)
    pl = ensure_kernel(G, "K") if lr_any else plan_of(G)
    if cfgs[0].get("autotune", False):
        pl.autotune(nb)
    psi = torch.stack([as_field(p0, G) for p0 in psi0s]).contiguous()
    sc = _solver_cfg(cfgs[0], verbose)
    ph = (_lib.Phys * nb)(*[phys_for(G, c["beta2"], c["beta3"], c["Omega"], c["G_C"]) for c in cfgs])
    res = (_lib.SolveOut * nb)()
    _lib.check(pl.lib.bec_ground_state_batch(pl.h, nb, ph, ctypes.byref(sc), _ptr(psi), res))
    return [psi[b] for b in range(nb)], [res[b] for b in range(nb)]


def solve_native_queue(cfgs, G, psi0s, slots=BATCH_MAX, verbose=False):
    """Run more solves than one batch holds on one grid as a queue
    (bec_ground_state_queue): `slots` instances advance together, and each
    slot whose solve has stopped takes the next waiting one, so the batch
    stays full until the queue drains.  cfgs share the solver knobs, as in
    solve_native_batch.  Returns (list of psi, list of raw results)."""
    nq = len(cfgs)
    if nq != len(psi0s):
        raise ValueError("one initial state per configuration")
    if nq < 1 or not 1 <= slots <= BATCH_MAX:
        raise ValueError(f"a queue needs solves and 1 to {BATCH_MAX} slots")
    lr_any = G["kernel"] != "none" and any(abs(float(c["G_C"])) > 1e-15 for c in cfgs)
    pl = ensure_kernel(G, "K") if lr_any else plan_of(G)
    if cfgs[0].get("autotune", False):
        pl.autotune(min(nq, slots))
    psi = torch.stack([as_field(p0, G) for p0 in psi0s]).contiguous()
    sc = _solver_cfg(cfgs[0], verbose)
    ph = (_lib.Phys * nq)(*[phys_for(G, c["beta2"], c["beta3"], c["Omega"], c["G_C"]) for c in cfgs])
    res = (_lib.SolveOut * nq)()
    _lib.check(pl.lib.bec_ground_state_queue(pl.h, int(slots), nq, ph, ctypes.byref(sc), _ptr(psi), res))
    return [psi[b] for b in range(nq)], [res[b] for b in range(nq)]


def _finish(psi, res, G, cfg):
    """Final residual and observables of a returned state (solvers.py:227-235)."""
    resid_abs, _mu, resid_rel = compute_residual(psi, G, cfg)
    converged = bool(res.converged or (resid_rel < float(cfg["res_tol"])))
    obs = observables(psi, G, cfg)
    obs.update(iters=res.iters, walltime=res.walltime, resid_abs=resid_abs, resid_rel=resid_rel,
               min_resid=res.best_resid, converged=converged,
               stop_reason=_lib.STOP_REASONS[res.stop_reason], preconditioner_shift=res.sigma)
    if res.warm_iters:
        obs["warm_iters"] = res.warm_iters
    return obs


# the configuration keys a solve reads besides the physics (_solver_cfg): one
# batched loop shares them, so configurations that differ in any of them are
# solved in separate loops
SOLVER_KEYS = ("maxit", "res_tol", "energy_tol", "conv_window", "step", "step_max", "linesearch_max",
               "linesearch_stall", "precond_shift_min", "cg_beta", "cg_restart", "check", "host_loop",
               "warm_fp32", "warm_res_tol", "line_search")


def _solver_key(c):
    sc = _solver_cfg(c, False)
    return tuple(getattr(sc, f) for f, _ in sc._fields_)


def ground_state_batch(cfgs, G=None, psi0s=None, verbose=False):
    """ground_state for many configurations of one grid at once: each entry of
    the result is what ground_state(cfgs[i], G, psi0s[i]) returns, computed
    by the batched device loop (64 instances per launch at most).  The
    physics (beta2, beta3, Omega, G_C) may differ per instance; instances
    whose solver settings differ (SOLVER_KEYS) run in loops of their own."""
    if G is None:
        G = make_grid(cfgs[0])
    if psi0s is None:
        psi0s = [make_seed(G, c) for c in cfgs]
    sets = {}
    for i, c in enumerate(cfgs):
        sets.setdefault(_solver_key(c), []).append(i)
    out = [None] * len(cfgs)
    for ks in sets.values():
        if len(ks) > BATCH_MAX and not cfgs[ks[0]].get("host_loop", False):
            # more solves than one batch: a queue that refills stopped slots
            psis, rs = solve_native_queue([cfgs[k] for k in ks], G, [psi0s[k] for k in ks], verbose=verbose)
            for k, p, r in zip(ks, psis, rs):
                out[k] = (p.detach(), G, _finish(p, r, G, cfgs[k]))
            continue
        for lo in range(0, len(ks), BATCH_MAX):
            kk = ks[lo:lo + BATCH_MAX]
            psis, rs = solve_native_batch([cfgs[k] for k in kk], G, [psi0s[k] for k in kk], verbose=verbose)
            for k, p, r in zip(kk, psis, rs):
                out[k] = (p.detach(), G, _finish(p, r, G, cfgs[k]))
    return out


def _plan_key(G):
    return (int(G["ndim"]), int(G["N"]), round(float(G["L"]), 14), int(G["pad"]), round(float(G["s"]), 14),
            round(float(G.get("trap_coeff", 0.5)), 14))


def _grid_key(G):
    return (int(G["ndim"]), int(G["N"]), round(float(G["L"]), 14), int(G["pad"]), round(float(G["s"]), 14),
            round(float(G.get("trap_coeff", 0.5)), 14), str(G["kernel"]), round(float(G.get("l_z", 0.0)), 14))


def uneven_schedule(grids):
    """Groups of instance indices sharing a grid (geometry, kernel, l_z), and
    whether their loops may run together in one native call.  Grids of one
    geometry share a native plan (plan_of), whose kernel spectrum, streams
    and workspaces one loop uses at a time, so groups that share a geometry
    (the same grid with and without long range, or two l_z) run one after
    the other; loops run together need a plan each."""
    sets = {}
    for i, G in enumerate(grids):
        sets.setdefault(_grid_key(G), []).append(i)
    groups = list(sets.values())
    plan_keys = [_plan_key(grids[ks[0]]) for ks in groups]
    return groups, len(groups) > 1 and len(set(plan_keys)) == len(plan_keys)


def solve_native_uneven(cfg_sets, grids, psi_sets, verbose=False):
    """Run the batches of several grids together (bec_ground_state_uneven):
    cfg_sets[g] and psi_sets[g] are the configurations and initial states of
    grid grids[g] (distinct native plans, at most 64 instances each, one set
    of solver settings per grid).  The steps of every grid are captured into
    one hipGraph, so each graph launch advances all instances of all grids.
    Returns, per grid, (list of psi, list of raw results)."""
    ng = len(grids)
    if not (ng == len(cfg_sets) == len(psi_sets)):
        raise ValueError("one configuration set and one state set per grid")
    if not 1 <= ng <= GRIDS_MAX:
        raise ValueError(f"number of grids must be in [1, {GRIDS_MAX}]")
    plans, psis, phs, scs, nbs = [], [], [], [], []
    for cs, G, p0s in zip(cfg_sets, grids, psi_sets):
        if len(cs) != len(p0s):
            raise ValueError("one initial state per configuration")
        if not 1 <= len(cs) <= BATCH_MAX:
            raise ValueError(f"batch size must be in [1, {BATCH_MAX}]")
        lr_any = G["kernel"] != "none" and any(abs(float(c["G_C"])) > 1e-15 for c in cs)
        pl = ensure_kernel(G, "K") if lr_any else plan_of(G)
        if cs[0].get("autotune", False):
            pl.autotune(len(cs))
        plans.append(pl)
        psis.append(torch.stack([as_field(p0, G) for p0 in p0s]).contiguous())
        phs += [phys_for(G, c["beta2"], c["beta3"], c["Omega"], c["G_C"]) for c in cs]
        scs.append(_solver_cfg(cs[0], verbose))
        nbs.append(len(cs))
    lib = plans[0].lib
    tot = sum(nbs)
    res = (_lib.SolveOut * tot)()
    _lib.check(lib.bec_ground_state_uneven(
        ng, (ctypes.c_void_p * ng)(*[pl.h.value for pl in plans]), (ctypes.c_int * ng)(*nbs),
        (_lib.Phys * tot)(*phs), (ctypes.POINTER(_lib.SolverCfg) * ng)(*[ctypes.pointer(sc) for sc in scs]),
        (ctypes.c_void_p * ng)(*[t.data_ptr() for t in psis]), res))
    out, o = [], 0
    for psi, nb in zip(psis, nbs):
        out.append(([psi[b] for b in range(nb)], [res[o + b] for b in range(nb)]))
        o += nb
    return out


def ground_state_uneven(cfgs, psi0s=None, grids=None, verbose=False):
    """ground_state for configurations on different grids (uneven sizes).

    Instances sharing a grid form batches as in ground_state_batch (at most
    64, split by solver settings); the batches of different grids are solved
    by one call of the native library, whose every graph launch advances all
    instances of all grids (solve_native_uneven).  Grids that share a native
    plan (the same geometry with another kernel or l_z), and host-driven
    loops, are solved one grid after another.  grids[i] (optional) is the
    grid of cfgs[i].  Returns, in input order, what ground_state(cfgs[i],
    grids[i], psi0s[i]) returns."""
    n = len(cfgs)
    if grids is None:
        grids = [make_grid(c) for c in cfgs]
    if psi0s is None:
        psi0s = [make_seed(G, c) for G, c in zip(grids, cfgs)]
    groups, together = uneven_schedule(grids)
    together = together and not any(c.get("host_loop", False) for c in cfgs)
    out = [None] * n
    if not together:
        for ks in groups:
            for k, r in zip(ks, ground_state_batch([cfgs[k] for k in ks], G=grids[ks[0]],
                                                   psi0s=[psi0s[k] for k in ks], verbose=verbose)):
                out[k] = r
        return out
    caps = [BATCH_MAX if batchable(grids[ks[0]]) else 1 for ks in groups]
    for part in uneven_calls(groups, [_solver_key(c) for c in cfgs], caps):
        got = solve_native_uneven([[cfgs[k] for k in kk] for kk in part], [grids[kk[0]] for kk in part],
                                  [[psi0s[k] for k in kk] for kk in part], verbose=verbose)
        for kk, (psis, rs) in zip(part, got):
            G = grids[kk[0]]
            for k, p, r_ in zip(kk, psis, rs):
                out[k] = (p.detach(), G, _finish(p, r_, G, cfgs[k]))
    return out


def uneven_calls(groups, keys, caps):
    """The native calls of ground_state_uneven: each grid's instances
    (groups[g]) split into batches of one solver setting (keys[i] of
    instance i) and at most caps[g] instances, in order; call r takes the r-th
    batch of every grid that has one (at most GRIDS_MAX grids per call, the
    rest in further calls).  Returns a list of calls, each a list of batches
    (lists of instance indices) on distinct grids."""
    queues = []
    for ks, cap in zip(groups, caps):
        sets = {}
        for k in ks:
            sets.setdefault(keys[k], []).append(k)
        queues.append([kk[lo:lo + cap] for kk in sets.values() for lo in range(0, len(kk), cap)])
    calls = []
    for r in range(max((len(q) for q in queues), default=0)):
        row = [q[r] for q in queues if r < len(q)]
        for lo in range(0, len(row), GRIDS_MAX):
            calls.append(row[lo:lo + GRIDS_MAX])
    return calls


def _seed_cfgs(cfg, n):
    """Per-seed configurations of ground_state_multiseed (solvers.py:244-251)."""
    rotating = abs(float(cfg.get("Omega", 0.0))) >= 0.3
    seed_cycle = ("triangular", "tf", "gaussian") if rotating else ("tf", "gaussian")
    cs = []
    for j in range(max(1, n)):
        c = dict(cfg)
        c["rng"] = int(cfg.get("rng", 0)) + j
        c["seed"] = seed_cycle[j % len(seed_cycle)]
        cs.append(c)
    return cs


def ground_state_multiseed(cfg, G=None, verbose=False):
    """nseeds solves with the reference's seed cycle; lowest energy wins
    (solvers.py:238-263).  The seeds are solved as one batch."""
    n = int(cfg.get("nseeds", 1))
    if G is None:
        G = make_grid(cfg)
    cs = _seed_cfgs(cfg, n)
    if len(cs) == 1 or not batchable(G):
        runs = [ground_state(c, G=G, verbose=verbose) for c in cs]
    else:
        runs = ground_state_batch(cs, G=G, verbose=verbose)
    best = None
    Es = []
    for c, (psi, _G, obs) in zip(cs, runs):
        Es.append(obs["E"])
        if best is None or obs["E"] < best[0]:
            best = (obs["E"], psi, obs, c)
    _, psi, obs, cbest = best
    obs["seed_energy_spread"] = (max(Es) - min(Es)) if len(Es) > 1 else 0.0
    obs["seed_used"] = cbest["seed"]
    obs["seed_energies"] = ";".join(f"{e:.10g}" for e in Es)
    return psi, G, obs
