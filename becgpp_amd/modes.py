"""run(cfg) and the modes of the ground-state path (becgpp/modes.py:30-500).

The dispatch, mode names, run ids and error behaviour are the reference's,
and all six of its modes run here on the GPU solver: ``single``
(ground_state / ground_state_multiseed), ``sweep`` (every fresh seed of every
point as one batched solve when continuation is off; with continuation each
point's fresh seeds and its continuation candidate form one batch),
``validate`` (the analytic gate, every convolution and energy by
libbecgpp_hip), ``smoke``, ``tf_only`` and ``convergence``.  Results are
returned as the reference returns them.  ``single`` and ``sweep`` resume from
and write the reference's checkpoints, CSV and per-run JSON (records.py) with
the diagnostic bundle (energies, virial, radii, Thomas-Fermi reference,
vortex count, LLL weight); figures are not produced.  ``zip_output`` archives
the output directory after the mode, as the reference does.
"""
import hashlib
import json
import math
import os

import numpy as np
import torch

from .config import CFG_DEFAULTS, canonical_key
from .constants import CODE_VERSION, DEV
from .grid import auto_grid, dV, geometry, make_grid, norm_of, resample_state, resolve_kernel
from .interactions import long_range_phi
from .operators import energy_components
from .solvers import batchable, ground_state, ground_state_batch, ground_state_multiseed
from .thomasfermi import extract_tf
from .diagnostics import diagnostics, diagnostics_batch
from .seeds import make_seed
from . import records

# keys that cannot change the computed state (becgpp/io.py:9-19)
_RID_EXCLUDE = frozenset({
    "mode", "tag", "maxit", "check", "record_trace", "mem_warn_gb",
    "want_vortices", "want_lll", "tf_cache", "validate_N",
    "sweep_param", "sweep_values", "sweep_autobox", "sweep_continuation",
    "conv_Ngrids", "conv_box_factors", "conv_pads",
    "show_inline", "save_figs", "save_ckpt", "zip_output",
})


def run_id(cfg):
    """tag_md5[:10] of the code version and every state-defining key (io.py:22-26)."""
    keys = sorted(k for k in CFG_DEFAULTS if k not in _RID_EXCLUDE)
    raw = json.dumps(dict(code=CODE_VERSION, **{k: cfg.get(k, CFG_DEFAULTS[k]) for k in keys}),
                     sort_keys=True, default=str)
    return f"{cfg.get('tag', 'becgpp')}_{hashlib.md5(raw.encode()).hexdigest()[:10]}"


def _record(psi, obs, G, cfg, rid, diag=None):
    """The run's record: the solver's observables and the diagnostic bundle of
    the state (diagnostics.py:18-47; given when computed for a batch of
    states), then the run's identity (modes.py:227-236)."""
    d = {k: v for k, v in obs.items() if k != "trace"}
    d.update(diagnostics(psi, G, cfg) if diag is None else diag)
    d.update(rid=rid, seed_used=obs.get("seed_used", cfg.get("seed")),
             seed_energy_spread=obs.get("seed_energy_spread", 0.0),
             dimension=cfg["dimension"], kernel=G["kernel"], L=cfg["L"], N=cfg["Ngrid"], s=cfg["s"],
             Omega=cfg["Omega"], beta2=cfg["beta2"], beta3=cfg["beta3"], G_C=cfg["G_C"],
             l_z=cfg.get("l_z", 0.0), device=str(DEV))
    return d


# -----------------------------------------------------------------------------
# single
# -----------------------------------------------------------------------------
def mode_single(cfg):
    """modes.py:207-262: one configuration; nseeds > 1 solves the seeds as one
    batch and keeps the lowest energy."""
    rid = run_id(cfg)
    G = make_grid(cfg)
    psi0 = records.load_checkpoint_psi(rid)             # modes.py:216-224
    if psi0 is not None:
        print(f"[resume] {records.ckpt_path(rid)}")
        psi0 = psi0.to(DEV)
    if int(cfg.get("nseeds", 1)) > 1 and psi0 is None:
        psi, G, obs = ground_state_multiseed(cfg, G=G, verbose=True)
    else:
        psi, G, obs = ground_state(cfg, G=G, psi0=psi0, verbose=True)
    d = _record(psi, obs, G, cfg, rid)
    if cfg.get("save_ckpt", True):                      # modes.py:234-247
        records.save_checkpoint(rid, psi, cfg, d)
    records.write_csv(os.path.join(records.out_dir(), "single_summary.csv"), [d])
    records.write_run_json(rid, cfg, d, DEV)
    print(f"[single] {cfg['dimension']} s={cfg['s']} Omega={cfg['Omega']} beta2={cfg['beta2']} "
          f"beta3={cfg['beta3']} G_C={cfg['G_C']} kernel={G['kernel']}")
    print(f"         E={d['E']:.8f} mu={d['mu']:.6f} Lz={d['Lz']:.5f} Nv={d['Nv']}")
    print(f"         R90={d['R90']:.4f} TF({d['tf_kind']}) R90={d['tf_R90']:.4f} "
          f"virial_rel={d['virial_rel']:.2e} w_LLL={d['w_LLL']}")
    print(f"         resid={d['resid_rel']:.2e} converged={d['converged']} it={d['iters']}")
    d["psi"] = psi
    return d


# -----------------------------------------------------------------------------
# sweep
# -----------------------------------------------------------------------------
def _sweep_args(cfg):
    param = canonical_key(cfg.get("sweep_param", "G_C"))
    tmpl = CFG_DEFAULTS.get(param)
    if param not in CFG_DEFAULTS or isinstance(tmpl, bool) or not isinstance(tmpl, (int, float)):
        numeric = sorted(k for k, v in CFG_DEFAULTS.items()
                         if isinstance(v, (int, float)) and not isinstance(v, bool))
        raise ValueError(f"sweep_param={cfg.get('sweep_param')!r} is not a numeric CFG key; "
                         f"choose one of {numeric}")
    values = list(cfg.get("sweep_values", []))
    bad = [v for v in values if isinstance(v, bool) or not isinstance(v, (int, float))]
    if bad:
        raise ValueError(f"sweep_values must be numbers, got {bad}")
    return param, values


def _seed_cfgs(c, nseeds):
    rotating = abs(float(c.get("Omega", 0.0))) >= 0.3
    cycle = ("triangular", "tf", "gaussian") if rotating else ("tf", "gaussian")
    out = []
    for j in range(max(1, nseeds)):
        cj = dict(c)
        cj["rng"] = int(c.get("rng", 0)) + j
        cj["seed"] = cycle[j % len(cycle)] if nseeds > 1 else c.get("seed", "tf")
        out.append(cj)
    return out


def _grid_key(c):
    return (geometry(c), float(c["L"]), int(c["Ngrid"]), int(c.get("pad", 2)), float(c["s"]),
            float(c.get("trap_coeff", 0.5)), resolve_kernel(c)[0], float(c.get("l_z", 0.0)))


def mode_sweep(cfg):
    """modes.py:295-387: vary one numeric key; each point keeps the lowest
    energy over its fresh seeds (and the previous point's state when
    sweep_continuation is on).  Without continuation all points' seeds are
    independent and are solved as batches of up to 64 per grid."""
    param, values = _sweep_args(cfg)
    use_cont = bool(cfg.get("sweep_continuation", True))
    nseeds = int(cfg.get("nseeds", 1))
    print(f"SWEEP over {param!r} = {values}  (continuation={use_cont}, nseeds={nseeds})")
    if not values:
        print("[sweep] sweep_values is empty: nothing to do")
        return []
    rot = [abs(float(v if param == "Omega" else cfg.get("Omega", 0.0))) for v in values]
    if max(rot) >= 0.5 and nseeds < 2:
        print("[hint] rotating sweep with nseeds=1: continuation alone can stay on a vortex-free")
        print("       (metastable) branch. Use nseeds>=2 so each point also tries a vortex-lattice seed.")
    pts = []
    for val in values:
        c = dict(cfg)
        c[param] = val
        if cfg.get("sweep_autobox", False):             # modes.py:343-345
            L, N, _ = auto_grid(c)
            c["L"], c["Ngrid"] = L, N
        c["tag"] = f"{cfg.get('tag', 'becgpp')}_{param}{val:g}"
        pts.append(c)
    rows = [None] * len(pts)
    if not use_cont:
        # every fresh seed of every point is independent: batch them per grid
        groups = {}
        for ip, c in enumerate(pts):
            groups.setdefault(_grid_key(c), []).append(ip)
        for ips in groups.values():
            G = make_grid(pts[ips[0]])
            jobs = [(ip, cj) for ip in ips for cj in _seed_cfgs(pts[ip], nseeds)]
            if batchable(G):
                res = ground_state_batch([cj for _, cj in jobs], G=G)
            else:
                res = [ground_state(cj, G=G, verbose=False) for _, cj in jobs]
            # the selected state of every point of the grid, and their
            # diagnostic bundle in one batch
            sel = {}
            for ip in ips:
                cands = [(f"seed:{cj['seed']}", r) for (jp, cj), r in zip(jobs, res) if jp == ip]
                sel[ip] = min(cands, key=lambda t: t[1][2]["E"])
            diags = diagnostics_batch([sel[ip][1][0] for ip in ips], G, [pts[ip] for ip in ips])
            for ip, dg in zip(ips, diags):
                cands = [(f"seed:{cj['seed']}", r) for (jp, cj), r in zip(jobs, res) if jp == ip]
                rows[ip] = _pick(cands, pts[ip], param, values[ip], dg)
        for ip in range(len(pts)):
            _sweep_record(cfg, param, pts, rows, ip)
        return rows
    prev, G_prev = None, None
    for ip, c in enumerate(pts):
        G = make_grid(c)
        # the point's candidates (modes.py:350-361): the continuation from the
        # previous point's state (resampled onto a new grid) and, with several
        # seeds or no continuation, the fresh seeds -- solved as one batch
        cont = None
        if prev is not None:
            cont = prev if tuple(prev.shape) == tuple(G["R2"].shape) else resample_state(prev, G_prev, G)
        seeds = _seed_cfgs(c, nseeds) if (nseeds > 1 or cont is None) else []
        if nseeds <= 1:
            seeds = [dict(c, seed=c.get("seed", "tf"))] if seeds else []
        jobs = ([("continuation", c, cont)] if cont is not None else []) + \
            [(f"seed:{cj['seed']}", cj, make_seed(G, cj)) for cj in seeds]
        if len(jobs) > 1 and batchable(G):
            res = ground_state_batch([cj for _, cj, _ in jobs], G=G, psi0s=[p0 for _, _, p0 in jobs])
        else:
            res = [ground_state(cj, G=G, psi0=p0, verbose=False) for _, cj, p0 in jobs]
        cands = []
        if cont is not None:
            cands.append(("continuation", res[0]))
            res = res[1:]
            jobs = jobs[1:]
        if res:
            best = min(range(len(res)), key=lambda j: res[j][2]["E"])
            cands.append((jobs[best][0], res[best]))
        rows[ip] = _pick(cands, c, param, values[ip])
        prev, G_prev = rows[ip]["psi"], G
        _sweep_record(cfg, param, pts, rows, ip)
    return rows


def _sweep_record(cfg, param, pts, rows, ip):
    """modes.py:370-380: point ip's checkpoint ckpt/<rid>.pt, and the sweep's
    CSV rewritten with the points done so far (sweep_<param>.csv, or
    sweep_<tag>_<param>.csv for a tagged sweep)."""
    tag0 = cfg.get("tag", "becgpp")
    stem = param if tag0 == "becgpp" else f"{tag0}_{param}"
    plain = [{k: v for k, v in r.items() if k != "psi"} for r in rows[:ip + 1] if r is not None]
    if cfg.get("save_ckpt", True):
        records.save_checkpoint(rows[ip]["rid"], rows[ip]["psi"], pts[ip], plain[-1])
    records.write_csv(os.path.join(records.out_dir(), f"sweep_{stem}.csv"), plain)


def _pick(cands, c, param, val, diag=None):
    cands = sorted(cands, key=lambda t: t[1][2]["E"])      # variational: lowest energy wins
    branch, (psi, G, obs) = cands[0]
    d = _record(psi, obs, G, c, run_id(c), diag)
    d.update({param: val}, branch=branch, walltime=sum(t[1][2]["walltime"] for t in cands),
             candidate_E=";".join(f"{t[0]}={t[1][2]['E']:.10g}" for t in cands), psi=psi)
    print(f"    {param}={val} [{branch}] E={d['E']:.8f} mu={d['mu']:.5f} Lz={d['Lz']:.4f} "
          f"res={d['resid_rel']:.2e} conv={d['converged']}")
    return d


# -----------------------------------------------------------------------------
# validate: the analytic gate (modes.py:30-201)
# -----------------------------------------------------------------------------
_EULER_GAMMA = 0.5772156649015329


def _gauss_rho(G, a):
    psi = torch.exp(-G["R2"] / (2 * a * a)).to(torch.complex128)
    psi = psi / norm_of(psi, G)
    return (psi.real ** 2 + psi.imag ** 2).contiguous()


def _center(phi):
    c = phi.shape[0] // 2
    return phi[(c,) * phi.dim()].item()


def _F(t):
    if abs(t - 1.0) < 1e-12:
        return 1.0
    return math.acos(t) / math.sqrt(1 - t * t) if t < 1 else math.acosh(t) / math.sqrt(t * t - 1)


def mode_validate(cfg):
    """Gaussian Coulomb integrals (2D 1/r, -ln r, quasi-2D K_eff; 3D 1/r),
    the LLL identities, the flat-top TF closed form and the kernel
    convergence orders, with the reference's thresholds."""
    rows = []
    Nv = int(cfg.get("validate_N", 256))

    def row(test, case, q, e):
        rows.append(dict(test=test, case=case, quantity=q, rel_error=e))
        return e

    G = make_grid(dict(cfg, dimension="2D", s=2, kernel="newton", G_C=1.0, L=12.0, Ngrid=Nv))
    dv = dV(G)
    mg = mp = 0.0
    for a in (0.5, 1.0, 2.0):
        rho = _gauss_rho(G, a)
        phi = long_range_phi(rho, G, 1.0)
        Eg = 0.5 * float((phi * rho).sum()) * dv
        ex = math.sqrt(math.pi / 8.0) / a
        mg = max(mg, row("2D_gauss", f"a={a}", "Egrav", abs(Eg + ex) / ex))
        ex0 = math.sqrt(math.pi) / a
        mp = max(mp, row("2D_gauss", f"a={a}", "phi0", abs(_center(phi) + ex0) / ex0))
    Gl = make_grid(dict(cfg, dimension="2D", s=2, kernel="log", G_C=1.0, L=12.0, Ngrid=Nv))
    mlg = mlp = 0.0
    for a in (0.5, 2.0, 3.0):
        rho = _gauss_rho(Gl, a)
        phi = long_range_phi(rho, Gl, 1.0)
        Eg = 0.5 * float((phi * rho).sum()) * dv
        Eg_ex = 0.5 * (math.log(math.sqrt(2.0) * a) - 0.5 * _EULER_GAMMA)
        p0_ex = math.log(a) - 0.5 * _EULER_GAMMA
        mlg = max(mlg, row("2D_log_gauss", f"a={a}", "Egrav", abs(Eg - Eg_ex) / abs(Eg_ex)))
        mlp = max(mlp, row("2D_log_gauss", f"a={a}", "phi0", abs(_center(phi) - p0_ex) / abs(p0_ex)))
    mq = 0.0
    for lz in (0.25, 1.0):
        Gq = make_grid(dict(cfg, dimension="quasi2D", s=2, kernel="q2d", l_z=lz, G_C=1.0, L=12.0, Ngrid=Nv))
        for a in (1.0, 2.0):
            rho = _gauss_rho(Gq, a)
            phi = long_range_phi(rho, Gq, 1.0)
            Eg = 0.5 * float((phi * rho).sum()) * dv
            sp = a / math.sqrt(2.0)
            p0_ex = -(2 / math.sqrt(math.pi)) / (math.sqrt(2) * sp) * _F(lz / sp)
            Eg_ex = -0.5 * (2 / math.sqrt(math.pi)) / (math.sqrt(2) * a) * _F(lz / a)
            mq = max(mq, row("q2D_gauss", f"lz={lz},a={a}", "Egrav", abs(Eg - Eg_ex) / abs(Eg_ex)),
                     row("q2D_gauss", f"lz={lz},a={a}", "phi0", abs(_center(phi) - p0_ex) / abs(p0_ex)))
    ml = 0.0
    for m in (0, 1, 10, 30):
        z = (G["X"] + 1j * G["Y"]).to(torch.complex128)
        psi = (z ** m) * torch.exp(-0.5 * G["R2"])
        psi = psi / norm_of(psi, G)
        comp = energy_components(psi, G, 0.0, 0.0, 1.0, 0.0)
        Esp = (comp["Ekin"] + comp["Etrap"] + comp["Erot"]).item()
        r2 = float((G["R2"] * (psi.real ** 2 + psi.imag ** 2)).sum()) * dv
        ml = max(ml, row("2D_LLL", f"m={m}", "max|E-1|,|Lz-m|,|r2-m-1|",
                         max(abs(Esp - 1.0), abs(comp["Lz"].item() - m), abs(r2 - (m + 1)))))
    mtf = 0.0
    for ndim in (2, 3):
        Gf = make_grid(dict(cfg, dimension=("2D" if ndim == 2 else "3D"), s=2, Omega=1.0, kernel="none",
                            beta2=-1.0, beta3=1.0, G_C=0.0, L=(20.0 if ndim == 2 else 8.0),
                            Ngrid=(192 if ndim == 2 else 96)))
        for b2, b3 in ((-0.5, 1.0), (-1.5, 1.0)):
            tf = extract_tf(dict(cfg, dimension=("2D" if ndim == 2 else "3D"), s=2, Omega=1.0, kernel="none",
                                 beta2=b2, beta3=b3, G_C=0.0), Gf)
            rho0 = -3 * b2 / (4 * b3)
            mu = b2 * rho0 + b3 * rho0 ** 2
            e = max(abs(tf["rho0"] - rho0) / rho0, abs(tf["mu"] - mu) / max(abs(mu), 1e-12))
            mtf = max(mtf, row(f"{ndim}D_flat_tf", f"b2={b2}", "rho0_mu", e))
    G3 = make_grid(dict(cfg, dimension="3D", s=2, kernel="newton", G_C=1.0, L=8.0, Ngrid=96))
    m3 = 0.0
    for sig in (1.0, 1.5):
        phi = long_range_phi(_gauss_rho(G3, sig), G3, 1.0)
        ex = -2.0 / (math.sqrt(math.pi) * sig)
        m3 = max(m3, row("3D_gauss", f"sig={sig}", "phi0", abs(_center(phi) - ex) / abs(ex)))
    kernel_orders = {}
    for name, ndim, kern, Ns, sig, L in (("2D", 2, "newton", (192, 256, 384, 512), 1.0, 10.0),
                                        ("2D_log", 2, "log", (96, 128, 192, 256), 1.0, 10.0),
                                        ("3D", 3, "newton", (64, 96, 128, 160), 1.0, 8.0)):
        ex = (math.log(sig) - 0.5 * _EULER_GAMMA) if kern == "log" else (
            (-math.sqrt(math.pi) / sig) if ndim == 2 else (-2.0 / (math.sqrt(math.pi) * sig)))
        dxs, errs = [], []
        for Ni in Ns:
            Gk = make_grid(dict(cfg, dimension=("2D" if ndim == 2 else "3D"), s=2, kernel=kern, G_C=1.0,
                                L=L, Ngrid=int(Ni)))
            phi = long_range_phi(_gauss_rho(Gk, sig), Gk, 1.0)
            dxs.append(Gk["dx"])
            errs.append(abs(_center(phi) - ex) / abs(ex))
        order = float(np.polyfit(np.log(dxs), np.log(np.maximum(errs, 1e-16)), 1)[0])
        kernel_orders[name] = order
        for dxi, ei in zip(dxs, errs):
            rows.append(dict(test=f"{name}_kernel_order", case=f"dx={dxi:.4f}", quantity="phi0_relerr",
                             rel_error=ei, order=order))
    passed = (mg < 5e-3 and mp < 5e-3 and mlg < 5e-3 and mlp < 5e-3 and mq < 5e-3
              and ml < 1e-4 and mtf < 1e-12 and m3 < 5e-3)
    order_ok = all(o > 1.7 for o in kernel_orders.values())
    records.write_csv(os.path.join(records.out_dir(), "validation.csv"), rows)     # modes.py:191
    print(f"[validate] 2D-Coulomb Eg={mg:.2e} phi0={mp:.2e} | 2D-log Eg={mlg:.2e} phi0={mlp:.2e} | "
          f"q2D={mq:.2e} | LLL={ml:.2e} | flatTF={mtf:.2e} | 3D-phi0={m3:.2e} | "
          f"kernel_order={kernel_orders} | PASS={passed and order_ok}")
    return dict(passed=passed and order_ok, mg=mg, mp=mp, mlg=mlg, mlp=mlp, mq=mq, ml=ml, mtf=mtf, m3=m3,
                kernel_orders=kernel_orders, rows=rows)


# -----------------------------------------------------------------------------
# tf_only: the Thomas-Fermi envelope alone
# -----------------------------------------------------------------------------
def mode_tf_only(cfg):
    """modes.py:264-291 (without its figures): the TF envelope of the
    configuration, appended to tf_only.csv; None when there is none."""
    rid = run_id(cfg) + "_tf"
    G = make_grid(cfg)
    tf = extract_tf(cfg, G)
    if tf is None:
        print("[tf_only] no TF reference for these parameters")
        return None
    print(f"[tf_only] kind={tf.get('kind')} R90={tf.get('R90'):.4f} mu={tf.get('mu'):.5f} "
          f"rho0={tf.get('rho0', float('nan'))} converged={tf.get('converged')}")
    records.write_csv(os.path.join(records.out_dir(), "tf_only.csv"),
                      [dict(rid=rid, kind=tf.get("kind"), R90=tf.get("R90"), R99=tf.get("R99"),
                            mu=tf.get("mu"), rho0=tf.get("rho0", float("nan")))])
    return tf


# -----------------------------------------------------------------------------
# convergence: refine N, box and padding at fixed physics
# -----------------------------------------------------------------------------
def mode_convergence(cfg):
    """modes.py:392-436: solves at each N of conv_Ngrids, at the boxes
    L conv_box_factors (N keeping dx, in multiples of 32, at least 96) and at
    each padding of conv_pads; the finest grid's state seeds the later
    variants of the same shape.  Rows carry the diagnostic bundle."""
    rows = []
    L0 = float(cfg["L"])
    dx0 = 2.0 * L0 / int(cfg["Ngrid"])
    variants = [("grid", L0, int(N), int(cfg.get("pad", 2))) for N in cfg.get("conv_Ngrids", [192, 256, 384])]
    for f in cfg.get("conv_box_factors", [1.0, 1.2]):
        L = L0 * f
        N = max(96, 32 * math.ceil((2 * L / dx0) / 32))
        variants.append(("box", L, N, int(cfg.get("pad", 2))))
    for pp in cfg.get("conv_pads", [2, 3]):
        variants.append(("pad", L0, int(cfg["Ngrid"]), int(pp)))
    print(f"CONVERGENCE  {cfg['dimension']} s={cfg['s']} Omega={cfg['Omega']} beta2={cfg['beta2']} "
          f"beta3={cfg['beta3']} G_C={cfg['G_C']}")
    seen = set()
    prev = None
    for axis, L, N, pad in variants:
        key = (round(L, 10), N, pad)
        if key in seen:
            continue
        seen.add(key)
        c = dict(cfg, L=float(L), Ngrid=int(N), pad=int(pad), tag=f"{cfg.get('tag', 'becgpp')}_{axis}_N{N}_p{pad}")
        G = make_grid(c)
        psi0 = prev if (prev is not None and tuple(prev.shape) == tuple(G["R2"].shape)) else None
        psi, G, obs = ground_state(c, G=G, psi0=psi0, verbose=False)
        d = diagnostics(psi, G, c)
        d.update(axis=axis, L=L, N=N, pad=pad, dx=G["dx"], resid_rel=obs["resid_rel"], iters=obs["iters"],
                 converged=obs["converged"], points_per_R90=d["R90"] / G["dx"])
        rows.append(d)
        if axis == "grid" and N == max(cfg.get("conv_Ngrids", [256])):
            prev = psi.detach()
        records.write_csv(os.path.join(records.out_dir(), "convergence.csv"), rows)
        print(f"    {axis} L={L:.4g} N={N} p={pad}: E={d['E']:.9g} R90={d['R90']:.6g} "
              f"vir={d['virial_rel']:.2e} ppR90={d['points_per_R90']:.1f} res={d['resid_rel']:.2e}")
    grid = sorted((r for r in rows if r["axis"] == "grid"), key=lambda r: r["N"])
    if len(grid) >= 2:
        base = grid[-1]                                 # finest grid (largest N)
        for r in grid:
            print(f"    N={r['N']}: dE_vs_finest={r['E'] - base['E']:+.3e} "
                  f"dR90_rel={(r['R90'] - base['R90']) / max(abs(base['R90']), 1e-30):+.3e}")
    return rows


# -----------------------------------------------------------------------------
# smoke
# -----------------------------------------------------------------------------
def mode_smoke(cfg):
    """modes.py:440-459: the gate at validate_N=256, one small run and the TF
    envelope of that run."""
    gate = mode_validate(dict(cfg, validate_N=256))
    small = dict(cfg, L=10.0, Ngrid=(128 if geometry(cfg)[0] == 2 else 72), maxit=1500, res_tol=3e-3,
                 nseeds=1, want_lll=True, tag="smoke")
    d = mode_single(small)
    mode_tf_only(small)
    run_ok = bool(d) and bool(d.get("converged", False))
    ok = bool(gate["passed"]) and run_ok
    print(f"[smoke] validation_passed={gate['passed']} run_converged={run_ok} OVERALL={ok}")
    return dict(gate=gate, run=d, ok=ok)


def zip_results(name="becgpp_results"):
    """modes.py:462-466: a zip archive of the output directory, beside it."""
    import shutil
    base = records.out_dir()
    parent = os.path.dirname(os.path.abspath(base))
    arch = shutil.make_archive(os.path.join(parent, name), "zip", root_dir=base)
    print(f"[zip] {arch} ({os.path.getsize(arch) / 1e6:.1f} MB)")
    return arch


DISPATCH = dict(validate=mode_validate, smoke=mode_smoke, single=mode_single,
                tf_only=mode_tf_only, sweep=mode_sweep, convergence=mode_convergence)


def run(cfg):
    """Dispatch a configuration to its mode and return the result (modes.py:473-500)."""
    mode = cfg.get("mode", "smoke")
    fn = DISPATCH.get(mode)
    if fn is None:
        raise ValueError(f"unknown mode {mode!r}; choose {sorted(DISPATCH)}")
    kkind = resolve_kernel(cfg)[0]
    requested = str(cfg.get("kernel", "auto")).lower()
    if requested in ("newton", "log") and kkind == "none" and abs(float(cfg.get("G_C", 0.0))) <= 1e-15:
        print(f"[warn] kernel={requested!r} is set but G_C=0, so the long-range term is OFF "
              f"(reported kernel='none'). Set G_C>0 to activate it, e.g. G_C=20.")
    print(f"becGPP-amd {CODE_VERSION} | device={DEV} | mode={mode} | dim={cfg.get('dimension')} kernel={kkind}")
    result = fn(cfg)
    if cfg.get("zip_output", False):            # modes.py:486-491
        try:
            zip_results()
        except Exception as exc:
            print(f"[zip-warn] {exc}")
    print(f"\nDone. Output in: {records.out_dir()}")
    return result
