"""Diagnostic bundle of a solved state (becgpp/diagnostics.py:10-47): the
observables, the mass radii R50/R90/R99, the 1% droplet edge, the vortex
counts, the lowest-Landau-level weight and criterion, the Thomas-Fermi
envelope's radii and chemical potential, and resolution figures."""
import math

import numpy as np
import torch

from .grid import as_field
from .operators import observables, observables_batch
from .thomasfermi import center_of_mass, center_of_mass_batch, extract_tf, mass_quantiles, mass_quantiles_batch
from .vortex import lll_weight, lll_weight_batch, vortex_diagnostic, vortex_diagnostic_batch


def lll_applicable(G, cfg):
    """diagnostics.py:10-15: the basis z^m exp(-r^2/2) belongs to the isotropic
    harmonic trap V = r^2/2 in 2D, and the weight is reported for rotating runs."""
    return (int(G["ndim"]) == 2 and abs(float(G["s"]) - 2.0) < 1e-12
            and abs(float(G.get("trap_coeff", 0.5)) - 0.5) < 1e-12
            and abs(float(cfg.get("Omega", 0.0))) > 0.0)


def radial_average(f, G, center=None, rmax=None, bins=None):
    """fields.py:40-56: bin centres and the mean of f over the grid points whose
    distance from center falls in each of `bins` equal bins on [0, rmax]
    (points beyond rmax go to the last bin; empty bins are NaN)."""
    if center is None:
        center = tuple(0.0 for _ in G["coords"])
    r = torch.sqrt(sum((c - center[i]) ** 2 for i, c in enumerate(G["coords"]))).reshape(-1)
    if rmax is None:
        rmax = float(G["L"])
    if bins is None:
        bins = int(min(400, max(20, rmax / (2.0 * float(G["dx"])))))
    edges = np.linspace(0.0, rmax, bins + 1)
    ctr = 0.5 * (edges[1:] + edges[:-1])
    e = torch.as_tensor(edges, dtype=r.dtype, device=r.device)
    idx = (torch.bucketize(r, e, right=True) - 1).clamp(0, bins - 1)
    ff = f.detach().real.reshape(-1).to(r.dtype)
    s = torch.bincount(idx, weights=ff, minlength=bins).to(r.dtype)
    c = torch.bincount(idx, minlength=bins).to(r.dtype)
    prof = torch.where(c > 0, s / c.clamp(min=1.0), torch.full_like(s, float("nan")))
    return ctr, prof.cpu().numpy()


def droplet_radius(rho, G, frac=1e-2):
    """fields.py:59-67: the outermost bin centre of the radial profile about the
    centre of mass still above frac of the profile's peak."""
    ctr, prof = radial_average(rho, G, center=center_of_mass(rho, G))
    pk = np.nanmax(prof)
    if not np.isfinite(pk) or pk <= 0:
        return float("nan")
    valid = prof > frac * pk
    return float(ctr[valid][-1]) if np.any(valid) else float("nan")


def _droplet_radii_batch(rho, G, centers, frac=1e-2):
    """droplet_radius of the rows of rho [nb, n] about their centres: one
    histogram over all rows (the bins of radial_average on [0, L])."""
    nb = rho.shape[0]
    rmax = float(G["L"])
    bins = int(min(400, max(20, rmax / (2.0 * float(G["dx"])))))
    edges = np.linspace(0.0, rmax, bins + 1)
    ctr = 0.5 * (edges[1:] + edges[:-1])
    dist = torch.sqrt(sum((c.reshape(1, -1) - centers[:, i:i + 1]) ** 2 for i, c in enumerate(G["coords"])))
    e = torch.as_tensor(edges, dtype=dist.dtype, device=dist.device)
    idx = (torch.bucketize(dist, e, right=True) - 1).clamp(0, bins - 1)
    idx = idx + bins * torch.arange(nb, device=idx.device).reshape(-1, 1)
    s = torch.bincount(idx.reshape(-1), weights=rho.reshape(-1), minlength=nb * bins).reshape(nb, bins)
    c = torch.bincount(idx.reshape(-1), minlength=nb * bins).reshape(nb, bins).to(s.dtype)
    prof = torch.where(c > 0, s / c.clamp(min=1.0), torch.full_like(s, float("nan"))).cpu().numpy()
    out = []
    for p in prof:
        pk = np.nanmax(p)
        if not np.isfinite(pk) or pk <= 0:
            out.append(float("nan"))
            continue
        valid = p > frac * pk
        out.append(float(ctr[valid][-1]) if np.any(valid) else float("nan"))
    return out


def _bundle(d, G, cfg, radii, r1, vort, wlll):
    """The rest of the bundle on top of the observables d (diagnostics.py:18-47)."""
    R50, R90, R99 = radii
    d.update(R50=R50, R90=R90, R99=R99, R_1pct=r1)
    d.update(vort if cfg.get("want_vortices", True) else dict(Nv=0, Nplus=0, Nminus=0, Nnet=0))
    lll_ok = bool(cfg.get("want_lll", True)) and lll_applicable(G, cfg)
    O = abs(float(cfg.get("Omega", 0.0)))
    d["w_LLL"] = wlll if lll_ok else float("nan")
    # the interaction energy beta2 n_peak against the gap (1 + Omega) to the
    # next Landau level in the rotating frame: << 1 in the LLL regime
    d["lll_param"] = (abs(float(cfg["beta2"])) * d["peak"] / (1.0 + O)) if lll_ok else float("nan")
    _tf_fields(d, G, cfg)
    return d


def diagnostics(psi, G, cfg):
    """diagnostics.py:18-47."""
    t = as_field(psi, G)
    d = observables(t, G, cfg)
    rho = t.real ** 2 + t.imag ** 2
    radii = mass_quantiles(rho, G, (0.50, 0.90, 0.99))
    vort = vortex_diagnostic(t, G, cfg.get("Omega", 0.0)) if cfg.get("want_vortices", True) else None
    lll_ok = bool(cfg.get("want_lll", True)) and lll_applicable(G, cfg)
    w = lll_weight(t, G, 60, float(cfg.get("Omega", 0.0))) if lll_ok else float("nan")
    return _bundle(d, G, cfg, radii, droplet_radius(rho, G), vort, w)


def diagnostics_batch(psis, G, cfgs):
    """diagnostics of several states of one grid, each kernel launch covering
    all of them: the observables (bec_observe_batch), the mass radii and the
    1% edge (one sort, one histogram over all states), the vortex counts
    (bec_vortex_count_batch) and the LLL weights (bec_lll_weight_batch).
    Returns one dict per state, as diagnostics(psis[i], G, cfgs[i])."""
    if not psis:
        return []
    t = torch.stack([as_field(p, G) for p in psis]).contiguous()
    nb = t.shape[0]
    obs = observables_batch(list(t), G, cfgs)
    rho = (t.real ** 2 + t.imag ** 2).reshape(nb, -1)
    com = center_of_mass_batch(rho, G)
    radii = mass_quantiles_batch(rho, G, (0.50, 0.90, 0.99), com).tolist()
    r1 = _droplet_radii_batch(rho, G, com)
    want_v = [k for k, c in enumerate(cfgs) if c.get("want_vortices", True)]
    vort = [None] * nb
    for k, v in zip(want_v, vortex_diagnostic_batch([t[k] for k in want_v], G,
                                                    [cfgs[k].get("Omega", 0.0) for k in want_v])):
        vort[k] = v
    want_l = [k for k, c in enumerate(cfgs) if c.get("want_lll", True) and lll_applicable(G, c)]
    wl = [float("nan")] * nb
    for k, w in zip(want_l, lll_weight_batch([t[k] for k in want_l], G, 60,
                                             [float(cfgs[k].get("Omega", 0.0)) for k in want_l])):
        wl[k] = w
    return [_bundle(obs[k], G, cfgs[k], radii[k], r1[k], vort[k], wl[k]) for k in range(nb)]


def _tf_fields(d, G, cfg):
    tf = extract_tf(cfg, G)
    if tf is not None:
        d.update(tf_kind=tf.get("kind", ""), tf_R90=tf.get("R90", float("nan")),
                 tf_R99=tf.get("R99", float("nan")), tf_mu=tf.get("mu", float("nan")),
                 tf_rho0=tf.get("rho0", float("nan")), tf_converged=bool(tf.get("converged", False)))
    else:
        d.update(tf_kind="", tf_R90=float("nan"), tf_R99=float("nan"), tf_mu=float("nan"),
                 tf_rho0=float("nan"), tf_converged=False)
    d["mem_estimate_gb"] = G.get("mem_estimate_gb", float("nan"))
    d["points_per_R90"] = d["R90"] / float(G["dx"]) if math.isfinite(d["R90"]) else float("nan")
