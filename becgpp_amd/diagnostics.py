"""Diagnostic bundle of a solved state (becgpp/diagnostics.py:10-47): the
observables, the mass radii R50/R90/R99, the 1% droplet edge, the vortex
counts, the lowest-Landau-level weight and criterion, the Thomas-Fermi
envelope's radii and chemical potential, and resolution figures."""
import math

import numpy as np
import torch

from .grid import as_field
from .operators import observables
from .thomasfermi import center_of_mass, extract_tf, mass_quantiles
from .vortex import lll_weight, vortex_diagnostic


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


def diagnostics(psi, G, cfg):
    """diagnostics.py:18-47."""
    t = as_field(psi, G)
    d = observables(t, G, cfg)
    rho = t.real ** 2 + t.imag ** 2
    R50, R90, R99 = mass_quantiles(rho, G, (0.50, 0.90, 0.99))
    d.update(R50=R50, R90=R90, R99=R99, R_1pct=droplet_radius(rho, G))
    if cfg.get("want_vortices", True):
        d.update(vortex_diagnostic(t, G, cfg.get("Omega", 0.0)))
    else:
        d.update(Nv=0, Nplus=0, Nminus=0, Nnet=0)
    lll_ok = bool(cfg.get("want_lll", True)) and lll_applicable(G, cfg)
    O = abs(float(cfg.get("Omega", 0.0)))
    d["w_LLL"] = lll_weight(t, G, 60, float(cfg.get("Omega", 0.0))) if lll_ok else float("nan")
    # the interaction energy beta2 n_peak against the gap (1 + Omega) to the
    # next Landau level in the rotating frame: << 1 in the LLL regime
    d["lll_param"] = (abs(float(cfg["beta2"])) * d["peak"] / (1.0 + O)) if lll_ok else float("nan")
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
    return d
