"""Thomas-Fermi reference densities for the current parameters
(the reference's thomasfermi.py:21-156), used by the tf and triangular seeds
and by the diagnostics.

Which reference exists depends on the active terms:

* long range on (kernel and G_C != 0) with some local repulsion: the
  compressed density of the nonlocal fixed point, run on the device by
  ``bec_tf_nonlocal`` (the convolution, the chemical potential's search and
  the relaxed update of each iteration stay there);
* no long range, attractive contact and repulsive three-body (b2 < 0 < b3):
  the zero-pressure flat top rho0 = -3 b2 / (4 b3) on a disc (ball) of unit
  mass;
* no long range, a confining V_eff and local repulsion: the inverted profile
  rho(mu - V_eff), with mu of unit norm from ``bec_tf_mu``;
* otherwise none (None).

V_eff = V - Omega^2 (x^2 + y^2) / 2 (the centrifugal term in-plane only).
Results are cached per grid and parameters unless cfg["tf_cache"] is False.
"""
import ctypes
import math

import torch

from . import _lib
from .grid import _ptr, as_field, dV, plan_of
from .interactions import ensure_kernel

_CACHE = {}


def center_of_mass(rho, G):
    """fields.py:17-22."""
    mass = rho.sum().item() * dV(G)
    if mass <= 0:
        return tuple(0.0 for _ in range(G["ndim"]))
    return tuple((c * rho).sum().item() * dV(G) / mass for c in G["coords"])


def mass_quantiles(rho, G, fractions=(0.50, 0.90, 0.99), center=None):
    """Radii enclosing the given mass fractions about the centre of mass
    (fields.py:25-38): the distance of the first point, in order of distance,
    at which the running mass reaches the fraction of the total."""
    r = rho.detach().real
    if center is None:
        center = center_of_mass(r, G)
    dist = torch.sqrt(sum((c - center[i]) ** 2 for i, c in enumerate(G["coords"]))).reshape(-1)
    order = torch.argsort(dist, stable=True)
    running = torch.cumsum((r.clamp(min=0.0) * dV(G)).reshape(-1)[order], dim=0)
    targets = torch.tensor([float(f) for f in fractions], dtype=running.dtype, device=running.device) * running[-1]
    at = torch.searchsorted(running, targets).clamp(max=running.numel() - 1)
    return dist[order[at]].tolist()


def mass_radius(rho, G, fraction=0.99, center=None):
    """fields.py:33-34: the radius enclosing the given mass fraction."""
    return mass_quantiles(rho, G, (fraction,), center=center)[0]


def tf_effective_potential(G, cfg):
    """V - Omega^2 (x^2 + y^2) / 2 (thomasfermi.py:21-30)."""
    om = float(cfg.get("Omega", 1.0))
    return G["V"] - 0.5 * om * om * (G["X"] ** 2 + G["Y"] ** 2)


def tf_density(mu, base, b2, b3):
    """rho(mu - base): the upper root rho >= 0 of b3 rho^2 + b2 rho = mu - base,
    0 where there is none (thomasfermi.py:33-45)."""
    rhs = mu - base
    if b3 > 1e-30:
        disc = b2 * b2 + 4.0 * b3 * rhs
        root = ((-b2 + torch.sqrt(disc.clamp(min=0.0))) / (2.0 * b3)).clamp(min=0.0)
        return torch.where(disc > 0.0, root, torch.zeros_like(rhs))
    if abs(b2) > 1e-30:
        return (rhs / b2).clamp(min=0.0)
    return torch.zeros_like(rhs)


def chemical_potential(base, b2, b3, G, lo, hi, steps=90):
    """mu with sum rho(mu - base) dV = 1, searched on the device from [lo, hi]
    (the upper end widened first when the norm at hi is below 1; NaN when 200
    widenings do not reach it): thomasfermi.py:48-62."""
    pl = plan_of(G)
    t = as_field(base, G, torch.float64)
    mu = ctypes.c_double()
    _lib.check(pl.lib.bec_tf_mu(pl.h, _ptr(t), float(b2), float(b3), float(lo), float(hi), int(steps),
                                ctypes.byref(mu)))
    return mu.value


# kept for callers of the reference's private name
_bisect_mu_for_norm = chemical_potential


def _with_radii(n, G, **fields):
    R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
    return dict(n=n, R90=R90, R99=R99, **fields)


def _compressed(cfg, G, veff, iters, relax):
    """The nonlocal fixed point (thomasfermi.py:87-119) on the device, from
    the normalised Gaussian of width 0.15 L."""
    b2, b3, gc = float(cfg["beta2"]), float(cfg["beta3"]), float(cfg["G_C"])
    width = 0.15 * float(G["L"])
    n = torch.exp(-G["R2"] / (2 * width ** 2))
    n = (n / (n.sum() * dV(G))).contiguous()
    pl = ensure_kernel(G, "K")
    v = as_field(veff, G, torch.float64)
    out = (ctypes.c_double * 4)()
    _lib.check(pl.lib.bec_tf_nonlocal(pl.h, _ptr(v), b2, b3, gc, int(iters), float(relax), 1e-9, _ptr(n), out))
    mu, dn, _its, conv = out[0], out[1], int(out[2]), bool(out[3])
    if not math.isfinite(mu):
        return None
    return _with_radii(n, G, kind="nonlocal_compressed", mu=mu, rho0=float("nan"), converged=conv, final_dn=dn)


def _flat_top(cfg, G):
    """Zero-pressure droplet: rho0 = -3 b2 / (4 b3), mu = b2 rho0 + b3 rho0^2,
    on the disc (2D) or ball (3D) of unit mass (thomasfermi.py:123-138)."""
    b2, b3 = float(cfg["beta2"]), float(cfg["beta3"])
    rho0 = -3.0 * b2 / (4.0 * b3)
    radius = (1.0 / math.sqrt(math.pi * rho0)) if G["ndim"] == 2 else (3.0 / (4.0 * math.pi * rho0)) ** (1.0 / 3.0)
    inside = torch.sqrt(G["R2"]) <= radius
    n = torch.where(inside, torch.full_like(G["R2"], rho0), torch.zeros_like(G["R2"]))
    n = n / (n.sum() * dV(G))
    return _with_radii(n, G, kind="local_flattop", mu=b2 * rho0 + b3 * rho0 ** 2, rho0=rho0, R=radius,
                       converged=True, final_dn=0.0)


def _inverted(cfg, G, veff):
    """Repulsive trapped gas: rho(mu - V_eff) of unit norm (thomasfermi.py:139-150)."""
    b2, b3 = float(cfg["beta2"]), float(cfg["beta3"])
    lo, hi = torch.stack(torch.aminmax(veff)).tolist()
    mu = chemical_potential(veff, b2, b3, G, lo, hi + abs(b2) + abs(b3) + 1.0)
    n = tf_density(mu, veff, b2, b3)
    n = n / (n.sum() * dV(G))
    kind = "repulsive_parabola" if b3 <= 0 else "local_trapped_repulsive"
    return _with_radii(n, G, kind=kind, mu=mu, rho0=float("nan"), converged=True, final_dn=0.0)


def _regime(cfg, G, veff):
    b2, b3, gc = float(cfg["beta2"]), float(cfg["beta3"]), float(cfg["G_C"])
    if G["kernel"] != "none" and abs(gc) > 1e-15:
        # attraction with no local pressure has no compressed-TF regime
        return "compressed" if (b2 > 0 or b3 > 0) else None
    if b2 < 0 and b3 > 0:
        return "flat_top"
    if torch.max(veff).item() > 1e-12 and (b2 > 0 or (b2 == 0 and b3 > 0)):
        return "inverted"
    return None


def extract_tf(cfg, G, iters=400, relax=0.4):
    """The TF reference of the current parameters (a dict with the density n,
    R90, R99, mu, kind, rho0, converged, final_dn), or None where none exists
    (thomasfermi.py:65-156)."""
    grid = (G["ndim"], G["N"], float(G["L"]), G["pad"], float(G["s"]), float(G.get("trap_coeff", 0.5)),
            G["kernel"], float(G.get("l_z", 0.0)))
    phys = tuple(float(cfg[k]) for k in ("beta2", "beta3", "G_C", "Omega"))
    key = (grid, phys, int(iters), float(relax))
    use_cache = cfg.get("tf_cache", True)
    if use_cache and key in _CACHE:
        return _CACHE[key]
    veff = tf_effective_potential(G, cfg)
    regime = _regime(cfg, G, veff)
    if regime == "compressed":
        tf = _compressed(cfg, G, veff, iters, relax)
    elif regime == "flat_top":
        tf = _flat_top(cfg, G)
    elif regime == "inverted":
        tf = _inverted(cfg, G, veff)
    else:
        tf = None
    if tf is not None and use_cache:
        _CACHE[key] = tf
    return tf


def center_of_mass_batch(rho, G):
    """Centres of mass of nb densities rho [nb, n] (flattened in the grid's
    order): [nb, ndim], zeros for a state without mass (fields.py:17-22)."""
    mass = rho.sum(dim=1) * dV(G)
    cs = [(rho * c.reshape(1, -1)).sum(dim=1) * dV(G) for c in G["coords"]]
    safe = torch.where(mass > 0, mass, torch.ones_like(mass))
    return torch.stack([torch.where(mass > 0, c / safe, torch.zeros_like(c)) for c in cs], dim=1)


def mass_quantiles_batch(rho, G, fractions, centers):
    """mass_quantiles of nb densities at once (rows of rho [nb, n], centres
    [nb, ndim]): [nb, len(fractions)] radii, one sort of all rows."""
    dist = torch.sqrt(sum((c.reshape(1, -1) - centers[:, i:i + 1]) ** 2 for i, c in enumerate(G["coords"])))
    order = torch.argsort(dist, dim=1, stable=True)
    running = torch.cumsum(torch.gather(rho.clamp(min=0.0) * dV(G), 1, order), dim=1)
    targets = running[:, -1:] * torch.tensor([float(f) for f in fractions], dtype=running.dtype,
                                             device=running.device).reshape(1, -1)
    at = torch.searchsorted(running.contiguous(), targets.contiguous()).clamp(max=running.shape[1] - 1)
    return torch.gather(dist, 1, torch.gather(order, 1, at))
