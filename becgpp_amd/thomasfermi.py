"""Thomas-Fermi envelopes used by the tf and triangular seeds
(becgpp/thomasfermi.py:21-156, fields.py:17-30).

One-time set-up work on device tensors; the chemical potential's bisection
runs on the device (bec_tf_mu), and the nonlocal (G_C != 0) fixed point
evaluates the long-range potential with the library's convolution.
"""
import ctypes
import math

import torch

from . import _lib
from .grid import _ptr, as_field, dV, plan_of
from .interactions import long_range_phi

_TF_CACHE = {}


def center_of_mass(rho, G):
    """fields.py:17-22."""
    mass = rho.sum().item() * dV(G)
    if mass <= 0:
        return tuple(0.0 for _ in range(G["ndim"]))
    return tuple((c * rho).sum().item() * dV(G) / mass for c in G["coords"])


def mass_quantiles(rho, G, fractions=(0.50, 0.90, 0.99), center=None):
    """Radii enclosing the given mass fractions (fields.py:25-38)."""
    r = rho.detach().real
    w = (r.clamp(min=0.0) * dV(G)).reshape(-1)
    if center is None:
        center = center_of_mass(r, G)
    rad2 = sum((c - center[i]) ** 2 for i, c in enumerate(G["coords"]))
    radii = torch.sqrt(rad2).reshape(-1)
    order = torch.argsort(radii, stable=True)
    cum = torch.cumsum(w[order], dim=0)
    out = []
    for f in fractions:
        target = torch.tensor([float(f)], dtype=cum.dtype, device=cum.device) * cum[-1]
        jj = int(torch.searchsorted(cum, target).clamp(max=cum.numel() - 1).item())
        out.append(radii[order[jj]].item())
    return out


def mass_radius(rho, G, fraction=0.99, center=None):
    """fields.py:33-34: the radius enclosing the given mass fraction."""
    return mass_quantiles(rho, G, (fraction,), center=center)[0]


def tf_effective_potential(G, cfg):
    """V - Omega^2 (x^2 + y^2)/2, centrifugal term in-plane only (thomasfermi.py:21-30)."""
    O = float(cfg.get("Omega", 1.0))
    return G["V"] - 0.5 * O * O * (G["X"] ** 2 + G["Y"] ** 2)


def _invert_local(mu, Veff, b2, b3):
    """Upper root rho >= 0 of b3 rho^2 + b2 rho = mu - Veff (thomasfermi.py:33-45)."""
    rhs = mu - Veff
    if b3 > 1e-30:
        disc = b2 * b2 + 4.0 * b3 * rhs
        root = (-b2 + torch.sqrt(torch.clamp(disc, min=0.0))) / (2.0 * b3)
        return torch.where(disc > 0.0, torch.clamp(root, min=0.0), torch.zeros_like(rhs))
    if abs(b2) > 1e-30:
        return torch.clamp(rhs / b2, min=0.0)
    return torch.zeros_like(rhs)


def _bisect_mu_for_norm(base, b2, b3, G, lo, hi, steps=90):
    """Chemical potential giving unit norm by bisection (thomasfermi.py:48-62)
    for rho = _invert_local(mu, base, b2, b3); the steps run on the device."""
    pl = plan_of(G)
    t = as_field(base, G, torch.float64)
    mu = ctypes.c_double()
    _lib.check(pl.lib.bec_tf_mu(pl.h, _ptr(t), float(b2), float(b3), float(lo), float(hi), int(steps),
                                ctypes.byref(mu)))
    return mu.value


def extract_tf(cfg, G, iters=400, relax=0.4):
    """TF reference for the current parameters, or None (thomasfermi.py:65-156)."""
    b2, b3, gc = float(cfg["beta2"]), float(cfg["beta3"]), float(cfg["G_C"])
    dv = dV(G)
    Veff = tf_effective_potential(G, cfg)
    nonlocal_on = (G["kernel"] != "none" and abs(gc) > 1e-15)
    key = ("tf", G["kernel"], round(b2, 10), round(b3, 10), round(gc, 10), round(float(cfg["Omega"]), 10),
           int(G["N"]), round(float(G["L"]), 10), int(G["ndim"]), round(float(G["s"]), 10),
           round(float(G.get("trap_coeff", 0.5)), 10), round(float(G.get("l_z", 0.0)), 10), int(G["pad"]))
    if cfg.get("tf_cache", True) and key in _TF_CACHE:
        return _TF_CACHE[key]
    tf = None
    if nonlocal_on and b2 <= 0 and b3 <= 0:
        tf = None
    elif nonlocal_on:
        sigma = 0.15 * G["L"]
        n = torch.exp(-G["R2"] / (2 * sigma ** 2))
        n = n / (n.sum() * dv)
        conv = False
        dn = float("inf")
        mu = float("nan")
        for _ in range(int(iters)):
            phi = long_range_phi(n, G, gc)
            base = Veff + phi
            bmin, bmax = torch.stack(torch.aminmax(base)).tolist()      # one synchronisation for both
            mu = _bisect_mu_for_norm(base, b2, b3, G, bmin, bmax + abs(b2) + abs(b3) + 1.0)
            if not math.isfinite(mu):
                break
            n_new = _invert_local(mu, base, b2, b3)
            s_t = n_new.sum() * dv
            # torch divides by a host scalar as a product with its reciprocal:
            # the same product here keeps the iterates bit for bit
            n_new = n_new * (1.0 / s_t)
            # the norm and the change fetched together; a non-positive norm
            # leaves n as it was, as when it is tested before the division
            ssum, dn_t = torch.stack([s_t, (n_new - n).abs().sum() * dv]).tolist()
            if ssum <= 0:
                break
            dn = dn_t
            n = (1 - relax) * n + relax * n_new
            n = n / (n.sum() * dv)
            if dn < 1e-9:
                conv = True
                break
        if math.isfinite(mu):
            R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
            tf = dict(kind="nonlocal_compressed", n=n, R90=R90, R99=R99, mu=mu,
                      rho0=float("nan"), converged=conv, final_dn=dn)
    else:
        confined = bool(torch.max(Veff).item() > 1e-12)
        if b2 < 0 and b3 > 0:
            rho0 = -3.0 * b2 / (4.0 * b3)
            mu = b2 * rho0 + b3 * rho0 ** 2
            R = 1.0 / math.sqrt(math.pi * rho0) if G["ndim"] == 2 else (3.0 / (4.0 * math.pi * rho0)) ** (1.0 / 3.0)
            rr = torch.sqrt(G["R2"])
            n = torch.where(rr <= R, torch.full_like(rr, rho0), torch.zeros_like(rr))
            n = n / (n.sum() * dv)
            R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
            tf = dict(kind="local_flattop", n=n, R90=R90, R99=R99, mu=mu, rho0=rho0, R=R,
                      converged=True, final_dn=0.0)
        elif confined and (b2 > 0 or (b2 == 0 and b3 > 0)):
            mu = _bisect_mu_for_norm(Veff, b2, b3, G, Veff.min().item(), Veff.max().item() + abs(b2) + abs(b3) + 1.0)
            n = _invert_local(mu, Veff, b2, b3)
            n = n / (n.sum() * dv)
            R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
            kind = "repulsive_parabola" if b3 <= 0 else "local_trapped_repulsive"
            tf = dict(kind=kind, n=n, R90=R90, R99=R99, mu=mu, rho0=float("nan"),
                      converged=True, final_dn=0.0)
    if tf is not None and cfg.get("tf_cache", True):
        _TF_CACHE[key] = tf
    return tf
