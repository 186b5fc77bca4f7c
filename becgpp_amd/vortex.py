"""Vortex count and lowest-Landau-level weight of a state (becgpp/vortex.py:10-84),
computed by libbecgpp_hip (bec_vortex_count, bec_lll_weight); 3D states are
counted on their z = 0 midplane."""
import ctypes
import math


from . import _lib
from .config import default_cfg
from .grid import _ptr, as_field, make_grid, plan_of
from .thomasfermi import center_of_mass, mass_quantiles


def _midplane(psi, G):
    """The 2D field and grid of a state: itself in 2D, the z = N/2 plane in 3D."""
    if int(G["ndim"]) == 2:
        return as_field(psi, G), G
    t = as_field(psi, G)[:, :, int(G["N"]) // 2].contiguous()
    G2 = make_grid(default_cfg(dimension="2D", Ngrid=int(G["N"]), L=float(G["L"]), pad=int(G["pad"]),
                               s=float(G["s"]), trap_coeff=float(G.get("trap_coeff", 0.5)), kernel="none", G_C=0.0))
    return t, G2


def vortex_diagnostic(psi, G, Omega=1.0, dens_thresh_frac=0.15):
    """Counts of +-1 phase windings inside the droplet (vortex.py:10-59):
    dict(Nv, Nplus, Nminus, Nnet)."""
    t, G2 = _midplane(psi, G)
    zero = dict(Nv=0, Nplus=0, Nminus=0, Nnet=0)
    rho = t.real ** 2 + t.imag ** 2
    dx = float(G2["dx"])
    if float(rho.sum()) * dx * dx <= 0:
        return zero
    xcm, ycm = center_of_mass(rho, G2)
    R = mass_quantiles(rho, G2, (0.99,), center=(xcm, ycm))[0]
    if not math.isfinite(R) or R <= 0:
        return zero
    ell = min(2.0 * math.sqrt(math.pi / max(abs(float(Omega)), 1e-9)), max(R / 3.0, 2.0 * dx))
    pl = plan_of(G2)
    out = (ctypes.c_int * 4)()
    _lib.check(pl.lib.bec_vortex_count(pl.h, _ptr(t), float(xcm), float(ycm), float(R), float(ell),
                                       float(dens_thresh_frac), out))
    return dict(Nv=out[0], Nplus=out[1], Nminus=out[2], Nnet=out[3])


def lll_weight(psi, G, Mmax=60, Omega=1.0):
    """Weight of psi in the lowest Landau level z^m exp(-r^2/2) (conj(z)^m for
    Omega < 0) of the rotating harmonic trap (vortex.py:61-84); NaN outside 2D."""
    if int(G["ndim"]) != 2:
        return float("nan")
    pl = plan_of(G)
    t = as_field(psi, G)
    w = ctypes.c_double()
    _lib.check(pl.lib.bec_lll_weight(pl.h, _ptr(t), float(Omega), int(Mmax), ctypes.byref(w), None))
    return w.value


def vortex_diagnostic_batch(psis, G, Omegas, dens_thresh_frac=0.15):
    """vortex_diagnostic of several states of one grid (up to 64 per call),
    every launch covering all of them (bec_vortex_count_batch)."""
    import torch
    from .thomasfermi import center_of_mass_batch, mass_quantiles_batch
    out = []
    for lo in range(0, len(psis), 64):
        chunk = [_midplane(p, G) for p in psis[lo:lo + 64]]
        G2 = chunk[0][1]
        t = torch.stack([c[0] for c in chunk]).contiguous()
        nb = t.shape[0]
        rho = (t.real ** 2 + t.imag ** 2).reshape(nb, -1)
        dx = float(G2["dx"])
        mass = (rho.sum(dim=1) * dx * dx).tolist()
        com = center_of_mass_batch(rho, G2)
        R = mass_quantiles_batch(rho, G2, (0.99,), com)[:, 0].tolist()
        com = com.tolist()
        prm, live = [], []
        for b in range(nb):
            if mass[b] <= 0 or not math.isfinite(R[b]) or R[b] <= 0:
                continue
            om = float(Omegas[lo + b])
            ell = min(2.0 * math.sqrt(math.pi / max(abs(om), 1e-9)), max(R[b] / 3.0, 2.0 * dx))
            prm += [com[b][0], com[b][1], R[b], ell]
            live.append(b)
        res = [dict(Nv=0, Nplus=0, Nminus=0, Nnet=0) for _ in range(nb)]
        if live:
            tl = t[live].contiguous()
            arr = (ctypes.c_double * len(prm))(*prm)
            o = (ctypes.c_int * (4 * len(live)))()
            pl = plan_of(G2)
            _lib.check(pl.lib.bec_vortex_count_batch(pl.h, len(live), _ptr(tl), arr, float(dens_thresh_frac), o))
            for k, b in enumerate(live):
                res[b] = dict(Nv=o[4 * k], Nplus=o[4 * k + 1], Nminus=o[4 * k + 2], Nnet=o[4 * k + 3])
        out += res
    return out


def lll_weight_batch(psis, G, Mmax, Omegas):
    """lll_weight of several 2D states of one grid (bec_lll_weight_batch)."""
    import torch
    if int(G["ndim"]) != 2:
        return [float("nan")] * len(psis)
    out = []
    pl = plan_of(G)
    for lo in range(0, len(psis), 64):
        t = torch.stack([as_field(p, G) for p in psis[lo:lo + 64]]).contiguous()
        nb = t.shape[0]
        om = (ctypes.c_double * nb)(*[float(x) for x in Omegas[lo:lo + nb]])
        w = (ctypes.c_double * nb)()
        _lib.check(pl.lib.bec_lll_weight_batch(pl.h, nb, _ptr(t), om, int(Mmax), w))
        out += list(w)
    return out
