"""Long-range interaction: zero-padded free-space convolution with a
cell-averaged kernel (becgpp/interactions.py:1-176).

The convolution runs in libbecgpp_hip as five (3D) or three (2D) LDS line
passes that pad on load and crop on store, so the (pad N)^d grid is never
materialised; the kernel spectrum is kept as its real-even half
[(pad N/2 + 1)^d].  Newton and log kernels are tabulated on the GPU; the
quasi-2D kernel K_eff and its virial kernel W are tabulated here with
``torch.special`` (the same library routine the reference calls) and handed to
the library as a real-space table.
"""
import math

import numpy as np
import torch

from . import _lib
from .constants import DEV
from .grid import _PLANS, _ptr, as_field, dV, plan_of

_GL_NODES = (-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526)
_GL_WTS = (0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538)


def _k0e(x):
    """exp(x) K_0(x) (interactions.py:46-52)."""
    return torch.special.scaled_modified_bessel_k0(x)


def _k1e(x):
    """exp(x) K_1(x) (interactions.py:55-61)."""
    return torch.special.scaled_modified_bessel_k1(x)


def q2d_kernel(u, l_z):
    """K_eff(u) = k0e(u^2 / 4 l_z^2) / (sqrt(2 pi) l_z) (interactions.py:64-67)."""
    x = u * u / (4.0 * l_z * l_z)
    return _k0e(x) / (math.sqrt(2.0 * math.pi) * l_z)


def q2d_virial_kernel(u, l_z):
    """W(u) = u dK_eff/du (interactions.py:70-74)."""
    x = u * u / (4.0 * l_z * l_z)
    xs = torch.clamp(x, min=1e-300)
    return 2.0 * xs * (_k0e(xs) - _k1e(xs)) / (math.sqrt(2.0 * math.pi) * l_z)


def _polar_self_cell_2d(f, dx, n=48):
    """Average of a radial kernel over the square self cell by polar
    Gauss-Legendre quadrature over its 8 triangles, r = R t^2
    (interactions.py:77-94)."""
    t, wt = np.polynomial.legendre.leggauss(n)
    th = 0.5 * (math.pi / 4.0) * (t + 1.0)
    wth = 0.5 * (math.pi / 4.0) * wt
    tt = 0.5 * (t + 1.0)
    wtt = torch.as_tensor(0.5 * wt, dtype=torch.float64)
    total = 0.0
    for thj, wj in zip(th, wth):
        R = 0.5 * dx / math.cos(thj)
        r = torch.as_tensor(R * tt ** 2, dtype=torch.float64)
        jac = torch.as_tensor(2.0 * R * R * tt ** 3, dtype=torch.float64)
        total += wj * float((f(r) * jac * wtt).sum())
    return 8.0 * total / (dx * dx)


def _q2d_table(G, which):
    """Real-space cell-averaged q2d kernel (or its virial kernel) on the
    wrap-around displacement grid, built on the host (interactions.py:117-135)."""
    N, dx, pad, lz = int(G["N"]), float(G["dx"]), int(G["pad"]), float(G["l_z"])
    M = pad * N
    j = torch.arange(M, dtype=torch.float64)
    dj = torch.where(j <= M // 2, j, j - M) * dx
    DX, DY = torch.meshgrid(dj, dj, indexing="ij")
    fn = q2d_virial_kernel if which == "W" else q2d_kernel
    ker = torch.zeros_like(DX)
    for u, wu in zip(_GL_NODES, _GL_WTS):
        for v, wv in zip(_GL_NODES, _GL_WTS):
            ker += 0.25 * wu * wv * fn(torch.sqrt((DX + 0.5 * dx * u) ** 2 + (DY + 0.5 * dx * v) ** 2), lz)
    ker[0, 0] = _polar_self_cell_2d(lambda r: fn(r, lz), dx)
    return ker


def ensure_kernel(G, which="K"):
    """Load the spectrum of G's kernel (or the q2d virial kernel W) into the
    plan, reusing it while the geometry and kernel are unchanged."""
    kind = G["kernel"]
    if kind == "none" and which == "K":
        raise ValueError("no long-range kernel on this grid")
    pl = plan_of(G)
    w = _lib.WHICH_W if which == "W" else _lib.WHICH_K
    key = (which, kind, round(float(G.get("l_z", 0.0)), 14))
    if pl.kernels.get(w) == key:
        return pl
    if which == "W" or kind == "q2d":
        if G["ndim"] != 2:
            raise ValueError("the quasi-2D kernel is two-dimensional only")
        tab = _q2d_table(G, which).to(DEV).contiguous()
        _lib.check(pl.lib.bec_plan_set_kernel_table(pl.h, w, _ptr(tab)))
        torch.cuda.synchronize()
    elif kind == "newton":
        _lib.check(pl.lib.bec_plan_build_kernel(pl.h, w, _lib.KERNEL_NEWTON))
    elif kind == "log":
        if G["ndim"] != 2:
            raise ValueError("kernel='log' is two-dimensional only")
        _lib.check(pl.lib.bec_plan_build_kernel(pl.h, w, _lib.KERNEL_LOG))
    else:
        raise ValueError(f"unknown kernel {kind!r}")
    pl.kernels[w] = key
    return pl


def _conv(rho, G, which, scale):
    pl = ensure_kernel(G, which)
    r = as_field(rho, G, torch.float64)
    out = torch.empty_like(r)
    _lib.check(pl.lib.bec_conv(pl.h, _lib.WHICH_W if which == "W" else _lib.WHICH_K, _ptr(r),
                               float(scale), _ptr(out)))
    return out


def kernel_convolution(rho, G, which="K"):
    """(K * rho)(r_i) = sum_j K_ij rho_j dV, free-space (interactions.py:154-164)."""
    return _conv(rho, G, which, 1.0)


def long_range_phi(rho, G, gc):
    """phi = -gc (K * rho); zeros when the kernel is off (interactions.py:167-171)."""
    if G["kernel"] == "none" or abs(gc) <= 1e-15:
        return torch.zeros_like(as_field(rho, G, torch.float64))
    return _conv(rho, G, "K", -float(gc))


def clear_kernel_cache():
    """Free the cached kernel spectra (interactions.py:174-176)."""
    for p in _PLANS.values():
        p.clear_kernels()


def dv_of(G):
    return dV(G)
