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
