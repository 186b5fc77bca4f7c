"""Energy functional, mean-field Hamiltonian, residual and observables
(becgpp/operators.py:10-103), computed by libbecgpp_hip.

The kinetic and rotation terms are applied axis by axis: along an axis with
wavenumber k, T - Omega L_z is the multiplier k^2/2 + Omega c k (c = -x on the
y axis, +y on the x axis), so H psi costs one forward and one inverse line
transform per axis, fused with the local terms; the energy needs forward line
transforms only (Parseval per line).  Results equal the reference's full
n-dimensional FFT formulation up to round-off.
"""
import ctypes
import math

import torch

from . import _lib
from .constants import DEV
from .grid import _ptr, as_field, dV, plan_of
from .interactions import ensure_kernel, kernel_convolution

_COMP_KEYS = ("E", "Ekin", "Etrap", "Econtact", "Ethree", "Erot", "Egrav", "Lz")


def _lr_on(G, gc):
    return G["kernel"] != "none" and abs(gc) > 1e-15


def _plan(G, gc):
    return ensure_kernel(G, "K") if _lr_on(G, gc) else plan_of(G)


def phys_for(G, b2, b3, O, gc):
    """Physics record for the C-ABI.  The long-range term is on only when the
    grid's kernel is not "none" and |G_C| > 1e-15 (operators.py:21-22,
    solvers.py:36): otherwise G_C goes down as 0, so a spectrum that a plan of
    the same geometry still holds from another grid is never applied."""
    return _lib.phys(b2, b3, O, gc if _lr_on(G, gc) else 0.0)


def _energy_host(t, G, beta2, beta3, Omega, G_C, want_phi):
    """The energy components as host floats (one synchronisation), and phi
    when asked for."""
    pl = _plan(G, G_C)
    phi = torch.empty(t.shape, dtype=torch.float64, device=DEV) if want_phi else None
    comps = (ctypes.c_double * 8)()
    ph = phys_for(G, beta2, beta3, Omega, G_C)
    _lib.check(pl.lib.bec_energy(pl.h, _ptr(t), ctypes.byref(ph), comps, _ptr(phi) if want_phi else None))
    return {k: comps[i] for i, k in enumerate(_COMP_KEYS)}, phi


def energy_components(p, G, beta2, beta3, Omega, G_C):
    """operators.py:10-29: E and its parts as 0-d tensors, plus phi."""
    t = as_field(p, G)
    comps, phi = _energy_host(t, G, beta2, beta3, Omega, G_C, True)
    out = {k: torch.tensor(v, dtype=torch.float64, device=DEV) for k, v in comps.items()}
    out["phi"] = phi
    return out


class _Energy(torch.autograd.Function):
    """E[psi] with the adjoint as its backward.  For real E of complex psi,
    PyTorch's gradient is 2 dE/d(conj psi), and dE/d(conj psi) = H psi dV with
    the mean-field H of apply_H (the long-range term's derivative is phi psi,
    since the kernel is symmetric), so the backward is 2 H psi dV.

    The physical parameters may be 0-d tensors that require grad; their
    partial derivatives at fixed psi are dE/dbeta2 = 1/2 sum rho^2 dV,
    dE/dbeta3 = 1/3 sum rho^3 dV, dE/dOmega = -L_z and dE/dG_C = Egrav / G_C
    (the long-range energy per unit coupling; 0 without a kernel).  At a
    converged ground state these are the Hellmann-Feynman derivatives of the
    ground-state energy."""

    @staticmethod
    def forward(ctx, psi, G, beta2, beta3, Omega, G_C):
        par = [float(x.detach()) if torch.is_tensor(x) else float(x) for x in (beta2, beta3, Omega, G_C)]
        t = as_field(psi.detach(), G)
        comps = energy_components(t, G, *par)
        ctx.save_for_backward(t, comps["phi"])
        ctx.G = G
        ctx.cfg = dict(beta2=par[0], beta3=par[1], Omega=par[2], G_C=par[3])
        ctx.comps = {k: float(comps[k]) for k in ("Lz", "Egrav")} if any(ctx.needs_input_grad[2:]) else None
        return comps["E"]

    @staticmethod
    def backward(ctx, gE):
        t, phi = ctx.saved_tensors
        gpsi = None
        if ctx.needs_input_grad[0]:
            h = apply_H(t, ctx.G, ctx.cfg, phi=phi if _lr_on(ctx.G, ctx.cfg["G_C"]) else None)
            gpsi = 2.0 * dV(ctx.G) * gE * h
        grads = [None, None, None, None]
        if any(ctx.needs_input_grad[2:]):
            dv = dV(ctx.G)
            rho = t.real ** 2 + t.imag ** 2
            gc = ctx.cfg["G_C"]
            d = [0.5 * (rho * rho).sum() * dv, (rho * rho * rho).sum() * dv / 3.0,
                 -ctx.comps["Lz"], (ctx.comps["Egrav"] / gc) if (_lr_on(ctx.G, gc) and gc != 0.0) else 0.0]
            for i in range(4):
                if ctx.needs_input_grad[2 + i]:
                    grads[i] = gE * torch.as_tensor(d[i], dtype=torch.float64, device=gE.device)
        return (gpsi, None, *grads)


def energy_t(p, G, beta2, beta3, Omega, G_C):
    """operators.py:32-33: the energy as a 0-d tensor that autograd
    differentiates with respect to psi (backward 2 H psi dV) and to any of
    beta2, beta3, Omega, G_C given as tensors that require grad."""
    tracked = (torch.is_tensor(p) and p.requires_grad) or any(
        torch.is_tensor(x) and x.requires_grad for x in (beta2, beta3, Omega, G_C))
    if not tracked:
        return energy_components(p, G, *[float(x) for x in (beta2, beta3, Omega, G_C)])["E"]
    return _Energy.apply(p, G, beta2, beta3, Omega, G_C)


def sensitivities(psi, G, cfg):
    """Hellmann-Feynman derivatives of the ground-state energy at a converged
    state: dE/dbeta2 = Econtact / beta2 (= 1/2 sum rho^2 dV), dE/dbeta3 =
    Ethree / beta3, dE/dG_C = Egrav / G_C, dE/dOmega = -L_z."""
    t = as_field(psi, G)
    rho = (t.real ** 2 + t.imag ** 2)
    dv = dV(G)
    c, _ = _energy_host(t, G, cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"], False)
    out = dict(dE_dbeta2=0.5 * float((rho * rho).sum()) * dv,
               dE_dbeta3=float((rho * rho * rho).sum()) * dv / 3.0,
               dE_dOmega=-c["Lz"])
    gc = float(cfg["G_C"])
    if _lr_on(G, gc):
        out["dE_dG_C"] = c["Egrav"] / gc
    else:
        out["dE_dG_C"] = 0.0 if G["kernel"] == "none" else -0.5 * float(
            (rho * kernel_convolution(rho, G)).sum()) * dv
    return out


def apply_H(psi, G, cfg, phi=None):
    """operators.py:36-45: mean-field H psi; phi reuses a long-range potential."""
    b2, b3, O, gc = cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"]
    pl = _plan(G, gc)
    t = as_field(psi, G)
    out = torch.empty_like(t)
    ph = phys_for(G, b2, b3, O, gc)
    if phi is not None and not (isinstance(phi, (int, float)) and phi == 0):
        f = as_field(phi, G, torch.float64)
        _lib.check(pl.lib.bec_apply_H(pl.h, _ptr(t), ctypes.byref(ph), _ptr(f), _ptr(out)))
    else:
        _lib.check(pl.lib.bec_apply_H(pl.h, _ptr(t), ctypes.byref(ph), None, _ptr(out)))
    return out


def compute_residual(psi, G, cfg):
    """operators.py:48-52 -> (resid_abs, mu, resid_rel)."""
    gc = cfg["G_C"]
    pl = _plan(G, gc)
    t = as_field(psi, G)
    out = (ctypes.c_double * 3)()
    ph = phys_for(G, cfg["beta2"], cfg["beta3"], cfg["Omega"], gc)
    _lib.check(pl.lib.bec_residual(pl.h, _ptr(t), ctypes.byref(ph), out))
    return out[0], out[1], out[2]


def moments(psi, G):
    """sum rho, sum r^2 rho, sum (x^2+y^2) rho, sum z^2 rho, max rho."""
    pl = plan_of(G)
    t = as_field(psi, G)
    out = (ctypes.c_double * 5)()
    _lib.check(pl.lib.bec_moments(pl.h, _ptr(t), out))
    return tuple(out)


def _observables_from(comp, mom, wsum, G, cfg):
    """observables (operators.py:55-103) from the energy components, the
    moments (sum rho, r^2 rho, (x^2+y^2) rho, z^2 rho, max rho before dV) and,
    for the q2d kernel, sum rho (W*rho) dV."""
    b2, b3, O, gc = cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"]
    d = float(G["ndim"])
    s = float(G["s"])
    dv = dV(G)
    E = comp["E"]
    msum, r2sum, rperp2, z2, peak = mom
    mass = msum * dv
    mu = E + comp["Econtact"] + 2.0 * comp["Ethree"] + comp["Egrav"]
    r2mean = r2sum * dv
    kind = G["kernel"]
    if kind == "none" or abs(gc) <= 1e-15:
        T_lr = 0.0
    elif kind == "newton":
        T_lr = comp["Egrav"]
    elif kind == "log":
        T_lr = -0.5 * gc * mass * mass
    else:
        T_lr = 0.5 * gc * wsum
    virial = (2.0 * comp["Ekin"] - s * comp["Etrap"] + d * comp["Econtact"] + 2.0 * d * comp["Ethree"]) + T_lr
    vscale = sum(abs(comp[k]) for k in ("Ekin", "Etrap", "Econtact", "Ethree", "Egrav")) + abs(T_lr)
    binding_ref = 1.0 if (abs(abs(O) - 1.0) < 1e-12 and s == 2 and G["ndim"] == 2) else float("nan")
    out = dict(E=E, mu=mu, Lz=comp["Lz"], rrms=math.sqrt(max(r2mean, 0.0)), peak=peak,
               E_bind=(E - binding_ref) if math.isfinite(binding_ref) else float("nan"),
               virial=virial, virial_rel=abs(virial) / max(1.0, vscale), virial_lr_term=T_lr)
    for k in ("Ekin", "Etrap", "Econtact", "Ethree", "Erot", "Egrav"):
        out[k] = comp[k]
    if G["ndim"] == 3:
        out["R_perp_rms"] = math.sqrt(max(rperp2 * dv, 0.0))
        out["R_z_rms"] = math.sqrt(max(z2 * dv, 0.0))
        out["oblateness"] = (out["R_perp_rms"] / math.sqrt(2.0)) / max(out["R_z_rms"], 1e-30)
    else:
        out["R_perp_rms"] = out["R_z_rms"] = out["oblateness"] = float("nan")
    return out


def observables(p, G, cfg):
    """operators.py:55-103: E, mu, Lz, radii, peak and the d-dimensional
    virial identity with the long-range term of each kernel."""
    b2, b3, O, gc = cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"]
    t = as_field(p, G)
    comp, _ = _energy_host(t, G, b2, b3, O, gc, False)
    wsum = float("nan")
    if G["kernel"] == "q2d" and abs(gc) > 1e-15:
        rho = (t.real ** 2 + t.imag ** 2).contiguous()
        wsum = (rho * kernel_convolution(rho, G, "W")).sum().item() * dV(G)
    return _observables_from(comp, moments(t, G), wsum, G, cfg)


def observables_batch(psis, G, cfgs):
    """observables of several states of one grid (up to 64 per call), every
    kernel launch covering all of them (bec_observe_batch)."""
    out = []
    for lo in range(0, len(cfgs), 64):
        cs = cfgs[lo:lo + 64]
        nb = len(cs)
        q2d = G["kernel"] == "q2d" and any(abs(float(c["G_C"])) > 1e-15 for c in cs)
        lr = G["kernel"] != "none" and any(abs(float(c["G_C"])) > 1e-15 for c in cs)
        pl = ensure_kernel(G, "K") if lr else plan_of(G)
        if q2d:
            ensure_kernel(G, "W")
        t = torch.stack([as_field(x, G) for x in psis[lo:lo + 64]]).contiguous()
        ph = (_lib.Phys * nb)(*[phys_for(G, c["beta2"], c["beta3"], c["Omega"], c["G_C"]) for c in cs])
        raw = (ctypes.c_double * (16 * nb))()
        _lib.check(pl.lib.bec_observe_batch(pl.h, nb, ph, _ptr(t), raw))
        for b, c in enumerate(cs):
            r = raw[16 * b:16 * b + 16]
            comp = {k: r[i] for i, k in enumerate(_COMP_KEYS)}
            out.append(_observables_from(comp, tuple(r[8:13]), r[13], G, c))
    return out


def preconditioner(grad, psi, G, cfg, phi=None):
    """solvers.py:28-45: (proj_tangent(IFFT(FFT(grad)/(k^2/2 + sigma))), sigma)."""
    gc = cfg["G_C"]
    pl = _plan(G, gc)
    g = as_field(grad, G)
    ps = as_field(psi, G)
    out = torch.empty_like(g)
    sig = ctypes.c_double()
    ph = phys_for(G, cfg["beta2"], cfg["beta3"], cfg["Omega"], gc)
    f = as_field(phi, G, torch.float64) if (phi is not None and _lr_on(G, gc)) else None
    _lib.check(pl.lib.bec_precondition(pl.h, _ptr(g), _ptr(ps), ctypes.byref(ph),
                                       _ptr(f) if f is not None else None,
                                       float(cfg.get("precond_shift_min", 0.5)), _ptr(out),
                                       ctypes.byref(sig)))
    return out, sig.value
