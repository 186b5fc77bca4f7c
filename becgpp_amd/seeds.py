"""Initial states on the device (the reference's seeds.py:11-103).

Every seed is one launch of ``bec_seed``: an envelope (the Gaussian
exp(-r^2/2), or the square root of the Thomas-Fermi density), optional
vortex cores and phases, amplitude and phase noise, a central winding, then
normalisation.  The three kinds of the reference:

* gaussian -- exp(-r^2/2) (1 + seed_noise xi), times e^{i m theta} for
  seed_winding m;
* tf -- sqrt(n_TF) (1 + seed_noise xi_a) e^{i seed_phase_noise xi_p} e^{i m theta},
  falling back to the Gaussian where no TF density exists;
* triangular (|Omega| >= 0.3, converged TF density) -- an Abrikosov lattice of
  vortex lines along z: spacing a_v = sqrt(2 pi / (sqrt(3) |Omega|)), rows
  sqrt(3) a_v / 2 apart, rotated by a random angle in [0, pi/3) and shifted by
  a random fraction of a_v in [-1/4, 1/4)^2 (numpy default_rng(rng), drawn as
  the reference draws them), kept inside R_v = max(a_v, 0.98 R_90,perp);
  sqrt(n_TF) prod_k tanh(|r - r_k| / (a_v/4)) (1 + seed_noise xi)
  e^{i sign(Omega) sum_k theta_k}.

The noise xi is the library's counter-based generator (``bec_noise``: a
function of rng, stream and point index), so a seed is the same on every run
and reproducible on the CPU; the reference draws it from torch's generator.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .constants import DEV
from .grid import _ptr, as_field, dV, norm_of, plan_of
from .thomasfermi import extract_tf


def _device_seed(G, cfg, env=None, sites=None, vortex_sign=1.0, core=1.0, phase_noise=0.0, winding=None):
    pl = plan_of(G)
    spec = _lib.SeedSpec(1 if env is not None else 0, float(cfg.get("seed_noise", 1e-2)), float(phase_noise),
                         int(cfg.get("seed_winding", 0)) if winding is None else int(winding), float(vortex_sign),
                         float(core), int(cfg.get("rng", 0)) & 0xFFFFFFFFFFFFFFFF)
    e = as_field(env, G, torch.float64) if env is not None else None
    v = None
    nv = 0
    if sites is not None and len(sites):
        v = torch.as_tensor(np.ascontiguousarray(sites, dtype=np.float64).reshape(-1), device=DEV)
        nv = len(sites)
    psi = torch.empty(tuple(G["R2"].shape), dtype=torch.complex128, device=DEV)
    _lib.check(pl.lib.bec_seed(pl.h, ctypes.byref(spec), _ptr(e) if e is not None else None,
                               _ptr(v) if v is not None else None, nv, _ptr(psi)))
    return psi


def gaussian_seed(G, cfg):
    """seeds.py:11-19."""
    return _device_seed(G, cfg)


def tf_seed(G, cfg):
    """seeds.py:22-34."""
    tf = extract_tf(cfg, G)
    if tf is None or tf.get("n") is None:
        return gaussian_seed(G, cfg)
    return _device_seed(G, cfg, env=tf["n"], phase_noise=float(cfg.get("seed_phase_noise", 5e-2)))


def inplane_radius(n, G, frac=0.90):
    """The in-plane (x-y) radius about the axis enclosing the mass fraction
    (seeds.py:37-44): the first point, in order of distance from the z axis,
    at which the running mass reaches frac of the total."""
    dist = torch.sqrt(G["X"] ** 2 + G["Y"] ** 2).reshape(-1)
    order = torch.argsort(dist, stable=True)
    running = torch.cumsum((torch.clamp(n, min=0.0) * dV(G)).reshape(-1)[order], dim=0)
    target = torch.tensor([float(frac)], dtype=running.dtype, device=running.device) * running[-1]
    at = int(torch.searchsorted(running, target).clamp(max=running.numel() - 1).item())
    return dist[order[at]].item()


def lattice_sites(Omega, radius, rng):
    """Triangular lattice of vortex positions for rotation Omega inside
    `radius`, rotated and shifted by the draws of `rng` (seeds.py:61-82), in
    row order (rows y, then x along a row)."""
    av = math.sqrt(2.0 * math.pi / (math.sqrt(3.0) * abs(Omega)))
    row = math.sqrt(3.0) * av / 2.0
    angle = float(rng.uniform(0.0, math.pi / 3.0))
    shift = rng.uniform(-0.25, 0.25, size=2) * av
    nrow = int(math.ceil(radius / row)) + 2
    ncol = int(math.ceil(radius / av)) + 3
    jj, ii = np.meshgrid(np.arange(-nrow, nrow + 1), np.arange(-ncol, ncol + 1), indexing="ij")
    x0 = (ii + 0.5 * (jj & 1)) * av + shift[0]
    y0 = jj * row + shift[1]
    ca, sa = math.cos(angle), math.sin(angle)
    xs = (ca * x0 - sa * y0).reshape(-1)
    ys = (sa * x0 + ca * y0).reshape(-1)
    keep = xs * xs + ys * ys < radius * radius
    return np.stack([xs[keep], ys[keep]], axis=1), av


def triangular_seed(G, cfg):
    """seeds.py:47-94."""
    om = float(cfg.get("Omega", 0.0))
    if abs(om) < 0.3:
        return tf_seed(G, cfg)
    tf = extract_tf(cfg, G)
    if tf is None or not tf.get("converged", False):
        return tf_seed(G, cfg)
    av = math.sqrt(2.0 * math.pi / (math.sqrt(3.0) * abs(om)))
    radius = max(av, 0.98 * inplane_radius(tf["n"], G, 0.90))
    sites, av = lattice_sites(om, radius, np.random.default_rng(int(cfg.get("rng", 0))))
    return _device_seed(G, cfg, env=tf["n"], sites=sites, vortex_sign=-1.0 if om < 0 else 1.0,
                        core=max(0.25 * av, 1e-12), winding=0)


_KINDS = {"triangular": triangular_seed, "tf": tf_seed}


def make_seed(G, cfg):
    """seeds.py:97-103: the seed kind cfg["seed"] (anything else: gaussian)."""
    return _KINDS.get(cfg.get("seed", "tf"), gaussian_seed)(G, cfg)


def noisy_gaussian(G, noise=1e-3, seed=0, width=1.0):
    """normalise(exp(-r^2/2w^2) (1 + noise (xi1 + i xi2))), xi from default_rng(seed):
    the seeded Gaussian input of the benchmark workloads."""
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((2,) + (int(G["N"]),) * int(G["ndim"]))
    env = torch.exp(-0.5 * G["R2"] / (width * width))
    z = torch.as_tensor(xi[0] + 1j * xi[1], dtype=torch.complex128, device=DEV)
    psi = env * (1.0 + noise * z)
    return psi / norm_of(psi, G)
