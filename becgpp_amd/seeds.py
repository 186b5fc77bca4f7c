"""Initial states: Gaussian, Thomas-Fermi envelope, triangular vortex lattice
(becgpp/seeds.py:11-103).

The envelopes follow the reference.  The random perturbations come from a
host ``numpy.random.default_rng(rng)`` stream (amplitude noise first, then
phase noise; the lattice draws its angle and shift before the noise), so a
seed is the same on every device; the reference's seed tensors depend on the
torch device RNG and are not reproducible across devices.
"""
import math

import numpy as np
import torch

from .constants import DEV
from .grid import dV, norm_of
from .thomasfermi import extract_tf


def _normal(rng, shape):
    return torch.as_tensor(rng.standard_normal(shape), dtype=torch.float64, device=DEV)


def _winding(psi, G, cfg):
    m = int(cfg.get("seed_winding", 0))
    if m != 0:
        psi = psi * torch.exp(1j * m * torch.atan2(G["Y"], G["X"]))
    return psi


def gaussian_seed(G, cfg):
    """seeds.py:11-19."""
    rng = np.random.default_rng(int(cfg.get("rng", 0)))
    env = torch.exp(-0.5 * G["R2"])
    noise = 1.0 + cfg.get("seed_noise", 1e-2) * _normal(rng, tuple(env.shape))
    psi = _winding((env * noise).to(torch.complex128), G, cfg)
    return psi / norm_of(psi, G)


def tf_seed(G, cfg):
    """seeds.py:22-34."""
    tf = extract_tf(cfg, G)
    if tf is None or tf.get("n") is None:
        return gaussian_seed(G, cfg)
    amp = torch.sqrt(torch.clamp(tf["n"], min=0.0))
    rng = np.random.default_rng(int(cfg.get("rng", 0)))
    noise = 1.0 + cfg.get("seed_noise", 1e-2) * _normal(rng, tuple(amp.shape))
    pn = cfg.get("seed_phase_noise", 5e-2) * _normal(rng, tuple(amp.shape))
    psi = _winding((amp * noise).to(torch.complex128) * torch.exp(1j * pn), G, cfg)
    return psi / norm_of(psi, G)


def _inplane_radius(n, G, frac=0.90):
    """In-plane radius enclosing a mass fraction (seeds.py:37-44)."""
    w = (torch.clamp(n, min=0.0) * dV(G)).reshape(-1)
    rperp = torch.sqrt(G["X"] ** 2 + G["Y"] ** 2).reshape(-1)
    order = torch.argsort(rperp, stable=True)
    cum = torch.cumsum(w[order], dim=0)
    target = torch.tensor([float(frac)], dtype=cum.dtype, device=cum.device) * cum[-1]
    jj = int(torch.searchsorted(cum, target).clamp(max=cum.numel() - 1).item())
    return rperp[order[jj]].item()


def triangular_seed(G, cfg):
    """Abrikosov lattice of straight vortex lines along z on the TF envelope,
    winding sign(Omega) (seeds.py:47-94)."""
    O = float(cfg.get("Omega", 0.0))
    if abs(O) < 0.3:
        return tf_seed(G, cfg)
    tf = extract_tf(cfg, G)
    if tf is None or not tf.get("converged", False):
        return tf_seed(G, cfg)
    amp = torch.sqrt(torch.clamp(tf["n"], min=0.0))
    wsign = -1.0 if O < 0 else 1.0
    O = abs(O)
    av = math.sqrt(2.0 * math.pi / (math.sqrt(3.0) * O))
    dy = math.sqrt(3.0) * av / 2.0
    Rperp = _inplane_radius(tf["n"], G, 0.90)
    Rv = max(av, 0.98 * Rperp)
    rng = np.random.default_rng(int(cfg.get("rng", 0)))
    angle = float(rng.uniform(0.0, math.pi / 3.0))
    shift = rng.uniform(-0.25, 0.25, size=2) * av
    ca, sa = math.cos(angle), math.sin(angle)
    pos = []
    jmax = int(math.ceil(Rv / dy)) + 2
    imax = int(math.ceil(Rv / av)) + 3
    for jj in range(-jmax, jmax + 1):
        for ii in range(-imax, imax + 1):
            x0 = (ii + 0.5 * (jj & 1)) * av + shift[0]
            y0 = jj * dy + shift[1]
            xx = ca * x0 - sa * y0
            yy = sa * x0 + ca * y0
            if xx * xx + yy * yy < Rv * Rv:
                pos.append((xx, yy))
    phase = torch.zeros_like(amp)
    core = torch.ones_like(amp)
    xi = 0.25 * av
    for xx, yy in pos:
        dxt = G["X"] - xx
        dyt = G["Y"] - yy
        phase = phase + wsign * torch.atan2(dyt, dxt)
        core = core * torch.tanh(torch.sqrt(dxt * dxt + dyt * dyt) / max(xi, 1e-12))
    noise = cfg.get("seed_noise", 1e-2) * _normal(rng, tuple(amp.shape))
    psi = (amp * core * (1.0 + noise)).to(torch.complex128) * torch.exp(1j * phase)
    return psi / norm_of(psi, G)


def make_seed(G, cfg):
    """seeds.py:97-103."""
    kind = cfg.get("seed", "tf")
    if kind == "triangular":
        return triangular_seed(G, cfg)
    if kind == "tf":
        return tf_seed(G, cfg)
    return gaussian_seed(G, cfg)


def noisy_gaussian(G, noise=1e-3, seed=0, width=1.0):
    """normalise(exp(-r^2/2w^2) (1 + noise (xi1 + i xi2))), xi from default_rng(seed):
    the seeded Gaussian input of the benchmark workloads."""
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((2,) + (int(G["N"]),) * int(G["ndim"]))
    env = torch.exp(-0.5 * G["R2"] / (width * width))
    z = torch.as_tensor(xi[0] + 1j * xi[1], dtype=torch.complex128, device=DEV)
    psi = env * (1.0 + noise * z)
    return psi / norm_of(psi, G)
