"""Physical (SI) inputs to the dimensionless couplings of a run, and its
outputs back to SI (becgpp/units.py:27-91).  Host arithmetic only: runs stay in
trap units (hbar = m = omega_perp = 1, unit norm), with the length a_ho =
sqrt(hbar / (m omega_perp)), the energy hbar omega_perp and the time
1 / omega_perp.  The mean-field maps:
  3D contact        beta2 = 4 pi N a_s / a_ho
  pancake contact   beta2 = sqrt(8 pi) N a_s / l_z,  l_z = sqrt(hbar / (m omega_z))
  self-gravity      G_C   = G m^2 N / (a_ho hbar omega_perp)   (one factor of N)
  pancake gravity   the same G_C with the q2d kernel at l_z / a_ho."""
import math

from .constants import G_NEWTON, HBAR

REGIMES = ("bec3d", "bec2d_pancake", "selfgrav3d", "selfgrav_pancake")


def physical_scales(m, omega_perp):
    """units.py:27-31: the SI length, energy (J and Hz) and time of trap units."""
    a_ho = math.sqrt(HBAR / (m * omega_perp))
    return dict(a_ho_m=a_ho, energy_J=HBAR * omega_perp, energy_Hz=omega_perp / (2.0 * math.pi),
                time_s=1.0 / omega_perp, m_kg=m, omega_perp=omega_perp)


def sim_params_from_physical(regime, m, omega_perp, N, a_s=0.0, omega_z=None, G=G_NEWTON, beta3=0.0,
                             verbose=True):
    """units.py:34-76: the couplings (beta2, beta3, G_C, Omega; dimension,
    kernel and l_z where the regime fixes them) and the scales of a regime in
    REGIMES, from m [kg], omega_perp [rad/s], N atoms, a_s [m] and, for the
    pancakes, omega_z [rad/s]."""
    if regime not in REGIMES:
        raise ValueError(f"unknown regime {regime!r}")
    sc = physical_scales(m, omega_perp)
    a_ho = sc["a_ho_m"]
    out = dict(beta2=0.0, beta3=float(beta3), G_C=0.0, Omega=0.0, scales=sc, regime=regime)
    pancake = regime in ("bec2d_pancake", "selfgrav_pancake")
    l_z = None
    if pancake:
        if omega_z is None:
            raise ValueError(f"{regime} needs omega_z")
        l_z = math.sqrt(HBAR / (m * omega_z))
        out["l_z_m"] = l_z
        out["dimension"] = "quasi2D"
    else:
        out["dimension"] = "3D"
    if regime.startswith("selfgrav"):
        out["G_C"] = G * m * m * N / (a_ho * HBAR * omega_perp)
        out["kernel"] = "q2d" if pancake else "newton"
        if pancake:
            out["l_z"] = l_z / a_ho
    if regime in ("bec3d", "bec2d_pancake") or a_s:
        out["beta2"] = (math.sqrt(8.0 * math.pi) * N * a_s / l_z) if pancake else (4.0 * math.pi * N * a_s / a_ho)
    if verbose:
        print(f"[units:{regime}] a_ho={a_ho:.3e} m  hbar*w={sc['energy_J']:.3e} J "
              f"({sc['energy_Hz']:.3g} Hz)  1/w={sc['time_s']:.3e} s")
        print(f"[units:{regime}] -> beta2={out['beta2']:.4g}  beta3={out['beta3']:.4g}  G_C={out['G_C']:.4g}")
    return out


def outputs_to_physical(scales, E=None, mu=None, R=None):
    """units.py:79-91: per-particle energies (x hbar omega_perp) and lengths
    (x a_ho) in SI."""
    res = {}
    if E is not None:
        res["E_J"] = E * scales["energy_J"]
        res["E_Hz"] = E * scales["energy_Hz"]
    if mu is not None:
        res["mu_J"] = mu * scales["energy_J"]
        res["mu_Hz"] = mu * scales["energy_Hz"]
    if R is not None:
        res["R_m"] = R * scales["a_ho_m"]
    return res
