"""numpy/scipy FP64 restatement of the reference ground-state path (oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

Citations are ``file:line`` in the reference (becGPP 1.2.1).  Arrays follow the
reference layout: C-contiguous, ``indexing="ij"`` (axis 0 = x), complex128 for
wavefunctions, float64 for densities and potentials.  FFTs are scipy's
pocketfft (``scipy.fft``) with ``workers`` threads; they are full complex
transforms exactly as the reference's ``torch.fft.fftn`` calls.

How each part is pinned (tests in ``tests/test_oracle_*.py``):

* grid / kinetic / L_z -- LLL identities E_sp=1, L_z=m, <r^2>=m+1 to 1e-12
  (becgpp/modes.py:108-122, tests/test_validation.py:48), exact
  non-interacting energies 1.0 / 1.5 / 0.5 / 1.0 (tests/test_audit.py:109-119),
  spectral derivative of band-limited plane waves.
* kernels + convolution -- the exact relative errors committed in
  benchmarks/results/verification/validation.csv:2-43 (N=512 2D gate, N=96
  3D, kernel-order ladders), brute-force O(n^2) direct sums on tiny grids,
  symmetry <a,K*b>=<K*a,b> (tests/test_audit.py:88-96), q2d limits
  (tests/test_kernels.py:53-63).
* energy / H / mu -- central finite differences of E vs 2 Re<H psi, v> dv over
  the six physics sets of tests/test_audit.py:55-85; mu = <H> (ibid. 99-106).
* solver -- trajectories the reference ran itself for eleven small cases
  (tests/golden/make_trajectory_golden.py, tests/test_oracle_trajectory.py):
  every accepted step's E and residual, the iteration count and stop reason;
  converged energies of benchmarks/results/itp_timing/timing.csv:2-3 and
  pcg_variants.csv; stall / maxit semantics (tests/test_audit.py:165-172,
  252-255); virial identities (tests/test_kernels.py:30-50).
* preconditioner (shift sigma and z) -- vectors the reference computed itself
  for nine cases (tests/golden/make_precond_golden.py: the six physics sets
  of tests/test_audit.py:55-62, the floor, a cubic-quintic and a 3D state;
  tests/test_oracle_precond.py), which fix the local stiffness, the bulk mask
  and the floor; the closed form P((k^2/2 + sigma) v) = proj_tangent(v); the
  lower median against a brute-force sort and a hand-computed density.
* diagnostics (observables, radii, vortex counts, LLL weight) and
  resampling -- the reference's own diagnostics bundle on eight states and
  three resampled states (tests/golden/make_diag_golden.py,
  tests/test_oracle_diag_bundle.py).
* Thomas-Fermi envelopes -- envelopes the reference computed itself for
  eight cases (tests/golden/make_tf_golden.py, tests/test_oracle_tf.py): the
  regime, mu, radii and density, the nonlocal fixed point included.
* seeds -- the recipes (envelopes, windings, the triangular lattice) by the
  reference's own seeds with the noise switched off (tests/golden/
  make_seed_golden.py, tests/test_oracle_seed_recipe.py); the noise is this
  build's counter-based generator
  (``counter_normals``: splitmix64 + Box-Muller; the splitmix64 stream is
  pinned by its published outputs for seed 1234567, the normals by their
  moments); the reference's seed tensors depend on the torch device RNG and
  cannot be reproduced, so the noise itself is parity unpinned against the
  reference by construction.  The envelopes, lattice positions (numpy
  default_rng angle and shift, as the reference draws them) and the order of
  the factors are the reference's; the noise-free seeds are pinned by closed
  forms (Gaussian envelope, TF parabola, vortex windings).
"""
import math
import os

import numpy as np
import scipy.fft as sfft
import scipy.special as ssp

WORKERS = max(1, int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1)))


def set_threads(n):
    global WORKERS
    WORKERS = max(1, int(n))


def fftn(a):
    return sfft.fftn(a, workers=WORKERS)


def ifftn(a):
    return sfft.ifftn(a, workers=WORKERS)


# ----------------------------------------------------------------------------
# grid  (becgpp/grid.py:10-116)
# ----------------------------------------------------------------------------
def geometry(cfg):
    """grid.py:10-20."""
    g = str(cfg.get("dimension", "2D")).strip().lower()
    if g == "2d":
        return 2, False
    if g == "3d":
        return 3, False
    if g == "quasi2d":
        return 2, True
    raise ValueError(f"unknown dimension {cfg.get('dimension')!r}")


def resolve_kernel(cfg):
    """grid.py:23-55."""
    ndim, quasi = geometry(cfg)
    k = str(cfg.get("kernel", "auto")).lower()
    if abs(float(cfg.get("G_C", 0.0))) <= 1e-15 or k == "none":
        return "none", 0
    lz = float(cfg.get("l_z", 0.0))
    if k == "auto":
        k = "newton" if ndim == 3 else (("q2d" if lz > 0 else "newton") if quasi else "log")
    if k == "newton":
        return "newton", 1
    if k == "log":
        if ndim != 2:
            raise ValueError("kernel='log' is two-dimensional only")
        return "log", 0
    if k == "q2d":
        if ndim != 2:
            raise ValueError("kernel='q2d' needs dimension '2D' or 'quasi2D'")
        if lz <= 0:
            raise ValueError("kernel='q2d' needs an axial width l_z > 0")
        return "q2d", 0
    raise ValueError(f"unknown kernel {k!r}")


def make_grid(cfg):
    """grid.py:69-101 (no memory warning)."""
    ndim, quasi = geometry(cfg)
    L, N, s = float(cfg["L"]), int(cfg["Ngrid"]), float(cfg["s"])
    kkind, kexp = resolve_kernel(cfg)
    dx = 2.0 * L / N
    x = (np.arange(N, dtype=np.float64) - N // 2) * dx
    k1 = 2 * math.pi * np.fft.fftfreq(N, d=dx)
    coords = np.meshgrid(*([x] * ndim), indexing="ij")
    kcoords = np.meshgrid(*([k1] * ndim), indexing="ij")
    R2 = sum(c ** 2 for c in coords)
    K2 = sum(c ** 2 for c in kcoords)
    tc = float(cfg.get("trap_coeff", 0.5))
    V = tc * R2 ** (s / 2.0) if s > 0 else np.zeros_like(R2)
    pad = max(2, int(cfg.get("pad", 2)))
    return dict(ndim=ndim, quasi=quasi, coords=tuple(coords), kcoords=tuple(kcoords),
                X=coords[0], Y=coords[1], KX=kcoords[0], KY=kcoords[1], R2=R2, V=V, K2=K2,
                s=s, trap_coeff=tc, dx=dx, N=N, L=L, kernel=kkind, kexp=kexp,
                l_z=float(cfg.get("l_z", 0.0)), pad=pad)


def dV(G):
    """grid.py:104-105."""
    return G["dx"] ** G["ndim"]


def norm_of(p, G):
    """grid.py:108-109."""
    return math.sqrt(float((np.abs(p) ** 2).sum()) * dV(G))


def Lz_op(p, G):
    """grid.py:112-116: -i (X d_y - Y d_x) psi with spectral derivatives."""
    fp = fftn(p)
    dpx = ifftn(1j * G["KX"] * fp)
    dpy = ifftn(1j * G["KY"] * fp)
    return -1j * (G["X"] * dpy - G["Y"] * dpx)


# ----------------------------------------------------------------------------
# long-range kernels  (becgpp/interactions.py:30-176)
# ----------------------------------------------------------------------------
GL_NODES = (-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526)
GL_WTS = (0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538)
SELF_NEWTON_2D = 4.0 * math.asinh(1.0)                       # interactions.py:41
SELF_NEWTON_3D = 2.3800773640                                # interactions.py:42
SELF_LOG_2D_CONST = 1.5 + 0.5 * math.log(2.0) - 0.25 * math.pi   # interactions.py:43

_KCACHE = {}


def clear_kernel_cache():
    _KCACHE.clear()


def q2d_kernel(u, l_z):
    """interactions.py:64-67: K_eff(u) = k0e(u^2/4l^2)/(sqrt(2 pi) l)."""
    x = np.asarray(u, dtype=np.float64) ** 2 / (4.0 * l_z * l_z)
    return ssp.k0e(x) / (math.sqrt(2.0 * math.pi) * l_z)


def q2d_virial_kernel(u, l_z):
    """interactions.py:70-74: W(u) = u K_eff'(u)."""
    x = np.asarray(u, dtype=np.float64) ** 2 / (4.0 * l_z * l_z)
    xs = np.maximum(x, 1e-300)
    return 2.0 * xs * (ssp.k0e(xs) - ssp.k1e(xs)) / (math.sqrt(2.0 * math.pi) * l_z)


def polar_self_cell_2d(f, dx, n=48):
    """interactions.py:77-94: average of radial f over the square self cell."""
    t, wt = np.polynomial.legendre.leggauss(n)
    th = 0.5 * (math.pi / 4.0) * (t + 1.0)
    wth = 0.5 * (math.pi / 4.0) * wt
    tt = 0.5 * (t + 1.0)
    wtt = 0.5 * wt
    total = 0.0
    for thj, wj in zip(th, wth):
        R = 0.5 * dx / math.cos(thj)
        r = R * tt ** 2
        jac = 2.0 * R * R * tt ** 3
        total += wj * float((f(r) * jac * wtt).sum())
    return 8.0 * total / (dx * dx)


def kernel_values(which, G, DXs):
    """interactions.py:97-109."""
    kind = G["kernel"]
    r2 = sum(d * d for d in DXs)
    if which == "W":
        return q2d_virial_kernel(np.sqrt(r2), G["l_z"])
    if kind == "newton":
        return 1.0 / np.sqrt(r2)
    if kind == "log":
        return -0.5 * np.log(r2)
    if kind == "q2d":
        return q2d_kernel(np.sqrt(r2), G["l_z"])
    raise ValueError(f"unknown kernel {kind!r}")


def kernel_real_space(G, which="K"):
    """interactions.py:117-147: the cell-averaged kernel on the (pad N)^d
    wrap-around displacement grid, self cell replaced by its exact average."""
    N, dx, ndim, pad = G["N"], G["dx"], G["ndim"], G["pad"]
    M = pad * N
    j = np.arange(M, dtype=np.float64)
    dj = np.where(j <= M // 2, j, j - M) * dx
    if ndim == 2:
        DX, DY = np.meshgrid(dj, dj, indexing="ij")
        ker = np.zeros_like(DX)
        for u, wu in zip(GL_NODES, GL_WTS):
            for v, wv in zip(GL_NODES, GL_WTS):
                ker += 0.25 * wu * wv * kernel_values(which, G, (DX + 0.5 * dx * u, DY + 0.5 * dx * v))
        if which == "W":
            ker[0, 0] = polar_self_cell_2d(lambda r: q2d_virial_kernel(r, G["l_z"]), dx)
        elif G["kernel"] == "newton":
            ker[0, 0] = SELF_NEWTON_2D / dx
        elif G["kernel"] == "log":
            ker[0, 0] = -math.log(dx) + SELF_LOG_2D_CONST
        else:
            ker[0, 0] = polar_self_cell_2d(lambda r: q2d_kernel(r, G["l_z"]), dx)
    else:
        if G["kernel"] != "newton":
            raise ValueError("only the Newton kernel is available in 3D")
        DX, DY, DZ = np.meshgrid(dj, dj, dj, indexing="ij")
        ker = np.zeros_like(DX)
        for u, wu in zip(GL_NODES, GL_WTS):
            for v, wv in zip(GL_NODES, GL_WTS):
                for w, ww in zip(GL_NODES, GL_WTS):
                    ker += (wu * wv * ww / 8.0) / np.sqrt(
                        (DX + 0.5 * dx * u) ** 2 + (DY + 0.5 * dx * v) ** 2 + (DZ + 0.5 * dx * w) ** 2)
        ker[0, 0, 0] = SELF_NEWTON_3D / dx
    return ker, M


def build_kernel_fft(G, which="K"):
    """interactions.py:112-151 (cached like the reference)."""
    key = (which, G["kernel"], G["ndim"], int(G["N"]), round(float(G["dx"]), 14),
           int(G["pad"]), round(float(G.get("l_z", 0.0)), 14))
    if key in _KCACHE:
        return _KCACHE[key]
    ker, M = kernel_real_space(G, which)
    Kf = fftn(ker.astype(np.complex128))
    _KCACHE[key] = (Kf, M)
    return Kf, M


def kernel_convolution(rho, G, which="K"):
    """interactions.py:154-164: dV * Re IFFT(FFT(rho_pad) Kf)[crop]."""
    N, ndim = G["N"], G["ndim"]
    Kf, M = build_kernel_fft(G, which)
    i0 = (M - N) // 2
    rho_pad = np.zeros((M,) * ndim, dtype=np.complex128)
    sl = tuple(slice(i0, i0 + N) for _ in range(ndim))
    rho_pad[sl] = rho
    conv = ifftn(fftn(rho_pad) * Kf).real
    return dV(G) * conv[sl]


def long_range_phi(rho, G, gc):
    """interactions.py:167-171."""
    if G["kernel"] == "none" or abs(gc) <= 1e-15:
        return np.zeros_like(rho)
    return -gc * kernel_convolution(rho, G, "K")


def direct_convolution(rho, G, which="K"):
    """Brute-force O(n^2) sum_j K(r_i - r_j) rho_j dV with the same cell-averaged
    kernel table (test aid for tiny grids)."""
    ker, M = kernel_real_space(G, which)
    N, ndim = G["N"], G["ndim"]
    out = np.zeros((N,) * ndim)
    idx = np.indices((N,) * ndim).reshape(ndim, -1).T
    flat_rho = rho.reshape(-1)
    for a, ia in enumerate(idx):
        d = (ia[None, :] - idx) % M
        out.reshape(-1)[a] = (ker[tuple(d.T)] * flat_rho).sum()
    return dV(G) * out


# ----------------------------------------------------------------------------
# operators  (becgpp/operators.py:10-103)
# ----------------------------------------------------------------------------
def energy_components(p, G, beta2, beta3, Omega, G_C):
    """operators.py:10-29."""
    dv = dV(G)
    rho = np.abs(p) ** 2
    Ntot = float(p.size)
    grad2 = float((G["K2"] * np.abs(fftn(p)) ** 2).sum()) * dv / Ntot
    Ekin = 0.5 * grad2
    Etrap = float((G["V"] * rho).sum()) * dv
    Econtact = 0.5 * beta2 * float((rho ** 2).sum()) * dv
    Ethree = (beta3 / 3.0) * float((rho ** 3).sum()) * dv
    Lz = float((p.conj() * Lz_op(p, G)).real.sum()) * dv
    Erot = -Omega * Lz
    if G["kernel"] != "none" and abs(G_C) > 1e-15:
        phi = long_range_phi(rho, G, G_C)
        Egrav = 0.5 * float((phi * rho).sum()) * dv
    else:
        phi = np.zeros_like(rho)
        Egrav = 0.0
    E = Ekin + Etrap + Econtact + Ethree + Erot + Egrav
    return dict(E=E, Ekin=Ekin, Etrap=Etrap, Econtact=Econtact, Ethree=Ethree,
                Erot=Erot, Egrav=Egrav, Lz=Lz, phi=phi)


def apply_H(psi, G, cfg, phi=None):
    """operators.py:36-45."""
    b2, b3, O, gc = cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"]
    kin = ifftn(0.5 * G["K2"] * fftn(psi))
    rho = np.abs(psi) ** 2
    if phi is None:
        phi = long_range_phi(rho, G, gc) if (G["kernel"] != "none" and abs(gc) > 1e-15) else 0.0
    return kin + (G["V"] + b2 * rho + b3 * rho ** 2 + phi) * psi - O * Lz_op(psi, G)


def compute_residual(psi, G, cfg):
    """operators.py:48-52 -> (resid_abs, mu, resid_rel)."""
    Hpsi = apply_H(psi, G, cfg)
    mu = float((psi.conj() * Hpsi).real.sum()) * dV(G)
    resid = math.sqrt(float((np.abs(Hpsi - mu * psi) ** 2).sum()) * dV(G))
    return resid, mu, resid / max(1.0, abs(mu))


def observables(p, G, cfg):
    """operators.py:55-103."""
    b2, b3, O, gc = cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"]
    d = float(G["ndim"])
    s = float(G["s"])
    dv = dV(G)
    comp = energy_components(p, G, b2, b3, O, gc)
    E = comp["E"]
    rho = np.abs(p) ** 2
    mass = float(rho.sum()) * dv
    mu = E + comp["Econtact"] + 2.0 * comp["Ethree"] + comp["Egrav"]
    r2mean = float((G["R2"] * rho).sum()) * dv
    kind = G["kernel"]
    if kind == "none" or abs(gc) <= 1e-15:
        T_lr = 0.0
    elif kind == "newton":
        T_lr = comp["Egrav"]
    elif kind == "log":
        T_lr = -0.5 * gc * mass * mass
    else:
        T_lr = 0.5 * gc * float((rho * kernel_convolution(rho, G, "W")).sum()) * dv
    virial = (2.0 * comp["Ekin"] - s * comp["Etrap"] + d * comp["Econtact"]
              + 2.0 * d * comp["Ethree"]) + T_lr
    vscale = sum(abs(comp[k]) for k in ("Ekin", "Etrap", "Econtact", "Ethree", "Egrav")) + abs(T_lr)
    binding_ref = 1.0 if (abs(abs(O) - 1.0) < 1e-12 and s == 2 and G["ndim"] == 2) else float("nan")
    out = dict(E=E, mu=mu, Lz=comp["Lz"], rrms=math.sqrt(max(r2mean, 0.0)), peak=float(rho.max()),
               E_bind=(E - binding_ref) if math.isfinite(binding_ref) else float("nan"),
               virial=virial, virial_rel=abs(virial) / max(1.0, vscale), virial_lr_term=T_lr)
    for k in ("Ekin", "Etrap", "Econtact", "Ethree", "Erot", "Egrav"):
        out[k] = comp[k]
    if G["ndim"] == 3:
        X, Y, Z = G["coords"]
        rperp2 = float(((X ** 2 + Y ** 2) * rho).sum()) * dv
        z2 = float((Z ** 2 * rho).sum()) * dv
        out["R_perp_rms"] = math.sqrt(max(rperp2, 0.0))
        out["R_z_rms"] = math.sqrt(max(z2, 0.0))
        out["oblateness"] = (out["R_perp_rms"] / math.sqrt(2.0)) / max(out["R_z_rms"], 1e-30)
    else:
        out["R_perp_rms"] = out["R_z_rms"] = out["oblateness"] = float("nan")
    return out


# ----------------------------------------------------------------------------
# solver  (becgpp/solvers.py:20-263)
# ----------------------------------------------------------------------------
def inner_re(u, v, dv):
    """solvers.py:20-21."""
    return float((u.conj() * v).real.sum()) * dv


def proj_tangent(v, psi, dv):
    """solvers.py:24-25."""
    return v - psi * ((psi.conj() * v).sum() * dv)


def lower_median(vals):
    """torch.median semantics: the lower of the two middle values."""
    s = np.sort(np.asarray(vals, dtype=np.float64).reshape(-1))
    return float(s[(s.size - 1) // 2])


def precond_shift(psi, G, cfg, phi=None):
    """solvers.py:33-42: sigma = max(lower median of the local potential over
    {rho > 0.05 max rho}, precond_shift_min)."""
    rho = np.abs(psi) ** 2
    b2, b3, gc = cfg["beta2"], cfg["beta3"], cfg["G_C"]
    local = G["V"] + 2.0 * b2 * rho + 3.0 * b3 * rho ** 2
    if G["kernel"] != "none" and abs(gc) > 1e-15:
        local = local + (phi if phi is not None else long_range_phi(rho, G, gc))
    mask = rho > 0.05 * rho.max()
    vals = local[mask]
    floor = float(cfg.get("precond_shift_min", 0.5))
    sigma = lower_median(vals) if vals.size else floor
    return max(sigma, floor)


def preconditioner(grad, psi, G, cfg, phi=None):
    """solvers.py:28-45."""
    sigma = precond_shift(psi, G, cfg, phi)
    denom = 0.5 * G["K2"] + sigma
    z = ifftn(fftn(grad) / denom)
    return proj_tangent(z, psi, dV(G)), sigma


def ground_state(cfg, G=None, psi0=None, verbose=False, callback=None):
    """solvers.py:48-235.  ``psi0`` is required (the seeds of this build are
    generated by the caller).  ``callback(it, E, resid_rel, trial)`` is called
    after every accepted step (test aid)."""
    if G is None:
        G = make_grid(cfg)
    if psi0 is None:
        psi0 = make_seed(G, cfg)
    dv = dV(G)
    psi = np.array(psi0, dtype=np.complex128)
    psi = psi / norm_of(psi, G)
    maxit = int(cfg["maxit"])
    res_tol = float(cfg["res_tol"])
    etol = float(cfg.get("energy_tol", 1e-9))
    alpha = float(cfg.get("step", 1.0))
    alpha_min, alpha_max = 1e-8, float(cfg.get("step_max", 3.0))
    ls_max = int(cfg.get("linesearch_max", 8))
    stall_max = max(1, int(cfg.get("linesearch_stall", 10)))
    c1 = 1e-4
    shrink, growth = 0.5, 1.1
    bb_min, bb_max, bb_mix = 0.05, alpha_max, 0.5
    restart_period = int(cfg.get("cg_restart", 30))
    cg_beta = str(cfg.get("cg_beta", "pr_precond")).lower()
    if cg_beta not in ("pr", "pr_precond", "none"):
        raise ValueError(f"cg_beta must be 'pr', 'pr_precond' or 'none', got {cg_beta!r}")
    conv_window = int(cfg.get("conv_window", 2000))
    it_ref = 0
    b2, b3, O, gc = cfg["beta2"], cfg["beta3"], cfg["Omega"], cfg["G_C"]

    def energy_phi(p):
        comp = energy_components(p, G, b2, b3, O, gc)
        return comp["E"], comp["phi"]

    def residual(p, phi=None):
        Hp = apply_H(p, G, cfg, phi=phi)
        mu = float((p.conj() * Hp).sum().real) * dv
        g = Hp - mu * p
        ra = math.sqrt(float((np.abs(g) ** 2).sum()) * dv)
        return g, mu, ra, ra / max(1.0, abs(mu))

    E, phi = energy_phi(psi)
    grad, mu, _, resid_rel = residual(psi, phi)
    z, sigma = preconditioner(grad, psi, G, cfg, phi)
    direction = z.copy()
    z_prev = z
    converged = False
    stop = "maxit"
    E_prev = E
    psi_prev = psi
    grad_prev = grad
    best_resid = resid_rel
    best_psi = psi.copy()
    E_ref = E
    n_fail = 0
    n_iter = 0
    n_energy = 1
    for it in range(maxit):
        n_iter = it + 1
        z, sigma = preconditioner(grad, psi, G, cfg, phi)
        if it == 0 or cg_beta == "none" or (restart_period > 0 and it % restart_period == 0):
            beta = 0.0
        else:
            y = grad - grad_prev
            if cg_beta == "pr_precond":
                den = max(inner_re(z_prev, grad_prev, dv), 1e-300)
                beta = max(0.0, inner_re(z, y, dv) / den)
            else:
                den = max(inner_re(grad_prev, grad_prev, dv), 1e-300)
                beta = max(0.0, inner_re(grad, y, dv) / den)
            if not math.isfinite(beta):
                beta = 0.0
        dirc = proj_tangent(z + beta * direction, psi, dv)
        gd = inner_re(grad, dirc, dv)
        if gd <= 0:
            dirc = z
            gd = inner_re(grad, dirc, dv)
        if gd <= 0:
            dirc = proj_tangent(grad, psi, dv)
            gd = inner_re(grad, dirc, dv)
        if it > 0:
            svec = psi - psi_prev
            yvec = grad - grad_prev
            sy = inner_re(svec, yvec, dv)
            if sy > 1e-30 and math.isfinite(sy):
                bb = inner_re(svec, svec, dv) / sy
                if math.isfinite(bb):
                    bb = min(max(bb, bb_min), bb_max)
                    alpha = (1.0 - bb_mix) * alpha + bb_mix * bb
        alpha = min(max(alpha, alpha_min), alpha_max)
        accepted = False
        trial = alpha
        psi_t, Et, phi_t = psi, E, phi
        for _ in range(ls_max):
            psi_t = psi - trial * dirc
            psi_t = psi_t / norm_of(psi_t, G)
            Et, phi_t = energy_phi(psi_t)
            n_energy += 1
            if math.isfinite(Et) and Et <= E - c1 * trial * gd:
                accepted = True
                break
            trial *= shrink
            if trial < alpha_min:
                break
        if not accepted:
            n_fail += 1
            alpha = max(alpha * 0.5, alpha_min)
            direction = z.copy()
            if alpha <= alpha_min or n_fail >= stall_max:
                stop = "line_search_stalled"
                break
            continue
        n_fail = 0
        psi_prev, grad_prev, z_prev, E_prev = psi, grad, z, E
        psi, E, phi = psi_t, Et, phi_t
        alpha = min(alpha_max, trial * growth)
        direction = dirc
        grad, mu, resid_abs, resid_rel = residual(psi, phi)
        if callback is not None:
            callback(it, E, resid_rel, trial)
        if resid_rel < best_resid:
            best_resid = resid_rel
            best_psi = psi.copy()
        if verbose and it % max(1, int(cfg.get("check", 200))) == 0:
            print(f"  [oracle] it={it} E={E:.12f} res={resid_rel:.3e}")
        if resid_rel < res_tol:
            converged = True
            stop = "converged"
            break
        if it - it_ref >= conv_window:
            dE_win = abs(E - E_ref) / max(1.0, abs(E))
            if dE_win < etol:
                converged = True
                stop = "energy_converged"
                break
            E_ref = E
            it_ref = it
    if best_resid < resid_rel:
        psi = best_psi
    resid_abs, mu_f, resid_rel = compute_residual(psi, G, cfg)
    converged = bool(converged or (resid_rel < res_tol))
    obs = observables(psi, G, cfg)
    obs.update(iters=n_iter, resid_abs=resid_abs, resid_rel=resid_rel, min_resid=best_resid,
               converged=converged, stop_reason=stop, preconditioner_shift=sigma,
               n_energy=n_energy)
    return psi, G, obs


def ground_state_multiseed(cfg, G=None, verbose=False):
    """solvers.py:238-263 (seeds from this build's RNG recipe, see make_seed)."""
    n = int(cfg.get("nseeds", 1))
    if G is None:
        G = make_grid(cfg)
    rotating = abs(float(cfg.get("Omega", 0.0))) >= 0.3
    seed_cycle = ("triangular", "tf", "gaussian") if rotating else ("tf", "gaussian")
    best = None
    Es = []
    for j in range(max(1, n)):
        c = dict(cfg)
        c["rng"] = int(cfg.get("rng", 0)) + j
        c["seed"] = seed_cycle[j % len(seed_cycle)]
        psi, G, obs = ground_state(c, G=G, psi0=make_seed(G, c), verbose=verbose)
        Es.append(obs["E"])
        if best is None or obs["E"] < best[0]:
            best = (obs["E"], psi, obs, c)
    _, psi, obs, cbest = best
    obs["seed_energy_spread"] = (max(Es) - min(Es)) if len(Es) > 1 else 0.0
    obs["seed_used"] = cbest["seed"]
    obs["seed_energies"] = ";".join(f"{e:.10g}" for e in Es)
    return psi, G, obs


# ----------------------------------------------------------------------------
# Thomas-Fermi envelopes  (becgpp/thomasfermi.py:21-156, fields.py:17-30)
# ----------------------------------------------------------------------------
def center_of_mass(rho, G):
    """fields.py:17-22."""
    mass = float(rho.sum()) * dV(G)
    if mass <= 0:
        return tuple(0.0 for _ in range(G["ndim"]))
    return tuple(float((c * rho).sum()) * dV(G) / mass for c in G["coords"])


def mass_quantiles(rho, G, fractions=(0.50, 0.90, 0.99), center=None):
    """fields.py:25-38 (stable argsort)."""
    w = (np.maximum(rho, 0.0) * dV(G)).reshape(-1)
    if center is None:
        center = center_of_mass(rho, G)
    rad2 = sum((c - center[i]) ** 2 for i, c in enumerate(G["coords"]))
    radii = np.sqrt(rad2).reshape(-1)
    order = np.argsort(radii, kind="stable")
    cum = np.cumsum(w[order])
    out = []
    for f in fractions:
        jj = min(int(np.searchsorted(cum, float(f) * cum[-1], side="left")), cum.size - 1)
        out.append(float(radii[order[jj]]))
    return out


# ----------------------------------------------------------------------------
# Resampling between grids  (becgpp/grid.py:140-161)
# ----------------------------------------------------------------------------
def resample_state(psi_old, G_old, G_new):
    """grid.py:140-161: torch grid_sample (bilinear / trilinear, align_corners
    = True, zero padding) restated: per axis t = ((g + 1) / 2) (n0 - 1) with
    g = 2 (c - x0) / (x1 - x0) - 1, corners floor(t), floor(t) + 1 with weights
    (floor(t) + 1 - t) and (t - floor(t)), multiplied last axis first and
    summed corner by corner (last axis fastest); renormalised; None when
    degenerate."""
    d = G_old["ndim"]
    n0, dx0 = int(G_old["N"]), float(G_old["dx"])
    x0 = -(n0 // 2) * dx0
    x1 = x0 + (n0 - 1) * dx0
    if not (x1 > x0):
        return None
    axes = []
    for c in G_new["coords"]:
        g = 2.0 * (c - x0) / (x1 - x0) - 1.0
        t = ((g + 1.0) / 2.0) * (n0 - 1)
        lo = np.floor(t)
        axes.append((lo.astype(np.int64), (lo + 1.0) - t, t - lo))
    out = np.zeros(G_new["coords"][0].shape, complex)
    for q in range(1 << d):
        bits = [(q >> (d - 1 - a)) & 1 for a in range(d)]        # axis 0 slowest
        idx, ok, w = [], np.ones(out.shape, bool), None
        for a in reversed(range(d)):                              # weights: last axis first
            lo, w0, w1 = axes[a]
            j = lo + bits[a]
            ok &= (j >= 0) & (j < n0)
            wa = w1 if bits[a] else w0
            w = wa if w is None else w * wa
        for a in range(d):
            idx.append(np.clip(axes[a][0] + bits[a], 0, n0 - 1))
        val = psi_old[tuple(idx)]
        out = out + np.where(ok, val.real * w, 0.0) + 1j * np.where(ok, val.imag * w, 0.0)
    nn = math.sqrt(float((np.abs(out) ** 2).sum()) * dV(G_new))
    if not np.isfinite(nn) or nn <= 1e-14:
        return None
    return out / nn


# ----------------------------------------------------------------------------
# Diagnostics of a state  (becgpp/vortex.py:10-84, fields.py:40-67)
# ----------------------------------------------------------------------------
def _wrap_pi(d):
    """(d + pi) mod 2 pi - pi, the modulus with the sign of the divisor."""
    return np.mod(d + math.pi, 2.0 * math.pi) - math.pi


def vortex_count_2d(psi, G, Omega=1.0, frac=0.15):
    """vortex.py:10-48: phase windings of the grid plaquettes inside the
    droplet (99% mass radius about the centre of mass) where the plaquette's
    mean density, smoothed by exp(-ell^2 K2/2), exceeds frac of its maximum."""
    rho = np.abs(psi) ** 2
    dx = float(G["dx"])
    X, Y = G["coords"][0], G["coords"][1]
    if float(rho.sum()) * dx * dx <= 0:
        return dict(Nv=0, Nplus=0, Nminus=0, Nnet=0)
    xcm, ycm = center_of_mass(rho, G)
    R = mass_quantiles(rho, G, (0.99,), center=(xcm, ycm))[0]
    if not np.isfinite(R) or R <= 0:
        return dict(Nv=0, Nplus=0, Nminus=0, Nnet=0)
    th = np.angle(psi)
    d1 = _wrap_pi(th[1:, :-1] - th[:-1, :-1])
    d2 = _wrap_pi(th[1:, 1:] - th[1:, :-1])
    d3 = _wrap_pi(th[:-1, 1:] - th[1:, 1:])
    d4 = _wrap_pi(th[:-1, :-1] - th[:-1, 1:])
    wind = np.round((d1 + d2 + d3 + d4) / (2.0 * math.pi))
    ell = min(2.0 * math.sqrt(math.pi / max(abs(Omega), 1e-9)), max(R / 3.0, 2.0 * dx))
    nbar = np.maximum(np.real(sfft.ifftn(sfft.fftn(rho) * np.exp(-0.5 * ell ** 2 * G["K2"]))), 0.0)
    cell = 0.25 * (nbar[:-1, :-1] + nbar[1:, :-1] + nbar[1:, 1:] + nbar[:-1, 1:])
    rc = np.sqrt((X[:-1, :-1] - xcm) ** 2 + (Y[:-1, :-1] - ycm) ** 2)
    mask = (cell > frac * nbar.max()) & (rc < R)
    npl, nmi = int(((wind >= 1) & mask).sum()), int(((wind <= -1) & mask).sum())
    return dict(Nv=npl + nmi, Nplus=npl, Nminus=nmi, Nnet=npl - nmi)


def droplet_radius(rho, G, frac=1e-2):
    """fields.py:40-67: the outermost centre of the bins of [0, L] (bins =
    int(min(400, max(20, L / (2 dx))))) whose mean rho about the centre of
    mass exceeds frac of the largest bin mean."""
    c = center_of_mass(rho, G)
    r = np.sqrt(sum((x - c[i]) ** 2 for i, x in enumerate(G["coords"]))).ravel()
    rmax = float(G["L"])
    bins = int(min(400, max(20, rmax / (2.0 * float(G["dx"])))))
    edges = np.linspace(0.0, rmax, bins + 1)
    ctr = 0.5 * (edges[1:] + edges[:-1])
    idx = np.clip(np.digitize(r, edges) - 1, 0, bins - 1)
    f = np.asarray(rho).ravel()
    prof = np.full(bins, np.nan)
    for i in range(bins):
        m = idx == i
        if np.any(m):
            prof[i] = f[m].mean()
    pk = np.nanmax(prof)
    if not np.isfinite(pk) or pk <= 0:
        return float("nan")
    valid = prof > frac * pk
    return float(ctr[valid][-1]) if np.any(valid) else float("nan")


def lll_weight(psi, G, Mmax=60, Omega=1.0):
    """vortex.py:61-84: sum over m <= Mcap of |<q_m/|q_m|, psi/|psi|>|^2,
    q_m = exp(m log r - r^2/2 - (log pi + lgamma(m+1))/2 + i m theta)."""
    if G["ndim"] != 2:
        return float("nan")
    dv = dV(G)
    pn = psi / norm_of(psi, G)
    X, Y = G["coords"][0], G["coords"][1]
    R2 = G["R2"]
    r = np.sqrt(R2)
    theta = np.arctan2(Y, X) * (1.0 if Omega >= 0 else -1.0)
    r2mean = float((R2 * np.abs(pn) ** 2).sum()) * dv
    est = int(math.ceil(r2mean + 10.0 * math.sqrt(max(r2mean, 1.0)) + 30.0))
    Mcap = min(max(int(Mmax), est), max(int(Mmax), int((0.80 * G["L"]) ** 2)))
    w = 0.0
    for m in range(Mcap + 1):
        lm = m * np.log(np.maximum(r, 1e-300)) - 0.5 * R2 - 0.5 * (math.log(math.pi) + math.lgamma(m + 1.0))
        q = np.exp(lm) * np.exp(1j * m * theta)
        q = q / norm_of(q, G)
        w += abs(complex((np.conj(q) * pn).sum()) * dv) ** 2
    return min(max(w, 0.0), 1.0)


def tf_effective_potential(G, cfg):
    """thomasfermi.py:21-30."""
    O = float(cfg.get("Omega", 1.0))
    return G["V"] - 0.5 * O * O * (G["X"] ** 2 + G["Y"] ** 2)


def invert_local(mu, Veff, b2, b3):
    """thomasfermi.py:33-45."""
    rhs = mu - Veff
    if b3 > 1e-30:
        disc = b2 * b2 + 4.0 * b3 * rhs
        root = (-b2 + np.sqrt(np.maximum(disc, 0.0))) / (2.0 * b3)
        return np.where(disc > 0.0, np.maximum(root, 0.0), 0.0)
    if abs(b2) > 1e-30:
        return np.maximum(rhs / b2, 0.0)
    return np.zeros_like(rhs)


def bisect_mu_for_norm(f_rho, G, lo, hi, steps=90):
    """thomasfermi.py:48-62."""
    dv = dV(G)
    guard = 0
    while float(f_rho(hi).sum()) * dv < 1.0:
        hi = lo + 2.0 * (hi - lo) + 1e-9
        guard += 1
        if guard > 200:
            return float("nan")
    for _ in range(steps):
        mid = 0.5 * (lo + hi)
        if float(f_rho(mid).sum()) * dv < 1.0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def extract_tf(cfg, G, iters=400, relax=0.4):
    """thomasfermi.py:65-156 (no cache)."""
    b2, b3, gc = float(cfg["beta2"]), float(cfg["beta3"]), float(cfg["G_C"])
    dv = dV(G)
    Veff = tf_effective_potential(G, cfg)
    nonlocal_on = (G["kernel"] != "none" and abs(gc) > 1e-15)
    if nonlocal_on and b2 <= 0 and b3 <= 0:
        return None
    if nonlocal_on:
        sigma = 0.15 * G["L"]
        n = np.exp(-G["R2"] / (2 * sigma ** 2))
        n = n / (float(n.sum()) * dv)
        conv = False
        dn = float("inf")
        mu = float("nan")
        for _ in range(int(iters)):
            phi = long_range_phi(n, G, gc)
            base = Veff + phi
            f = lambda m: invert_local(m, base, b2, b3)  # noqa: E731
            mu = bisect_mu_for_norm(f, G, float(base.min()), float(base.max()) + abs(b2) + abs(b3) + 1.0)
            if not math.isfinite(mu):
                break
            n_new = f(mu)
            ssum = float(n_new.sum()) * dv
            if ssum <= 0:
                break
            n_new = n_new / ssum
            dn = float(np.abs(n_new - n).sum()) * dv
            n = (1 - relax) * n + relax * n_new
            n = n / (float(n.sum()) * dv)
            if dn < 1e-9:
                conv = True
                break
        if not math.isfinite(mu):
            return None
        R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
        return dict(kind="nonlocal_compressed", n=n, R90=R90, R99=R99, mu=mu,
                    rho0=float("nan"), converged=conv, final_dn=dn)
    confined = bool(float(np.max(Veff)) > 1e-12)
    if b2 < 0 and b3 > 0:
        rho0 = -3.0 * b2 / (4.0 * b3)
        mu = b2 * rho0 + b3 * rho0 ** 2
        R = 1.0 / math.sqrt(math.pi * rho0) if G["ndim"] == 2 else (3.0 / (4.0 * math.pi * rho0)) ** (1.0 / 3.0)
        rr = np.sqrt(G["R2"])
        n = np.where(rr <= R, rho0, 0.0)
        n = n / (float(n.sum()) * dv)
        R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
        return dict(kind="local_flattop", n=n, R90=R90, R99=R99, mu=mu, rho0=rho0, R=R,
                    converged=True, final_dn=0.0)
    if confined and (b2 > 0 or (b2 == 0 and b3 > 0)):
        f = lambda m: invert_local(m, Veff, b2, b3)  # noqa: E731
        mu = bisect_mu_for_norm(f, G, float(Veff.min()), float(Veff.max()) + abs(b2) + abs(b3) + 1.0)
        n = f(mu)
        n = n / (float(n.sum()) * dv)
        R90, R99 = mass_quantiles(n, G, (0.90, 0.99))
        kind = "repulsive_parabola" if b3 <= 0 else "local_trapped_repulsive"
        return dict(kind=kind, n=n, R90=R90, R99=R99, mu=mu, rho0=float("nan"),
                    converged=True, final_dn=0.0)
    return None


# ----------------------------------------------------------------------------
# seeds (becgpp/seeds.py:11-103), with this build's counter-based noise
# ----------------------------------------------------------------------------
_SM_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _sm_mix(z):
    """splitmix64's output function on uint64 arrays (wrapping arithmetic)."""
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix64_outputs(state, k):
    """Outputs k (0-based, uint64 array) of the splitmix64 sequence seeded at state."""
    with np.errstate(over="ignore"):
        return _sm_mix(np.uint64(state) + (np.asarray(k, dtype=np.uint64) + np.uint64(1)) * _SM_GAMMA)


def counter_normals(rng, stream, count, first=0):
    """Standard normals of points first .. first+count-1 of (rng, stream): Box-Muller
    on outputs 2i, 2i+1 of the splitmix64 sequence seeded at output 0 of the
    sequence seeded at (rng mod 2^32) + 2^32 stream -- the seeds' noise."""
    key = int(splitmix64_outputs((int(rng) & 0xFFFFFFFF) | (int(stream) << 32), 0))
    i = np.arange(first, first + count, dtype=np.uint64)
    a = splitmix64_outputs(key, np.uint64(2) * i)
    c = splitmix64_outputs(key, np.uint64(2) * i + np.uint64(1))
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (c >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def _noise(cfg, stream, shape):
    return counter_normals(int(cfg.get("rng", 0)), stream, int(np.prod(shape))).reshape(shape)


def _winding(psi, G, cfg):
    m = int(cfg.get("seed_winding", 0))
    if m != 0:
        psi = psi * np.exp(1j * m * np.arctan2(G["Y"], G["X"]))
    return psi


def gaussian_seed(G, cfg):
    """seeds.py:11-19; noise xi ~ N(0,1): counter_normals(rng, stream 0)."""
    env = np.exp(-0.5 * G["R2"])
    noise = 1.0 + cfg.get("seed_noise", 1e-2) * _noise(cfg, 0, env.shape)
    psi = _winding((env * noise).astype(np.complex128), G, cfg)
    return psi / norm_of(psi, G)


def tf_seed(G, cfg):
    """seeds.py:22-34; amplitude noise from stream 0, phase noise from stream 1."""
    tf = extract_tf(cfg, G)
    if tf is None or tf.get("n") is None:
        return gaussian_seed(G, cfg)
    amp = np.sqrt(np.maximum(tf["n"], 0.0))
    noise = 1.0 + cfg.get("seed_noise", 1e-2) * _noise(cfg, 0, amp.shape)
    pn = cfg.get("seed_phase_noise", 5e-2) * _noise(cfg, 1, amp.shape)
    psi = _winding((amp * noise).astype(np.complex128) * np.exp(1j * pn), G, cfg)
    return psi / norm_of(psi, G)


def inplane_radius(n, G, frac=0.90):
    """seeds.py:37-44."""
    w = (np.maximum(n, 0.0) * dV(G)).reshape(-1)
    rperp = np.sqrt(G["X"] ** 2 + G["Y"] ** 2).reshape(-1)
    order = np.argsort(rperp, kind="stable")
    cum = np.cumsum(w[order])
    jj = min(int(np.searchsorted(cum, float(frac) * cum[-1], side="left")), cum.size - 1)
    return float(rperp[order[jj]])


def triangular_seed(G, cfg):
    """seeds.py:47-94; angle and shift from numpy default_rng(rng) as in the
    reference, amplitude noise from stream 0."""
    O = float(cfg.get("Omega", 0.0))
    if abs(O) < 0.3:
        return tf_seed(G, cfg)
    tf = extract_tf(cfg, G)
    if tf is None or not tf.get("converged", False):
        return tf_seed(G, cfg)
    amp = np.sqrt(np.maximum(tf["n"], 0.0))
    wsign = -1.0 if O < 0 else 1.0
    O = abs(O)
    av = math.sqrt(2.0 * math.pi / (math.sqrt(3.0) * O))
    dy = math.sqrt(3.0) * av / 2.0
    Rperp = inplane_radius(tf["n"], G, 0.90)
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
    phase = np.zeros_like(amp)
    core = np.ones_like(amp)
    xi = 0.25 * av
    for xx, yy in pos:
        dxt = G["X"] - xx
        dyt = G["Y"] - yy
        phase = phase + wsign * np.arctan2(dyt, dxt)
        core = core * np.tanh(np.sqrt(dxt * dxt + dyt * dyt) / max(xi, 1e-12))
    noise = cfg.get("seed_noise", 1e-2) * _noise(cfg, 0, amp.shape)
    psi = (amp * core * (1.0 + noise)).astype(np.complex128) * np.exp(1j * phase)
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
    """Bench/parity input: normalise(exp(-r^2/2w^2) (1 + noise (xi1 + i xi2)))."""
    rng = np.random.default_rng(seed)
    env = np.exp(-0.5 * G["R2"] / (width * width))
    xi = rng.standard_normal((2,) + env.shape)
    psi = env * (1.0 + noise * (xi[0] + 1j * xi[1]))
    return psi / norm_of(psi, G)


# ----------------------------------------------------------------------------
# analytic gate numerics  (becgpp/modes.py:30-201)
# ----------------------------------------------------------------------------
EULER_GAMMA = 0.5772156649015329


class _OracleOps:
    """The numeric operations gate_rows evaluates; the defaults are the oracle's
    own.  A test may pass another implementation (numpy in, numpy out)."""

    def __init__(self):
        self.make_grid = make_grid
        self.long_range_phi = long_range_phi
        self.energy_components = energy_components
        self.clear_kernel_cache = clear_kernel_cache


def _gauss_rho(G, a):
    psi = np.exp(-G["R2"] / (2 * a * a)).astype(np.complex128)
    psi = psi / norm_of(psi, G)
    return np.abs(psi) ** 2


def gate_rows(Nv=256, order_sizes=None, ops=None):
    """Recompute the rows of modes.mode_validate (modes.py:30-201) as
    (test, case, quantity) -> rel_error.  ``order_sizes`` optionally restricts
    the kernel-order ladders; ``ops`` supplies make_grid, long_range_phi,
    energy_components and clear_kernel_cache (default: this oracle)."""
    ops = ops or _OracleOps()
    make_grid, long_range_phi, energy_components = ops.make_grid, ops.long_range_phi, ops.energy_components
    clear_kernel_cache = ops.clear_kernel_cache
    rows = {}
    base = dict(s=2, trap_coeff=0.5, Omega=0.0, beta2=0.0, beta3=0.0, pad=2)
    G = make_grid(dict(base, dimension="2D", kernel="newton", G_C=1.0, L=12.0, Ngrid=Nv))
    dv = dV(G)
    c = Nv // 2
    for a in (0.5, 1.0, 2.0):
        rho = _gauss_rho(G, a)
        phi = long_range_phi(rho, G, 1.0)
        Eg = 0.5 * float((phi * rho).sum()) * dv
        ex = math.sqrt(math.pi / 8.0) / a
        rows[("2D_gauss", f"a={a}", "Egrav")] = abs(Eg + ex) / ex
        ex0 = math.sqrt(math.pi) / a
        rows[("2D_gauss", f"a={a}", "phi0")] = abs(phi[c, c] + ex0) / ex0
    Gl = make_grid(dict(base, dimension="2D", kernel="log", G_C=1.0, L=12.0, Ngrid=Nv))
    for a in (0.5, 2.0, 3.0):
        rho = _gauss_rho(Gl, a)
        phi = long_range_phi(rho, Gl, 1.0)
        Eg = 0.5 * float((phi * rho).sum()) * dv
        Eg_ex = 0.5 * (math.log(math.sqrt(2.0) * a) - 0.5 * EULER_GAMMA)
        p0_ex = math.log(a) - 0.5 * EULER_GAMMA
        rows[("2D_log_gauss", f"a={a}", "Egrav")] = abs(Eg - Eg_ex) / abs(Eg_ex)
        rows[("2D_log_gauss", f"a={a}", "phi0")] = abs(phi[c, c] - p0_ex) / abs(p0_ex)

    def _F(t):
        if abs(t - 1.0) < 1e-12:
            return 1.0
        return math.acos(t) / math.sqrt(1 - t * t) if t < 1 else math.acosh(t) / math.sqrt(t * t - 1)
    for lz in (0.25, 1.0):
        Gq = make_grid(dict(base, dimension="quasi2D", kernel="q2d", l_z=lz, G_C=1.0, L=12.0, Ngrid=Nv))
        for a in (1.0, 2.0):
            rho = _gauss_rho(Gq, a)
            phi = long_range_phi(rho, Gq, 1.0)
            Eg = 0.5 * float((phi * rho).sum()) * dv
            sp = a / math.sqrt(2.0)
            p0_ex = -(2 / math.sqrt(math.pi)) / (math.sqrt(2) * sp) * _F(lz / sp)
            Eg_ex = -0.5 * (2 / math.sqrt(math.pi)) / (math.sqrt(2) * a) * _F(lz / a)
            rows[("q2D_gauss", f"lz={lz},a={a}", "Egrav")] = abs(Eg - Eg_ex) / abs(Eg_ex)
            rows[("q2D_gauss", f"lz={lz},a={a}", "phi0")] = abs(phi[c, c] - p0_ex) / abs(p0_ex)
    for m in (0, 1, 10, 30):
        z = (G["X"] + 1j * G["Y"]).astype(np.complex128)
        psi = (z ** m) * np.exp(-0.5 * G["R2"])
        psi = psi / norm_of(psi, G)
        comp = energy_components(psi, G, 0.0, 0.0, 1.0, 0.0)
        Esp = comp["Ekin"] + comp["Etrap"] + comp["Erot"]
        r2 = float((G["R2"] * np.abs(psi) ** 2).sum()) * dv
        rows[("2D_LLL", f"m={m}", "max|E-1|,|Lz-m|,|r2-m-1|")] = max(abs(Esp - 1.0), abs(comp["Lz"] - m),
                                                                     abs(r2 - (m + 1)))
    G3 = make_grid(dict(base, dimension="3D", kernel="newton", G_C=1.0, L=8.0, Ngrid=96))
    for sig in (1.0, 1.5):
        rho = _gauss_rho(G3, sig)
        phi = long_range_phi(rho, G3, 1.0)
        p0 = phi[48, 48, 48]
        ex = -2.0 / (math.sqrt(math.pi) * sig)
        rows[("3D_gauss", f"sig={sig}", "phi0")] = abs(p0 - ex) / abs(ex)
    ladders = (("2D", 2, "newton", (192, 256, 384, 512), 1.0, 10.0),
               ("2D_log", 2, "log", (96, 128, 192, 256), 1.0, 10.0),
               ("3D", 3, "newton", (64, 96, 128, 160), 1.0, 8.0))
    for name, ndim, kern, Ns, sig, L in ladders:
        if order_sizes is not None:
            Ns = tuple(n for n in Ns if n in order_sizes.get(name, ()))
        ex = (math.log(sig) - 0.5 * EULER_GAMMA) if kern == "log" else (
            (-math.sqrt(math.pi) / sig) if ndim == 2 else (-2.0 / (math.sqrt(math.pi) * sig)))
        for Ni in Ns:
            Gk = make_grid(dict(base, dimension=("2D" if ndim == 2 else "3D"), kernel=kern, G_C=1.0,
                                L=L, Ngrid=int(Ni)))
            phi = long_range_phi(_gauss_rho(Gk, sig), Gk, 1.0)
            cc = Gk["N"] // 2
            p0 = phi[(cc,) * ndim]
            rows[(f"{name}_kernel_order", f"dx={Gk['dx']:.4f}", "phi0_relerr")] = abs(p0 - ex) / abs(ex)
            clear_kernel_cache()
    return rows
