"""Grid construction and the device plan behind it (becgpp/grid.py:10-161).

``make_grid`` returns the reference's dictionary (same keys and formulas,
tensors on the GPU) so code written against the reference keeps working.  The
arithmetic of the solve path never reads those tensors: the HIP kernels
rebuild x, k and V analytically from (N, L, s, trap_coeff), so the dictionary
is for the caller.  Each grid geometry owns one native plan (FFT twiddles,
workspaces, kernel spectra), cached here and freed by ``clear_kernel_cache``.
"""
import ctypes
import math

import torch

from . import _lib
from .constants import DEV


def geometry(cfg):
    """(ndim, quasi) of "2D" | "3D" | "quasi2D", case-insensitive (grid.py:10-20)."""
    g = str(cfg.get("dimension", "2D")).strip().lower()
    table = {"2d": (2, False), "3d": (3, False), "quasi2d": (2, True)}
    if g not in table:
        raise ValueError(f"unknown dimension {cfg.get('dimension')!r}; use '2D', '3D' or 'quasi2D'")
    return table[g]


def resolve_kernel(cfg):
    """(kind, homogeneity degree) of the long-range kernel (grid.py:23-55)."""
    ndim, quasi = geometry(cfg)
    k = str(cfg.get("kernel", "auto")).lower()
    if abs(float(cfg.get("G_C", 0.0))) <= 1e-15 or k == "none":
        return "none", 0
    lz = float(cfg.get("l_z", 0.0))
    if k == "auto":
        if ndim == 3:
            k = "newton"
        elif quasi:
            k = "q2d" if lz > 0 else "newton"
        else:
            k = "log"
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


def _memory_estimate_gb(ndim, N, pad, kernel):
    """Peak HBM of one solve on this implementation: 9 complex fields + 4 real
    fields + conv spectra (N (pN/2+1) [N] complex, twice in 3D)."""
    n = float(N) ** ndim
    b = 9 * 16 * n + 4 * 8 * n
    if kernel != "none":
        M = pad * N
        H = M // 2 + 1
        b += 16 * (N * H if ndim == 2 else N * N * H + N * M * H) + 8 * H ** ndim
    return b / 1e9


def make_grid(cfg):
    """grid.py:69-101: x = (i - N//2) dx, k = 2 pi fftfreq(N, dx), V = tc r^s."""
    ndim, quasi = geometry(cfg)
    L, N, s = float(cfg["L"]), int(cfg["Ngrid"]), float(cfg["s"])
    kkind, kexp = resolve_kernel(cfg)
    dx = 2.0 * L / N
    x = (torch.arange(N, device=DEV, dtype=torch.float64) - N // 2) * dx
    k1 = 2 * math.pi * torch.fft.fftfreq(N, d=dx, dtype=torch.float64).to(DEV)
    coords = torch.meshgrid(*([x] * ndim), indexing="ij")
    kcoords = torch.meshgrid(*([k1] * ndim), indexing="ij")
    R2 = sum(c ** 2 for c in coords)
    K2 = sum(c ** 2 for c in kcoords)
    tc = float(cfg.get("trap_coeff", 0.5))
    V = tc * R2 ** (s / 2.0) if s > 0 else torch.zeros_like(R2)
    pad = max(2, int(cfg.get("pad", 2)))
    mem = _memory_estimate_gb(ndim, N, pad, kkind)
    if mem > float(cfg.get("mem_warn_gb", 12.0)):
        print(f"[warn] estimated peak memory ~{mem:.0f} GB for N={N} ({ndim}D, kernel={kkind}, pad={pad})")
    return dict(ndim=ndim, quasi=quasi, coords=tuple(coords), kcoords=tuple(kcoords),
                X=coords[0], Y=coords[1], KX=kcoords[0], KY=kcoords[1], R2=R2, V=V, K2=K2,
                s=s, trap_coeff=tc, dx=dx, N=N, L=L, kernel=kkind, kexp=kexp,
                l_z=float(cfg.get("l_z", 0.0)), pad=pad, mem_estimate_gb=mem)


def dV(G):
    return G["dx"] ** G["ndim"]


# ----------------------------------------------------------------------------
# native plans
# ----------------------------------------------------------------------------
class NativePlan:
    """One libbecgpp_hip plan: FFT twiddles, workspace and kernel spectra for
    a grid geometry."""

    def __init__(self, ndim, N, L, pad, s, tc):
        lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("becgpp_amd needs a GPU: no HIP device is visible")
        h = ctypes.c_void_p()
        _lib.check(lib.bec_plan_create(int(ndim), int(N), float(L), int(pad), float(s), float(tc),
                                       ctypes.byref(h)))
        self.h = h
        self.lib = lib
        self.ndim, self.N, self.L, self.pad, self.s, self.tc = ndim, N, L, pad, s, tc
        self.kernels = {}           # which -> key of the spectrum currently loaded

    def autotune(self, nb):
        """Choose this plan's FFT stage orders and lines per workgroup by timing
        the line passes of a solver step over nb instances (once per plan)."""
        if getattr(self, "schedule", None) is None:
            best, dflt = ctypes.c_double(), ctypes.c_double()
            rep = ctypes.create_string_buffer(256)
            _lib.check(self.lib.bec_plan_autotune(self.h, int(nb), ctypes.byref(best), ctypes.byref(dflt), rep, 256))
            self.schedule = dict(schedule=rep.value.decode(), step_passes_ms=best.value,
                                 default_ms=dflt.value, tuned_for_instances=int(nb))
        return self.schedule

    def bind_stream(self):
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.bec_plan_set_stream(self.h, ctypes.c_void_p(st)))
        return self

    def clear_kernels(self):
        _lib.check(self.lib.bec_plan_clear_kernels(self.h))
        self.kernels.clear()

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.bec_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_PLANS = {}


def plan_of(G):
    """The native plan of grid G (created on first use, then cached)."""
    key = (int(G["ndim"]), int(G["N"]), round(float(G["L"]), 14), int(G["pad"]),
           round(float(G["s"]), 14), round(float(G.get("trap_coeff", 0.5)), 14))
    p = _PLANS.get(key)
    if p is None:
        p = NativePlan(*key)
        _PLANS[key] = p
    return p.bind_stream()


def release_plans():
    for p in _PLANS.values():
        p.close()
    _PLANS.clear()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def as_field(p, G, dtype=torch.complex128):
    """Contiguous device tensor of the grid's shape (inputs are never mutated)."""
    t = torch.as_tensor(p)
    if t.device.type != "cuda" or t.dtype != dtype or not t.is_contiguous():
        t = t.to(device=DEV, dtype=dtype).contiguous()
    shape = (int(G["N"]),) * int(G["ndim"])
    if tuple(t.shape) != shape:
        raise ValueError(f"field shape {tuple(t.shape)} does not match the grid {shape}")
    return t


def norm_of(p, G):
    """sqrt(sum |p|^2 dV) (grid.py:108-109), as a 0-d tensor."""
    pl = plan_of(G)
    t = as_field(p, G)
    out = ctypes.c_double()
    _lib.check(pl.lib.bec_norm2(pl.h, _ptr(t), ctypes.byref(out)))
    return torch.tensor(math.sqrt(out.value * dV(G)), dtype=torch.float64, device=DEV)


def Lz_op(p, G):
    """-i (X d_y - Y d_x) p with spectral derivatives (grid.py:112-116)."""
    pl = plan_of(G)
    t = as_field(p, G)
    out = torch.empty_like(t)
    _lib.check(pl.lib.bec_Lz(pl.h, _ptr(t), _ptr(out)))
    return out


def auto_grid(cfg):
    """Half-box L and grid N from a radius estimate (grid.py:119-137)."""
    b2, b3, gc = float(cfg["beta2"]), float(cfg["beta3"]), float(cfg["G_C"])
    ndim, _ = geometry(cfg)
    if gc > 1e-15 and b2 > 0:
        R_est = 0.5 * b2 / gc
    elif b2 < 0 and b3 > 0:
        rho0 = max(-3 * b2 / (4 * b3), 1e-6)
        R_est = 1.0 / math.sqrt(math.pi * rho0) if ndim == 2 else (3 / (4 * math.pi * rho0)) ** (1 / 3)
    else:
        R_est = 4.0
    L = max(6.0, 2.5 * R_est)
    dxt = min(0.10, max(R_est / 20.0, 0.01))
    Nn = int(math.ceil(2.0 * L / dxt))
    cap = 512 if ndim == 2 else 224
    N = int(min(cap, max(128, 32 * math.ceil(Nn / 32))))
    return float(L), N, R_est


def resample_state(psi_old, G_old, G_new):
    """Multilinear interpolation of a state between two physical grids,
    renormalised on the new grid (grid.py:140-161); None when the
    interpolation is degenerate.  Computed by bec_resample on the device."""
    d = int(G_old["ndim"])
    if int(G_new["ndim"]) != d:
        raise ValueError("resample_state needs grids of the same dimension")
    src = as_field(psi_old, G_old)
    out = torch.empty((int(G_new["N"]),) * d, dtype=torch.complex128, device=src.device)
    nn = ctypes.c_double()
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.bec_resample(d, int(G_old["N"]), float(G_old["dx"]), _ptr(src), int(G_new["N"]),
                                float(G_new["dx"]), _ptr(out), ctypes.byref(nn), ctypes.c_void_p(st)))
    if not math.isfinite(nn.value) or nn.value <= 1e-14:
        return None
    return out
