"""bec_resample_noalloc against bec_resample through the C-ABI.

The allocation-free entry point runs the same kernels with its partial rows
in a device array of the library; its output and norm must equal the
allocating entry point's bit for bit, on two calls in a row (the second
smaller, so stale rows of the first call would show), in 2D and 3D.
"""
import ctypes

import numpy as np
import pytest
import torch

from becgpp_amd import _lib

pytestmark = pytest.mark.gpu


def _state(ndim, n0, seed):
    rng = np.random.default_rng(seed)
    shape = (n0,) * ndim
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return torch.as_tensor(a.astype(np.complex128)).cuda().contiguous()


def _call(fn, ndim, n0, dx0, psi, n1, dx1):
    out = torch.full((n1,) * ndim, complex(float("nan"), 0.0), dtype=torch.complex128, device="cuda")
    nrm = ctypes.c_double(-1.0)
    _lib.check(fn(ndim, n0, dx0, ctypes.c_void_p(psi.data_ptr()), n1, dx1, ctypes.c_void_p(out.data_ptr()),
                  ctypes.byref(nrm), None))
    torch.cuda.synchronize()
    return out, nrm.value


@pytest.mark.parametrize("ndim,cases", [
    (2, [(96, 0.125, 128, 0.1), (40, 0.3, 52, 0.25)]),
    (3, [(40, 0.25, 48, 0.2), (24, 0.5, 20, 0.55)]),
])
def test_noalloc_resample_equals_allocating_entry_point(ndim, cases):
    lib = _lib.load()
    for k, (n0, dx0, n1, dx1) in enumerate(cases):
        print(f"[noalloc] ndim={ndim} call {k}: n0={n0} dx0={dx0} -> n1={n1} dx1={dx1}", flush=True)
        psi = _state(ndim, n0, seed=10 * ndim + k)
        ref, nref = _call(lib.bec_resample, ndim, n0, dx0, psi, n1, dx1)
        got, ngot = _call(lib.bec_resample_noalloc, ndim, n0, dx0, psi, n1, dx1)
        assert ngot == nref and nref > 0.0
        assert torch.equal(got, ref)


def test_noalloc_resample_degenerate_and_errors():
    lib = _lib.load()
    psi = _state(2, 1, seed=3)
    out, nrm = _call(lib.bec_resample_noalloc, 2, 1, 0.5, psi, 8, 0.5)
    assert nrm == 0.0                       # one-point grid: no interpolation, no state
    with pytest.raises(ValueError):
        _call(lib.bec_resample_noalloc, 4, 8, 0.5, psi, 8, 0.5)
