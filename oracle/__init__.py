"""CPU FP64 oracle for the becGPP ground-state solve path.

TEST INFRASTRUCTURE ONLY.  This package restates, in numpy/scipy, the
arithmetic of the reference's ground-state path (grid, cell-averaged kernels,
zero-padded convolution, energy, H.psi, residual, Sobolev preconditioner,
Riemannian PCG with Barzilai-Borwein and Armijo, multi-seed selection, seeds
and the Thomas-Fermi envelopes).  It is imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``, and
only as the checker.  The product (``becgpp_amd``) never imports it.

Every function cites the reference file:line it follows.  How each part is
pinned (golden vectors, closed forms, brute force) is listed in
``oracle/ref.py``'s module docstring.
"""
from .ref import *  # noqa: F401,F403
