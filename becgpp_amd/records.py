"""On-disk records of a run, in the reference's formats so that resumes
interoperate (becgpp/paths.py:16-33, io.py:29-36, modes.py:216-247):

* the output directory is ``$GPP_OUTDIR`` or ``./becgpp_out``, with ``ckpt/``;
* a checkpoint ``ckpt/<run_id>.pt`` holds ``{psi, cfg, obs}`` (psi on the CPU);
* ``<name>.csv`` holds one row per record, columns the sorted union of keys;
* ``<run_id>.json`` holds the configuration and every diagnostic of one run.

Checkpoints are read with ``torch.load(weights_only=True)``: a file that needs
anything beyond tensors and plain containers is not resumed from.
"""
import csv
import json
import os

import torch

from .constants import CODE_VERSION
from . import paths


def out_dir():
    base = paths.base_dir()
    os.makedirs(os.path.join(base, "ckpt"), exist_ok=True)
    return base


def ckpt_path(rid):
    return os.path.join(out_dir(), "ckpt", f"{rid}.pt")


def _numpy_scalar_globals():
    """The classes a pickled numpy scalar refers to (numpy 1.x and 2.x names)."""
    import numpy as np
    out = [np.dtype]
    for mod in ("numpy._core.multiarray", "numpy.core.multiarray"):
        try:
            m = __import__(mod, fromlist=["scalar"])
            out.append(m.scalar)
        except (ImportError, AttributeError):
            pass
    try:
        out.append(type(np.dtype(np.float64)))         # numpy.dtypes.Float64DType
    except TypeError:
        pass
    return out


def load_checkpoint_psi(rid):
    """psi of ckpt/<rid>.pt, or None (missing or not loadable safely)."""
    path = ckpt_path(rid)
    if not os.path.isfile(path):
        return None
    try:
        # still a weights-only load: the reference's saved observables hold
        # numpy float64 scalars (its droplet_radius), which need these globals
        with torch.serialization.safe_globals(_numpy_scalar_globals()):
            return torch.load(path, map_location="cpu", weights_only=True)["psi"]
    except Exception as exc:                   # the reference warns and starts fresh
        print(f"[resume-warn] {exc}")
        return None


def _plain(v):
    if torch.is_tensor(v):
        return v.item() if v.numel() == 1 else v.detach().cpu().tolist()
    if isinstance(v, (list, tuple)):
        return [_plain(x) for x in v]
    if isinstance(v, dict):
        return {k: _plain(x) for k, x in v.items()}
    return v


def save_checkpoint(rid, psi, cfg, obs):
    torch.save(dict(psi=psi.detach().cpu(), cfg=_plain(dict(cfg)), obs=_plain(dict(obs))), ckpt_path(rid))


def write_csv(path, rows):
    """io.py:29-36: header = sorted union of the rows' keys."""
    if not rows:
        return
    fields = sorted({k for r in rows for k in r.keys()})
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields, extrasaction="ignore")
        w.writeheader()
        w.writerows(rows)


def write_run_json(rid, cfg, diag, device):
    with open(os.path.join(out_dir(), f"{rid}.json"), "w") as f:
        json.dump(dict(code_version=CODE_VERSION, device=str(device), cfg=_plain(dict(cfg)),
                       diagnostics=_plain(dict(diag))), f, indent=2, default=str)
