"""Process-wide settings (becgpp/constants.py:9-12).

FP64 is the default real dtype, as in the reference: the long-range coupling
can be many orders of magnitude weaker than the contact term.  The compute
device is the GPU; this package has no CPU path.
"""
import torch

torch.set_default_dtype(torch.float64)

DEV = "cuda"
CODE_VERSION = "becGPP_1.2.1"

# CODATA 2018 (becgpp/constants.py:15-16)
HBAR = 1.054571817e-34       # J s
G_NEWTON = 6.67430e-11       # m^3 kg^-1 s^-2
