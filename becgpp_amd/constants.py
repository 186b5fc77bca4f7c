"""Process-wide settings (becgpp/constants.py:9-12).

FP64 is the default real dtype, as in the reference: the long-range coupling
can be many orders of magnitude weaker than the contact term.  The compute
device is the GPU; this package has no CPU path.
"""
import torch

torch.set_default_dtype(torch.float64)

DEV = "cuda"
CODE_VERSION = "becGPP_1.2.1"
