"""Output directory of the run records (becgpp/paths.py): GPP_OUTDIR when set,
else /kaggle/working/becgpp on Kaggle, else ./becgpp_out; configure(base)
switches it at run time.  BASE and CKPT_DIR are read as module attributes and
follow the current choice."""
import os

_configured = None


def base_dir():
    if _configured:
        return _configured
    env = os.environ.get("GPP_OUTDIR")
    if env:
        return env
    if os.path.isdir("/kaggle/working"):
        return "/kaggle/working/becgpp"
    return "./becgpp_out"


def configure(base):
    """Switch the output directory to base (created with its ckpt/ if missing)."""
    global _configured
    _configured = str(base)
    os.makedirs(os.path.join(_configured, "ckpt"), exist_ok=True)
    return _configured


def __getattr__(name):
    if name == "BASE":
        return base_dir()
    if name == "CKPT_DIR":
        return os.path.join(base_dir(), "ckpt")
    raise AttributeError(name)
