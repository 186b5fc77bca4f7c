import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
# run records (checkpoints, CSV, JSON) of the tests go to a scratch directory
if not os.environ.get("GPP_OUTDIR"):
    import tempfile
    os.environ["GPP_OUTDIR"] = tempfile.mkdtemp(prefix="becgpp_test_out_")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")
